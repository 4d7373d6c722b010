// The RTGL1 C API (include/RTGL1/RTGL1.h) on MI355X.
//
// Mirrors the reference's API shim and frame orchestration: Source/RTGL1.cpp
// (singleton instance, exception -> RgResult), VulkanDevice::StartFrame /
// DrawFrame / FillUniform / Render (Source/VulkanDevice.cpp:142-893), with
// the hot segment of Render (:540-588) replaced by the HIP passes: scene
// assembly, LBVH build, ray-traced passes, A-SVGF, eye adaptation.
#include "rg_bvh.h"
#include "rg_denoise.h"
#include "rg_frame.h"
#include "rg_framebuffers.h"
#include "rg_internal.h"
#include "rg_kernels.h"
#include "rg_pathtrace.h"
#include "rg_scene.h"

#include <cmath>
#include <memory>

namespace rg
{
namespace
{

enum Stage
{
    ST_BEGIN,
    ST_UPLOADED,
    ST_ASSEMBLED,
    ST_BVH,
    ST_TRACED,
    ST_DENOISED,
    ST_END,
    ST_COUNT
};

} // namespace

struct Instance
{
    hipStream_t                                 stream = nullptr;
    RgInstanceCreateInfo                        info{};
    SceneCollector                              scene;
    DeviceScene                                 dscene;
    TextureManager                              textures;
    BvhObject                                   bvh;
    TraceContext                                tc;
    FrameGpu                                    g;
    CameraState                                 cam;
    std::unique_ptr< RgExtFramebuffers_T >      fb;
    RgExtFrameStats                             stats{};
    bool                                        frameStarted = false;
    bool                                        staticGeometry = false;
    hipEvent_t                                  ev[ ST_COUNT ] = {};
    // Two frames in flight (MAX_FRAMES_IN_FLIGHT, Source/Common.h:38):
    // rgDrawFrame returns once its frame is enqueued; frameDone[k] marks the
    // end of the last frame drawn in slot k = frame count mod 2, and a frame
    // waits for the one two frames back before it is enqueued.
    hipEvent_t                                  frameDone[ 2 ] = {};
    uint32_t                                    framesDrawn  = 0;
    // The GPU times, traversal counts and adapted luminance of the last
    // frame are read once it has finished, when they are first asked for.
    bool                                        statsPending = false;

    Instance( const RgInstanceCreateInfo& ci )
    {
        info = ci;
        for( int i = 0; i < 3; i++ )
        {
            cam.worldUp[ i ] = ci.worldUp.data[ i ];
        }
        scene.textures          = &textures;
        textures.forceMinLinear = ci.textureSamplerForceMinificationFilterLinear != 0;
        int n = 0;
        if( hipGetDeviceCount( &n ) != hipSuccess || n == 0 )
        {
            throw Error( RG_RESULT_CANT_FIND_SUPPORTED_PHYSICAL_DEVICE, "no HIP device" );
        }
        RG_HIP_CHECK( hipStreamCreateWithFlags( &stream, hipStreamNonBlocking ) );
        for( auto& e : ev )
        {
            RG_HIP_CHECK( hipEventCreate( &e ) );
        }
        for( auto& e : frameDone )
        {
            RG_HIP_CHECK( hipEventCreateWithFlags( &e, hipEventDisableTiming ) );
        }
        g.Init( stream );
        RG_HIP_CHECK( hipStreamSynchronize( stream ) );
    }
    ~Instance()
    {
        if( stream )
        {
            ( void )hipStreamSynchronize( stream );
        }
        for( auto& e : ev )
        {
            if( e )
            {
                ( void )hipEventDestroy( e );
            }
        }
        for( auto& e : frameDone )
        {
            if( e )
            {
                ( void )hipEventDestroy( e );
            }
        }
        fb.reset();
        if( stream )
        {
            ( void )hipStreamDestroy( stream );
        }
    }

    void StartFrame()
    {
        if( frameStarted )
        {
            throw Error( RG_RESULT_FRAME_WASNT_ENDED, "rgStartFrame was called twice without rgDrawFrame" );
        }
        frameStarted = true;
        if( framesDrawn > 0 )
        {
            // the last frame's uploads read the pinned staging arrays, the
            // light list and the geometry map that this frame refills
            RG_HIP_CHECK( hipEventSynchronize( ev[ ST_UPLOADED ] ) );
        }
        scene.BeginFrame();
    }

    // Waits for every enqueued frame (before host-side readback of device
    // memory, texture changes and teardown).
    void Quiesce() { RG_HIP_CHECK( hipStreamSynchronize( stream ) ); }

    // Completes the statistics of the last drawn frame (waits for it).
    const RgExtFrameStats& Stats()
    {
        if( !statsPending )
        {
            return stats;
        }
        RG_HIP_CHECK( hipEventSynchronize( ev[ ST_END ] ) );
        auto ms = [ & ]( Stage a, Stage b ) {
            float v = 0.0f;
            ( void )hipEventElapsedTime( &v, ev[ a ], ev[ b ] );
            return v;
        };
        stats.gpuMsUpload       = ms( ST_BEGIN, ST_UPLOADED );
        stats.gpuMsAssemble     = ms( ST_UPLOADED, ST_ASSEMBLED );
        stats.gpuMsTraversal    = g.pt.TraversalMs();
        stats.traversalLaunches = g.pt.traceLaunches;
        g.pt.TraversalCounts( &stats.nodeVisits, &stats.triangleTests, &stats.traversalErrors );
        stats.gpuMsBvh     = ms( ST_ASSEMBLED, ST_BVH );
        stats.gpuMsTrace   = ms( ST_BVH, ST_TRACED );
        stats.gpuMsDenoise = ms( ST_TRACED, ST_DENOISED );
        stats.gpuMsTonemap = ms( ST_DENOISED, ST_END );
        stats.gpuMsTotal   = ms( ST_BEGIN, ST_END );
        RG_HIP_CHECK( hipMemcpy( &stats.avgLuminance, g.avgLum.ptr, 4, hipMemcpyDeviceToHost ) );
        statsPending = false;
        return stats;
    }

    void DrawFrame( const RgDrawFrameInfo& di )
    {
        if( !frameStarted )
        {
            throw Error( RG_RESULT_FRAME_WASNT_STARTED, "rgDrawFrame was called without rgStartFrame" );
        }
        frameStarted = false;
        uint32_t w, h;
        RenderSizeOf( di, &w, &h );
        scene.FinishLights();
        const auto* illum = FindParams< RgDrawFrameIlluminationParams >( di, RG_STRUCTURE_TYPE_ILLUMINATION );
        FillFrameParams( cam, di, w, h, uint32_t( scene.lights.size() ), scene.prevLightCount, scene.hasDirectional,
                         info.rayCullBackFacingTriangles != 0, info.allowGeometryWithSkyFlag != 0,
                         scene.LightIndexOf( illum ? illum->lightUniqueIdIgnoreFirstPersonViewerShadows : nullptr ) );
        {
            // styles for the lights of the next frame's uploads
            const auto* il = FindParams< RgDrawFrameIlluminationParams >( di, RG_STRUCTURE_TYPE_ILLUMINATION );
            if( il && il->pLightstyleValues && il->lightstyleValuesCount > 0 )
            {
                scene.lightstyles.assign( il->pLightstyleValues, il->pLightstyleValues + il->lightstyleValuesCount );
            }
            else
            {
                scene.lightstyles.clear();
            }
        }

        const auto*    pathExt  = FindParams< RgExtDrawFramePathParams >( di, RG_EXT_STRUCTURE_TYPE_PATH_PARAMS );
        const uint32_t prec     = pathExt ? pathExt->precisionFlags : 0u;
        RG_REQUIRE( ( prec & ~( RG_EXT_PRECISION_FP16_STORAGE | RG_EXT_PRECISION_FP16_NODES ) ) == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                    "unknown precisionFlags bits" );
        const uint32_t storage  = ( prec & RG_EXT_PRECISION_FP16_STORAGE ) ? RG_EXT_STORAGE_FP16 : RG_EXT_STORAGE_FP32;
        const uint32_t nodeFmt  = ( prec & RG_EXT_PRECISION_FP16_NODES ) ? RG_EXT_BVH_NODE_FP16 : RG_EXT_BVH_NODE_FP32;
        // at most two frames in flight: the frame two back has finished
        const uint32_t slot = framesDrawn % 2u;
        RG_HIP_CHECK( hipEventSynchronize( frameDone[ slot ] ) );
        if( !fb || fb->fb.width != w || fb->fb.height != h || fb->fb.storage != storage )
        {
            Quiesce(); // the last frame may still use the old planes
            fb = std::make_unique< RgExtFramebuffers_T >();
            fb->fb.Create( w, h, storage );
            cam.firstFrame = true;
        }
        else
        {
            fb->fb.Swap();
        }

        hipStream_t s = stream;
        const auto* ext = FindParams< RgExtDrawFramePathParams >( di, RG_EXT_STRUCTURE_TYPE_PATH_PARAMS );
        g.pt.collectStats = ext && ext->collectTraversalStats;
        RG_HIP_CHECK( hipEventRecord( ev[ ST_BEGIN ], s ) );
        g.pt.UploadLights( scene, s );
        // static geometry: a frame without primitive uploads keeps the resident scene
        const bool keep = staticGeometry && scene.triCount == 0 && dscene.triCount > 0;
        if( !keep )
        {
            UploadScene( scene, dscene, s );
        }
        else
        {
            scene.geomUids = scene.prevGeomUids; // the resident geometry is last frame's
        }
        {
            std::vector< uint32_t > geomMap;
            scene.GeomPrevToCur( geomMap );
            g.pt.UploadGeomMap( geomMap, s );
        }
        RG_HIP_CHECK( hipEventRecord( ev[ ST_UPLOADED ], s ) );
        const bool changed = keep ? false : AssembleScene( dscene, s );
        RG_HIP_CHECK( hipEventRecord( ev[ ST_ASSEMBLED ], s ) );
        uint32_t rebuilt = 0;
        if( dscene.triCount > 0 &&
            ( changed || !dscene.bvhValid || bvh.nodeFormat != nodeFmt ||
              std::memcmp( dscene.builtClassTris, dscene.classTris, sizeof( dscene.classTris ) ) != 0 ) )
        {
            // one triangle per leaf: measured fastest at 1M triangles (more
            // nodes, fewer triangle tests per leaf visit)
            RgExtBvhBuildParams bp = { 1, nodeFmt };
            BuildFrameBvh( bvh, dscene.tris.as< float >(), dscene.classTris, GC_COUNT, bp, s );
            dscene.bvhValid      = true;
            dscene.builtTriCount = dscene.triCount;
            std::memcpy( dscene.builtClassTris, dscene.classTris, sizeof( dscene.classTris ) );
            rebuilt              = 1;
        }
        RG_HIP_CHECK( hipEventRecord( ev[ ST_BVH ], s ) );
        dscene.tex          = textures.View( s );
        const uint64_t rays = EncodeFrame( g, tc, cam.params, bvh, dscene, fb->fb, di, s, ev[ ST_TRACED ], ev[ ST_DENOISED ] );
        RG_HIP_CHECK( hipEventRecord( ev[ ST_END ], s ) );
        RG_HIP_CHECK( hipEventRecord( frameDone[ slot ], s ) );
        framesDrawn++;

        // what the host knows now; the GPU figures follow in Stats()
        stats               = RgExtFrameStats{};
        stats.frameId       = cam.frameId;
        stats.renderWidth   = w;
        stats.renderHeight  = h;
        stats.triangleCount = dscene.triCount;
        stats.bvhNodeCount  = bvh.nodeCount;
        stats.raysTraced    = rays;
        stats.bvhRebuilt    = rebuilt;
        statsPending        = true;
        cam.frameId++;
        cam.firstFrame = false;
    }
};

} // namespace rg

using namespace rg;

namespace
{
std::unique_ptr< Instance > g_instance;

Instance& Get( RgInstance h )
{
    if( !g_instance || h != reinterpret_cast< RgInstance >( g_instance.get() ) )
    {
        throw Error( RG_RESULT_WRONG_INSTANCE, "wrong RgInstance" );
    }
    return *g_instance;
}
} // namespace

extern "C" {

RgResult rgCreateInstance( const RgInstanceCreateInfo* pInfo, RgInstance* pResult )
{
    return Guard( [ & ] {
        RG_REQUIRE( pResult, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "pResult is null" );
        *pResult = nullptr;
        if( g_instance )
        {
            throw Error( RG_RESULT_ALREADY_INITIALIZED, "an instance already exists" );
        }
        RgInstanceCreateInfo ci{};
        if( pInfo )
        {
            ci = *pInfo;
        }
        g_instance = std::make_unique< Instance >( ci );
        *pResult   = reinterpret_cast< RgInstance >( g_instance.get() );
    } );
}

RgResult rgDestroyInstance( RgInstance instance )
{
    return Guard( [ & ] {
        Get( instance );
        g_instance.reset();
    } );
}

RgResult rgUploadMeshPrimitive( RgInstance instance, const RgMeshInfo* pMesh, const RgMeshPrimitiveInfo* pPrimitive )
{
    return Guard( [ & ] {
        Instance& in = Get( instance );
        RG_REQUIRE( pMesh && pPrimitive, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "Argument is null" );
        in.scene.AddPrimitive( *pMesh, *pPrimitive );
    } );
}

RgResult rgUploadNonWorldPrimitive( RgInstance instance, const RgMeshPrimitiveInfo*, const float*, const RgViewport* )
{
    return Guard( [ & ] { Get( instance ); } ); // rasterised HUD: not on the ray-traced path
}

RgResult rgUploadDecal( RgInstance instance, const RgDecalUploadInfo* )
{
    return Guard( [ & ] { Get( instance ); } );
}

RgResult rgUploadLensFlare( RgInstance instance, const RgLensFlareUploadInfo* )
{
    return Guard( [ & ] { Get( instance ); } );
}

RgResult rgUploadDirectionalLight( RgInstance instance, const RgDirectionalLightUploadInfo* pInfo )
{
    return Guard( [ & ] {
        Instance& in = Get( instance );
        RG_REQUIRE( pInfo, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "Argument is null" );
        in.scene.AddDirectional( *pInfo );
    } );
}

RgResult rgUploadSphericalLight( RgInstance instance, const RgSphericalLightUploadInfo* pInfo )
{
    return Guard( [ & ] {
        Instance& in = Get( instance );
        RG_REQUIRE( pInfo, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "Argument is null" );
        in.scene.AddSphere( *pInfo );
    } );
}

RgResult rgUploadSpotLight( RgInstance instance, const RgSpotLightUploadInfo* pInfo )
{
    return Guard( [ & ] {
        Instance& in = Get( instance );
        RG_REQUIRE( pInfo, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "Argument is null" );
        in.scene.AddSpot( *pInfo );
    } );
}

RgResult rgUploadPolygonalLight( RgInstance instance, const RgPolygonalLightUploadInfo* pInfo )
{
    return Guard( [ & ] {
        Instance& in = Get( instance );
        RG_REQUIRE( pInfo, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "Argument is null" );
        in.scene.AddPolygonal( *pInfo );
    } );
}

RgResult rgProvideOriginalTexture( RgInstance instance, const RgOriginalTextureInfo* pInfo )
{
    return Guard( [ & ] {
        Instance& in = Get( instance );
        RG_REQUIRE( pInfo, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "Argument is null" );
        in.Quiesce(); // frames in flight may sample the texture table
        in.textures.Provide( *pInfo, in.stream );
    } );
}

RgResult rgProvideOriginalCubemapTexture( RgInstance instance, const RgOriginalCubemapInfo* )
{
    return Guard( [ & ] { Get( instance ); } );
}

RgResult rgMarkOriginalTextureAsDeleted( RgInstance instance, const char* pTextureName )
{
    return Guard( [ & ] {
        Instance& in = Get( instance );
        in.Quiesce();
        in.textures.MarkDeleted( pTextureName );
    } );
}

RgResult rgStartFrame( RgInstance instance, const RgStartFrameInfo* )
{
    return Guard( [ & ] { Get( instance ).StartFrame(); } );
}

RgResult rgDrawFrame( RgInstance instance, const RgDrawFrameInfo* pInfo )
{
    return Guard( [ & ] {
        Instance& in = Get( instance );
        RG_REQUIRE( pInfo, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "Argument is null" );
        in.DrawFrame( *pInfo );
    } );
}

// ---------------------------------------------------------- utilities

RgPrimitiveVertex* rgUtilScratchAllocForVertices( RgInstance, uint32_t vertexCount )
{
    return vertexCount ? new( std::nothrow ) RgPrimitiveVertex[ vertexCount ]() : nullptr;
}

void rgUtilScratchFree( RgInstance, const RgPrimitiveVertex* pPointer )
{
    delete[] pPointer;
}

void rgUtilScratchGetIndices( RgInstance, RgUtilImScratchTopology, uint32_t, const uint32_t** ppOutIndices, uint32_t* pOutIndexCount )
{
    if( ppOutIndices )
    {
        *ppOutIndices = nullptr;
    }
    if( pOutIndexCount )
    {
        *pOutIndexCount = 0;
    }
}

void rgUtilImScratchClear( RgInstance ) {}
void rgUtilImScratchStart( RgInstance, RgUtilImScratchTopology ) {}
void rgUtilImScratchVertex( RgInstance, float, float, float ) {}
void rgUtilImScratchNormal( RgInstance, float, float, float ) {}
void rgUtilImScratchTexCoord( RgInstance, float, float ) {}
void rgUtilImScratchTexCoord_Layer1( RgInstance, float, float ) {}
void rgUtilImScratchTexCoord_Layer2( RgInstance, float, float ) {}
void rgUtilImScratchTexCoord_Layer3( RgInstance, float, float ) {}
void rgUtilImScratchColor( RgInstance, RgColor4DPacked32 ) {}
void rgUtilImScratchEnd( RgInstance ) {}
void rgUtilImScratchSetToPrimitive( RgInstance, RgMeshPrimitiveInfo* ) {}

RgBool32 rgUtilIsUpscaleTechniqueAvailable( RgInstance, RgRenderUpscaleTechnique technique )
{
    return technique == RG_RENDER_UPSCALE_TECHNIQUE_LINEAR || technique == RG_RENDER_UPSCALE_TECHNIQUE_NEAREST;
}

const char* rgUtilGetResultDescription( RgResult result )
{
    switch( result )
    {
        case RG_RESULT_SUCCESS: return "Success";
        case RG_RESULT_SUCCESS_FOUND_MESH: return "Success. Found mesh";
        case RG_RESULT_SUCCESS_FOUND_TEXTURE: return "Success. Found texture";
        case RG_RESULT_WRONG_INSTANCE: return "Wrong instance";
        case RG_RESULT_ALREADY_INITIALIZED: return "Instance was already initialized";
        case RG_RESULT_GRAPHICS_API_ERROR: return "Graphics API error";
        case RG_RESULT_INTERNAL_ERROR: return "Internal error";
        case RG_RESULT_CANT_FIND_SUPPORTED_PHYSICAL_DEVICE: return "Can't find a supported device";
        case RG_RESULT_FRAME_WASNT_STARTED: return "Frame wasn't started";
        case RG_RESULT_FRAME_WASNT_ENDED: return "Frame wasn't ended";
        case RG_RESULT_WRONG_FUNCTION_CALL: return "Wrong function call";
        case RG_RESULT_WRONG_FUNCTION_ARGUMENT: return "Wrong function argument";
        case RG_RESULT_ERROR_CANT_FIND_HARDCODED_RESOURCES: return "Can't find hardcoded resources";
        case RG_RESULT_ERROR_CANT_FIND_SHADER: return "Can't find shader";
        default: return "Unknown RgResult";
    }
}

RgColor4DPacked32 rgUtilPackColorByte4D( uint8_t r, uint8_t g, uint8_t b, uint8_t a )
{
    return ( uint32_t( a ) << 24 ) | ( uint32_t( b ) << 16 ) | ( uint32_t( g ) << 8 ) | uint32_t( r );
}

RgColor4DPacked32 rgUtilPackColorFloat4D( float r, float g, float b, float a )
{
    auto q = []( float x ) { return uint8_t( std::clamp( x, 0.0f, 1.0f ) * 255.0f ); };
    return rgUtilPackColorByte4D( q( r ), q( g ), q( b ), q( a ) );
}

void rgUtilExportAsTGA( RgInstance, const void*, uint32_t, uint32_t, const char* ) {}

// ------------------------------------------------------ ext: instance

RgResult rgExtGetFrameStats( RgInstance instance, RgExtFrameStats* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        *pOut = Get( instance ).Stats();
    } );
}

RgResult rgExtGetFramebuffers( RgInstance instance, RgExtFramebuffers* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        Instance& in = Get( instance );
        RG_REQUIRE( in.fb, RG_RESULT_WRONG_FUNCTION_CALL, "no frame was drawn yet" );
        in.Quiesce(); // the planes are read and written by the host after this
        *pOut = in.fb.get();
    } );
}

RgResult rgExtGetFrameParams( RgInstance instance, RgExtFrameParams* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        *pOut = Get( instance ).cam.params;
    } );
}

RgResult rgExtReadbackHistogram( RgInstance instance, uint32_t* pOut256 )
{
    return Guard( [ & ] {
        RG_REQUIRE( pOut256, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        Instance& in = Get( instance );
        in.Quiesce();
        RG_HIP_CHECK( hipMemcpy( pOut256, in.g.histCopy.ptr, 256 * 4, hipMemcpyDeviceToHost ) );
    } );
}

RgResult rgExtSetStaticGeometry( RgInstance instance, RgBool32 enable )
{
    return Guard( [ & ] { Get( instance ).staticGeometry = enable != 0; } );
}

RgResult rgExtGetStream( RgInstance instance, void** pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        *pOut = Get( instance ).stream;
    } );
}

} // extern "C"
