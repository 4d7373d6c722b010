// The RTGL1 C API (include/RTGL1/RTGL1.h) on MI355X.
//
// Mirrors the reference's API shim and frame orchestration: Source/RTGL1.cpp
// (singleton instance, exception -> RgResult), VulkanDevice::StartFrame /
// DrawFrame / FillUniform / Render (Source/VulkanDevice.cpp:142-893), with
// the hot segment of Render (:540-588) replaced by the HIP passes: scene
// assembly, LBVH build, ray-traced passes, A-SVGF, eye adaptation.
#include "rg_bvh.h"
#include "rg_denoise.h"
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

// Column-major 4x4 inverse in FP64 (cofactor expansion).
void Inverse4( const float* m, float* out )
{
    double a[ 16 ], inv[ 16 ];
    for( int i = 0; i < 16; i++ )
    {
        a[ i ] = m[ i ];
    }
    inv[ 0 ]  = a[ 5 ] * a[ 10 ] * a[ 15 ] - a[ 5 ] * a[ 11 ] * a[ 14 ] - a[ 9 ] * a[ 6 ] * a[ 15 ] + a[ 9 ] * a[ 7 ] * a[ 14 ] + a[ 13 ] * a[ 6 ] * a[ 11 ] - a[ 13 ] * a[ 7 ] * a[ 10 ];
    inv[ 4 ]  = -a[ 4 ] * a[ 10 ] * a[ 15 ] + a[ 4 ] * a[ 11 ] * a[ 14 ] + a[ 8 ] * a[ 6 ] * a[ 15 ] - a[ 8 ] * a[ 7 ] * a[ 14 ] - a[ 12 ] * a[ 6 ] * a[ 11 ] + a[ 12 ] * a[ 7 ] * a[ 10 ];
    inv[ 8 ]  = a[ 4 ] * a[ 9 ] * a[ 15 ] - a[ 4 ] * a[ 11 ] * a[ 13 ] - a[ 8 ] * a[ 5 ] * a[ 15 ] + a[ 8 ] * a[ 7 ] * a[ 13 ] + a[ 12 ] * a[ 5 ] * a[ 11 ] - a[ 12 ] * a[ 7 ] * a[ 9 ];
    inv[ 12 ] = -a[ 4 ] * a[ 9 ] * a[ 14 ] + a[ 4 ] * a[ 10 ] * a[ 13 ] + a[ 8 ] * a[ 5 ] * a[ 14 ] - a[ 8 ] * a[ 6 ] * a[ 13 ] - a[ 12 ] * a[ 5 ] * a[ 10 ] + a[ 12 ] * a[ 6 ] * a[ 9 ];
    inv[ 1 ]  = -a[ 1 ] * a[ 10 ] * a[ 15 ] + a[ 1 ] * a[ 11 ] * a[ 14 ] + a[ 9 ] * a[ 2 ] * a[ 15 ] - a[ 9 ] * a[ 3 ] * a[ 14 ] - a[ 13 ] * a[ 2 ] * a[ 11 ] + a[ 13 ] * a[ 3 ] * a[ 10 ];
    inv[ 5 ]  = a[ 0 ] * a[ 10 ] * a[ 15 ] - a[ 0 ] * a[ 11 ] * a[ 14 ] - a[ 8 ] * a[ 2 ] * a[ 15 ] + a[ 8 ] * a[ 3 ] * a[ 14 ] + a[ 12 ] * a[ 2 ] * a[ 11 ] - a[ 12 ] * a[ 3 ] * a[ 10 ];
    inv[ 9 ]  = -a[ 0 ] * a[ 9 ] * a[ 15 ] + a[ 0 ] * a[ 11 ] * a[ 13 ] + a[ 8 ] * a[ 1 ] * a[ 15 ] - a[ 8 ] * a[ 3 ] * a[ 13 ] - a[ 12 ] * a[ 1 ] * a[ 11 ] + a[ 12 ] * a[ 3 ] * a[ 9 ];
    inv[ 13 ] = a[ 0 ] * a[ 9 ] * a[ 14 ] - a[ 0 ] * a[ 10 ] * a[ 13 ] - a[ 8 ] * a[ 1 ] * a[ 14 ] + a[ 8 ] * a[ 2 ] * a[ 13 ] + a[ 12 ] * a[ 1 ] * a[ 10 ] - a[ 12 ] * a[ 2 ] * a[ 9 ];
    inv[ 2 ]  = a[ 1 ] * a[ 6 ] * a[ 15 ] - a[ 1 ] * a[ 7 ] * a[ 14 ] - a[ 5 ] * a[ 2 ] * a[ 15 ] + a[ 5 ] * a[ 3 ] * a[ 14 ] + a[ 13 ] * a[ 2 ] * a[ 7 ] - a[ 13 ] * a[ 3 ] * a[ 6 ];
    inv[ 6 ]  = -a[ 0 ] * a[ 6 ] * a[ 15 ] + a[ 0 ] * a[ 7 ] * a[ 14 ] + a[ 4 ] * a[ 2 ] * a[ 15 ] - a[ 4 ] * a[ 3 ] * a[ 14 ] - a[ 12 ] * a[ 2 ] * a[ 7 ] + a[ 12 ] * a[ 3 ] * a[ 6 ];
    inv[ 10 ] = a[ 0 ] * a[ 5 ] * a[ 15 ] - a[ 0 ] * a[ 7 ] * a[ 13 ] - a[ 4 ] * a[ 1 ] * a[ 15 ] + a[ 4 ] * a[ 3 ] * a[ 13 ] + a[ 12 ] * a[ 1 ] * a[ 7 ] - a[ 12 ] * a[ 3 ] * a[ 5 ];
    inv[ 14 ] = -a[ 0 ] * a[ 5 ] * a[ 14 ] + a[ 0 ] * a[ 6 ] * a[ 13 ] + a[ 4 ] * a[ 1 ] * a[ 14 ] - a[ 4 ] * a[ 2 ] * a[ 13 ] - a[ 12 ] * a[ 1 ] * a[ 6 ] + a[ 12 ] * a[ 2 ] * a[ 5 ];
    inv[ 3 ]  = -a[ 1 ] * a[ 6 ] * a[ 11 ] + a[ 1 ] * a[ 7 ] * a[ 10 ] + a[ 5 ] * a[ 2 ] * a[ 11 ] - a[ 5 ] * a[ 3 ] * a[ 10 ] - a[ 9 ] * a[ 2 ] * a[ 7 ] + a[ 9 ] * a[ 3 ] * a[ 6 ];
    inv[ 7 ]  = a[ 0 ] * a[ 6 ] * a[ 11 ] - a[ 0 ] * a[ 7 ] * a[ 10 ] - a[ 4 ] * a[ 2 ] * a[ 11 ] + a[ 4 ] * a[ 3 ] * a[ 10 ] + a[ 8 ] * a[ 2 ] * a[ 7 ] - a[ 8 ] * a[ 3 ] * a[ 6 ];
    inv[ 11 ] = -a[ 0 ] * a[ 5 ] * a[ 11 ] + a[ 0 ] * a[ 7 ] * a[ 9 ] + a[ 4 ] * a[ 1 ] * a[ 11 ] - a[ 4 ] * a[ 3 ] * a[ 9 ] - a[ 8 ] * a[ 1 ] * a[ 7 ] + a[ 8 ] * a[ 3 ] * a[ 5 ];
    inv[ 15 ] = a[ 0 ] * a[ 5 ] * a[ 10 ] - a[ 0 ] * a[ 6 ] * a[ 9 ] - a[ 4 ] * a[ 1 ] * a[ 10 ] + a[ 4 ] * a[ 2 ] * a[ 9 ] + a[ 8 ] * a[ 1 ] * a[ 6 ] - a[ 8 ] * a[ 2 ] * a[ 5 ];
    double det = a[ 0 ] * inv[ 0 ] + a[ 1 ] * inv[ 4 ] + a[ 2 ] * inv[ 8 ] + a[ 3 ] * inv[ 12 ];
    det        = det != 0.0 ? 1.0 / det : 0.0;
    for( int i = 0; i < 16; i++ )
    {
        out[ i ] = float( inv[ i ] * det );
    }
}

// Matrix::MakeProjectionMatrix (Source/Matrix.cpp:380-395)
void MakeProjection( float* m, float aspect, float fovY, float zn, float zf )
{
    const float t = tanf( fovY * 0.5f );
    std::memset( m, 0, 16 * sizeof( float ) );
    m[ 0 ]  = 1.0f / ( aspect * t );
    m[ 5 ]  = -1.0f / t;
    m[ 10 ] = zf / ( zn - zf );
    m[ 11 ] = -1.0f;
    m[ 14 ] = ( zf * zn ) / ( zn - zf );
}

template< typename T >
const T* FindParams( const RgDrawFrameInfo& info, RgStructureType type )
{
    for( auto* n = static_cast< const RgDrawFrameTonemappingParams* >( info.pParams ); n;
         n       = static_cast< const RgDrawFrameTonemappingParams* >( n->pNext ) )
    {
        if( n->sType == type )
        {
            return reinterpret_cast< const T* >( n );
        }
    }
    return nullptr;
}

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
    BvhObject                                   bvh;
    TraceContext                                tc;
    PathTracer                                  pt;
    std::unique_ptr< RgExtFramebuffers_T >      fb;
    DeviceBuffer                                hist, histCopy, avgLum;
    RgExtFrameParams                            params{};
    RgExtFrameStats                             stats{};
    bool                                        frameStarted = false;
    bool                                        staticGeometry = false;
    bool                                        firstFrame   = true;
    uint32_t                                    frameId      = 0;
    double                                      prevTime     = 0.0;
    hipEvent_t                                  ev[ ST_COUNT ] = {};

    Instance( const RgInstanceCreateInfo& ci )
    {
        info = ci;
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
        hist.Reserve( 256 * 4 );
        histCopy.Reserve( 256 * 4 );
        avgLum.Reserve( 4 );
        RG_HIP_CHECK( hipMemsetAsync( hist.ptr, 0, 256 * 4, stream ) );
        RG_HIP_CHECK( hipMemsetAsync( avgLum.ptr, 0, 4, stream ) );
        pt.Init( stream );
        RgExtFrameParams d;
        rgExtFrameParamsDefaults( 1280, 720, &d );
        params = d;
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
        scene.BeginFrame();
    }

    // FillUniform (Source/VulkanDevice.cpp:142-482), hot-path fields only.
    void FillParams( const RgDrawFrameInfo& di, uint32_t w, uint32_t h )
    {
        RgExtFrameParams& p = params;
        std::memcpy( p.viewPrev, p.view, sizeof( p.view ) );
        std::memcpy( p.projectionPrev, p.projection, sizeof( p.projection ) );
        std::memcpy( p.cameraPositionPrev, p.cameraPosition, sizeof( p.cameraPosition ) );
        std::memcpy( p.view, di.view, sizeof( p.view ) );
        MakeProjection( p.projection, float( w ) / float( h ), di.fovYRadians, di.cameraNear, di.cameraFar );
        Inverse4( p.view, p.invView );
        Inverse4( p.projection, p.invProjection );
        p.cameraPosition[ 0 ] = p.invView[ 12 ];
        p.cameraPosition[ 1 ] = p.invView[ 13 ];
        p.cameraPosition[ 2 ] = p.invView[ 14 ];
        p.cameraPosition[ 3 ] = 0.0f;
        if( firstFrame )
        {
            std::memcpy( p.viewPrev, p.view, sizeof( p.view ) );
            std::memcpy( p.projectionPrev, p.projection, sizeof( p.projection ) );
            std::memcpy( p.cameraPositionPrev, p.cameraPosition, sizeof( p.cameraPosition ) );
        }
        p.renderWidth  = w;
        p.renderHeight = h;
        p.frameId      = frameId;
        p.timeDelta    = float( std::max( di.currentTime - prevTime, 0.001 ) );
        prevTime       = di.currentTime;
        p.jitterX = p.jitterY = 0.0f;

        const auto* tm      = FindParams< RgDrawFrameTonemappingParams >( di, RG_STRUCTURE_TYPE_TONEMAPPING );
        const float ev100Min = tm ? tm->ev100Min : 0.0f, ev100Max = tm ? tm->ev100Max : 10.0f;
        p.stopEyeAdaptation  = tm ? tm->disableEyeAdaptation : 0;
        p.minLogLuminance    = std::log2( std::exp2( ev100Min ) * 12.5f / 100.0f );
        p.maxLogLuminance    = std::log2( std::exp2( ev100Max ) * 12.5f / 100.0f );

        p.lightCount             = uint32_t( scene.lights.size() ) - 1u;
        p.lightCountPrev         = scene.prevLightCount > 0 ? scene.prevLightCount - 1u : 0u;
        p.directionalLightExists = scene.hasDirectional ? 1u : 0u;

        const auto* sky = FindParams< RgDrawFrameSkyParams >( di, RG_STRUCTURE_TYPE_SKY );
        if( sky )
        {
            for( int i = 0; i < 3; i++ )
            {
                p.skyColorDefault[ i ] = sky->skyColorDefault.data[ i ];
            }
            p.skyColorMultiplier = std::max( 0.0f, sky->skyColorMultiplier );
            p.skyColorSaturation = std::max( 0.0f, sky->skyColorSaturation );
        }
        const auto* il = FindParams< RgDrawFrameIlluminationParams >( di, RG_STRUCTURE_TYPE_ILLUMINATION );
        p.maxBounceShadowsLights   = il ? il->maxBounceShadows : 2u;
        p.polyLightSpotlightFactor = il ? std::max( 0.0f, il->polygonalLightSpotlightFactor ) : 2.0f;
        p.indirSecondBounce        = il ? !!il->enableSecondBounceForIndirect : 1u;
        p.cellWorldSize            = il ? std::max( il->cellWorldSize, 0.001f ) : 1.0f;
        p.gradientMultDiffuse      = il ? std::clamp( il->directDiffuseSensitivityToChange, 0.0f, 1.0f ) : 0.5f;
        p.gradientMultIndirect     = il ? std::clamp( il->indirectDiffuseSensitivityToChange, 0.0f, 1.0f ) : 0.2f;
        p.gradientMultSpecular     = il ? std::clamp( il->specularSensitivityToChange, 0.0f, 1.0f ) : 0.5f;
        const auto* tx             = FindParams< RgDrawFrameTexturesParams >( di, RG_STRUCTURE_TYPE_TEXTURES );
        p.minRoughness             = tx ? std::clamp( tx->minRoughness, 0.0f, 1.0f ) : 0.0f;
        p.emissionMapBoost         = tx ? std::max( tx->emissionMapBoost, 0.0f ) : 100.0f;

        p.rayCullBackFaces  = info.rayCullBackFacingTriangles ? 1u : 0u;
        p.rayLength         = std::clamp( di.rayLength, 0.1f, 10000.0f );
        p.primaryRayMinDist = std::clamp( di.cameraNear, 0.001f, p.rayLength );
        p.antiFireflyEnabled = 1;

        const auto* ext = FindParams< RgExtDrawFramePathParams >( di, RG_EXT_STRUCTURE_TYPE_PATH_PARAMS );
        p.atrousIterations        = ext && ext->atrousIterations ? ext->atrousIterations : 4u;
        p.restirInitialCandidates = ext && ext->restirInitialCandidates ? ext->restirInitialCandidates : 8u;
        p.restirTemporalSamples   = ext && ext->restirTemporalSamples != 0xFFFFFFFFu ? ext->restirTemporalSamples : 1u;
        p.restirSpatialSamples    = ext && ext->restirSpatialSamples != 0xFFFFFFFFu ? ext->restirSpatialSamples : 8u;
        RG_REQUIRE( p.atrousIterations == 4 || p.atrousIterations == 5, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "atrousIterations must be 4 or 5" );
    }

    void RenderSize( const RgDrawFrameInfo& di, uint32_t* w, uint32_t* h )
    {
        *w      = 1280;
        *h      = 720;
        auto rr = FindParams< RgDrawFrameRenderResolutionParams >( di, RG_STRUCTURE_TYPE_RENDER_RESOLUTION );
        if( rr && rr->resolutionMode == RG_RENDER_RESOLUTION_MODE_CUSTOM && rr->customRenderSize.width && rr->customRenderSize.height )
        {
            *w = rr->customRenderSize.width;
            *h = rr->customRenderSize.height;
        }
        auto ext = FindParams< RgExtDrawFramePathParams >( di, RG_EXT_STRUCTURE_TYPE_PATH_PARAMS );
        if( ext && ext->renderWidth && ext->renderHeight )
        {
            *w = ext->renderWidth;
            *h = ext->renderHeight;
        }
        *w &= ~1u; // render width must be even (VulkanDevice.cpp:186-187)
        RG_REQUIRE( *w >= 2 && *h >= 1, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "render size too small" );
    }

    void DrawFrame( const RgDrawFrameInfo& di )
    {
        if( !frameStarted )
        {
            throw Error( RG_RESULT_FRAME_WASNT_STARTED, "rgDrawFrame was called without rgStartFrame" );
        }
        frameStarted = false;
        uint32_t w, h;
        RenderSize( di, &w, &h );
        scene.FinishLights();
        FillParams( di, w, h );
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
        if( !fb || fb->fb.width != w || fb->fb.height != h || fb->fb.storage != storage )
        {
            fb = std::make_unique< RgExtFramebuffers_T >();
            fb->fb.Create( w, h, storage );
            firstFrame = true;
        }
        else
        {
            fb->fb.Swap();
        }

        hipStream_t s = stream;
        const auto* ext = FindParams< RgExtDrawFramePathParams >( di, RG_EXT_STRUCTURE_TYPE_PATH_PARAMS );
        pt.collectStats = ext && ext->collectTraversalStats;
        RG_HIP_CHECK( hipEventRecord( ev[ ST_BEGIN ], s ) );
        pt.UploadLights( scene, s );
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
            pt.UploadGeomMap( geomMap, s );
        }
        RG_HIP_CHECK( hipEventRecord( ev[ ST_UPLOADED ], s ) );
        const bool changed = keep ? false : AssembleScene( dscene, s );
        RG_HIP_CHECK( hipEventRecord( ev[ ST_ASSEMBLED ], s ) );
        stats.bvhRebuilt = 0;
        if( dscene.triCount > 0 && ( changed || !dscene.bvhValid || bvh.nodeFormat != nodeFmt ) )
        {
            // one triangle per leaf: measured fastest at 1M triangles (more
            // nodes, fewer triangle tests per leaf visit)
            RgExtBvhBuildParams bp = { 1, nodeFmt };
            BuildBvh( bvh, dscene.tris.as< float >(), dscene.triCount, bp, s );
            dscene.bvhValid      = true;
            dscene.builtTriCount = dscene.triCount;
            stats.bvhRebuilt     = 1;
        }
        RG_HIP_CHECK( hipEventRecord( ev[ ST_BVH ], s ) );
        const uint32_t flags = params.rayCullBackFaces ? RG_EXT_TRACE_CULL_FRONT_FACE : 0u;
        const uint64_t rays  = pt.Trace( params, bvh, tc, dscene, fb->fb, flags, s, nullptr );
        RG_HIP_CHECK( hipEventRecord( ev[ ST_TRACED ], s ) );
        Denoise( fb->fb, params, s );
        RG_HIP_CHECK( hipEventRecord( ev[ ST_DENOISED ], s ) );
        LuminanceHistogram( reinterpret_cast< const float* >( fb->fb.Cur( RG_EXT_PLANE_HISTOGRAM_INPUT ) ), w * h, params.minLogLuminance,
                            params.maxLogLuminance, hist.as< uint32_t >(), s );
        RG_HIP_CHECK( hipMemcpyAsync( histCopy.ptr, hist.ptr, 256 * 4, hipMemcpyDeviceToDevice, s ) );
        LuminanceAverage( hist.as< uint32_t >(), w * h, params.minLogLuminance, params.maxLogLuminance, params.timeDelta,
                          params.stopEyeAdaptation != 0, avgLum.as< float >(), s );
        {
            // ImageComposition::Finalize (Source/ImageComposition.cpp:94-156)
            const auto* tm     = FindParams< RgDrawFrameTonemappingParams >( di, RG_STRUCTURE_TYPE_TONEMAPPING );
            float       sat[ 3 ] = { 0.0f, 0.0f, 0.0f }, ct[ 3 ] = { 1.0f, 1.0f, 1.0f }; // DrawFrameInfo.h:104-105
            if( tm )
            {
                std::memcpy( sat, tm->saturation.data, sizeof( sat ) );
                std::memcpy( ct, tm->crosstalk.data, sizeof( ct ) );
            }
            ComposeFinal( fb->fb, MakeLpmConsts( sat, ct ), avgLum.as< float >(), s );
        }
        RG_HIP_CHECK( hipEventRecord( ev[ ST_END ], s ) );
        // The staging arrays are reused by the next frame's uploads.
        RG_HIP_CHECK( hipStreamSynchronize( s ) );

        auto ms = [ & ]( Stage a, Stage b ) {
            float v = 0.0f;
            ( void )hipEventElapsedTime( &v, ev[ a ], ev[ b ] );
            return v;
        };
        stats.frameId        = frameId;
        stats.renderWidth    = w;
        stats.renderHeight   = h;
        stats.triangleCount  = dscene.triCount;
        stats.bvhNodeCount   = bvh.nodeCount;
        stats.raysTraced     = rays;
        stats.gpuMsUpload    = ms( ST_BEGIN, ST_UPLOADED );
        stats.gpuMsAssemble  = ms( ST_UPLOADED, ST_ASSEMBLED );
        stats.gpuMsTraversal = pt.TraversalMs();
        stats.traversalLaunches = pt.traceLaunches;
        pt.TraversalCounts( &stats.nodeVisits, &stats.triangleTests, &stats.traversalErrors );
        stats.gpuMsBvh       = ms( ST_ASSEMBLED, ST_BVH );
        stats.gpuMsTrace     = ms( ST_BVH, ST_TRACED );
        stats.gpuMsDenoise   = ms( ST_TRACED, ST_DENOISED );
        stats.gpuMsTonemap   = ms( ST_DENOISED, ST_END );
        stats.gpuMsTotal     = ms( ST_BEGIN, ST_END );
        RG_HIP_CHECK( hipMemcpy( &stats.avgLuminance, avgLum.ptr, 4, hipMemcpyDeviceToHost ) );
        frameId++;
        firstFrame = false;
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

RgResult rgProvideOriginalTexture( RgInstance instance, const RgOriginalTextureInfo* )
{
    return Guard( [ & ] { Get( instance ); } );
}

RgResult rgProvideOriginalCubemapTexture( RgInstance instance, const RgOriginalCubemapInfo* )
{
    return Guard( [ & ] { Get( instance ); } );
}

RgResult rgMarkOriginalTextureAsDeleted( RgInstance instance, const char* )
{
    return Guard( [ & ] { Get( instance ); } );
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
        *pOut = Get( instance ).stats;
    } );
}

RgResult rgExtGetFramebuffers( RgInstance instance, RgExtFramebuffers* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        Instance& in = Get( instance );
        RG_REQUIRE( in.fb, RG_RESULT_WRONG_FUNCTION_CALL, "no frame was drawn yet" );
        *pOut = in.fb.get();
    } );
}

RgResult rgExtGetFrameParams( RgInstance instance, RgExtFrameParams* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        *pOut = Get( instance ).params;
    } );
}

RgResult rgExtReadbackHistogram( RgInstance instance, uint32_t* pOut256 )
{
    return Guard( [ & ] {
        RG_REQUIRE( pOut256, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        Instance& in = Get( instance );
        RG_HIP_CHECK( hipMemcpy( pOut256, in.histCopy.ptr, 256 * 4, hipMemcpyDeviceToHost ) );
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
