// extern "C" entry points of RTGL1_ext.h: device memory helpers, timers, the
// LBVH and traversal stages, and eye adaptation.
#include <algorithm>
#include "rg_bvh.h"
#include "rg_denoise.h"
#include "rg_internal.h"
#include "rg_kernels.h"
#include "rg_pathtrace.h"

#include <cmath>
#include <memory>

namespace rg
{
namespace
{
thread_local std::string g_lastError;
}

void SetLastError( const char* msg )
{
    g_lastError = msg ? msg : "";
}
} // namespace rg

using namespace rg;

struct RgExtTimer_T
{
    hipEvent_t start = nullptr;
    hipEvent_t stop  = nullptr;
};

struct RgExtBvh_T
{
    BvhObject    bvh;
    TraceContext trace;
    bool         lastNoAlloc = false; // the last trace was rgExtTraceRaysNoAlloc
};

namespace
{
// Plain 16-byte-per-lane copy: reads and writes each byte once.
__global__ __launch_bounds__( 256 ) void k_copy16( uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n )
{
    const size_t stride = size_t( gridDim.x ) * blockDim.x;
    for( size_t i = size_t( blockIdx.x ) * blockDim.x + threadIdx.x; i < n; i += stride )
    {
        dst[ i ] = src[ i ];
    }
}
} // namespace

extern "C" {

// ------------------------------------------------------------ memory

RgResult rgExtDeviceAlloc( size_t bytes, void** pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( pOut != nullptr, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "pOut is null" );
        *pOut = nullptr;
        RG_HIP_CHECK( hipMalloc( pOut, bytes > 0 ? bytes : 1 ) );
    } );
}

RgResult rgExtDeviceFree( void* d_ptr )
{
    return Guard( [ & ] {
        if( d_ptr )
        {
            RG_HIP_CHECK( hipFree( d_ptr ) );
        }
    } );
}

RgResult rgExtCopyHostToDevice( void* d_dst, const void* src, size_t bytes )
{
    return Guard( [ & ] {
        RG_REQUIRE( ( d_dst && src ) || bytes == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null pointer" );
        if( bytes )
        {
            // after the work of every stream (an instance keeps frames in
            // flight on its own non-blocking stream)
            RG_HIP_CHECK( hipDeviceSynchronize() );
            RG_HIP_CHECK( hipMemcpy( d_dst, src, bytes, hipMemcpyHostToDevice ) );
        }
    } );
}

RgResult rgExtCopyDeviceToHost( void* dst, const void* d_src, size_t bytes )
{
    return Guard( [ & ] {
        RG_REQUIRE( ( dst && d_src ) || bytes == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null pointer" );
        if( bytes )
        {
            // after the work of every stream (an instance keeps frames in
            // flight on its own non-blocking stream)
            RG_HIP_CHECK( hipDeviceSynchronize() );
            RG_HIP_CHECK( hipMemcpy( dst, d_src, bytes, hipMemcpyDeviceToHost ) );
        }
    } );
}

RgResult rgExtCopyDeviceToDevice( void* d_dst, const void* d_src, size_t bytes, void* stream )
{
    return Guard( [ & ] {
        RG_REQUIRE( ( d_dst && d_src ) || bytes == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null pointer" );
        RG_REQUIRE( bytes % 16 == 0 && ( reinterpret_cast< uintptr_t >( d_dst ) % 16 ) == 0 &&
                        ( reinterpret_cast< uintptr_t >( d_src ) % 16 ) == 0,
                    RG_RESULT_WRONG_FUNCTION_ARGUMENT, "pointers and size must be 16-byte aligned" );
        if( bytes == 0 )
        {
            return;
        }
        const size_t n      = bytes / 16;
        const size_t blocks = ( n + 255 ) / 256 < 8192 ? ( n + 255 ) / 256 : 8192;
        k_copy16<<< uint32_t( blocks ), 256, 0, AsStream( stream ) >>>( static_cast< uint4* >( d_dst ), static_cast< const uint4* >( d_src ), n );
        CheckLaunch( "k_copy16" );
    } );
}

RgResult rgExtMemsetDevice( void* d_dst, int value, size_t bytes )
{
    return Guard( [ & ] {
        RG_REQUIRE( d_dst || bytes == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null pointer" );
        if( bytes )
        {
            // after the work of every stream (an instance keeps frames in
            // flight on its own non-blocking stream)
            RG_HIP_CHECK( hipDeviceSynchronize() );
            RG_HIP_CHECK( hipMemset( d_dst, value, bytes ) );
        }
    } );
}

RgResult rgExtSynchronize( void )
{
    return Guard( [ & ] { RG_HIP_CHECK( hipDeviceSynchronize() ); } );
}

int rgExtGetDeviceCount( void )
{
    int n = 0;
    if( hipGetDeviceCount( &n ) != hipSuccess )
    {
        return 0;
    }
    return n;
}

const char* rgExtGetLastErrorMessage( void )
{
    return g_lastError.c_str();
}

// ------------------------------------------------------------- timers

RgResult rgExtTimerCreate( RgExtTimer* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "pOut is null" );
        auto t = std::make_unique< RgExtTimer_T >();
        RG_HIP_CHECK( hipEventCreate( &t->start ) );
        RG_HIP_CHECK( hipEventCreate( &t->stop ) );
        *pOut = t.release();
    } );
}

RgResult rgExtTimerDestroy( RgExtTimer timer )
{
    return Guard( [ & ] {
        if( timer )
        {
            ( void )hipEventDestroy( timer->start );
            ( void )hipEventDestroy( timer->stop );
            delete timer;
        }
    } );
}

RgResult rgExtTimerStart( RgExtTimer timer, void* stream )
{
    return Guard( [ & ] {
        RG_REQUIRE( timer, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "timer is null" );
        RG_HIP_CHECK( hipEventRecord( timer->start, AsStream( stream ) ) );
    } );
}

RgResult rgExtTimerStop( RgExtTimer timer, void* stream )
{
    return Guard( [ & ] {
        RG_REQUIRE( timer, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "timer is null" );
        RG_HIP_CHECK( hipEventRecord( timer->stop, AsStream( stream ) ) );
    } );
}

RgResult rgExtTimerElapsedMs( RgExtTimer timer, float* pMs )
{
    return Guard( [ & ] {
        RG_REQUIRE( timer && pMs, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        RG_HIP_CHECK( hipEventSynchronize( timer->stop ) );
        RG_HIP_CHECK( hipEventElapsedTime( pMs, timer->start, timer->stop ) );
    } );
}

// --------------------------------------------------------------- LBVH

RgResult rgExtBvhCreate( RgExtBvh* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "pOut is null" );
        *pOut = new RgExtBvh_T();
    } );
}

RgResult rgExtBvhDestroy( RgExtBvh bvh )
{
    return Guard( [ & ] { delete bvh; } );
}

RgResult rgExtBvhBuild( RgExtBvh bvh, const float* d_triangles, uint32_t triangleCount, const RgExtBvhBuildParams* pParams, void* stream )
{
    return Guard( [ & ] {
        RG_REQUIRE( bvh && d_triangles, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        RgExtBvhBuildParams p = { 3, RG_EXT_BVH_NODE_FP32 };
        if( pParams )
        {
            p = *pParams;
        }
        BuildBvh( bvh->bvh, d_triangles, triangleCount, p, AsStream( stream ) );
    } );
}

RgResult rgExtBvhGetView( RgExtBvh bvh, RgExtBvhView* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( bvh && pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        const BvhObject& b = bvh->bvh;
        RG_REQUIRE( b.nodeCount > 0, RG_RESULT_WRONG_FUNCTION_CALL, "BVH is not built" );
        pOut->triangleCount     = b.n;
        pOut->nodeCount         = b.nodeCount;
        pOut->nodeFormat        = b.nodeFormat;
        pOut->nodeStrideBytes   = b.NodeStride();
        pOut->d_nodes           = b.nodes.ptr;
        pOut->d_sortedTriangles = reinterpret_cast< const float* >( b.sortedTris.ptr );
        pOut->d_mortonCodes     = reinterpret_cast< const uint64_t* >( b.codes.ptr );
        pOut->d_sortedIndices   = b.sortedIdx;
        pOut->d_binaryChildren  = reinterpret_cast< const uint32_t* >( b.children.ptr );
        pOut->d_binaryBounds    = reinterpret_cast< const float* >( b.nodeBounds.ptr );
        for( int k = 0; k < 3; k++ )
        {
            pOut->sceneCentroidMin[ k ] = b.sceneMin[ k ];
            pOut->sceneCentroidMax[ k ] = b.sceneMax[ k ];
        }
    } );
}

// ---------------------------------------------------------- traversal

RgResult rgExtTraceRays( RgExtBvh bvh, const RgExtRay* d_rays, uint32_t rayCount, RgExtHit* d_hits, uint32_t flags, void* stream )
{
    return Guard( [ & ] {
        RG_REQUIRE( bvh && ( ( d_rays && d_hits ) || rayCount == 0 ), RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        RG_REQUIRE( ( flags & ~7u ) == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "unknown trace flags" );
        bvh->lastNoAlloc = false;
        TraceRays( bvh->bvh, bvh->trace, d_rays, rayCount, d_hits, flags, AsStream( stream ) );
    } );
}

RgResult rgExtTraceRaysNoAlloc( RgExtBvh bvh, const RgExtRay* d_rays, uint32_t rayCount, RgExtHit* d_hits, uint32_t flags, void* stream )
{
    return Guard( [ & ] {
        RG_REQUIRE( bvh && ( ( d_rays && d_hits ) || rayCount == 0 ), RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        RG_REQUIRE( ( flags & ~7u ) == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "unknown trace flags" );
        TraceRaysNoAlloc( bvh->bvh, d_rays, rayCount, d_hits, flags, AsStream( stream ) );
        bvh->lastNoAlloc = true;
    } );
}

RgResult rgExtBvhBuildBatch( RgExtBvh bvh, const float* d_triangles, const uint32_t* sceneTriangleOffsets, uint32_t sceneCount,
                             const RgExtBvhBuildParams* pParams, void* stream )
{
    return Guard( [ & ] {
        RG_REQUIRE( bvh && d_triangles && sceneTriangleOffsets, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        RgExtBvhBuildParams p = { 3, RG_EXT_BVH_NODE_FP32 };
        if( pParams )
        {
            p = *pParams;
        }
        BuildBvhBatch( bvh->bvh, d_triangles, sceneTriangleOffsets, sceneCount, p, AsStream( stream ) );
    } );
}

RgResult rgExtBvhGetSceneRoots( RgExtBvh bvh, uint32_t* pRoots, uint32_t capacity, uint32_t* pSceneCount )
{
    return Guard( [ & ] {
        RG_REQUIRE( bvh && pSceneCount && ( pRoots || capacity == 0 ), RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        const auto& r = bvh->bvh.hostSceneRoots;
        *pSceneCount  = bvh->bvh.sceneCount;
        for( uint32_t i = 0; i < capacity && i < bvh->bvh.sceneCount; i++ )
        {
            pRoots[ i ] = r[ i ];
        }
    } );
}

RgResult rgExtTraceRaysBatch( RgExtBvh bvh, const RgExtRay* d_rays, const uint32_t* rayOffsets, RgExtHit* d_hits, uint32_t flags, void* stream )
{
    return Guard( [ & ] {
        RG_REQUIRE( bvh && rayOffsets, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        RG_REQUIRE( ( flags & ~7u ) == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "unknown trace flags" );
        RG_REQUIRE( ( d_rays && d_hits ) || rayOffsets[ bvh->bvh.sceneCount ] == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        TraceRaysBatch( bvh->bvh, bvh->trace, d_rays, rayOffsets, d_hits, flags, AsStream( stream ) );
    } );
}

RgResult rgExtPrimaryRays( const RgExtFrameParams* pParams, RgExtRay* d_rays, void* stream )
{
    return Guard( [ & ] {
        RG_REQUIRE( pParams && d_rays, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        PrimaryRays( *pParams, d_rays, AsStream( stream ) );
    } );
}

RgResult rgExtTraceAutotune( RgExtBvh bvh, const RgExtRay* d_rays, uint32_t rayCount, const uint32_t* rayOffsets, RgExtHit* d_hits, uint32_t flags,
                             void* stream, RgExtTraceTuning* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( bvh && d_rays && d_hits, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        RG_REQUIRE( ( flags & ~7u ) == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "unknown trace flags" );
        AutotuneTrace( bvh->bvh, bvh->trace, d_rays, rayCount, rayOffsets, d_hits, flags, AsStream( stream ), pOut );
    } );
}

RgResult rgExtBvhGetTraceStats( RgExtBvh bvh, uint64_t* pNodeVisits, uint64_t* pTriangleTests )
{
    return Guard( [ & ] {
        RG_REQUIRE( bvh && pNodeVisits && pTriangleTests, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        if( bvh->lastNoAlloc )
        {
            ReadTraceNoAllocWords( nullptr, pNodeVisits, pTriangleTests );
            return;
        }
        ReadTraceStats( bvh->trace, pNodeVisits, pTriangleTests );
    } );
}

RgResult rgExtBvhGetTraceErrors( RgExtBvh bvh, uint32_t* pErrors )
{
    return Guard( [ & ] {
        RG_REQUIRE( bvh && pErrors, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        if( bvh->lastNoAlloc )
        {
            ReadTraceNoAllocWords( pErrors, nullptr, nullptr );
            return;
        }
        *pErrors = ReadTraceErrors( bvh->trace );
    } );
}

// ------------------------------------------------------------ tonemap

RgResult rgExtLuminanceHistogram( const float* d_luminance, uint32_t width, uint32_t height, float minLogLuminance, float maxLogLuminance, uint32_t* d_histogram, void* stream )
{
    return Guard( [ & ] {
        RG_REQUIRE( d_histogram && ( d_luminance || width * height == 0 ), RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        LuminanceHistogram( d_luminance, width * height, minLogLuminance, maxLogLuminance, d_histogram, AsStream( stream ) );
    } );
}

RgResult rgExtComposeFinalEx( RgExtFramebuffers fbs, const RgDrawFrameTonemappingParams* tonemapping, const RgDrawFrameTexturesParams* textures,
                              const float* d_avgLuminance, void* stream )
{
    return Guard( [ & ] {
        RG_REQUIRE( fbs && d_avgLuminance, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        float sat[ 3 ] = { 0.0f, 0.0f, 0.0f }, ct[ 3 ] = { 1.0f, 1.0f, 1.0f };
        if( tonemapping )
        {
            std::memcpy( sat, tonemapping->saturation.data, sizeof( sat ) );
            std::memcpy( ct, tonemapping->crosstalk.data, sizeof( ct ) );
        }
        const float emis = textures ? std::max( textures->emissionMaxScreenColor, 0.0f ) : 1.5f; // VulkanDevice.cpp:278
        ComposeFinal( fbs->fb, MakeLpmConsts( sat, ct ), emis, d_avgLuminance, AsStream( stream ) );
    } );
}

RgResult rgExtComposeFinal( RgExtFramebuffers fbs, const RgDrawFrameTonemappingParams* tonemapping, const float* d_avgLuminance, void* stream )
{
    return rgExtComposeFinalEx( fbs, tonemapping, nullptr, d_avgLuminance, stream );
}

RgResult rgExtLuminanceAverage( uint32_t* d_histogram, uint32_t pixelCount, float minLogLuminance, float maxLogLuminance, float timeDelta, RgBool32 stopEyeAdaptation, float* d_avgLuminance, void* stream )
{
    return Guard( [ & ] {
        RG_REQUIRE( d_histogram && d_avgLuminance, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        LuminanceAverage( d_histogram, pixelCount, minLogLuminance, maxLogLuminance, timeDelta, stopEyeAdaptation != 0, d_avgLuminance, AsStream( stream ) );
    } );
}

// ------------------------------------------------------- framebuffers

RgResult rgExtFramebuffersCreate( uint32_t width, uint32_t height, RgExtFramebuffers* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "pOut is null" );
        auto f = std::make_unique< RgExtFramebuffers_T >();
        f->fb.Create( width, height );
        *pOut = f.release();
    } );
}

RgResult rgExtFramebuffersCreateWithStorage( uint32_t width, uint32_t height, uint32_t storage, RgExtFramebuffers* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "pOut is null" );
        auto f = std::make_unique< RgExtFramebuffers_T >();
        f->fb.Create( width, height, storage );
        *pOut = f.release();
    } );
}

RgResult rgExtFramebuffersDestroy( RgExtFramebuffers fb )
{
    return Guard( [ & ] { delete fb; } );
}

RgResult rgExtFramebuffersGetPlane( RgExtFramebuffers fb, RgExtPlane plane, RgBool32 prev, void** d_pOut, size_t* pBytes )
{
    return Guard( [ & ] {
        RG_REQUIRE( fb && d_pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        RG_REQUIRE( int( plane ) >= 0 && plane < RG_EXT_PLANE_COUNT, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "unknown plane" );
        *d_pOut = prev ? fb->fb.Prev( plane ) : fb->fb.Cur( plane );
        if( pBytes )
        {
            *pBytes = fb->fb.PlaneBytes( plane );
        }
    } );
}

RgResult rgExtFramebuffersSwap( RgExtFramebuffers fb )
{
    return Guard( [ & ] {
        RG_REQUIRE( fb, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        fb->fb.Swap();
    } );
}

RgResult rgExtFrameParamsDefaults( uint32_t width, uint32_t height, RgExtFrameParams* p )
{
    return Guard( [ & ] {
        RG_REQUIRE( p, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        std::memset( p, 0, sizeof( *p ) );
        p->renderWidth  = width;
        p->renderHeight = height;
        p->timeDelta    = 1.0f / 60.0f;
        for( int i = 0; i < 4; i++ )
        {
            p->view[ i * 5 ] = p->invView[ i * 5 ] = p->viewPrev[ i * 5 ] = 1.0f;
            p->projection[ i * 5 ] = p->invProjection[ i * 5 ] = p->projectionPrev[ i * 5 ] = 1.0f;
        }
        // Tonemapping defaults ev100 in [0, 10] (DrawFrameInfo.h:95-108),
        // converted as FillUniform does (VulkanDevice.cpp:202-212).
        p->minLogLuminance          = std::log2( std::exp2( 0.0f ) * 12.5f / 100.0f );
        p->maxLogLuminance          = std::log2( std::exp2( 10.0f ) * 12.5f / 100.0f );
        p->antiFireflyEnabled       = 1;
        p->primaryRayMinDist        = 0.1f;
        p->rayLength                = 10000.0f;
        p->skyColorDefault[ 0 ]     = 199.0f / 255.0f;
        p->skyColorDefault[ 1 ]     = 233.0f / 255.0f;
        p->skyColorDefault[ 2 ]     = 255.0f / 255.0f;
        p->skyColorMultiplier       = 1000.0f;
        p->skyColorSaturation       = 1.0f;
        p->cellWorldSize            = 1.0f;
        p->maxBounceShadowsLights   = 2;
        p->gradientMultDiffuse      = 0.5f;
        p->gradientMultIndirect     = 0.2f;
        p->gradientMultSpecular     = 0.5f;
        p->indirSecondBounce        = 1;
        p->polyLightSpotlightFactor = 2.0f;
        p->atrousIterations         = 4;
        p->restirInitialCandidates  = 8;
        p->restirTemporalSamples    = 1;
        p->restirSpatialSamples     = 8;
        p->emissionMapBoost         = 100.0f;
        // RgDrawFrameReflectRefractParams defaults (DrawFrameInfo.h:125-146)
        p->reflectRefractMaxDepth   = 2;
        p->cameraMediaType          = RG_MEDIA_TYPE_VACUUM;
        p->indexOfRefractionGlass   = 1.52f;
        p->indexOfRefractionWater   = 1.33f;
        p->thinMediaWidth           = 0.1f;
        const float water[ 4 ] = { 0.3f, 0.73f, 0.63f, 0.0f }, acid[ 4 ] = { 0.0f, 0.66f, 0.55f, 10.0f };
        std::memcpy( p->waterColorAndDensity, water, sizeof( water ) );
        std::memcpy( p->acidColorAndDensity, acid, sizeof( acid ) );
        p->worldUp[ 1 ] = 1.0f;
        p->normalMapStrength = 1.0f; // DrawFrameInfo.h:175
        p->waterWaveStrength = 1.0f; // DrawFrameInfo.h:138
        p->emissionMaxScreenColor = 1.5f; // DrawFrameInfo.h:177
        p->rayCullMaskWorld           = 7u; // every world
        p->rayCullMaskWorldShadow     = 7u;
        p->lightIndexIgnoreFPVShadows = 32767u; // LIGHT_INDEX_NONE
        p->skyGeometry                = 0u;
    } );
}

// ----------------------------------------------------------- denoiser

#define RG_DENOISE_ENTRY( expr )                                                                 \
    return Guard( [ & ] {                                                                        \
        RG_REQUIRE( fb && p, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );               \
        expr;                                                                                    \
    } )

RgResult rgExtDenoiseGradientAtrous( RgExtFramebuffers fb, const RgExtFrameParams* p, uint32_t iteration, void* stream )
{
    RG_DENOISE_ENTRY( DenoiseGradientAtrous( fb->fb, *p, iteration, AsStream( stream ) ) );
}

RgResult rgExtDenoiseTemporal( RgExtFramebuffers fb, const RgExtFrameParams* p, void* stream )
{
    RG_DENOISE_ENTRY( DenoiseTemporal( fb->fb, *p, AsStream( stream ) ) );
}

RgResult rgExtDenoiseAntiFirefly( RgExtFramebuffers fb, const RgExtFrameParams* p, void* stream )
{
    RG_DENOISE_ENTRY( DenoiseAntiFirefly( fb->fb, *p, AsStream( stream ) ) );
}

RgResult rgExtDenoiseVariance( RgExtFramebuffers fb, const RgExtFrameParams* p, void* stream )
{
    RG_DENOISE_ENTRY( DenoiseVariance( fb->fb, *p, AsStream( stream ) ) );
}

RgResult rgExtDenoiseAtrous( RgExtFramebuffers fb, const RgExtFrameParams* p, uint32_t iteration, void* stream )
{
    RG_DENOISE_ENTRY( DenoiseAtrous( fb->fb, *p, iteration, AsStream( stream ) ) );
}

RgResult rgExtDenoise( RgExtFramebuffers fb, const RgExtFrameParams* p, void* stream )
{
    RG_DENOISE_ENTRY( Denoise( fb->fb, *p, AsStream( stream ) ) );
}

RgResult rgExtFramebuffersSetAtrousTile( RgExtFramebuffers fb, uint32_t tileWidth )
{
    return Guard( [ & ] {
        RG_REQUIRE( fb, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        SetAtrousTile( fb->fb, tileWidth );
    } );
}

RgResult rgExtDenoiseAutotune( RgExtFramebuffers fb, const RgExtFrameParams* p, void* stream, uint32_t* pTileWidth, float* pMs )
{
    RG_DENOISE_ENTRY( {
        const uint32_t w = AutotuneAtrous( fb->fb, *p, AsStream( stream ), pMs );
        if( pTileWidth )
        {
            *pTileWidth = w;
        }
    } );
}

} // extern "C"
