// extern "C" entry points of RTGL1_ext.h: device memory helpers, timers, the
// LBVH and traversal stages, and eye adaptation.
#include "rg_bvh.h"
#include "rg_internal.h"
#include "rg_kernels.h"

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
};

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
            RG_HIP_CHECK( hipMemcpy( dst, d_src, bytes, hipMemcpyDeviceToHost ) );
        }
    } );
}

RgResult rgExtMemsetDevice( void* d_dst, int value, size_t bytes )
{
    return Guard( [ & ] {
        RG_REQUIRE( d_dst || bytes == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null pointer" );
        if( bytes )
        {
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
        RG_REQUIRE( ( flags & ~3u ) == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "unknown trace flags" );
        TraceRays( bvh->bvh, bvh->trace, d_rays, rayCount, d_hits, flags, AsStream( stream ) );
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

RgResult rgExtLuminanceAverage( uint32_t* d_histogram, uint32_t pixelCount, float minLogLuminance, float maxLogLuminance, float timeDelta, RgBool32 stopEyeAdaptation, float* d_avgLuminance, void* stream )
{
    return Guard( [ & ] {
        RG_REQUIRE( d_histogram && d_avgLuminance, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        LuminanceAverage( d_histogram, pixelCount, minLogLuminance, maxLogLuminance, timeDelta, stopEyeAdaptation != 0, d_avgLuminance, AsStream( stream ) );
    } );
}

} // extern "C"
