// Luminance histogram and adaptive average luminance (eye adaptation).
//
// CmLuminanceHistogram.comp:41-86 builds a 256-bin histogram per 16x16
// group in shared memory and adds it to the global one; here a persistent
// grid of 256-thread blocks walks the image with a grid stride, so each block
// folds many pixels into its LDS bins before one global atomic per bin.
// CmLuminanceAvg.comp:39-90 is a single 256-thread group.
#include "rg_device_math.h"
#include "rg_internal.h"
#include "rg_kernels.h"

namespace rg
{
namespace
{
constexpr int HIST_THREADS = 256;
constexpr int HIST_BINS    = 256;

// getColorBin (CmLuminanceHistogram.comp:41-54), with a deterministic log2.
// NaN luminance goes to bin 0, +inf to bin 255.
__device__ __forceinline__ uint32_t ColorBin( float lum, float minLog, float invRange )
{
    if( !( lum >= 0.001f ) )
    {
        return 0;
    }
    if( lum > 3.0e38f )
    {
        return 255;
    }
    const float l = clampf( ( rgLog2( lum ) - minLog ) * invRange, 0.0f, 1.0f );
    return uint32_t( l * 254.0f + 1.0f );
}

__global__ __launch_bounds__( HIST_THREADS ) void k_lum_histogram( const float* lum, uint32_t count, float minLog, float maxLog, uint32_t* hist )
{
    __shared__ uint32_t s_bins[ HIST_BINS ];
    s_bins[ threadIdx.x ] = 0;
    __syncthreads();
    const float invRange = 1.0f / ( maxLog - minLog );
    for( uint32_t i = blockIdx.x * HIST_THREADS + threadIdx.x; i < count; i += gridDim.x * HIST_THREADS )
    {
        atomicAdd( &s_bins[ ColorBin( lum[ i ], minLog, invRange ) ], 1u );
    }
    __syncthreads();
    const uint32_t c = s_bins[ threadIdx.x ];
    if( c != 0 )
    {
        atomicAdd( &hist[ threadIdx.x ], c );
    }
}

__global__ __launch_bounds__( HIST_BINS ) void k_lum_average( uint32_t* hist,
                                                              uint32_t  pixelCount,
                                                              float     minLog,
                                                              float     maxLog,
                                                              float     timeDelta,
                                                              uint32_t  stop,
                                                              float*    avgLum )
{
    __shared__ uint32_t s[ HIST_BINS ];
    const uint32_t      i     = threadIdx.x;
    const uint32_t      count = hist[ i ];
    s[ i ]                    = count * i;
    __syncthreads();
    hist[ i ] = 0; // cleared for the next frame
    for( uint32_t cutoff = HIST_BINS >> 1; cutoff > 0; cutoff >>= 1 )
    {
        if( i < cutoff )
        {
            s[ i ] += s[ i + cutoff ];
        }
        __syncthreads();
    }
    if( i == 0 && stop == 0 )
    {
        const float range     = maxLog - minLog;
        const float pixels    = float( pixelCount );
        const float black     = float( count );
        const float weighted  = ( float( s[ 0 ] ) / fmaxf( pixels - black, 1.0f ) ) - 1.0f;
        const float avg       = rgExp2( weighted / 254.0f * range + minLog );
        const float last      = *avgLum;
        // t = 1 - exp(-dt * 1.1), exp(x) = 2^(x log2 e)
        const float t         = 1.0f - rgExp2( -timeDelta * 1.1f * 1.44269504f );
        const float adapted   = last + ( avg - last ) * t;
        *avgLum               = fminf( adapted, rgExp2( maxLog ) );
    }
}

int HistogramGrid()
{
    static int blocks = 0;
    if( blocks == 0 )
    {
        int dev = 0, cus = 0;
        ( void )hipGetDevice( &dev );
        ( void )hipDeviceGetAttribute( &cus, hipDeviceAttributeMultiprocessorCount, dev );
        blocks = ( cus > 0 ? cus : 64 ) * 4;
    }
    return blocks;
}
} // namespace

void LuminanceHistogram( const float* d_lum, uint32_t count, float minLog, float maxLog, uint32_t* d_hist, hipStream_t stream )
{
    RG_REQUIRE( maxLog > minLog, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "maxLogLuminance must exceed minLogLuminance" );
    if( count == 0 )
    {
        return;
    }
    const uint32_t blocks = DivUp( count, HIST_THREADS ) < uint32_t( HistogramGrid() ) ? DivUp( count, HIST_THREADS ) : uint32_t( HistogramGrid() );
    k_lum_histogram<<< blocks, HIST_THREADS, 0, stream >>>( d_lum, count, minLog, maxLog, d_hist );
    CheckLaunch( "k_lum_histogram" );
}

void LuminanceAverage( uint32_t* d_hist, uint32_t pixelCount, float minLog, float maxLog, float timeDelta, bool stop, float* d_avg, hipStream_t stream )
{
    k_lum_average<<< 1, HIST_BINS, 0, stream >>>( d_hist, pixelCount, minLog, maxLog, timeDelta, stop ? 1u : 0u, d_avg );
    CheckLaunch( "k_lum_average" );
}

} // namespace rg
