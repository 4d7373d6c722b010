// Device-wide exclusive prefix sum of uint32 arrays (reduce / scan partials /
// downsweep), used by the radix sort offsets and the BVH8 node numbering.
#include "rg_internal.h"
#include "rg_scan.h"

namespace rg
{

namespace
{
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS   = 8;
constexpr int SCAN_TILE    = SCAN_THREADS * SCAN_ITEMS;

// Inclusive scan of one value per thread over the block; returns the
// exclusive prefix and writes the block total.
__device__ uint32_t BlockExclusiveScan( uint32_t v, uint32_t* total )
{
    __shared__ uint32_t s_wave[ SCAN_THREADS / 64 ];
    const int           lane = threadIdx.x & 63;
    const int           wave = threadIdx.x >> 6;

    uint32_t incl = v;
    for( int off = 1; off < 64; off <<= 1 )
    {
        uint32_t n = __shfl_up( incl, off, 64 );
        if( lane >= off )
        {
            incl += n;
        }
    }
    if( lane == 63 )
    {
        s_wave[ wave ] = incl;
    }
    __syncthreads();
    uint32_t waveBase = 0, sum = 0;
    for( int w = 0; w < SCAN_THREADS / 64; w++ )
    {
        if( w < wave )
        {
            waveBase += s_wave[ w ];
        }
        sum += s_wave[ w ];
    }
    __syncthreads();
    *total = sum;
    return waveBase + incl - v;
}

__global__ __launch_bounds__( SCAN_THREADS ) void k_scan_reduce( const uint32_t* in, uint32_t n, uint32_t* partials )
{
    const uint64_t base = uint64_t( blockIdx.x ) * SCAN_TILE;
    uint32_t       s    = 0;
    for( int k = 0; k < SCAN_ITEMS; k++ )
    {
        uint64_t i = base + uint64_t( k ) * SCAN_THREADS + threadIdx.x;
        if( i < n )
        {
            s += in[ i ];
        }
    }
    uint32_t total;
    ( void )BlockExclusiveScan( s, &total );
    if( threadIdx.x == 0 )
    {
        partials[ blockIdx.x ] = total;
    }
}

// Single block: exclusive scan of the partials in place; *grandTotal gets the sum.
__global__ __launch_bounds__( SCAN_THREADS ) void k_scan_partials( uint32_t* partials, uint32_t count, uint32_t* grandTotal )
{
    uint32_t carry = 0;
    for( uint32_t base = 0; base < count; base += SCAN_THREADS )
    {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < count ? partials[ i ] : 0u;
        uint32_t total;
        uint32_t ex = BlockExclusiveScan( v, &total );
        if( i < count )
        {
            partials[ i ] = carry + ex;
        }
        carry += total;
    }
    if( threadIdx.x == 0 && grandTotal )
    {
        *grandTotal = carry;
    }
}

// Each thread owns SCAN_ITEMS consecutive elements of the tile.
__global__ __launch_bounds__( SCAN_THREADS ) void k_scan_downsweep( const uint32_t* in,
                                                                    uint32_t        n,
                                                                    const uint32_t* partials,
                                                                    uint32_t*       out )
{
    const uint64_t base = uint64_t( blockIdx.x ) * SCAN_TILE + uint64_t( threadIdx.x ) * SCAN_ITEMS;
    uint32_t       vals[ SCAN_ITEMS ];
    uint32_t       s = 0;
#pragma unroll
    for( int k = 0; k < SCAN_ITEMS; k++ )
    {
        uint64_t i = base + k;
        vals[ k ]  = i < n ? in[ i ] : 0u;
        s += vals[ k ];
    }
    uint32_t total;
    uint32_t run = BlockExclusiveScan( s, &total ) + partials[ blockIdx.x ];
#pragma unroll
    for( int k = 0; k < SCAN_ITEMS; k++ )
    {
        uint64_t i = base + k;
        if( i < n )
        {
            out[ i ] = run;
        }
        run += vals[ k ];
    }
}
} // namespace

size_t ScanScratchBytes( uint32_t n )
{
    return size_t( DivUp( n, SCAN_TILE ) + 1 ) * sizeof( uint32_t );
}

void ExclusiveScan( const uint32_t* d_in, uint32_t n, uint32_t* d_out, uint32_t* d_total, uint32_t* d_scratch, hipStream_t stream )
{
    if( n == 0 )
    {
        if( d_total )
        {
            RG_HIP_CHECK( hipMemsetAsync( d_total, 0, sizeof( uint32_t ), stream ) );
        }
        return;
    }
    const uint32_t blocks = DivUp( n, SCAN_TILE );
    k_scan_reduce<<< blocks, SCAN_THREADS, 0, stream >>>( d_in, n, d_scratch );
    CheckLaunch( "k_scan_reduce" );
    k_scan_partials<<< 1, SCAN_THREADS, 0, stream >>>( d_scratch, blocks, d_total );
    CheckLaunch( "k_scan_partials" );
    k_scan_downsweep<<< blocks, SCAN_THREADS, 0, stream >>>( d_in, n, d_scratch, d_out );
    CheckLaunch( "k_scan_downsweep" );
}

} // namespace rg
