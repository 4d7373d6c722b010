// Linear BVH build for gfx950: FP64 bounds and Morton quantisation, stable
// LSD radix sort with wave-ballot ranking, Karras (2012) hierarchy, bottom-up
// refit with agent-scope acquire/release counters, and a deterministic
// collapse into 8-wide nodes.
//
// This replaces the driver-side acceleration-structure build of the reference
// (ASBuilder::BuildBottomLevel / BuildTopLevel, Source/ASBuilder.cpp:141-206);
// the reference has no BVH code of its own, so the build follows the
// published algorithms named above.
#include "rg_bvh.h"
#include "rg_device_math.h"
#include "rg_scan.h"

#include <cfloat>
#include <cmath>

namespace rg
{
namespace
{

constexpr int      THREADS     = 256;
constexpr int      RS_ITEMS    = 16;
constexpr int      RS_TILE     = THREADS * RS_ITEMS;
constexpr uint32_t MORTON_BITS = 21;
constexpr double   MORTON_CELLS = 2097152.0; // 2^21

// ------------------------------------------------------------- bounds

__device__ __forceinline__ void LoadTri( const float* tris, uint32_t i, v3& a, v3& b, v3& c )
{
    const float* t = tris + size_t( i ) * 9;
    a              = mk3( t[ 0 ], t[ 1 ], t[ 2 ] );
    b              = mk3( t[ 3 ], t[ 4 ], t[ 5 ] );
    c              = mk3( t[ 6 ], t[ 7 ], t[ 8 ] );
}

// Centroid in FP64: ((v0 + v1) + v2) / 3 per axis.
__device__ __forceinline__ void Centroid( const float* tris, uint32_t i, double c[ 3 ] )
{
    const float* t = tris + size_t( i ) * 9;
    for( int k = 0; k < 3; k++ )
    {
        c[ k ] = ( ( double( t[ k ] ) + double( t[ 3 + k ] ) ) + double( t[ 6 + k ] ) ) / 3.0;
    }
}

__global__ __launch_bounds__( THREADS ) void k_centroid_bounds( const float* tris, uint32_t n, double* partials )
{
    __shared__ double s[ 6 ][ THREADS ];
    double            lo[ 3 ] = { DBL_MAX, DBL_MAX, DBL_MAX };
    double            hi[ 3 ] = { -DBL_MAX, -DBL_MAX, -DBL_MAX };
    for( uint32_t i = blockIdx.x * THREADS + threadIdx.x; i < n; i += gridDim.x * THREADS )
    {
        double c[ 3 ];
        Centroid( tris, i, c );
        for( int k = 0; k < 3; k++ )
        {
            lo[ k ] = fmin( lo[ k ], c[ k ] );
            hi[ k ] = fmax( hi[ k ], c[ k ] );
        }
    }
    for( int k = 0; k < 3; k++ )
    {
        s[ k ][ threadIdx.x ]     = lo[ k ];
        s[ 3 + k ][ threadIdx.x ] = hi[ k ];
    }
    __syncthreads();
    for( int stride = THREADS / 2; stride > 0; stride >>= 1 )
    {
        if( int( threadIdx.x ) < stride )
        {
            for( int k = 0; k < 3; k++ )
            {
                s[ k ][ threadIdx.x ]     = fmin( s[ k ][ threadIdx.x ], s[ k ][ threadIdx.x + stride ] );
                s[ 3 + k ][ threadIdx.x ] = fmax( s[ 3 + k ][ threadIdx.x ], s[ 3 + k ][ threadIdx.x + stride ] );
            }
        }
        __syncthreads();
    }
    if( threadIdx.x < 6 )
    {
        partials[ blockIdx.x * 6 + threadIdx.x ] = s[ threadIdx.x ][ 0 ];
    }
}

// One block of THREADS: strided partial reductions, then an LDS tree.
__global__ __launch_bounds__( THREADS ) void k_bounds_final( const double* partials, uint32_t blocks, double* out )
{
    __shared__ double s[ 6 ][ THREADS ];
    double            v[ 6 ] = { DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX };
    for( uint32_t b = threadIdx.x; b < blocks; b += THREADS )
    {
        for( int k = 0; k < 3; k++ )
        {
            v[ k ]     = fmin( v[ k ], partials[ b * 6 + k ] );
            v[ 3 + k ] = fmax( v[ 3 + k ], partials[ b * 6 + 3 + k ] );
        }
    }
    for( int k = 0; k < 6; k++ )
    {
        s[ k ][ threadIdx.x ] = v[ k ];
    }
    __syncthreads();
    for( int stride = THREADS / 2; stride > 0; stride >>= 1 )
    {
        if( int( threadIdx.x ) < stride )
        {
            for( int k = 0; k < 3; k++ )
            {
                s[ k ][ threadIdx.x ]     = fmin( s[ k ][ threadIdx.x ], s[ k ][ threadIdx.x + stride ] );
                s[ 3 + k ][ threadIdx.x ] = fmax( s[ 3 + k ][ threadIdx.x ], s[ 3 + k ][ threadIdx.x + stride ] );
            }
        }
        __syncthreads();
    }
    if( threadIdx.x < 6 )
    {
        out[ threadIdx.x ] = s[ threadIdx.x ][ 0 ];
    }
}

// ------------------------------------------------------------- morton

__device__ __forceinline__ uint64_t ExpandBits21( uint32_t v )
{
    uint64_t x = v & 0x1fffffu;
    x          = ( x | x << 32 ) & 0x1f00000000ffffull;
    x          = ( x | x << 16 ) & 0x1f0000ff0000ffull;
    x          = ( x | x << 8 ) & 0x100f00f00f00f00full;
    x          = ( x | x << 4 ) & 0x10c30c30c30c30c3ull;
    x          = ( x | x << 2 ) & 0x1249249249249249ull;
    return x;
}

__device__ __forceinline__ uint32_t Quantize21( double c, double lo, double hi )
{
    const double extent = hi - lo;
    if( !( extent > 0.0 ) )
    {
        return 0;
    }
    const double f = floor( ( ( c - lo ) / extent ) * MORTON_CELLS );
    return f >= MORTON_CELLS - 1.0 ? ( 1u << MORTON_BITS ) - 1u : ( f <= 0.0 ? 0u : uint32_t( f ) );
}

__global__ __launch_bounds__( THREADS ) void k_morton( const float* tris, uint32_t n, const double* bounds, uint64_t* codes, uint64_t* keys, uint32_t* vals )
{
    uint32_t i = blockIdx.x * THREADS + threadIdx.x;
    if( i >= n )
    {
        return;
    }
    double c[ 3 ];
    Centroid( tris, i, c );
    uint32_t q[ 3 ];
    for( int k = 0; k < 3; k++ )
    {
        q[ k ] = Quantize21( c[ k ], bounds[ k ], bounds[ 3 + k ] );
    }
    uint64_t code = ( ExpandBits21( q[ 0 ] ) << 2 ) | ( ExpandBits21( q[ 1 ] ) << 1 ) | ExpandBits21( q[ 2 ] );
    codes[ i ]    = code;
    keys[ i ]     = code;
    vals[ i ]     = i;
}

// ------------------------------------------------------ batched scenes
// S scenes concatenated: scene s owns triangles [offsets[s], offsets[s+1]).

__device__ __forceinline__ uint32_t SceneOf( const uint32_t* offsets, uint32_t S, uint32_t i )
{
    uint32_t lo = 0, hi = S; // last scene with offsets[s] <= i
    while( hi - lo > 1 )
    {
        const uint32_t mid = ( lo + hi ) / 2;
        if( offsets[ mid ] <= i )
        {
            lo = mid;
        }
        else
        {
            hi = mid;
        }
    }
    return lo;
}

// Block b of scene s (blockIdx.x = s * B + b) reduces part of that scene.
__global__ __launch_bounds__( THREADS ) void k_centroid_bounds_seg( const float* tris, const uint32_t* offsets, uint32_t B, double* partials )
{
    __shared__ double s[ 6 ][ THREADS ];
    const uint32_t    scene = blockIdx.x / B, b = blockIdx.x % B;
    const uint32_t    first = offsets[ scene ], last = offsets[ scene + 1 ];
    double            lo[ 3 ] = { DBL_MAX, DBL_MAX, DBL_MAX };
    double            hi[ 3 ] = { -DBL_MAX, -DBL_MAX, -DBL_MAX };
    for( uint32_t i = first + b * THREADS + threadIdx.x; i < last; i += B * THREADS )
    {
        double c[ 3 ];
        Centroid( tris, i, c );
        for( int k = 0; k < 3; k++ )
        {
            lo[ k ] = fmin( lo[ k ], c[ k ] );
            hi[ k ] = fmax( hi[ k ], c[ k ] );
        }
    }
    for( int k = 0; k < 3; k++ )
    {
        s[ k ][ threadIdx.x ]     = lo[ k ];
        s[ 3 + k ][ threadIdx.x ] = hi[ k ];
    }
    __syncthreads();
    for( int stride = THREADS / 2; stride > 0; stride >>= 1 )
    {
        if( int( threadIdx.x ) < stride )
        {
            for( int k = 0; k < 3; k++ )
            {
                s[ k ][ threadIdx.x ]     = fmin( s[ k ][ threadIdx.x ], s[ k ][ threadIdx.x + stride ] );
                s[ 3 + k ][ threadIdx.x ] = fmax( s[ 3 + k ][ threadIdx.x ], s[ 3 + k ][ threadIdx.x + stride ] );
            }
        }
        __syncthreads();
    }
    if( threadIdx.x < 6 )
    {
        partials[ size_t( blockIdx.x ) * 6 + threadIdx.x ] = s[ threadIdx.x ][ 0 ];
    }
}

// Block s combines scene s's B partials.
__global__ void k_bounds_final_seg( const double* partials, uint32_t B, double* out )
{
    if( threadIdx.x < 6 )
    {
        const bool isMin = threadIdx.x < 3;
        double     v     = isMin ? DBL_MAX : -DBL_MAX;
        for( uint32_t b = 0; b < B; b++ )
        {
            const double p = partials[ ( size_t( blockIdx.x ) * B + b ) * 6 + threadIdx.x ];
            v              = isMin ? fmin( v, p ) : fmax( v, p );
        }
        out[ size_t( blockIdx.x ) * 6 + threadIdx.x ] = v;
    }
}

// Morton code in the scene's own bounds; the sort key carries the scene in
// its top 7 bits above the code's top 57 bits (19 per axis), so one sort
// groups the scenes and orders each by its Morton curve, and the Karras
// hierarchy over the keys splits at scene boundaries first.
__global__ __launch_bounds__( THREADS ) void k_morton_seg( const float* tris, uint32_t n, const uint32_t* offsets, uint32_t S,
                                                           const double* bounds, uint64_t* codes, uint64_t* keys, uint32_t* vals )
{
    uint32_t i = blockIdx.x * THREADS + threadIdx.x;
    if( i >= n )
    {
        return;
    }
    const uint32_t scene = SceneOf( offsets, S, i );
    const double*  bb    = bounds + size_t( scene ) * 6;
    double         c[ 3 ];
    Centroid( tris, i, c );
    uint32_t q[ 3 ];
    for( int k = 0; k < 3; k++ )
    {
        q[ k ] = Quantize21( c[ k ], bb[ k ], bb[ 3 + k ] );
    }
    const uint64_t code = ( ExpandBits21( q[ 0 ] ) << 2 ) | ( ExpandBits21( q[ 1 ] ) << 1 ) | ExpandBits21( q[ 2 ] );
    codes[ i ]          = code;
    keys[ i ]           = ( uint64_t( scene ) << 57 ) | ( code >> 6 );
    vals[ i ]           = i;
}

// The internal node whose range is exactly scene s (each scene has >= 2
// triangles, and its keys differ from its neighbours' in the scene bits).
__global__ __launch_bounds__( THREADS ) void k_scene_roots( uint32_t internal, const uint32_t* rf, const uint32_t* rl, const uint32_t* offsets,
                                                            uint32_t S, uint32_t* roots )
{
    const uint32_t i = blockIdx.x * THREADS + threadIdx.x;
    if( i >= internal )
    {
        return;
    }
    const uint32_t scene = SceneOf( offsets, S, rf[ i ] );
    if( rf[ i ] == offsets[ scene ] && rl[ i ] + 1u == offsets[ scene + 1 ] )
    {
        roots[ scene ] = i;
    }
}

// Seeds the collapse with every scene root; a one-triangle scene (frame
// segments only) has no internal node and is left out.
__global__ void k_seed_frontier_seg( const uint32_t* roots, uint32_t S, uint32_t* frontier, uint32_t* counts )
{
    if( threadIdx.x == 0 )
    {
        uint32_t k = 0;
        for( uint32_t s = 0; s < S; s++ )
        {
            if( roots[ s ] != 0xFFFFFFFFu )
            {
                frontier[ k++ ] = roots[ s ];
            }
        }
        counts[ 0 ] = k;
    }
}

// 8-wide root of each scene; a one-triangle scene's root is the leaf of its
// triangle, which the sort left at the scene's first position.
__global__ void k_gather_roots8( const uint32_t* roots, const uint32_t* offsets, uint32_t S, const uint32_t* node8Index, uint32_t* roots8 )
{
    for( uint32_t s = threadIdx.x; s < S; s += blockDim.x )
    {
        roots8[ s ] = roots[ s ] != 0xFFFFFFFFu ? node8Index[ roots[ s ] ] : ( RG_EXT_BVH_CHILD_LEAF_BIT | ( offsets[ s ] << 3 ) );
    }
}

// Box of each frame segment's root: the union of an 8-wide root's child
// boxes, or a one-triangle root's triangle box.  Out: 8 floats per segment
// (lo xyz, hi xyz, 2 unused).
__global__ void k_segment_boxes( const BvhNode8* nodes, const float4* sortedTris, const uint32_t* roots8, uint32_t S, float* out )
{
    const uint32_t s = threadIdx.x;
    if( s >= S )
    {
        return;
    }
    float          lo[ 3 ] = { INFINITY, INFINITY, INFINITY }, hi[ 3 ] = { -INFINITY, -INFINITY, -INFINITY };
    const uint32_t r       = roots8[ s ];
    if( r & RG_EXT_BVH_CHILD_LEAF_BIT )
    {
        const uint32_t first = ( r >> 3 ) & 0x0FFFFFFFu;
        for( uint32_t k = 0; k <= ( r & 7u ); k++ )
        {
            for( int v = 0; v < 3; v++ )
            {
                const float4 q = sortedTris[ size_t( first + k ) * 3 + v ];
                lo[ 0 ] = fminf( lo[ 0 ], q.x ), hi[ 0 ] = fmaxf( hi[ 0 ], q.x );
                lo[ 1 ] = fminf( lo[ 1 ], q.y ), hi[ 1 ] = fmaxf( hi[ 1 ], q.y );
                lo[ 2 ] = fminf( lo[ 2 ], q.z ), hi[ 2 ] = fmaxf( hi[ 2 ], q.z );
            }
        }
    }
    else
    {
        const BvhNode8& nd = nodes[ r ];
        for( int j = 0; j < 8; j++ )
        {
            if( nd.child[ j ] != RG_EXT_BVH_CHILD_EMPTY )
            {
                lo[ 0 ] = fminf( lo[ 0 ], nd.lo_x[ j ] ), hi[ 0 ] = fmaxf( hi[ 0 ], nd.hi_x[ j ] );
                lo[ 1 ] = fminf( lo[ 1 ], nd.lo_y[ j ] ), hi[ 1 ] = fmaxf( hi[ 1 ], nd.hi_y[ j ] );
                lo[ 2 ] = fminf( lo[ 2 ], nd.lo_z[ j ] ), hi[ 2 ] = fmaxf( hi[ 2 ], nd.hi_z[ j ] );
            }
        }
    }
    for( int k = 0; k < 3; k++ )
    {
        out[ s * 8 + k ]     = lo[ k ];
        out[ s * 8 + 3 + k ] = hi[ k ];
    }
    out[ s * 8 + 6 ] = 0.0f;
    out[ s * 8 + 7 ] = 0.0f;
}

// --------------------------------------------------------- radix sort

__global__ __launch_bounds__( THREADS ) void k_radix_hist( const uint64_t* keys, uint32_t n, int shift, uint32_t* hist, uint32_t numBlocks )
{
    __shared__ uint32_t cnt[ 256 ];
    cnt[ threadIdx.x ] = 0;
    __syncthreads();
    const uint64_t base = uint64_t( blockIdx.x ) * RS_TILE;
    for( int k = 0; k < RS_ITEMS; k++ )
    {
        uint64_t i = base + uint64_t( k ) * THREADS + threadIdx.x;
        if( i < n )
        {
            atomicAdd( &cnt[ uint32_t( keys[ i ] >> shift ) & 255u ], 1u );
        }
    }
    __syncthreads();
    hist[ threadIdx.x * numBlocks + blockIdx.x ] = cnt[ threadIdx.x ];
}

// Stable scatter: keys of a tile are ranked in (round, thread) order, which is
// their input order; equal digits are ranked within a wave by ballot matching.
__global__ __launch_bounds__( THREADS ) void k_radix_scatter( const uint64_t* keysIn,
                                                              const uint32_t* valsIn,
                                                              uint64_t*       keysOut,
                                                              uint32_t*       valsOut,
                                                              uint32_t        n,
                                                              int             shift,
                                                              const uint32_t* offsets,
                                                              uint32_t        numBlocks )
{
    constexpr int       WAVES = THREADS / 64;
    __shared__ uint32_t s_base[ 256 ];
    __shared__ uint32_t s_waveCnt[ WAVES ][ 256 ];

    const int      tid  = threadIdx.x;
    const int      lane = tid & 63;
    const int      wave = tid >> 6;
    const uint64_t ltMask = ( lane == 0 ) ? 0ull : ( ~0ull >> ( 64 - lane ) );

    s_base[ tid ] = offsets[ tid * numBlocks + blockIdx.x ];
    for( int w = 0; w < WAVES; w++ )
    {
        s_waveCnt[ w ][ tid ] = 0;
    }
    __syncthreads();

    const uint64_t base = uint64_t( blockIdx.x ) * RS_TILE;
    for( int round = 0; round < RS_ITEMS; round++ )
    {
        const uint64_t i     = base + uint64_t( round ) * THREADS + tid;
        const bool     valid = i < n;
        uint64_t       key   = valid ? keysIn[ i ] : 0ull;
        uint32_t       val   = valid ? valsIn[ i ] : 0u;
        uint32_t       digit = uint32_t( key >> shift ) & 255u;

        uint64_t peers = __ballot( valid );
#pragma unroll
        for( int b = 0; b < 8; b++ )
        {
            const bool     bit = ( digit >> b ) & 1u;
            const uint64_t m   = __ballot( bit );
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = __popcll( peers & ltMask );
        if( valid && rank == 0 )
        {
            s_waveCnt[ wave ][ digit ] = __popcll( peers );
        }
        __syncthreads();
        if( valid )
        {
            uint32_t pos = s_base[ digit ] + rank;
            for( int w = 0; w < wave; w++ )
            {
                pos += s_waveCnt[ w ][ digit ];
            }
            keysOut[ pos ] = key;
            valsOut[ pos ] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for( int w = 0; w < WAVES; w++ )
        {
            add += s_waveCnt[ w ][ tid ];
            s_waveCnt[ w ][ tid ] = 0;
        }
        s_base[ tid ] += add;
        __syncthreads();
    }
}

// ------------------------------------------------------------- Karras

__device__ __forceinline__ int Delta( const uint64_t* keys, int64_t n, int64_t i, int64_t j )
{
    if( j < 0 || j >= n )
    {
        return -1;
    }
    const uint64_t a = keys[ i ], b = keys[ j ];
    if( a == b )
    {
        return 64 + __clz( uint32_t( i ) ^ uint32_t( j ) );
    }
    return __clzll( a ^ b );
}

// Internal node i of n-1; child ids >= n-1 denote leaves (sorted position id-(n-1)).
__global__ __launch_bounds__( THREADS ) void k_karras( const uint64_t* keys,
                                                       uint32_t        n,
                                                       uint32_t*       children,
                                                       uint32_t*       parents,
                                                       uint32_t*       rangeFirst,
                                                       uint32_t*       rangeLast )
{
    const int64_t i = int64_t( blockIdx.x ) * THREADS + threadIdx.x;
    const int64_t N = n;
    if( i >= N - 1 )
    {
        return;
    }
    const int d    = ( Delta( keys, N, i, i + 1 ) - Delta( keys, N, i, i - 1 ) ) >= 0 ? 1 : -1;
    const int dmin = Delta( keys, N, i, i - d );

    int64_t lmax = 2;
    while( Delta( keys, N, i, i + lmax * d ) > dmin )
    {
        lmax <<= 1;
    }
    int64_t l = 0;
    for( int64_t t = lmax >> 1; t >= 1; t >>= 1 )
    {
        if( Delta( keys, N, i, i + ( l + t ) * d ) > dmin )
        {
            l += t;
        }
    }
    const int64_t j     = i + l * d;
    const int     dnode = Delta( keys, N, i, j );

    int64_t s = 0, t = l;
    do
    {
        t = ( t + 1 ) >> 1;
        if( Delta( keys, N, i, i + ( s + t ) * d ) > dnode )
        {
            s += t;
        }
    } while( t > 1 );

    const int64_t gamma = i + s * d + ( d < 0 ? d : 0 );
    const int64_t lo = i < j ? i : j, hi = i < j ? j : i;

    const uint32_t left  = uint32_t( lo == gamma ? ( N - 1 ) + gamma : gamma );
    const uint32_t right = uint32_t( hi == gamma + 1 ? ( N - 1 ) + gamma + 1 : gamma + 1 );

    children[ i * 2 + 0 ] = left;
    children[ i * 2 + 1 ] = right;
    parents[ left ]       = uint32_t( i );
    parents[ right ]      = uint32_t( i );
    rangeFirst[ i ]       = uint32_t( lo );
    rangeLast[ i ]        = uint32_t( hi );
    if( i == 0 )
    {
        parents[ 0 ] = 0xFFFFFFFFu;
    }
}

// --------------------------------------------------------------- refit

__device__ __forceinline__ void StoreBox( float* b, v3 lo, v3 hi )
{
    b[ 0 ] = lo.x;
    b[ 1 ] = lo.y;
    b[ 2 ] = lo.z;
    b[ 3 ] = hi.x;
    b[ 4 ] = hi.y;
    b[ 5 ] = hi.z;
}

// Level-synchronous refit: the depth of every binary node by pointer
// jumping over the parent links (7 rounds reach depth 128), a counting sort
// of the internal nodes by depth, then one launch per level from the deepest
// up, each node the union of its two children.  No atomics on the boxes and
// no cross-wave ordering: min/max unions give the same boxes in any order.
constexpr int      REFIT_JUMP_ROUNDS = 7;
constexpr uint32_t REFIT_MAX_DEPTH   = 128;

__global__ __launch_bounds__( THREADS ) void k_leaf_boxes( const float* tris, const uint32_t* sortedIdx, uint32_t n, float* bounds )
{
    const uint32_t k = blockIdx.x * THREADS + threadIdx.x;
    if( k >= n )
    {
        return;
    }
    v3 a, b, c;
    LoadTri( tris, sortedIdx[ k ], a, b, c );
    StoreBox( bounds + size_t( ( n - 1 ) + k ) * 6, vmin( vmin( a, b ), c ), vmax( vmax( a, b ), c ) );
}

// dist[i] = 1 and jump[i] = parent for every node but the root (0, itself)
__global__ __launch_bounds__( THREADS ) void k_depth_init( const uint32_t* parents, uint32_t count, uint32_t* dist, uint32_t* jump )
{
    const uint32_t i = blockIdx.x * THREADS + threadIdx.x;
    if( i >= count )
    {
        return;
    }
    const uint32_t p    = parents[ i ];
    const bool     root = i == 0 || p == 0xFFFFFFFFu;
    dist[ i ]           = root ? 0u : 1u;
    jump[ i ]           = root ? 0u : p;
}

__global__ __launch_bounds__( THREADS ) void k_depth_jump( uint32_t count, const uint32_t* distIn, const uint32_t* jumpIn, uint32_t* distOut,
                                                           uint32_t* jumpOut )
{
    const uint32_t i = blockIdx.x * THREADS + threadIdx.x;
    if( i >= count )
    {
        return;
    }
    const uint32_t j = jumpIn[ i ];
    distOut[ i ]     = distIn[ i ] + distIn[ j ];
    jumpOut[ i ]     = jumpIn[ j ];
}

// Histogram of internal-node depths, privatised per block in LDS (one
// global atomic per bin per block); counts[REFIT_MAX_DEPTH] flags a deeper
// tree.  Also writes the (depth, node) pairs one radix pass groups by depth.
__global__ __launch_bounds__( THREADS ) void k_depth_keys( uint32_t internal, const uint32_t* depth, uint32_t* counts, uint64_t* keys,
                                                           uint32_t* vals )
{
    __shared__ uint32_t s[ REFIT_MAX_DEPTH + 1 ];
    for( uint32_t b = threadIdx.x; b <= REFIT_MAX_DEPTH; b += THREADS )
    {
        s[ b ] = 0;
    }
    __syncthreads();
    const uint32_t i = blockIdx.x * THREADS + threadIdx.x;
    if( i < internal )
    {
        const uint32_t d = depth[ i ] < REFIT_MAX_DEPTH ? depth[ i ] : REFIT_MAX_DEPTH;
        atomicAdd( &s[ d ], 1u );
        keys[ i ] = d;
        vals[ i ] = i;
    }
    __syncthreads();
    for( uint32_t b = threadIdx.x; b <= REFIT_MAX_DEPTH; b += THREADS )
    {
        if( s[ b ] )
        {
            atomicAdd( &counts[ b ], s[ b ] );
        }
    }
}

__global__ __launch_bounds__( THREADS ) void k_refit_level( const uint32_t* order, uint32_t first, uint32_t count, const uint32_t* children,
                                                            float* bounds )
{
    const uint32_t t = blockIdx.x * THREADS + threadIdx.x;
    if( t >= count )
    {
        return;
    }
    const uint32_t node = order[ first + t ];
    const float*   l    = bounds + size_t( children[ node * 2 + 0 ] ) * 6;
    const float*   r    = bounds + size_t( children[ node * 2 + 1 ] ) * 6;
    const v3       lo   = vmin( mk3( l[ 0 ], l[ 1 ], l[ 2 ] ), mk3( r[ 0 ], r[ 1 ], r[ 2 ] ) );
    const v3       hi   = vmax( mk3( l[ 3 ], l[ 4 ], l[ 5 ] ), mk3( r[ 3 ], r[ 4 ], r[ 5 ] ) );
    StoreBox( bounds + size_t( node ) * 6, lo, hi );
}

// ------------------------------------------------------------ collapse

__device__ __forceinline__ float HalfArea( const float* b )
{
    const float dx = b[ 3 ] - b[ 0 ], dy = b[ 4 ] - b[ 1 ], dz = b[ 5 ] - b[ 2 ];
    return dx * dy + dy * dz + dz * dx;
}

__device__ __forceinline__ uint32_t NodeCount( uint32_t id, uint32_t n, const uint32_t* rf, const uint32_t* rl )
{
    return id >= n - 1 ? 1u : rl[ id ] - rf[ id ] + 1u;
}

constexpr uint32_t OPEN_BIT = 0x80000000u;

// Per internal binary node: its two children, each with OPEN_BIT set when it
// may be opened (internal, more than leafMax triangles), and their half
// areas (-1 for the others).  The collapse then reads one 16-byte record per
// opening instead of a chain of child, range and box loads.
__global__ __launch_bounds__( THREADS ) void k_expand_info( uint32_t internal, uint32_t n, uint32_t leafMax, const uint32_t* children,
                                                            const uint32_t* rf, const uint32_t* rl, const float* bounds, uint4* info )
{
    const uint32_t b = blockIdx.x * THREADS + threadIdx.x;
    if( b >= internal )
    {
        return;
    }
    const uint32_t l  = children[ b * 2 + 0 ], r = children[ b * 2 + 1 ];
    const bool     el = l < n - 1 && NodeCount( l, n, rf, rl ) > leafMax;
    const bool     er = r < n - 1 && NodeCount( r, n, rf, rl ) > leafMax;
    const float    al = el ? HalfArea( bounds + size_t( l ) * 6 ) : -1.0f;
    const float    ar = er ? HalfArea( bounds + size_t( r ) * 6 ) : -1.0f;
    info[ b ]         = make_uint4( l | ( el ? OPEN_BIT : 0u ), r | ( er ? OPEN_BIT : 0u ), __float_as_uint( al ), __float_as_uint( ar ) );
}

// Expands binary node b into up to 8 children by repeatedly opening the
// child with the largest surface area (first one on ties) among those that
// are internal and hold more than leafMax triangles.
__global__ __launch_bounds__( THREADS ) void k_collapse_level( const uint32_t* frontierIn,
                                                               const uint32_t* countIn,
                                                               uint32_t*       frontierOut,
                                                               uint32_t*       countOut,
                                                               uint32_t        n,
                                                               uint32_t        leafMax,
                                                               const uint4*    info,
                                                               const uint32_t* rf,
                                                               const uint32_t* rl,
                                                               uint32_t*       expansion,
                                                               uint32_t*       isNode8 )
{
    const uint32_t count = *countIn;
    for( uint32_t t = blockIdx.x * THREADS + threadIdx.x; t < count; t += gridDim.x * THREADS )
    {
        const uint32_t b = frontierIn[ t ];
        // The node's children so far, each with whether it may be opened
        // (internal, more than leafMax triangles) and its half area.  Every
        // index into these arrays is a compile-time constant (the loops are
        // unrolled and the updates predicated), so they stay in registers.
        uint32_t ch[ 8 ];
        float    ar[ 8 ];
        bool     ex[ 8 ];
#pragma unroll
        for( int k = 0; k < 8; k++ )
        {
            ch[ k ] = 0xFFFFFFFFu;
            ar[ k ] = -1.0f;
            ex[ k ] = false;
        }
        int nch = 0;
        if( b >= n - 1 || NodeCount( b, n, rf, rl ) <= leafMax )
        {
            ch[ 0 ] = b; // a root that is small enough is a single leaf
            nch     = 1;
        }
        else
        {
            const uint4 q = info[ b ];
            ch[ 0 ]       = q.x & ~OPEN_BIT;
            ch[ 1 ]       = q.y & ~OPEN_BIT;
            ex[ 0 ]       = ( q.x & OPEN_BIT ) != 0u;
            ex[ 1 ]       = ( q.y & OPEN_BIT ) != 0u;
            ar[ 0 ]       = __uint_as_float( q.z );
            ar[ 1 ]       = __uint_as_float( q.w );
            nch           = 2;
            while( nch < 8 )
            {
                // the first child of largest area among those that open
                int   best     = -1;
                float bestArea = -1.0f;
#pragma unroll
                for( int k = 0; k < 8; k++ )
                {
                    if( k < nch && ex[ k ] && ar[ k ] > bestArea )
                    {
                        bestArea = ar[ k ];
                        best     = k;
                    }
                }
                if( best < 0 )
                {
                    break;
                }
                uint32_t c = 0;
#pragma unroll
                for( int k = 0; k < 8; k++ )
                {
                    c = k == best ? ch[ k ] : c;
                }
                const uint4    q   = info[ c ];
                const uint32_t l   = q.x & ~OPEN_BIT, r = q.y & ~OPEN_BIT;
                const bool     el  = ( q.x & OPEN_BIT ) != 0u, er = ( q.y & OPEN_BIT ) != 0u;
                const float    al  = __uint_as_float( q.z ), arr = __uint_as_float( q.w );
                // ch[best + 2 ..] = old ch[best + 1 .. nch - 1]; then the two
                // children take slots best and best + 1
#pragma unroll
                for( int k = 7; k >= 1; k-- )
                {
                    if( k <= nch && k > best + 1 )
                    {
                        ch[ k ] = ch[ k - 1 ];
                        ar[ k ] = ar[ k - 1 ];
                        ex[ k ] = ex[ k - 1 ];
                    }
                }
#pragma unroll
                for( int k = 0; k < 8; k++ )
                {
                    if( k == best )
                    {
                        ch[ k ] = l;
                        ar[ k ] = al;
                        ex[ k ] = el;
                    }
                    else if( k == best + 1 )
                    {
                        ch[ k ] = r;
                        ar[ k ] = arr;
                        ex[ k ] = er;
                    }
                }
                nch++;
            }
        }
        const uint32_t slot = b < n - 1 ? b : 0;
#pragma unroll
        for( int k = 0; k < 8; k++ )
        {
            expansion[ size_t( slot ) * 8 + k ] = k < nch ? ch[ k ] : 0xFFFFFFFFu;
        }
        isNode8[ slot ] = 1u;
        if( nch > 1 )
        {
#pragma unroll
            for( int k = 0; k < 8; k++ )
            {
                if( k < nch && ex[ k ] )
                {
                    frontierOut[ atomicAdd( countOut, 1u ) ] = ch[ k ];
                }
            }
        }
    }
}

__device__ __forceinline__ uint32_t LeafRef( uint32_t first, uint32_t count )
{
    return RG_EXT_BVH_CHILD_LEAF_BIT | ( first << 3 ) | ( count - 1u );
}

__global__ __launch_bounds__( THREADS ) void k_write_nodes( uint32_t        slots,
                                                            uint32_t        n,
                                                            uint32_t        leafMax,
                                                            const uint32_t* isNode8,
                                                            const uint32_t* node8Index,
                                                            const uint32_t* expansion,
                                                            const uint32_t* rf,
                                                            const uint32_t* rl,
                                                            const float*    bounds,
                                                            BvhNode8*       nodes )
{
    // one thread per (slot, child): the 8 threads of a node write its SoA
    // arrays as contiguous 32-byte runs
    const uint64_t t = uint64_t( blockIdx.x ) * THREADS + threadIdx.x;
    const uint32_t b = uint32_t( t >> 3 );
    const int      k = int( t & 7u );
    if( b >= slots || !isNode8[ b ] )
    {
        return;
    }
    BvhNode8& node = nodes[ node8Index[ b ] ];
    {
        const uint32_t c = expansion[ size_t( b ) * 8 + k ];
        if( c == 0xFFFFFFFFu )
        {
            node.lo_x[ k ] = node.lo_y[ k ] = node.lo_z[ k ] = INFINITY;
            node.hi_x[ k ] = node.hi_y[ k ] = node.hi_z[ k ] = -INFINITY;
            node.child[ k ] = RG_EXT_BVH_CHILD_EMPTY;
            node.pad[ k ]   = 0;
            return;
        }
        const float* bb = bounds + size_t( c ) * 6;
        node.lo_x[ k ]  = bb[ 0 ];
        node.lo_y[ k ]  = bb[ 1 ];
        node.lo_z[ k ]  = bb[ 2 ];
        node.hi_x[ k ]  = bb[ 3 ];
        node.hi_y[ k ]  = bb[ 4 ];
        node.hi_z[ k ]  = bb[ 5 ];
        node.pad[ k ]   = 0;
        if( c >= n - 1 )
        {
            node.child[ k ] = LeafRef( c - ( n - 1 ), 1 );
        }
        else if( NodeCount( c, n, rf, rl ) <= leafMax || isNode8[ c ] == 0u || c == b )
        {
            node.child[ k ] = LeafRef( rf[ c ], rl[ c ] - rf[ c ] + 1u );
        }
        else
        {
            node.child[ k ] = node8Index[ c ];
        }
    }
}

// Outward rounding to half: the half nearest to v - origin, stepped down
// (lo) or up (hi) until origin + float(h), in FP32, is on the outside of v.
__device__ __forceinline__ uint16_t HalfBitsOf( _Float16 h )
{
    return __builtin_bit_cast( uint16_t, h );
}
__device__ __forceinline__ _Float16 HalfOfBits( uint16_t b )
{
    return __builtin_bit_cast( _Float16, b );
}
__device__ _Float16 HalfStep( _Float16 h, bool up )
{
    const uint16_t b = HalfBitsOf( h );
    if( ( b & 0x7C00u ) == 0x7C00u )
    {
        return h; // inf (or NaN) stays
    }
    if( ( b & 0x7FFFu ) == 0u )
    {
        return HalfOfBits( up ? 0x0001u : 0x8001u ); // +-0 -> smallest subnormal
    }
    const bool neg = ( b & 0x8000u ) != 0u;
    return HalfOfBits( uint16_t( ( neg != up ) ? b + 1u : b - 1u ) );
}
__device__ uint16_t EncodeOffset( float v, float origin, bool up )
{
    _Float16 h = _Float16( v - origin );
    if( !up && HalfBitsOf( h ) == 0x7C00u )
    {
        h = HalfOfBits( 0x7BFFu ); // beyond the half range: the largest finite offset is still below v
    }
    for( int i = 0; i < 64; i++ )
    {
        const float d = origin + float( h );
        if( up ? !( d < v ) : !( d > v ) )
        {
            break;
        }
        h = HalfStep( h, up );
    }
    return HalfBitsOf( h );
}

__global__ __launch_bounds__( THREADS ) void k_half_nodes( const BvhNode8* in, uint32_t count, BvhNode8H* out )
{
    const uint32_t i = blockIdx.x * THREADS + threadIdx.x;
    if( i >= count )
    {
        return;
    }
    const BvhNode8& a = in[ i ];
    BvhNode8H       h;
    float           o[ 3 ] = { INFINITY, INFINITY, INFINITY };
    for( int k = 0; k < 8; k++ )
    {
        if( a.child[ k ] != RG_EXT_BVH_CHILD_EMPTY )
        {
            o[ 0 ] = fminf( o[ 0 ], a.lo_x[ k ] );
            o[ 1 ] = fminf( o[ 1 ], a.lo_y[ k ] );
            o[ 2 ] = fminf( o[ 2 ], a.lo_z[ k ] );
        }
    }
    for( int c = 0; c < 3; c++ )
    {
        h.origin[ c ] = o[ c ] < INFINITY ? o[ c ] : 0.0f;
    }
    h.pad = 0;
    for( int k = 0; k < 8; k++ )
    {
        h.child[ k ] = a.child[ k ];
        if( a.child[ k ] == RG_EXT_BVH_CHILD_EMPTY )
        {
            h.lo_x[ k ] = h.lo_y[ k ] = h.lo_z[ k ] = 0x7C00u; // +inf
            h.hi_x[ k ] = h.hi_y[ k ] = h.hi_z[ k ] = 0xFC00u; // -inf
            continue;
        }
        h.lo_x[ k ] = EncodeOffset( a.lo_x[ k ], h.origin[ 0 ], false );
        h.lo_y[ k ] = EncodeOffset( a.lo_y[ k ], h.origin[ 1 ], false );
        h.lo_z[ k ] = EncodeOffset( a.lo_z[ k ], h.origin[ 2 ], false );
        h.hi_x[ k ] = EncodeOffset( a.hi_x[ k ], h.origin[ 0 ], true );
        h.hi_y[ k ] = EncodeOffset( a.hi_y[ k ], h.origin[ 1 ], true );
        h.hi_z[ k ] = EncodeOffset( a.hi_z[ k ], h.origin[ 2 ], true );
    }
    out[ i ] = h;
}

__global__ __launch_bounds__( THREADS ) void k_sorted_tris( const float* tris, const uint32_t* sortedIdx, uint32_t n, float4* out )
{
    const uint32_t k = blockIdx.x * THREADS + threadIdx.x;
    if( k >= n )
    {
        return;
    }
    const uint32_t id = sortedIdx[ k ];
    v3             a, b, c;
    LoadTri( tris, id, a, b, c );
    out[ size_t( k ) * 3 + 0 ] = make_float4( a.x, a.y, a.z, __uint_as_float( id ) );
    out[ size_t( k ) * 3 + 1 ] = make_float4( b.x, b.y, b.z, 0.0f );
    out[ size_t( k ) * 3 + 2 ] = make_float4( c.x, c.y, c.z, 0.0f );
}

__global__ void k_seed_frontier( uint32_t* frontier, uint32_t* counts )
{
    frontier[ 0 ] = 0;
    counts[ 0 ]   = 1;
}

} // namespace

namespace
{
// Shared by the single-scene and the batched build; `batch` (host offsets,
// S + 1 of them) switches steps 1-2 to per-scene bounds and keys and seeds
// the collapse with every scene's root.
void BuildBvhImpl( BvhObject& bvh, const float* d_tris, uint32_t n, const RgExtBvhBuildParams& params, const uint32_t* batchOffsets,
                   uint32_t sceneCount, hipStream_t stream, bool segmentBoxes = false )
{
    RG_REQUIRE( n > 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "triangle count must be > 0" );
    RG_REQUIRE( n < ( 1u << 28 ), RG_RESULT_WRONG_FUNCTION_ARGUMENT, "triangle count must be < 2^28" );
    RG_REQUIRE( params.maxLeafSize >= 1 && params.maxLeafSize <= 8, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "maxLeafSize must be in [1, 8]" );
    RG_REQUIRE( params.nodeFormat == RG_EXT_BVH_NODE_FP32 || params.nodeFormat == RG_EXT_BVH_NODE_FP16, RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                "unknown node format" );

    bvh.n          = n;
    bvh.leafMax    = params.maxLeafSize;
    bvh.nodeFormat = params.nodeFormat;

    const uint32_t internal   = n - 1;
    const uint32_t slots      = internal > 0 ? internal : 1;
    const uint32_t rsBlocks   = DivUp( n, RS_TILE );
    const uint32_t bndBlocks  = DivUp( n, THREADS ) < 1024 ? DivUp( n, THREADS ) : 1024;
    const uint32_t histCount  = 256 * rsBlocks;
    const uint32_t scanElems  = histCount > slots ? histCount : slots;

    bvh.sortedTris.Reserve( size_t( n ) * 48 );
    bvh.codes.Reserve( size_t( n ) * 8 );
    bvh.keysA.Reserve( size_t( n ) * 8 );
    bvh.keysB.Reserve( size_t( n ) * 8 );
    bvh.valsA.Reserve( size_t( n ) * 4 );
    bvh.valsB.Reserve( size_t( n ) * 4 );
    bvh.hist.Reserve( size_t( histCount ) * 4 );
    bvh.scanScratch.Reserve( ScanScratchBytes( scanElems ) );
    bvh.boundsPartials.Reserve( size_t( bndBlocks + 1 ) * 6 * 8 );
    bvh.children.Reserve( size_t( slots ) * 8 );
    bvh.parents.Reserve( size_t( 2 * n - 1 ) * 4 );
    bvh.rangeFirst.Reserve( size_t( slots ) * 4 );
    bvh.rangeLast.Reserve( size_t( slots ) * 4 );
    bvh.nodeBounds.Reserve( size_t( 2 * n - 1 ) * 24 );
    bvh.depthA.Reserve( size_t( 2 * n - 1 ) * 4 );
    bvh.depthB.Reserve( size_t( 2 * n - 1 ) * 4 );
    bvh.jumpA.Reserve( size_t( 2 * n - 1 ) * 4 );
    bvh.jumpB.Reserve( size_t( 2 * n - 1 ) * 4 );
    bvh.levelHist.Reserve( size_t( REFIT_MAX_DEPTH + 1 ) * 4 );
    bvh.levelOrder.Reserve( size_t( slots ) * 4 );
    if( !bvh.hostLevels )
    {
        RG_HIP_CHECK( hipHostMalloc( reinterpret_cast< void** >( &bvh.hostLevels ), ( REFIT_MAX_DEPTH + 1 ) * 4, hipHostMallocDefault ) );
    }
    bvh.expansion.Reserve( size_t( slots ) * 32 );
    bvh.isNode8.Reserve( size_t( slots ) * 4 );
    bvh.node8Index.Reserve( size_t( slots ) * 4 );
    bvh.frontier[ 0 ].Reserve( size_t( slots ) * 4 );
    bvh.frontier[ 1 ].Reserve( size_t( slots ) * 4 );
    bvh.levelCounts.Reserve( ( BVH_MAX_LEVELS + 2 ) * 4 );
    if( !bvh.hostCounts )
    {
        RG_HIP_CHECK( hipHostMalloc( reinterpret_cast< void** >( &bvh.hostCounts ), ( BVH_MAX_LEVELS + 2 ) * 4, hipHostMallocDefault ) );
    }

    double* partials    = bvh.boundsPartials.as< double >();
    double* sceneBounds = partials + size_t( bndBlocks ) * 6;
    const bool batched  = batchOffsets != nullptr;
    bvh.sceneCount      = batched ? sceneCount : 0u;

    if( !batched )
    {
        // 1. FP64 centroid bounds.
        k_centroid_bounds<<< bndBlocks, THREADS, 0, stream >>>( d_tris, n, partials );
        CheckLaunch( "k_centroid_bounds" );
        k_bounds_final<<< 1, THREADS, 0, stream >>>( partials, bndBlocks, sceneBounds );
        CheckLaunch( "k_bounds_final" );

        // 2. Morton codes (input order) and the identity permutation.
        k_morton<<< DivUp( n, THREADS ), THREADS, 0, stream >>>( d_tris, n, sceneBounds, bvh.codes.as< uint64_t >(), bvh.keysA.as< uint64_t >(), bvh.valsA.as< uint32_t >() );
        CheckLaunch( "k_morton" );
    }
    else
    {
        // 1-2 per scene: bounds of each scene's centroids, then keys with the
        // scene in the top bits
        const uint32_t B = 16;
        bvh.sceneOffsets.Reserve( size_t( sceneCount + 1 ) * 4 );
        bvh.sceneRoots.Reserve( size_t( sceneCount ) * 4 );
        bvh.sceneRoots8.Reserve( size_t( sceneCount ) * 4 );
        bvh.sceneBounds.Reserve( size_t( sceneCount ) * 6 * 8 );
        bvh.scenePartials.Reserve( size_t( sceneCount ) * B * 6 * 8 );
        bvh.hostSceneOffsets.assign( batchOffsets, batchOffsets + sceneCount + 1 );
        RG_HIP_CHECK( hipMemcpyAsync( bvh.sceneOffsets.ptr, bvh.hostSceneOffsets.data(), size_t( sceneCount + 1 ) * 4, hipMemcpyHostToDevice, stream ) );
        const uint32_t* offs = bvh.sceneOffsets.as< uint32_t >();
        k_centroid_bounds_seg<<< sceneCount * B, THREADS, 0, stream >>>( d_tris, offs, B, bvh.scenePartials.as< double >() );
        CheckLaunch( "k_centroid_bounds_seg" );
        k_bounds_final_seg<<< sceneCount, 64, 0, stream >>>( bvh.scenePartials.as< double >(), B, bvh.sceneBounds.as< double >() );
        CheckLaunch( "k_bounds_final_seg" );
        k_morton_seg<<< DivUp( n, THREADS ), THREADS, 0, stream >>>( d_tris, n, offs, sceneCount, bvh.sceneBounds.as< double >(),
                                                                     bvh.codes.as< uint64_t >(), bvh.keysA.as< uint64_t >(), bvh.valsA.as< uint32_t >() );
        CheckLaunch( "k_morton_seg" );
        // the whole-batch bounds are not used; keep the view's fields defined
        k_centroid_bounds<<< bndBlocks, THREADS, 0, stream >>>( d_tris, n, partials );
        CheckLaunch( "k_centroid_bounds" );
        k_bounds_final<<< 1, THREADS, 0, stream >>>( partials, bndBlocks, sceneBounds );
        CheckLaunch( "k_bounds_final" );
    }

    // 3. Stable LSD radix sort, 8 passes of 8 bits.
    uint64_t* kin  = bvh.keysA.as< uint64_t >();
    uint64_t* kout = bvh.keysB.as< uint64_t >();
    uint32_t* vin  = bvh.valsA.as< uint32_t >();
    uint32_t* vout = bvh.valsB.as< uint32_t >();
    for( int pass = 0; pass < 8; pass++ )
    {
        const int shift = pass * 8;
        k_radix_hist<<< rsBlocks, THREADS, 0, stream >>>( kin, n, shift, bvh.hist.as< uint32_t >(), rsBlocks );
        CheckLaunch( "k_radix_hist" );
        ExclusiveScan( bvh.hist.as< uint32_t >(), histCount, bvh.hist.as< uint32_t >(), nullptr, bvh.scanScratch.as< uint32_t >(), stream );
        k_radix_scatter<<< rsBlocks, THREADS, 0, stream >>>( kin, vin, kout, vout, n, shift, bvh.hist.as< uint32_t >(), rsBlocks );
        CheckLaunch( "k_radix_scatter" );
        std::swap( kin, kout );
        std::swap( vin, vout );
    }
    bvh.sortedKeys = kin;
    bvh.sortedIdx  = vin;

    // 4. Karras hierarchy.
    uint32_t* children = bvh.children.as< uint32_t >();
    uint32_t* parents  = bvh.parents.as< uint32_t >();
    uint32_t* rf       = bvh.rangeFirst.as< uint32_t >();
    uint32_t* rl       = bvh.rangeLast.as< uint32_t >();
    if( internal > 0 )
    {
        k_karras<<< DivUp( internal, THREADS ), THREADS, 0, stream >>>( kin, n, children, parents, rf, rl );
        CheckLaunch( "k_karras" );
    }

    // 5. Refit.
    float* bounds = bvh.nodeBounds.as< float >();
    k_leaf_boxes<<< DivUp( n, THREADS ), THREADS, 0, stream >>>( d_tris, vin, n, bounds );
    CheckLaunch( "k_leaf_boxes" );
    if( internal > 0 )
    {
        const uint32_t all  = 2 * n - 1;
        uint32_t*      dA   = bvh.depthA.as< uint32_t >();
        uint32_t*      dB   = bvh.depthB.as< uint32_t >();
        uint32_t*      jA   = bvh.jumpA.as< uint32_t >();
        uint32_t*      jB   = bvh.jumpB.as< uint32_t >();
        k_depth_init<<< DivUp( all, THREADS ), THREADS, 0, stream >>>( parents, all, dA, jA );
        CheckLaunch( "k_depth_init" );
        for( int r = 0; r < REFIT_JUMP_ROUNDS; r++ )
        {
            k_depth_jump<<< DivUp( all, THREADS ), THREADS, 0, stream >>>( all, dA, jA, dB, jB );
            CheckLaunch( "k_depth_jump" );
            std::swap( dA, dB );
            std::swap( jA, jB );
        }
        uint32_t* hist = bvh.levelHist.as< uint32_t >();
        RG_HIP_CHECK( hipMemsetAsync( hist, 0, ( REFIT_MAX_DEPTH + 1 ) * 4, stream ) );
        // (depth, node) pairs into the free key buffer and dB, then one
        // stable 8-bit radix pass groups the nodes by depth into levelOrder
        k_depth_keys<<< DivUp( internal, THREADS ), THREADS, 0, stream >>>( internal, dA, hist, kout, dB );
        CheckLaunch( "k_depth_keys" );
        RG_HIP_CHECK( hipMemcpyAsync( bvh.hostLevels, hist, ( REFIT_MAX_DEPTH + 1 ) * 4, hipMemcpyDeviceToHost, stream ) );
        const uint32_t lvBlocks = DivUp( internal, RS_TILE );
        k_radix_hist<<< lvBlocks, THREADS, 0, stream >>>( kout, internal, 0, bvh.hist.as< uint32_t >(), lvBlocks );
        CheckLaunch( "k_radix_hist" );
        ExclusiveScan( bvh.hist.as< uint32_t >(), 256 * lvBlocks, bvh.hist.as< uint32_t >(), nullptr, bvh.scanScratch.as< uint32_t >(), stream );
        k_radix_scatter<<< lvBlocks, THREADS, 0, stream >>>( kout, dB, kin, bvh.levelOrder.as< uint32_t >(), internal, 0, bvh.hist.as< uint32_t >(),
                                                             lvBlocks );
        CheckLaunch( "k_radix_scatter" );
        bvh.sortedKeys = nullptr; // the sorted Morton keys were overwritten
        RG_HIP_CHECK( hipStreamSynchronize( stream ) );
        RG_REQUIRE( bvh.hostLevels[ REFIT_MAX_DEPTH ] == 0, RG_RESULT_INTERNAL_ERROR, "BVH deeper than the refit level limit" );
        uint32_t first[ REFIT_MAX_DEPTH ], sum = 0;
        int      deepest = 0;
        for( uint32_t d = 0; d < REFIT_MAX_DEPTH; d++ )
        {
            first[ d ] = sum;
            sum += bvh.hostLevels[ d ];
            deepest = bvh.hostLevels[ d ] ? int( d ) : deepest;
        }
        for( int d = deepest; d >= 0; d-- )
        {
            const uint32_t c = bvh.hostLevels[ d ];
            if( c )
            {
                k_refit_level<<< DivUp( c, THREADS ), THREADS, 0, stream >>>( bvh.levelOrder.as< uint32_t >(), first[ d ], c, children, bounds );
                CheckLaunch( "k_refit_level" );
            }
        }
    }

    // 6. Collapse to 8-wide nodes, level by level.
    bvh.expandInfo.Reserve( size_t( internal > 0 ? internal : 1 ) * sizeof( uint4 ) );
    if( internal > 0 )
    {
        k_expand_info<<< DivUp( internal, THREADS ), THREADS, 0, stream >>>( internal, n, bvh.leafMax, children, rf, rl, bounds,
                                                                            bvh.expandInfo.as< uint4 >() );
        CheckLaunch( "k_expand_info" );
    }
    uint32_t* counts = bvh.levelCounts.as< uint32_t >();
    RG_HIP_CHECK( hipMemsetAsync( counts, 0, ( BVH_MAX_LEVELS + 2 ) * 4, stream ) );
    RG_HIP_CHECK( hipMemsetAsync( bvh.isNode8.ptr, 0, size_t( slots ) * 4, stream ) );
    if( !batched )
    {
        k_seed_frontier<<< 1, 1, 0, stream >>>( bvh.frontier[ 0 ].as< uint32_t >(), counts );
        CheckLaunch( "k_seed_frontier" );
    }
    else
    {
        RG_HIP_CHECK( hipMemsetAsync( bvh.sceneRoots.ptr, 0xFF, size_t( sceneCount ) * 4, stream ) );
        k_scene_roots<<< DivUp( internal, THREADS ), THREADS, 0, stream >>>( internal, rf, rl, bvh.sceneOffsets.as< uint32_t >(), sceneCount,
                                                                           bvh.sceneRoots.as< uint32_t >() );
        CheckLaunch( "k_scene_roots" );
        k_seed_frontier_seg<<< 1, 256, 0, stream >>>( bvh.sceneRoots.as< uint32_t >(), sceneCount, bvh.frontier[ 0 ].as< uint32_t >(), counts );
        CheckLaunch( "k_seed_frontier_seg" );
    }

    const uint32_t maxFrontier   = n / ( bvh.leafMax + 1 ) + 1;
    const uint32_t collapseGrid  = DivUp( maxFrontier, THREADS ) < 4096 ? DivUp( maxFrontier, THREADS ) : 4096;
    uint32_t       level         = 0;
    bool           done          = false;
    while( !done )
    {
        constexpr uint32_t BATCH = 8;
        for( uint32_t k = 0; k < BATCH; k++, level++ )
        {
            RG_REQUIRE( level + 1 < BVH_MAX_LEVELS, RG_RESULT_INTERNAL_ERROR, "BVH deeper than the level limit" );
            k_collapse_level<<< collapseGrid, THREADS, 0, stream >>>( bvh.frontier[ level & 1 ].as< uint32_t >(),
                                                                      counts + level,
                                                                      bvh.frontier[ ( level + 1 ) & 1 ].as< uint32_t >(),
                                                                      counts + level + 1,
                                                                      n,
                                                                      bvh.leafMax,
                                                                      bvh.expandInfo.as< uint4 >(),
                                                                      rf,
                                                                      rl,
                                                                      bvh.expansion.as< uint32_t >(),
                                                                      bvh.isNode8.as< uint32_t >() );
            CheckLaunch( "k_collapse_level" );
        }
        RG_HIP_CHECK( hipMemcpyAsync( bvh.hostCounts, counts + level, 4, hipMemcpyDeviceToHost, stream ) );
        RG_HIP_CHECK( hipStreamSynchronize( stream ) );
        done = bvh.hostCounts[ 0 ] == 0;
    }

    uint32_t* nodeTotal = counts + BVH_MAX_LEVELS;
    ExclusiveScan( bvh.isNode8.as< uint32_t >(), slots, bvh.node8Index.as< uint32_t >(), nodeTotal, bvh.scanScratch.as< uint32_t >(), stream );
    RG_HIP_CHECK( hipMemcpyAsync( bvh.hostCounts, nodeTotal, 4, hipMemcpyDeviceToHost, stream ) );
    RG_HIP_CHECK( hipStreamSynchronize( stream ) );
    bvh.nodeCount = bvh.hostCounts[ 0 ];
    // a batch of one-triangle frame segments has leaves for roots and no node
    RG_REQUIRE( bvh.nodeCount >= 1 || segmentBoxes, RG_RESULT_INTERNAL_ERROR, "BVH collapse produced no nodes" );

    const bool    half    = params.nodeFormat == RG_EXT_BVH_NODE_FP16;
    DeviceBuffer& fp32Out = half ? bvh.nodes32 : bvh.nodes;
    fp32Out.Reserve( size_t( bvh.nodeCount ) * sizeof( BvhNode8 ) );
    k_write_nodes<<< DivUp( slots * 8u, THREADS ), THREADS, 0, stream >>>( slots, n, bvh.leafMax, bvh.isNode8.as< uint32_t >(), bvh.node8Index.as< uint32_t >(), bvh.expansion.as< uint32_t >(), rf, rl, bounds, fp32Out.as< BvhNode8 >() );
    CheckLaunch( "k_write_nodes" );
    k_sorted_tris<<< DivUp( n, THREADS ), THREADS, 0, stream >>>( d_tris, vin, n, bvh.sortedTris.as< float4 >() );
    CheckLaunch( "k_sorted_tris" );
    if( batched )
    {
        k_gather_roots8<<< 1, 256, 0, stream >>>( bvh.sceneRoots.as< uint32_t >(), bvh.sceneOffsets.as< uint32_t >(), sceneCount,
                                                  bvh.node8Index.as< uint32_t >(), bvh.sceneRoots8.as< uint32_t >() );
        CheckLaunch( "k_gather_roots8" );
        bvh.hostSceneRoots.resize( sceneCount );
        RG_HIP_CHECK( hipMemcpyAsync( bvh.hostSceneRoots.data(), bvh.sceneRoots8.ptr, size_t( sceneCount ) * 4, hipMemcpyDeviceToHost, stream ) );
        if( segmentBoxes )
        {
            RG_REQUIRE( sceneCount <= 64, RG_RESULT_INTERNAL_ERROR, "too many frame segments" );
            bvh.segBoxes.Reserve( size_t( sceneCount ) * 8 * sizeof( float ) );
            k_segment_boxes<<< 1, 64, 0, stream >>>( fp32Out.as< BvhNode8 >(), bvh.sortedTris.as< float4 >(), bvh.sceneRoots8.as< uint32_t >(), sceneCount,
                                                     bvh.segBoxes.as< float >() );
            CheckLaunch( "k_segment_boxes" );
            bvh.hostSegBoxes.resize( size_t( sceneCount ) * 8 );
            RG_HIP_CHECK( hipMemcpyAsync( bvh.hostSegBoxes.data(), bvh.segBoxes.ptr, size_t( sceneCount ) * 8 * sizeof( float ), hipMemcpyDeviceToHost,
                                          stream ) );
        }
    }
    if( half && bvh.nodeCount > 0 )
    {
        bvh.nodes.Reserve( size_t( bvh.nodeCount ) * sizeof( BvhNode8H ) );
        k_half_nodes<<< DivUp( bvh.nodeCount, THREADS ), THREADS, 0, stream >>>( bvh.nodes32.as< BvhNode8 >(), bvh.nodeCount, bvh.nodes.as< BvhNode8H >() );
        CheckLaunch( "k_half_nodes" );
    }

    RG_HIP_CHECK( hipMemcpyAsync( bvh.sceneMin, sceneBounds, 3 * sizeof( double ), hipMemcpyDeviceToHost, stream ) );
    RG_HIP_CHECK( hipMemcpyAsync( bvh.sceneMax, sceneBounds + 3, 3 * sizeof( double ), hipMemcpyDeviceToHost, stream ) );
    RG_HIP_CHECK( hipStreamSynchronize( stream ) );
}
} // namespace

void BuildBvh( BvhObject& bvh, const float* d_tris, uint32_t n, const RgExtBvhBuildParams& params, hipStream_t stream )
{
    BuildBvhImpl( bvh, d_tris, n, params, nullptr, 0, stream );
    bvh.frameSegs.clear();
}

void BuildBvhBatch( BvhObject& bvh, const float* d_tris, const uint32_t* offsets, uint32_t sceneCount, const RgExtBvhBuildParams& params,
                    hipStream_t stream )
{
    RG_REQUIRE( offsets && sceneCount >= 1 && sceneCount <= 128, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "1..128 scenes" );
    RG_REQUIRE( offsets[ 0 ] == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "scene offsets must start at 0" );
    for( uint32_t s = 0; s < sceneCount; s++ )
    {
        RG_REQUIRE( offsets[ s + 1 ] >= offsets[ s ] + 2, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "every scene needs at least 2 triangles" );
    }
    BuildBvhImpl( bvh, d_tris, offsets[ sceneCount ], params, offsets, sceneCount, stream );
    bvh.frameSegs.clear();
    for( uint32_t s = 0; s < sceneCount; s++ )
    {
        RG_REQUIRE( bvh.hostSceneRoots[ s ] < bvh.nodeCount, RG_RESULT_INTERNAL_ERROR, "a scene has no root node" );
    }
}

void BuildFrameBvh( BvhObject& bvh, const float* d_tris, const uint32_t* classTris, uint32_t classCount, const RgExtBvhBuildParams& params,
                    hipStream_t stream )
{
    std::vector< uint32_t > offsets( 1, 0u ), cls;
    for( uint32_t c = 0; c < classCount; c++ )
    {
        if( classTris[ c ] > 0 )
        {
            offsets.push_back( offsets.back() + classTris[ c ] );
            cls.push_back( c );
        }
    }
    const uint32_t S = uint32_t( cls.size() ), n = offsets.back();
    RG_REQUIRE( S > 0 && S <= 8, RG_RESULT_INTERNAL_ERROR, "a frame BVH has 1..8 segments" );
    bvh.frameSegs.assign( S, BvhObject::FrameSegment{} );
    if( S == 1 )
    {
        // one class: the single-scene build (21 Morton bits per axis); rays
        // whose mask admits it start at its root
        BuildBvhImpl( bvh, d_tris, n, params, nullptr, 0, stream );
        bvh.frameSegs[ 0 ].cls  = cls[ 0 ];
        bvh.frameSegs[ 0 ].root = 0;
        for( int k = 0; k < 3; k++ )
        {
            bvh.frameSegs[ 0 ].box[ k ]     = -INFINITY;
            bvh.frameSegs[ 0 ].box[ 3 + k ] = INFINITY;
        }
        return;
    }
    // several classes: one batched build, each class a segment with its own
    // root (a leaf for a one-triangle class) and root box
    BuildBvhImpl( bvh, d_tris, n, params, offsets.data(), S, stream, true );
    for( uint32_t s = 0; s < S; s++ )
    {
        const uint32_t r = bvh.hostSceneRoots[ s ];
        RG_REQUIRE( ( r & RG_EXT_BVH_CHILD_LEAF_BIT ) ? ( ( r >> 3 ) & 0x0FFFFFFFu ) == offsets[ s ] : r < bvh.nodeCount, RG_RESULT_INTERNAL_ERROR,
                    "a frame segment has no root" );
        bvh.frameSegs[ s ].cls  = cls[ s ];
        bvh.frameSegs[ s ].root = r;
        std::memcpy( bvh.frameSegs[ s ].box, &bvh.hostSegBoxes[ size_t( s ) * 8 ], 6 * sizeof( float ) );
    }
    bvh.sceneCount = 0; // traced as a frame (TraceRaysFrame), not as a batch of scenes
}

} // namespace rg
