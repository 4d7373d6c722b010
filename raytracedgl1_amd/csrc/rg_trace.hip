// Software ray traversal of the 8-wide LBVH on gfx950.
//
// Stands in for traceRayEXT (Source/Shaders/RaygenCommon.h:217-269 closest
// hit; :322-359 shadow rays with gl_RayFlagsSkipClosestHitShaderEXT).
//
// Execution model: a persistent grid of 256-thread blocks.  Each lane owns
// one ray at a time; whenever lanes of a wave go idle they take the next rays
// of the wave's current chunk, so long rays do not hold the wave's SIMD lanes
// empty.  Chunks of 64 rays come from 32 counters 256 bytes apart (block b
// takes chunks j * 32 + b % 32), one atomic per chunk: a refill atomic per
// wave iteration on a single address serialises at the memory side and made
// the kernel 5.5x slower at 1080p.  The traversal stack is a short LDS
// stack (24 entries per lane, interleaved by lane to avoid bank conflicts)
// backed by a per-lane overflow area in global memory.
//
// Closest hit is order-independent: a hit replaces the best one if its t is
// smaller, or equal with a smaller primitive id, so the result equals a
// brute-force loop over all triangles with the same intersection routine.
#include "rg_bvh.h"
#include "rg_trace_common.h"
#include "rg_textures.h"

#include <mutex>

namespace rg
{
namespace
{

constexpr int      TR_THREADS    = 256;
constexpr int      STACK_TOTAL   = 128; // entries per lane: LDS part + global overflow part
constexpr int      STACK_OVF_STRIDE = 112; // global entries per lane, enough for the smallest LDS part (16)
constexpr uint32_t MAX_STEPS_PER_RAY = 1u << 22;
constexpr uint32_t RAY_CHUNK_DEFAULT = 64; // rays per counter atomic
// A wave refills its idle lanes once at least this many are idle (or none
// is busy): the refill block (ballots, the chunk atomic, the ray set-up with
// its six divisions) then runs for many lanes at once instead of for the
// one or two lanes that finish in a typical iteration.  Measured on the
// 1080p frame: 1 -> 3.77 ms of traversal, 8 -> 3.64, 16 -> 3.57, 32 -> 3.60.
constexpr uint32_t REFILL_MIN_LANES = 16;
constexpr uint32_t RAY_PARTITIONS    = 32;  // ray counters
constexpr uint32_t COUNTER_STRIDE    = 64;  // u32 between counters (256 B)

struct TraceArgs
{
    const void*     nodes; // BvhNode8 or BvhNode8H
    const float4*   tris;
    const RgExtRay* rays;
    RgExtHit*       hits;
    uint32_t        count;
    uint32_t        flags;
    uint32_t        partitions;
    uint32_t*       rayCounter;
    uint32_t*       overflowStack;
    uint32_t        ovfStride;  // global stack entries per lane
    uint32_t        stackTotal; // LDS + global stack entries per lane
    uint32_t*       errorFlag;
    uint64_t*       stats; // [0] node visits, [1] triangle tests (RG_EXT_TRACE_COLLECT_STATS)
    // rays of several scenes of a batched BVH (segCount > 0): rays of scene s
    // are [segRays[s], segRays[s+1]), traversed from segRoots[s]; hits report
    // primitive ids relative to segTris[s]
    const uint32_t* segRays;
    const uint32_t* segRoots;
    const uint32_t* segTris;
    uint32_t        segCount;
    uint32_t        chunk; // rays per counter atomic (tuned)
    // alpha test (the ALPHA kernels): per-triangle records by global id and
    // the texture table
    const AlphaRec* alpha;
    TexView         tex;
    // frame launches (rootCount > 0): a ray starts at the entry roots whose
    // instance mask shares a bit with its cull mask (rayMasks[ray], or
    // defaultMask); with several, their boxes are slab-tested first, like
    // the children of a node, without counting a node visit
    const uint8_t*  rayMasks;
    uint32_t        defaultMask;
    uint32_t        rootCount;
    uint32_t        rootMask[ 8 ];
    uint32_t        rootRef[ 8 ];
    float           rootBox[ 8 ][ 6 ];
};

// RtAlphaTest.rahit: the layer-0 colour times the texture at the base level
// (an any-hit stage has no derivatives); the hit is ignored when
// (r+g+b)/3 * a + a < 0.5.
__device__ __forceinline__ bool AlphaPasses( const TexView& tv, const AlphaRec& ar, float u, float v )
{
    const float  b0 = 1.0f - u - v;
    const float  tu = ar.uv[ 0 ] * b0 + ar.uv[ 2 ] * u + ar.uv[ 4 ] * v;
    const float  tv_ = ar.uv[ 1 ] * b0 + ar.uv[ 3 ] * u + ar.uv[ 5 ] * v;
    const float4 t  = SampleLod( tv, ar.tex, tu, tv_, 0.0f );
    const float  r  = t.x * float( ar.color & 0xFFu ) * ( 1.0f / 255.0f );
    const float  g  = t.y * float( ( ar.color >> 8 ) & 0xFFu ) * ( 1.0f / 255.0f );
    const float  b  = t.z * float( ( ar.color >> 16 ) & 0xFFu ) * ( 1.0f / 255.0f );
    const float  al = t.w * float( ( ar.color >> 24 ) & 0xFFu ) * ( 1.0f / 255.0f );
    return ( r + g + b ) / 3.0f * al + al >= 0.5f;
}

__device__ __forceinline__ uint32_t SegmentOf( const uint32_t* offsets, uint32_t S, uint32_t i )
{
    uint32_t lo = 0, hi = S; // last segment with offsets[s] <= i
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

// Children are ordered by a 64-bit key: the distance (+0 for -0) mapped to
// an unsigned integer with the same order, then the child reference.  A
// total order, so the visit order (and the node / triangle counts) does not
// depend on the sorting network.
__device__ __forceinline__ uint64_t SortKey( float d, uint32_t ref )
{
    const uint32_t u = __float_as_uint( d + 0.0f );
    const uint32_t s = ( u & 0x80000000u ) ? ~u : ( u | 0x80000000u );
    return ( uint64_t( s ) << 32 ) | ref;
}

__device__ __forceinline__ void CmpSwap( uint64_t& a, uint64_t& b )
{
    const uint64_t lo = a < b ? a : b;
    const uint64_t hi = a < b ? b : a;
    a                 = lo;
    b                 = hi;
}

// Batcher odd-even merge sort network for 8 keys (19 compare-exchanges).
__device__ __forceinline__ void Sort8( uint64_t k[ 8 ] )
{
#define CS( a, b ) CmpSwap( k[ a ], k[ b ] )
    CS( 0, 1 ); CS( 2, 3 ); CS( 4, 5 ); CS( 6, 7 );
    CS( 0, 2 ); CS( 1, 3 ); CS( 4, 6 ); CS( 5, 7 );
    CS( 1, 2 ); CS( 5, 6 );
    CS( 0, 4 ); CS( 1, 5 ); CS( 2, 6 ); CS( 3, 7 );
    CS( 2, 4 ); CS( 3, 5 );
    CS( 1, 2 ); CS( 3, 4 ); CS( 5, 6 );
#undef CS
}

// Child boxes of an FP16 node as FP32, the planes the ray enters through
// (near[axis][child]) and leaves through (far), plus the child references:
// origin + float(offset), the sum its build rounded outward against.  Row
// q[1 + 2 * axis] holds the lo offsets, q[2 + 2 * axis] the hi ones; `sgn`
// bit `axis` is set where the ray's direction is negative on that axis, so
// hi is its entering side.  9 16-byte loads.
__device__ __forceinline__ void LoadHalfNode( const void* nodes, uint32_t node, uint32_t sgn, float nearP[ 3 ][ 8 ], float farP[ 3 ][ 8 ],
                                              uint32_t child[ 8 ] )
{
    typedef _Float16 half8 __attribute__( ( ext_vector_type( 8 ) ) );
    const float4* np = reinterpret_cast< const float4* >( reinterpret_cast< const BvhNode8H* >( nodes ) + node );
    const float4  q0 = np[ 0 ];
    const float   o[ 3 ] = { q0.x, q0.y, q0.z };
#pragma unroll
    for( int ax = 0; ax < 3; ax++ )
    {
        const uint32_t s  = ( sgn >> ax ) & 1u;
        const half8    hn = __builtin_bit_cast( half8, np[ 1 + 2 * ax + s ] );
        const half8    hf = __builtin_bit_cast( half8, np[ 2 + 2 * ax - s ] );
#pragma unroll
        for( int k = 0; k < 8; k++ )
        {
            nearP[ ax ][ k ] = o[ ax ] + float( hn[ k ] );
            farP[ ax ][ k ]  = o[ ax ] + float( hf[ k ] );
        }
    }
    const float4 c0 = np[ 7 ], c1 = np[ 8 ];
    const uint32_t c[ 8 ] = { __float_as_uint( c0.x ), __float_as_uint( c0.y ), __float_as_uint( c0.z ), __float_as_uint( c0.w ),
                              __float_as_uint( c1.x ), __float_as_uint( c1.y ), __float_as_uint( c1.z ), __float_as_uint( c1.w ) };
#pragma unroll
    for( int k = 0; k < 8; k++ )
    {
        child[ k ] = c[ k ];
    }
}

template< bool H, int STACK_LDS, bool ALPHA >
__global__ __launch_bounds__( TR_THREADS ) void k_trace( TraceArgs a )
{
    __shared__ uint32_t s_stack[ STACK_LDS * TR_THREADS ];
    // The frame's entry roots, staged in LDS once per block: read from the
    // kernel arguments inside the loop they would hold 64 scalar registers,
    // which the compiler spills to vector-register lanes and reloads.
    __shared__ uint32_t s_rootMask[ 8 ], s_rootRef[ 8 ];
    __shared__ float    s_rootBox[ 8 ][ 6 ];

    const int      tid       = threadIdx.x;
    if( tid < 8 )
    {
        s_rootMask[ tid ] = a.rootMask[ tid ];
        s_rootRef[ tid ]  = a.rootRef[ tid ];
#pragma unroll
        for( int j = 0; j < 6; j++ )
        {
            s_rootBox[ tid ][ j ] = a.rootBox[ tid ][ j ];
        }
    }
    __syncthreads();
    const int      lane      = tid & 63;
    const uint64_t ltMask    = lane == 0 ? 0ull : ( ~0ull >> ( 64 - lane ) );
    const uint32_t gthread   = blockIdx.x * TR_THREADS + tid;
    uint32_t*      ovf       = a.overflowStack + size_t( gthread ) * a.ovfStride;
    const int      stackTotal = int( a.stackTotal );
    const bool     anyHit    = ( a.flags & RG_EXT_TRACE_ANY_HIT ) != 0;
    const bool     cullFront = ( a.flags & RG_EXT_TRACE_CULL_FRONT_FACE ) != 0;

    bool     active = false, exhausted = false;
    uint32_t rayId  = 0;
    RayPre   ray;
    float    tmin = 0.0f, bestT = 0.0f, bestU = 0.0f, bestV = 0.0f;
    uint32_t bestPrim = RG_EXT_MISS;
    uint32_t node     = 0;
    uint32_t sgn      = 0; // bit a: the ray's direction is negative on axis a
    uint32_t triBase  = 0; // first triangle of the ray's scene (batched BVHs)
    int      sp       = 0;
    uint32_t steps    = 0;
    uint32_t nodeVisits = 0, triTests = 0;
    const uint32_t part = blockIdx.x % a.partitions; // chunks j * partitions + part
    uint32_t chunkNext = 0, chunkEnd = 0; // this wave's rays [chunkNext, chunkEnd), wave-uniform
    bool     drained   = false;           // the global counter has passed the last ray

    auto push = [ & ]( uint32_t v ) __attribute__( ( always_inline ) ) {
        if( sp < STACK_LDS )
        {
            s_stack[ sp * TR_THREADS + tid ] = v;
        }
        else if( sp < stackTotal )
        {
            ovf[ sp - STACK_LDS ] = v;
        }
        else
        {
            atomicOr( a.errorFlag, RG_EXT_TRACE_ERROR_STACK_OVERFLOW ); // dropped entry
            return;
        }
        sp++;
    };

    while( true )
    {
        // Refill idle lanes.  Each wave owns a chunk of a.chunk consecutive
        // rays taken with one atomic on its partition's counter; idle lanes
        // take rays from the chunk (wave-uniform bookkeeping, no atomic).
        const bool     need = !active && !exhausted;
        const uint64_t mask = __ballot( need );
        if( mask != 0ull && ( __popcll( mask ) >= REFILL_MIN_LANES || __ballot( active ) == 0ull ) )
        {
            const uint32_t cnt    = __popcll( mask );
            const uint32_t rank   = __popcll( mask & ltMask );
            const uint32_t left   = chunkEnd - chunkNext;
            const int      leader = __ffsll( ( long long )mask ) - 1;
            uint32_t       newBase = 0;
            const bool     grab    = left < cnt && !drained;
            if( grab )
            {
                uint32_t b = 0;
                if( lane == leader )
                {
                    const uint32_t j = atomicAdd( a.rayCounter + part * COUNTER_STRIDE, 1u );
                    const uint64_t g = uint64_t( j ) * a.partitions + part;
                    b                = g * a.chunk < a.count ? uint32_t( g * a.chunk ) : a.count;
                }
                newBase = __shfl( b, leader, 64 );
            }
            if( need )
            {
                uint32_t my = a.count;
                if( rank < left )
                {
                    my = chunkNext + rank;
                }
                else if( grab && newBase < a.count )
                {
                    my = newBase + ( rank - left );
                }
                if( my < a.count )
                {
                    const RgExtRay r = a.rays[ my ];
                    rayId            = my;
                    ray              = MakeRayPre( mk3( r.origin[ 0 ], r.origin[ 1 ], r.origin[ 2 ] ),
                                                   mk3( r.direction[ 0 ], r.direction[ 1 ], r.direction[ 2 ] ) );
                    sgn              = ( ray.invd.x < 0.0f ? 1u : 0u ) | ( ray.invd.y < 0.0f ? 2u : 0u ) | ( ray.invd.z < 0.0f ? 4u : 0u );
                    tmin             = r.tmin;
                    bestT            = r.tmax;
                    bestU = bestV = 0.0f;
                    bestPrim      = RG_EXT_MISS;
                    node          = 0; // root
                    triBase       = 0;
                    if( a.segCount )
                    {
                        const uint32_t seg = SegmentOf( a.segRays, a.segCount, my );
                        node               = a.segRoots[ seg ];
                        triBase            = a.segTris[ seg ];
                    }
                    sp            = 0;
                    steps         = 0;
                    active        = true;
                    if( a.rootCount )
                    {
                        // frame launch: the entry roots the ray's cull mask admits
                        const uint32_t rm      = a.rayMasks ? uint32_t( a.rayMasks[ my ] ) : a.defaultMask;
                        uint32_t       admit   = 0, only = RG_EXT_BVH_CHILD_EMPTY;
#pragma unroll
                        for( int k = 0; k < 8; k++ )
                        {
                            if( uint32_t( k ) < a.rootCount && ( s_rootMask[ k ] & rm ) != 0u )
                            {
                                admit++;
                                only = s_rootRef[ k ];
                            }
                        }
                        node = only;
                        if( admit > 1u )
                        {
                            // several segments: those whose box the ray
                            // meets, in segment order (closest hit does
                            // not depend on the order)
                            node = RG_EXT_BVH_CHILD_EMPTY;
#pragma unroll
                            for( int k = 7; k >= 0; k-- )
                            {
                                float tn;
                                if( uint32_t( k ) < a.rootCount && ( s_rootMask[ k ] & rm ) != 0u &&
                                    SlabTest( ray, s_rootBox[ k ][ 0 ], s_rootBox[ k ][ 1 ], s_rootBox[ k ][ 2 ], s_rootBox[ k ][ 3 ],
                                              s_rootBox[ k ][ 4 ], s_rootBox[ k ][ 5 ], tmin, bestT, &tn ) )
                                {
                                    if( node != RG_EXT_BVH_CHILD_EMPTY )
                                    {
                                        push( node );
                                    }
                                    node = s_rootRef[ k ];
                                }
                            }
                        }
                        if( node == RG_EXT_BVH_CHILD_EMPTY )
                        {
                            // no segment to visit: a miss
                            RgExtHit h;
                            h.t         = bestT;
                            h.u         = 0.0f;
                            h.v         = 0.0f;
                            h.primitiveId = RG_EXT_MISS;
                            a.hits[ my ] = h;
                            active       = false;
                        }
                    }
                }
                else
                {
                    exhausted = true;
                }
            }
            if( !grab )
            {
                chunkNext += cnt < left ? cnt : left;
            }
            else if( newBase < a.count )
            {
                chunkEnd  = a.count - newBase > a.chunk ? newBase + a.chunk : a.count;
                chunkNext = newBase + ( cnt - left ) < chunkEnd ? newBase + ( cnt - left ) : chunkEnd;
            }
            else
            {
                drained   = true;
                chunkNext = chunkEnd;
            }
        }
        if( __ballot( active ) == 0ull )
        {
            if( __ballot( !exhausted ) != 0ull )
            {
                continue; // rays that ended at the frame roots: take the next ones
            }
            if( a.stats )
            {
                uint64_t nv = nodeVisits, tt = triTests;
                for( int off = 32; off > 0; off >>= 1 )
                {
                    nv += __shfl_xor( nv, off, 64 );
                    tt += __shfl_xor( tt, off, 64 );
                }
                if( lane == 0 )
                {
                    atomicAdd( reinterpret_cast< unsigned long long* >( &a.stats[ 0 ] ), ( unsigned long long )nv );
                    atomicAdd( reinterpret_cast< unsigned long long* >( &a.stats[ 1 ] ), ( unsigned long long )tt );
                }
            }
            break;
        }
        if( !active )
        {
            continue;
        }

        bool finished = false;
        if( ++steps > MAX_STEPS_PER_RAY )
        {
            // Only a malformed tree can get here; end the ray instead of spinning.
            atomicOr( a.errorFlag, RG_EXT_TRACE_ERROR_STEP_LIMIT );
            node = RG_EXT_BVH_CHILD_EMPTY;
            sp   = 0;
        }
        else if( ( node & RG_EXT_BVH_CHILD_LEAF_BIT ) == 0u )
        {
            // Interior node: slab-test the 8 children, visit near to far.
            // An FP32 node is 14 16-byte loads (boxes + child refs), an FP16
            // node 9.
            nodeVisits++;
            uint64_t key[ 8 ];
            int      hitCount = 0;
            if constexpr( !H )
            {
                // rows of 8 floats: lo_x hi_x lo_y hi_y lo_z hi_z child; each
                // axis's entering row (hi where the direction is negative)
                // and leaving row are loaded in that order, so the slab test
                // needs no min / max to order them
                const float4*  np = reinterpret_cast< const float4* >( reinterpret_cast< const BvhNode8* >( a.nodes ) + node );
                const uint32_t sx = sgn & 1u, sy = ( sgn >> 1 ) & 1u, sz = ( sgn >> 2 ) & 1u;
                // two halves of four children (7 loads each) keep fewer
                // registers live than the whole node
#pragma unroll
                for( int half = 0; half < 2; half++ )
                {
                    const float4 qnx = np[ 2 * ( 0 + sx ) + half ], qfx = np[ 2 * ( 1 - sx ) + half ];
                    const float4 qny = np[ 2 * ( 2 + sy ) + half ], qfy = np[ 2 * ( 3 - sy ) + half ];
                    const float4 qnz = np[ 2 * ( 4 + sz ) + half ], qfz = np[ 2 * ( 5 - sz ) + half ];
                    const float4 qc  = np[ 12 + half ];
                    const float  nx[ 4 ] = { qnx.x, qnx.y, qnx.z, qnx.w }, fx[ 4 ] = { qfx.x, qfx.y, qfx.z, qfx.w };
                    const float  ny[ 4 ] = { qny.x, qny.y, qny.z, qny.w }, fy[ 4 ] = { qfy.x, qfy.y, qfy.z, qfy.w };
                    const float  nz[ 4 ] = { qnz.x, qnz.y, qnz.z, qnz.w }, fz[ 4 ] = { qfz.x, qfz.y, qfz.z, qfz.w };
                    const float  cf[ 4 ] = { qc.x, qc.y, qc.z, qc.w };
#pragma unroll
                    for( int j = 0; j < 4; j++ )
                    {
                        const int      k = half * 4 + j;
                        const uint32_t c = __float_as_uint( cf[ j ] );
                        float          tn;
                        const bool     inBox = SlabTestOrdered( ray, nx[ j ], ny[ j ], nz[ j ], fx[ j ], fy[ j ], fz[ j ], tmin, bestT, &tn );
                        const bool     h     = inBox && c != RG_EXT_BVH_CHILD_EMPTY;
                        key[ k ] = h ? SortKey( tn, c ) : ~0ull;
                        hitCount += h ? 1 : 0;
                    }
                }
            }
            else
            {
                float    nearP[ 3 ][ 8 ], farP[ 3 ][ 8 ];
                uint32_t child[ 8 ];
                LoadHalfNode( a.nodes, node, sgn, nearP, farP, child );
#pragma unroll
                for( int k = 0; k < 8; k++ )
                {
                    const uint32_t c = child[ k ];
                    float          tn;
                    const bool     inBox = SlabTestOrdered( ray, nearP[ 0 ][ k ], nearP[ 1 ][ k ], nearP[ 2 ][ k ], farP[ 0 ][ k ], farP[ 1 ][ k ],
                                                            farP[ 2 ][ k ], tmin, bestT, &tn );
                    const bool     h     = inBox && c != RG_EXT_BVH_CHILD_EMPTY;
                    key[ k ] = h ? SortKey( tn, c ) : ~0ull;
                    hitCount += h ? 1 : 0;
                }
            }
            if( hitCount == 0 )
            {
                node = RG_EXT_BVH_CHILD_EMPTY;
            }
            else if( !anyHit )
            {
                // closest hit: near to far, so nearer hits shrink the interval early
                Sort8( key );
                const int np = hitCount - 1; // key[np] .. key[1] are pushed, farthest first
                if( sp + 7 <= STACK_LDS )
                {
                    // room in LDS for any node: predicated stores, no
                    // per-entry overflow branches
#pragma unroll
                    for( int k = 7; k >= 1; k-- )
                    {
                        if( k <= np )
                        {
                            s_stack[ ( sp + np - k ) * TR_THREADS + tid ] = uint32_t( key[ k ] );
                        }
                    }
                    sp += np;
                }
                else
                {
#pragma unroll
                    for( int k = 7; k >= 1; k-- )
                    {
                        if( k < hitCount )
                        {
                            push( uint32_t( key[ k ] ) );
                        }
                    }
                }
                node = uint32_t( key[ 0 ] );
            }
            else
            {
                // any hit: the interval never shrinks before the ray ends, so
                // the order does not change which nodes an unoccluded ray
                // visits; children go in slot order, unsorted
                uint32_t next = RG_EXT_BVH_CHILD_EMPTY;
                if( sp + 7 <= STACK_LDS )
                {
                    int pos = sp;
#pragma unroll
                    for( int k = 7; k >= 0; k-- )
                    {
                        if( key[ k ] != ~0ull )
                        {
                            if( next != RG_EXT_BVH_CHILD_EMPTY )
                            {
                                s_stack[ pos * TR_THREADS + tid ] = next;
                                pos++;
                            }
                            next = uint32_t( key[ k ] );
                        }
                    }
                    sp = pos;
                }
                else
                {
#pragma unroll
                    for( int k = 7; k >= 0; k-- )
                    {
                        if( key[ k ] != ~0ull )
                        {
                            if( next != RG_EXT_BVH_CHILD_EMPTY )
                            {
                                push( next );
                            }
                            next = uint32_t( key[ k ] );
                        }
                    }
                }
                node = next;
            }
        }
        else
        {
            const uint32_t first = ( node >> 3 ) & 0x0FFFFFFFu;
            const uint32_t cnt   = ( node & 7u ) + 1u;
            for( uint32_t k = 0; k < cnt; k++ )
            {
                triTests++;
                const float4 q0 = a.tris[ size_t( first + k ) * 3 + 0 ];
                const float4 q1 = a.tris[ size_t( first + k ) * 3 + 1 ];
                const float4 q2 = a.tris[ size_t( first + k ) * 3 + 2 ];
                float        t, u, v;
                if( IntersectTri( ray, mk3( q0.x, q0.y, q0.z ), mk3( q1.x, q1.y, q1.z ), mk3( q2.x, q2.y, q2.z ), tmin, bestT, cullFront, &t, &u, &v ) )
                {
                    const uint32_t prim = __float_as_uint( q0.w );
                    if constexpr( ALPHA )
                    {
                        const AlphaRec& ar = a.alpha[ prim ];
                        if( TexValid( a.tex, ar.tex ) && !AlphaPasses( a.tex, ar, u, v ) )
                        {
                            continue; // ignoreIntersectionEXT
                        }
                    }
                    if( t < bestT || prim < bestPrim )
                    {
                        bestT    = t;
                        bestU    = u;
                        bestV    = v;
                        bestPrim = prim;
                        if( anyHit )
                        {
                            finished = true;
                            break;
                        }
                    }
                }
            }
            node = RG_EXT_BVH_CHILD_EMPTY;
        }

        if( !finished && node == RG_EXT_BVH_CHILD_EMPTY )
        {
            if( sp == 0 )
            {
                finished = true;
            }
            else
            {
                sp--;
                if( sp < STACK_LDS )
                {
                    node = s_stack[ sp * TR_THREADS + tid ];
                }
                else
                {
                    // Keeps the rare overflow read a separate global load
                    // (no generic-address load on the common LDS path).
                    asm volatile( "" ::: "memory" );
                    node = ovf[ sp - STACK_LDS ];
                }
            }
        }
        if( finished )
        {
            RgExtHit h;
            h.t                = bestT;
            h.u                = bestU;
            h.v                = bestV;
            h.primitiveId      = bestPrim == RG_EXT_MISS ? RG_EXT_MISS : bestPrim - triBase;
            a.hits[ rayId ]    = h;
            active             = false;
        }
    }
}

} // namespace

void ReadTraceStats( const TraceContext& s, uint64_t* nodeVisits, uint64_t* triTests )
{
    uint64_t v[ 2 ] = { 0, 0 };
    if( s.stats.ptr )
    {
        RG_HIP_CHECK( hipDeviceSynchronize() ); // the trace may run on any stream
        RG_HIP_CHECK( hipMemcpy( v, s.stats.ptr, sizeof( v ), hipMemcpyDeviceToHost ) );
    }
    *nodeVisits = v[ 0 ];
    *triTests   = v[ 1 ];
}

uint32_t ReadTraceErrors( const TraceContext& s )
{
    uint32_t e = 0;
    if( s.counter.ptr )
    {
        RG_HIP_CHECK( hipDeviceSynchronize() );
        RG_HIP_CHECK( hipMemcpy( &e, s.counter.as< uint32_t >() + RAY_PARTITIONS * COUNTER_STRIDE, sizeof( e ), hipMemcpyDeviceToHost ) );
    }
    return e;
}

namespace
{
// LDS stack depths the autotuner chooses from (entries per lane; the rest of
// the 128 entries are in global memory).
constexpr int TRACE_STACK_DEPTH[ TRACE_STACK_VARIANTS ] = { 16, 24, 32 };

using TraceKernelFn = void ( * )( TraceArgs );

TraceKernelFn TraceKernel( int half, int v, bool alpha = false )
{
    if( alpha )
    {
        static const TraceKernelFn ka[ 2 ][ TRACE_STACK_VARIANTS ] = {
            { k_trace< false, 16, true >, k_trace< false, 24, true >, k_trace< false, 32, true > },
            { k_trace< true, 16, true >, k_trace< true, 24, true >, k_trace< true, 32, true > },
        };
        return ka[ half ][ v ];
    }
    static const TraceKernelFn k[ 2 ][ TRACE_STACK_VARIANTS ] = {
        { k_trace< false, 16, false >, k_trace< false, 24, false >, k_trace< false, 32, false > },
        { k_trace< true, 16, false >, k_trace< true, 24, false >, k_trace< true, 32, false > },
    };
    return k[ half ][ v ];
}

// Resident blocks of every variant; the overflow stacks are sized for the
// largest grid, so any variant may run on a context.
void InitGrids( TraceContext& s )
{
    if( s.gridBlocks != 0 )
    {
        return;
    }
    int dev = 0, cus = 0;
    RG_HIP_CHECK( hipGetDevice( &dev ) );
    RG_HIP_CHECK( hipDeviceGetAttribute( &cus, hipDeviceAttributeMultiprocessorCount, dev ) );
    uint32_t maxGrid = 0;
    for( int h = 0; h < 2; h++ )
    {
        for( int v = 0; v < TRACE_STACK_VARIANTS; v++ )
        {
            int perCu = 0;
            RG_HIP_CHECK( hipOccupancyMaxActiveBlocksPerMultiprocessor( &perCu, TraceKernel( h, v ), TR_THREADS, 0 ) );
            s.grid[ h ][ v ] = uint32_t( cus * ( perCu > 0 ? perCu : 1 ) );
            maxGrid          = s.grid[ h ][ v ] > maxGrid ? s.grid[ h ][ v ] : maxGrid;
        }
    }
    s.overflow.Reserve( size_t( maxGrid ) * TR_THREADS * STACK_OVF_STRIDE * sizeof( uint32_t ) );
    s.gridBlocks = maxGrid;
}

// Entry roots of a frame launch (rootCount 0: none).
struct FrameRoots
{
    uint32_t       rootCount = 0;
    uint32_t       mask[ 8 ] = {};
    uint32_t       ref[ 8 ]  = {};
    float          box[ 8 ][ 6 ] = {};
    const uint8_t* rayMasks    = nullptr;
    uint32_t       defaultMask = 0;
};

void TraceRaysImpl( const BvhObject& bvh, TraceContext& s, const RgExtRay* d_rays, uint32_t count, RgExtHit* d_hits, uint32_t flags,
                    hipStream_t stream, hipEvent_t evBefore, hipEvent_t evAfter, const TraceAccum* accum, const uint32_t* d_segRays,
                    const FrameRoots* frame = nullptr, const AlphaRec* alpha = nullptr, const TexView* tex = nullptr )
{
    RG_REQUIRE( bvh.nodeCount > 0 || ( frame && frame->rootCount > 0 ), RG_RESULT_WRONG_FUNCTION_CALL, "BVH is not built" );
    RG_REQUIRE( bvh.nodeFormat == RG_EXT_BVH_NODE_FP32 || bvh.nodeFormat == RG_EXT_BVH_NODE_FP16, RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                "unsupported node format" );
    // the chunk counter runs up to count + (waves + 1) * chunk without wrapping
    RG_REQUIRE( count <= 0xF0000000u, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "too many rays in one launch" );
    if( count == 0 )
    {
        return;
    }
    InitGrids( s );
    const int      h      = bvh.nodeFormat == RG_EXT_BVH_NODE_FP16 ? 1 : 0;
    const int      v      = int( s.stackVariant < TRACE_STACK_VARIANTS ? s.stackVariant : 1u );
    const uint32_t grid   = s.grid[ h ][ v ];
    const uint32_t blocks = DivUp( count, TR_THREADS ) < grid ? DivUp( count, TR_THREADS ) : grid;
    s.counter.Reserve( ( RAY_PARTITIONS * COUNTER_STRIDE + 1 ) * sizeof( uint32_t ) );
    s.stats.Reserve( 2 * sizeof( uint64_t ) );
    RG_HIP_CHECK( hipMemsetAsync( s.counter.ptr, 0, ( RAY_PARTITIONS * COUNTER_STRIDE + 1 ) * sizeof( uint32_t ), stream ) );

    TraceArgs a;
    a.nodes         = bvh.nodes.ptr;
    a.tris          = reinterpret_cast< const float4* >( bvh.sortedTris.ptr );
    a.rays          = d_rays;
    a.hits          = d_hits;
    a.count         = count;
    a.flags         = flags;
    a.rayCounter    = s.counter.as< uint32_t >();
    a.errorFlag     = s.counter.as< uint32_t >() + RAY_PARTITIONS * COUNTER_STRIDE;
    a.partitions    = blocks < RAY_PARTITIONS ? blocks : RAY_PARTITIONS;
    a.overflowStack = s.overflow.as< uint32_t >();
    a.ovfStride     = STACK_OVF_STRIDE;
    a.stackTotal    = STACK_TOTAL;
    a.stats         = nullptr;
    a.segRays       = d_segRays;
    a.segRoots      = d_segRays ? bvh.sceneRoots8.as< uint32_t >() : nullptr;
    a.segTris       = d_segRays ? bvh.sceneOffsets.as< uint32_t >() : nullptr;
    a.segCount      = d_segRays ? bvh.sceneCount : 0u;
    a.rayMasks    = nullptr;
    a.defaultMask = 0;
    a.rootCount   = 0;
    for( int k = 0; k < 8; k++ )
    {
        a.rootMask[ k ] = 0;
        a.rootRef[ k ]  = RG_EXT_BVH_CHILD_EMPTY;
        for( int j = 0; j < 6; j++ )
        {
            a.rootBox[ k ][ j ] = 0.0f;
        }
    }
    if( frame )
    {
        a.rayMasks    = frame->rayMasks;
        a.defaultMask = frame->defaultMask;
        a.rootCount   = frame->rootCount;
        for( uint32_t k = 0; k < frame->rootCount && k < 8; k++ )
        {
            a.rootMask[ k ] = frame->mask[ k ];
            a.rootRef[ k ]  = frame->ref[ k ];
            for( int j = 0; j < 6; j++ )
            {
                a.rootBox[ k ][ j ] = frame->box[ k ][ j ];
            }
        }
    }
    a.chunk         = s.rayChunk >= 16 && s.rayChunk <= 1024 ? s.rayChunk : RAY_CHUNK_DEFAULT;
    a.alpha         = alpha && tex ? alpha : nullptr;
    a.tex           = tex ? *tex : TexView{};
    if( accum )
    {
        RG_REQUIRE( accum->errors, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "accumulator without an error word" );
        a.errorFlag = accum->errors;
        a.stats     = ( flags & RG_EXT_TRACE_COLLECT_STATS ) ? accum->stats : nullptr;
    }
    else if( flags & RG_EXT_TRACE_COLLECT_STATS )
    {
        RG_HIP_CHECK( hipMemsetAsync( s.stats.ptr, 0, 2 * sizeof( uint64_t ), stream ) );
        a.stats = s.stats.as< uint64_t >();
    }
    if( evBefore )
    {
        RG_HIP_CHECK( hipEventRecord( evBefore, stream ) );
    }
    hipLaunchKernelGGL( TraceKernel( h, v, a.alpha != nullptr ), dim3( blocks ), dim3( TR_THREADS ), 0, stream, a );
    CheckLaunch( "k_trace" );
    if( evAfter )
    {
        RG_HIP_CHECK( hipEventRecord( evAfter, stream ) );
    }
}
} // namespace

void TraceRays( const BvhObject& bvh, TraceContext& s, const RgExtRay* d_rays, uint32_t count, RgExtHit* d_hits, uint32_t flags, hipStream_t stream,
                hipEvent_t evBefore, hipEvent_t evAfter, const TraceAccum* accum )
{
    RG_REQUIRE( bvh.sceneCount == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "a batched BVH is traced with TraceRaysBatch" );
    TraceRaysImpl( bvh, s, d_rays, count, d_hits, flags, stream, evBefore, evAfter, accum, nullptr );
}

// ------------------------------------------------ allocation-free traversal
//
// TraceRaysNoAlloc keeps its scratch in one file-scope device array per GPU
// instead of a TraceContext: the 32 ray-chunk counters (256 bytes apart),
// the error word, the statistics and the global overflow stacks of a
// persistent grid of at most NOALLOC_MAX_BLOCKS blocks, NOALLOC_OVF entries
// per lane behind the 24 in LDS (a 64-entry stack; deeper entries set
// RG_EXT_TRACE_ERROR_STACK_OVERFLOW).  Every call on a GPU uses that GPU's
// array, so calls that overlap on one GPU race.
namespace
{
constexpr uint32_t NOALLOC_MAX_BLOCKS = 1024;
constexpr uint32_t NOALLOC_OVF        = 40;
constexpr int      NOALLOC_VARIANT    = 1; // the 24-entry LDS stack
constexpr uint32_t NOALLOC_ERROR      = RAY_PARTITIONS * COUNTER_STRIDE; // u32 index of the error word
constexpr uint32_t NOALLOC_STATS      = NOALLOC_ERROR + 2;               // u32 index of 2 x u64
constexpr uint32_t NOALLOC_HEADER     = NOALLOC_ERROR + 64;              // u32 index of the overflow stacks
constexpr size_t   NOALLOC_WORDS      = NOALLOC_HEADER + size_t( NOALLOC_MAX_BLOCKS ) * TR_THREADS * NOALLOC_OVF;
static_assert( ( NOALLOC_STATS * 4 ) % 8 == 0, "statistics alignment" );
static_assert( ( NOALLOC_HEADER * 4 ) % 256 == 0, "stack alignment" );
static_assert( TRACE_STACK_DEPTH[ NOALLOC_VARIANT ] == 24, "no-alloc stack variant" );

// Clears the counters, the error word and the statistics.
__global__ __launch_bounds__( 64 ) void k_noalloc_clear( uint32_t* hdr )
{
    const uint32_t t = threadIdx.x;
    if( t < RAY_PARTITIONS )
    {
        hdr[ t * COUNTER_STRIDE ] = 0u;
    }
    if( t < 6 )
    {
        hdr[ NOALLOC_ERROR + t ] = 0u; // error word, pad, 2 x u64 statistics
    }
}

struct NoAllocDevice
{
    uint32_t* scratch = nullptr; // the device address of `scratch` on this GPU
    uint32_t  grid[ 2 ] = {};    // persistent blocks per node format (FP32, FP16)
};
constexpr int MAX_DEVICES = 64;
std::mutex    g_noAllocMutex;
NoAllocDevice g_noAlloc[ MAX_DEVICES ];
} // namespace

static __device__ __align__( 256 ) uint32_t scratch[ NOALLOC_WORDS ];

namespace
{
// Set-up once per device, with that device current: the array's device
// address (the host-side name `scratch` is not one) and the grid sizes.
const NoAllocDevice& NoAllocFor( int dev )
{
    RG_REQUIRE( dev >= 0 && dev < MAX_DEVICES, RG_RESULT_INTERNAL_ERROR, "device index out of range" );
    std::lock_guard< std::mutex > lock( g_noAllocMutex );
    NoAllocDevice&                d = g_noAlloc[ dev ];
    if( !d.scratch )
    {
        void* ptr = nullptr;
        RG_HIP_CHECK( hipGetSymbolAddress( &ptr, HIP_SYMBOL( scratch ) ) );
        int cus = 0;
        RG_HIP_CHECK( hipDeviceGetAttribute( &cus, hipDeviceAttributeMultiprocessorCount, dev ) );
        for( int h = 0; h < 2; h++ )
        {
            int perCu = 0;
            RG_HIP_CHECK( hipOccupancyMaxActiveBlocksPerMultiprocessor( &perCu, TraceKernel( h, NOALLOC_VARIANT ), TR_THREADS, 0 ) );
            const uint32_t g = uint32_t( cus * ( perCu > 0 ? perCu : 1 ) );
            d.grid[ h ]      = g < NOALLOC_MAX_BLOCKS ? g : NOALLOC_MAX_BLOCKS;
        }
        d.scratch = static_cast< uint32_t* >( ptr );
    }
    return d;
}
} // namespace

void TraceRaysNoAlloc( const BvhObject& bvh, const RgExtRay* d_rays, uint32_t count, RgExtHit* d_hits, uint32_t flags, hipStream_t stream )
{
    RG_REQUIRE( bvh.sceneCount == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "a batched BVH is traced with TraceRaysBatch" );
    RG_REQUIRE( bvh.nodeCount > 0, RG_RESULT_WRONG_FUNCTION_CALL, "BVH is not built" );
    RG_REQUIRE( bvh.nodeFormat == RG_EXT_BVH_NODE_FP32 || bvh.nodeFormat == RG_EXT_BVH_NODE_FP16, RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                "unsupported node format" );
    RG_REQUIRE( count <= RG_EXT_TRACE_NOALLOC_MAX_RAYS, RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                "too many rays for one call (at most RG_EXT_TRACE_NOALLOC_MAX_RAYS)" );
    int dev = 0;
    RG_HIP_CHECK( hipGetDevice( &dev ) );
    const NoAllocDevice& nd = NoAllocFor( dev );
    if( count == 0 )
    {
        return;
    }
    const int      h      = bvh.nodeFormat == RG_EXT_BVH_NODE_FP16 ? 1 : 0;
    const uint32_t blocks = DivUp( count, TR_THREADS ) < nd.grid[ h ] ? DivUp( count, TR_THREADS ) : nd.grid[ h ];
    RG_REQUIRE( size_t( blocks ) * TR_THREADS * NOALLOC_OVF <= NOALLOC_WORDS - NOALLOC_HEADER, RG_RESULT_INTERNAL_ERROR,
                "no-alloc traversal grid exceeds its scratch" );
    k_noalloc_clear<<< 1, 64, 0, stream >>>( nd.scratch );
    CheckLaunch( "k_noalloc_clear" );

    TraceArgs a{};
    a.nodes         = bvh.nodes.ptr;
    a.tris          = reinterpret_cast< const float4* >( bvh.sortedTris.ptr );
    a.rays          = d_rays;
    a.hits          = d_hits;
    a.count         = count;
    a.flags         = flags;
    a.partitions    = blocks < RAY_PARTITIONS ? blocks : RAY_PARTITIONS;
    a.rayCounter    = nd.scratch;
    a.errorFlag     = nd.scratch + NOALLOC_ERROR;
    a.overflowStack = nd.scratch + NOALLOC_HEADER;
    a.ovfStride     = NOALLOC_OVF;
    a.stackTotal    = TRACE_STACK_DEPTH[ NOALLOC_VARIANT ] + NOALLOC_OVF;
    a.stats         = ( flags & RG_EXT_TRACE_COLLECT_STATS ) ? reinterpret_cast< uint64_t* >( nd.scratch + NOALLOC_STATS ) : nullptr;
    a.segCount      = 0;
    a.rootCount     = 0;
    for( int k = 0; k < 8; k++ )
    {
        a.rootRef[ k ] = RG_EXT_BVH_CHILD_EMPTY;
    }
    a.chunk = RAY_CHUNK_DEFAULT;
    a.alpha = nullptr;
    a.tex   = TexView{};
    hipLaunchKernelGGL( TraceKernel( h, NOALLOC_VARIANT, false ), dim3( blocks ), dim3( TR_THREADS ), 0, stream, a );
    CheckLaunch( "k_trace" );
}

void ReadTraceNoAllocWords( uint32_t* errors, uint64_t* nodeVisits, uint64_t* triTests )
{
    int dev = 0;
    RG_HIP_CHECK( hipGetDevice( &dev ) );
    const NoAllocDevice& nd = NoAllocFor( dev );
    uint32_t             w[ 6 ] = {};
    RG_HIP_CHECK( hipDeviceSynchronize() );
    RG_HIP_CHECK( hipMemcpy( w, nd.scratch + NOALLOC_ERROR, sizeof( w ), hipMemcpyDeviceToHost ) );
    if( errors )
    {
        *errors = w[ 0 ];
    }
    uint64_t st[ 2 ];
    std::memcpy( st, w + 2, sizeof( st ) );
    if( nodeVisits )
    {
        *nodeVisits = st[ 0 ];
    }
    if( triTests )
    {
        *triTests = st[ 1 ];
    }
}

void TraceRaysFrame( const BvhObject& bvh, TraceContext& s, const RgExtRay* d_rays, const uint8_t* d_rayMasks, uint32_t defaultMask, uint32_t count,
                     RgExtHit* d_hits, uint32_t flags, const uint32_t* classMasks, uint32_t classCount, hipStream_t stream, hipEvent_t evBefore,
                     hipEvent_t evAfter, const TraceAccum* accum, const AlphaRec* alpha, const TexView* tex )
{
    RG_REQUIRE( !bvh.frameSegs.empty(), RG_RESULT_WRONG_FUNCTION_CALL, "not a frame BVH" );
    FrameRoots fr;
    fr.rayMasks    = d_rayMasks;
    fr.defaultMask = defaultMask;
    for( const auto& seg : bvh.frameSegs )
    {
        const uint32_t m = seg.cls < classCount ? classMasks[ seg.cls ] : 0u;
        if( m == 0u )
        {
            continue; // not in this frame's TLAS
        }
        fr.mask[ fr.rootCount ] = m;
        fr.ref[ fr.rootCount ]  = seg.root;
        for( int j = 0; j < 6; j++ )
        {
            fr.box[ fr.rootCount ][ j ] = seg.box[ j ];
        }
        fr.rootCount++;
    }
    if( fr.rootCount == 0 || count == 0 )
    {
        // nothing in the TLAS: every ray misses
        if( evBefore )
        {
            RG_HIP_CHECK( hipEventRecord( evBefore, stream ) );
        }
        if( count )
        {
            RG_HIP_CHECK( hipMemsetAsync( d_hits, 0xFF, size_t( count ) * sizeof( RgExtHit ), stream ) );
        }
        if( evAfter )
        {
            RG_HIP_CHECK( hipEventRecord( evAfter, stream ) );
        }
        return;
    }
    TraceRaysImpl( bvh, s, d_rays, count, d_hits, flags, stream, evBefore, evAfter, accum, nullptr, &fr, alpha, tex );
}

void TraceRaysBatch( const BvhObject& bvh, TraceContext& s, const RgExtRay* d_rays, const uint32_t* rayOffsets, RgExtHit* d_hits, uint32_t flags,
                     hipStream_t stream, hipEvent_t evBefore, hipEvent_t evAfter, const TraceAccum* accum )
{
    const uint32_t S = bvh.sceneCount;
    RG_REQUIRE( S > 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "the BVH was not built as a batch" );
    RG_REQUIRE( rayOffsets && rayOffsets[ 0 ] == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "ray offsets must start at 0" );
    for( uint32_t k = 0; k < S; k++ )
    {
        RG_REQUIRE( rayOffsets[ k + 1 ] >= rayOffsets[ k ], RG_RESULT_WRONG_FUNCTION_ARGUMENT, "ray offsets must not decrease" );
    }
    s.segHost.assign( rayOffsets, rayOffsets + S + 1 ); // kept alive until the copy has run
    s.segRays.Reserve( size_t( S + 1 ) * 4 );
    RG_HIP_CHECK( hipMemcpyAsync( s.segRays.ptr, s.segHost.data(), size_t( S + 1 ) * 4, hipMemcpyHostToDevice, stream ) );
    TraceRaysImpl( bvh, s, d_rays, rayOffsets[ S ], d_hits, flags, stream, evBefore, evAfter, accum, s.segRays.as< uint32_t >() );
}

void AutotuneTrace( const BvhObject& bvh, TraceContext& s, const RgExtRay* d_rays, uint32_t count, const uint32_t* rayOffsets, RgExtHit* d_hits,
                    uint32_t flags, hipStream_t stream, RgExtTraceTuning* out )
{
    RG_REQUIRE( count > 0 || rayOffsets, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "no rays to tune on" );
    static const uint32_t chunks[ 3 ] = { 32, 64, 128 };
    hipEvent_t            e0, e1;
    RG_HIP_CHECK( hipEventCreate( &e0 ) );
    RG_HIP_CHECK( hipEventCreate( &e1 ) );
    auto run = [ & ]() {
        if( rayOffsets )
        {
            TraceRaysBatch( bvh, s, d_rays, rayOffsets, d_hits, flags, stream, e0, e1 );
        }
        else
        {
            TraceRays( bvh, s, d_rays, count, d_hits, flags, stream, e0, e1 );
        }
        RG_HIP_CHECK( hipEventSynchronize( e1 ) );
        float ms = 0.0f;
        RG_HIP_CHECK( hipEventElapsedTime( &ms, e0, e1 ) );
        return ms;
    };
    float    best = 1e30f, def = 0.0f;
    uint32_t bestV = 1, bestC = 64;
    for( uint32_t v = 0; v < TRACE_STACK_VARIANTS; v++ )
    {
        for( uint32_t c : chunks )
        {
            s.stackVariant = v;
            s.rayChunk     = c;
            run(); // warm
            float t = 1e30f;
            for( int r = 0; r < 3; r++ )
            {
                const float ms = run();
                t              = ms < t ? ms : t;
            }
            if( v == 1 && c == 64 )
            {
                def = t;
            }
            if( t < best )
            {
                best  = t;
                bestV = v;
                bestC = c;
            }
        }
    }
    ( void )hipEventDestroy( e0 );
    ( void )hipEventDestroy( e1 );
    s.stackVariant = bestV;
    s.rayChunk     = bestC;
    if( out )
    {
        out->stackLdsEntries = uint32_t( TRACE_STACK_DEPTH[ bestV ] );
        out->rayChunk        = bestC;
        out->msPerLaunch     = best;
        out->msDefault       = def;
    }
}

} // namespace rg
