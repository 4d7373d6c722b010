// LBVH object shared by the build (rg_lbvh.hip) and traversal (rg_trace.hip).
#pragma once

#include "rg_internal.h"

#include <vector>

namespace rg
{

// 8-wide node with FP32 child boxes, 256 bytes (SoA over the 8 children so
// one child's slab test reads 6 floats at a fixed stride).
struct alignas( 256 ) BvhNode8
{
    float    lo_x[ 8 ];
    float    hi_x[ 8 ];
    float    lo_y[ 8 ];
    float    hi_y[ 8 ];
    float    lo_z[ 8 ];
    float    hi_z[ 8 ];
    uint32_t child[ 8 ];
    uint32_t pad[ 8 ];
};
static_assert( sizeof( BvhNode8 ) == 256, "node size" );

// 8-wide node with FP16 child boxes, 144 bytes: an FP32 origin (the min
// corner of the node's own box) and each child box as IEEE-half offsets from
// it, rounded outward so that origin + float(offset), evaluated in FP32 as
// the traversal does, contains the FP32 box.  Offsets keep 11 significant
// bits relative to the node's extent, where absolute FP16 coordinates would
// lose all precision far from the scene origin.
struct alignas( 16 ) BvhNode8H
{
    float    origin[ 3 ];
    uint32_t pad;
    uint16_t lo_x[ 8 ];
    uint16_t hi_x[ 8 ];
    uint16_t lo_y[ 8 ];
    uint16_t hi_y[ 8 ];
    uint16_t lo_z[ 8 ];
    uint16_t hi_z[ 8 ];
    uint32_t child[ 8 ];
};
static_assert( sizeof( BvhNode8H ) == 144, "node size" );

constexpr uint32_t BVH_MAX_LEVELS = 128;

struct BvhObject
{
    uint32_t n          = 0;
    uint32_t nodeCount  = 0;
    uint32_t nodeFormat = RG_EXT_BVH_NODE_FP32;
    uint32_t leafMax    = 3;
    double   sceneMin[ 3 ] = { 0, 0, 0 };
    double   sceneMax[ 3 ] = { 0, 0, 0 };

    DeviceBuffer sortedTris;   // n * 3 float4
    DeviceBuffer codes;        // n u64, input order
    DeviceBuffer keysA, keysB; // n u64
    DeviceBuffer valsA, valsB; // n u32
    DeviceBuffer hist;         // 256 * radix blocks u32
    DeviceBuffer scanScratch;
    DeviceBuffer boundsPartials; // 6 doubles per block + 6 final
    DeviceBuffer children;     // (n-1) * 2 u32
    DeviceBuffer parents;      // (2n-1) u32
    DeviceBuffer rangeFirst;   // (n-1) u32
    DeviceBuffer rangeLast;    // (n-1) u32
    DeviceBuffer nodeBounds;   // (2n-1) * 6 f32
    DeviceBuffer depthA, depthB, jumpA, jumpB; // (2n-1) u32: node depths by pointer jumping
    DeviceBuffer levelHist;    // per-depth counts, then offsets and cursors
    DeviceBuffer levelOrder;   // (n-1) internal nodes grouped by depth
    DeviceBuffer expansion;    // max(1,n-1) * 8 u32
    DeviceBuffer expandInfo;   // max(1,n-1) uint4: children with open bits, their half areas
    DeviceBuffer isNode8;      // max(1,n-1) u32
    DeviceBuffer node8Index;   // max(1,n-1) u32
    DeviceBuffer frontier[ 2 ];
    DeviceBuffer levelCounts;  // BVH_MAX_LEVELS + 2 u32 (last two: node count, error flag)
    DeviceBuffer nodes;        // nodeCount * stride
    DeviceBuffer nodes32;      // FP32 nodes an FP16 build converts from
    uint32_t*    hostCounts = nullptr; // pinned
    uint32_t*    hostLevels = nullptr; // pinned per-depth node counts
    // batched build: per-scene offsets, binary and 8-wide roots, bounds
    uint32_t                sceneCount = 0;
    DeviceBuffer            sceneOffsets, sceneRoots, sceneRoots8, sceneBounds, scenePartials;
    std::vector< uint32_t > hostSceneOffsets, hostSceneRoots;

    // Index of the sorted-order permutation after the sort (keysA/valsA or B).
    const uint32_t* sortedIdx = nullptr;
    const uint64_t* sortedKeys = nullptr;

    ~BvhObject()
    {
        if( hostCounts )
        {
            ( void )hipHostFree( hostCounts );
        }
        if( hostLevels )
        {
            ( void )hipHostFree( hostLevels );
        }
    }

    uint32_t NodeStride() const { return nodeFormat == RG_EXT_BVH_NODE_FP16 ? uint32_t( sizeof( BvhNode8H ) ) : uint32_t( sizeof( BvhNode8 ) ); }

    // Frame BVH (BuildFrameBvh): one segment per non-empty class of the
    // frame's triangles, in class order.  A ray starts at the roots of the
    // segments its cull mask admits, the way a TLAS instance is visited
    // when (instance mask & cull mask) != 0.
    struct FrameSegment
    {
        uint32_t cls  = 0;       // the class (index into BuildFrameBvh's counts)
        uint32_t root = 0;       // 8-wide node, or the leaf of a one-triangle segment
        float    box[ 6 ] = {};  // lo x, y, z, hi x, y, z (infinite for a single segment)
    };
    std::vector< FrameSegment > frameSegs;
    DeviceBuffer                segBoxes;     // 8 floats per segment
    std::vector< float >        hostSegBoxes;
};

void BuildBvh( BvhObject& bvh, const float* d_tris, uint32_t n, const RgExtBvhBuildParams& params, hipStream_t stream );
// Several scenes in one BVH object: triangles of scene s are
// [offsets[s], offsets[s+1]) of d_tris (host offsets); each scene gets its
// own root (hostSceneRoots) and never shares a node with another scene.
void BuildBvhBatch( BvhObject& bvh, const float* d_tris, const uint32_t* offsets, uint32_t sceneCount, const RgExtBvhBuildParams& params,
                    hipStream_t stream );

// The frame's BVH over the triangles of classTris[c] triangles per class,
// classes in order.  One class: a single-scene build.  Several: one batched
// build with a segment per non-empty class (1..8 of them).
void BuildFrameBvh( BvhObject& bvh, const float* d_tris, const uint32_t* classTris, uint32_t classCount, const RgExtBvhBuildParams& params,
                    hipStream_t stream );

constexpr int TRACE_STACK_VARIANTS = 3;

// Per-caller traversal scratch: the ray counter, an overflow-flag word and
// the global overflow stacks of the persistent grid.  Launches that share a
// context must not overlap on the device.
struct TraceContext
{
    DeviceBuffer counter;
    DeviceBuffer overflow;
    DeviceBuffer stats; // 2 x u64: node visits, triangle tests of the last collecting launch
    DeviceBuffer            segRays; // ray offsets of a batched launch
    std::vector< uint32_t > segHost;
    uint32_t     gridBlocks = 0;                        // the largest grid (sizes the overflow stacks)
    uint32_t     grid[ 2 ][ TRACE_STACK_VARIANTS ] = {}; // resident blocks per node format and stack variant
    uint32_t     stackVariant = 1;                      // LDS stack depth 16, 24 (default) or 32
    uint32_t     rayChunk     = 64;                     // rays per counter atomic
};

// Device accumulators a caller clears once and lets several launches add to.
struct TraceAccum
{
    uint64_t* stats  = nullptr; // node visits, triangle tests (with RG_EXT_TRACE_COLLECT_STATS)
    uint32_t* errors = nullptr; // OR of RG_EXT_TRACE_ERROR_* bits
};

// evBefore / evAfter (optional) are recorded right around the kernel launch.
// Without accum, the statistics and error bits of this launch alone are kept
// in ctx (read back by ReadTraceStats / ReadTraceErrors).
void TraceRays( const BvhObject& bvh, TraceContext& ctx, const RgExtRay* d_rays, uint32_t count, RgExtHit* d_hits, uint32_t flags, hipStream_t stream,
                hipEvent_t evBefore = nullptr, hipEvent_t evAfter = nullptr, const TraceAccum* accum = nullptr );
// TraceRays over a frame BVH (BuildFrameBvh).  classMasks[c] is the
// instance mask of class c, 0 when the class is not in the frame's TLAS; a
// ray sees the segments whose mask shares a bit with its own cull mask,
// rayMasks[i] (or defaultMask for every ray when rayMasks is null).
// alpha (with tex): per-triangle alpha-test records; the hits of
// alpha-tested triangles whose texture is transparent there are ignored.
void TraceRaysFrame( const BvhObject& bvh, TraceContext& ctx, const RgExtRay* d_rays, const uint8_t* d_rayMasks, uint32_t defaultMask, uint32_t count,
                     RgExtHit* d_hits, uint32_t flags, const uint32_t* classMasks, uint32_t classCount, hipStream_t stream, hipEvent_t evBefore,
                     hipEvent_t evAfter, const TraceAccum* accum, const struct AlphaRec* alpha = nullptr, const struct TexView* tex = nullptr );
// One persistent launch over the rays of every scene of a batched BVH: rays
// of scene s are [rayOffsets[s], rayOffsets[s+1]) (host array, sceneCount + 1
// entries); primitive ids in the hits are relative to the scene.
void TraceRaysBatch( const BvhObject& bvh, TraceContext& ctx, const RgExtRay* d_rays, const uint32_t* rayOffsets, RgExtHit* d_hits, uint32_t flags,
                     hipStream_t stream, hipEvent_t evBefore = nullptr, hipEvent_t evAfter = nullptr, const TraceAccum* accum = nullptr );
// Times every LDS stack depth x ray chunk on these rays (a batched BVH takes
// rayOffsets, else count) and keeps the fastest on ctx for later launches.
void AutotuneTrace( const BvhObject& bvh, TraceContext& ctx, const RgExtRay* d_rays, uint32_t count, const uint32_t* rayOffsets, RgExtHit* d_hits,
                    uint32_t flags, hipStream_t stream, RgExtTraceTuning* out );
// Synchronous reads of the last un-accumulated launch's results.
// rgExtTraceRays without allocation: scratch in a fixed per-GPU device
// array (rg_trace.hip); ReadTraceNoAllocWords reads the last such call's
// error bits and counts on the current device (synchronises it).
void TraceRaysNoAlloc( const BvhObject& bvh, const RgExtRay* d_rays, uint32_t count, RgExtHit* d_hits, uint32_t flags, hipStream_t stream );
void ReadTraceNoAllocWords( uint32_t* errors, uint64_t* nodeVisits, uint64_t* triTests );
void     ReadTraceStats( const TraceContext& ctx, uint64_t* nodeVisits, uint64_t* triTests );
uint32_t ReadTraceErrors( const TraceContext& ctx );

} // namespace rg
