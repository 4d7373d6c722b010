// Ray-traced passes of a frame: primary G-buffer, light grid, ReSTIR DI
// (initial reservoirs, temporal + spatial reuse, shadow rays) and indirect
// bounces, scheduled as wavefronts around the shared traversal kernel.
#pragma once

#include "rg_bvh.h"
#include "rg_framebuffers.h"
#include "rg_scene.h"

namespace rg
{

struct FrameTimings
{
    float assembleMs = 0, bvhMs = 0, primaryMs = 0, directMs = 0, indirectMs = 0, denoiseMs = 0, tonemapMs = 0, totalMs = 0;
    uint64_t rays = 0;
};

// Primary rays of a camera (p's inverse view / projection, render size,
// jitter, primaryRayMinDist, rayLength), row-major over regular pixels.
void PrimaryRays( const RgExtFrameParams& p, RgExtRay* rays, hipStream_t s );

// Ray, hit and per-pixel state buffers between the wavefronts of one frame.
// Nothing in them survives a frame, so path tracers whose frames run one after
// another on one stream can share one set.
struct Wavefront
{
    DeviceBuffer rays, hits, rays2, hits2; // W*H
    DeviceBuffer pendA, pendB, pendC;      // W*H x (6 float4 / ray / hit)
    DeviceBuffer masks, masks2, masksB;    // W*H cull masks of rays, rays2, pendB's rays
    size_t       capPx = 0;

    void          Reserve( size_t pixels );
    static size_t BytesFor( size_t pixels );
};

struct PathTracer
{
    Wavefront    own;
    Wavefront*   wf = &own;                // set to a shared set before the first frame
    DeviceBuffer lights, prevToCur;
    DeviceBuffer geomMap;                  // last frame's geometry index -> this frame's
    std::vector< uint32_t > geomMapHost;
    uint32_t     geomMapCount = 0;
    DeviceBuffer grid[ 2 ];                // 16^3 * 128 light-grid reservoirs, ping-pong
    const uint32_t* blueNoise = nullptr;   // 128 layers of 128x128 RGBA8, one table per process
    DeviceBuffer frameAccum;               // node visits, triangle tests (u64), error bits (u32) of the frame
    static constexpr uint32_t TRACE_EVENT_PAIRS = 32; // traversal launches timed per frame
    hipEvent_t   traceEv[ 2 * TRACE_EVENT_PAIRS ] = {};
    uint32_t     traceLaunches = 0;
    bool         collectStats  = false;
    int          gridParity = 0;
    uint32_t     lightCount = 0, lightArraySize = 0, prevToCurCount = 0;
    uint32_t     W = 0, H = 0;

    void Init( hipStream_t s );
    // Device bytes a path tracer owns apart from its wavefront buffers.
    static size_t OwnBytes();
    ~PathTracer();
    // After the stream has finished the frame: traversal time summed over the
    // frame's launches, and the node / triangle counts when collected.
    float    TraversalMs() const;
    void     TraversalCounts( uint64_t* nodes, uint64_t* tris, uint32_t* errors ) const;
    void Resize( uint32_t w, uint32_t h );
    void UploadLights( const SceneCollector& sc, hipStream_t s );
    void UploadGeomMap( const std::vector< uint32_t >& map, hipStream_t s );

    // One frame of ray tracing into fb (G-buffer, unfiltered direct /
    // specular / indirect, reservoirs).  Returns the number of rays traced.
    uint64_t Trace( const RgExtFrameParams& p, const BvhObject& bvh, TraceContext& tc, const DeviceScene& scene,
                    FramebufferSet& fb, uint32_t traceFlags, hipStream_t s, FrameTimings* t );
};

// One frame of a batch: its path tracer (light grid, reservoir history),
// frame constants and framebuffers.
struct TraceSlot
{
    PathTracer*             pt;
    const RgExtFrameParams* p;
    FramebufferSet*         fb;
    const DeviceScene*      scene = nullptr; // its scene, when the frames' scenes differ
    uint32_t                group = 0;       // its scene's index in the grouped BVH
};
constexpr uint32_t MAX_TRACE_SLOTS = 16;

// Frames that can share traversal launches: the same ray cull masks,
// reflection / refraction depth, face culling and light kinds (their scene
// must be the same too).
bool TraceCompatible( const RgExtFrameParams& a, const RgExtFrameParams& b );

// Scenes whose frames can share traversal launches although they differ:
// every triangle opaque world-0 geometry, no alpha test, no reflecting or
// refracting geometry (one BVH segment, no per-ray cull masks to honour).
bool GroupableScene( const DeviceScene& scene );

// The traced passes of B frames, stage by stage: each stage's shading
// kernels run per frame, and each traversal covers every frame's rays of
// that stage in one launch (frame k's rays are a slice of wf).  With
// `scene`, every frame is of that scene and bvh is its frame BVH; with a
// null scene, frame k is of slots[k].scene (GroupableScene), bvh is one
// batched BVH over those scenes (BuildBvhBatch) in which that scene is
// number slots[k].group, and the slots come in non-decreasing group order.
// Every frame's result equals the frame traced alone.  The first slot's
// path tracer times and counts the launches.  Returns the rays traced.
uint64_t TraceFrames( const TraceSlot* slots, uint32_t B, const BvhObject& bvh, TraceContext& tc, const DeviceScene* scene, Wavefront& wf,
                      uint32_t traceFlags, hipStream_t s );

} // namespace rg
