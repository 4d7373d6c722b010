// Frame orchestration shared by the RTGL1 instance and the render server:
// frame constants from an RgDrawFrameInfo (FillUniform,
// Source/VulkanDevice.cpp:142-482) and the GPU passes of one frame after the
// scene is resident (VulkanDevice::Render, :540-588, as rebuilt here).
#pragma once

#include "rg_bvh.h"
#include "rg_framebuffers.h"
#include "rg_pathtrace.h"
#include "rg_scene.h"

namespace rg
{

// Finds a structure of the RgDrawFrameInfo::pParams chain by its sType.
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

// Render size of a frame: RENDER_RESOLUTION custom size or the path
// params' size, else 1280 x 720; the width is made even.
void RenderSizeOf( const RgDrawFrameInfo& di, uint32_t* w, uint32_t* h );

// What a camera carries between frames on the host.
struct CameraState
{
    RgExtFrameParams params{};
    uint32_t         frameId    = 0;
    double           prevTime   = 0.0;
    bool             firstFrame = true;
    float            worldUp[ 3 ] = { 0.0f, 1.0f, 0.0f }; // RgInstanceCreateInfo::worldUp

    CameraState();
};

// FillUniform's hot-path fields for this frame; the previous frame's camera
// becomes the *Prev fields.  allowSkyGeometry: RgInstanceCreateInfo::
// allowGeometryWithSkyFlag; ignoreFpvLightIndex: this frame's index of the
// light whose shadows skip first-person-viewer geometry (LIGHT_INDEX_NONE).
void FillFrameParams( CameraState& cs, const RgDrawFrameInfo& di, uint32_t w, uint32_t h, uint32_t lightArraySize, uint32_t prevLightArraySize,
                      bool hasDirectional, bool cullBackFaces, bool allowSkyGeometry = false, uint32_t ignoreFpvLightIndex = LIGHT_INDEX_NONE );

// Device state of one camera's frames besides its framebuffers.
struct FrameGpu
{
    PathTracer   pt;
    DeviceBuffer hist, histCopy, avgLum; // 256-bin histogram, its last copy, adapted luminance

    void Init( hipStream_t s );
    static size_t OwnBytes() { return PathTracer::OwnBytes() + 2 * 256 * 4 + 4; }
};

// Ray-traced passes, A-SVGF, eye adaptation and final compose of one frame,
// enqueued on s without host synchronisation.  evTraced / evDenoised
// (optional) are recorded after the traced passes and after the denoiser.
// Returns the number of rays traced.
uint64_t EncodeFrame( FrameGpu& g, TraceContext& tc, const RgExtFrameParams& p, const BvhObject& bvh, const DeviceScene& scene,
                      FramebufferSet& fb, const RgDrawFrameInfo& di, hipStream_t s, hipEvent_t evTraced = nullptr,
                      hipEvent_t evDenoised = nullptr );

// One frame of a batch for EncodeFrames.
struct EncodeSlot
{
    FrameGpu*               g;
    const RgExtFrameParams* p;
    FramebufferSet*         fb;
    const RgDrawFrameInfo*  di;
    const DeviceScene*      scene = nullptr; // its scene, when the frames' scenes differ
    uint32_t                group = 0;       // its scene's index in the grouped BVH
};

// B frames: their traced passes share each stage's traversal launch
// (TraceFrames over the wavefront wf, which must hold every frame's pixels);
// A-SVGF, eye adaptation and compose then run frame by frame.  With `scene`
// every frame is of that scene and bvh is its frame BVH; with a null scene
// the frames are of their slots' plain scenes, in group order, and bvh is
// the batched BVH of the group.  A frame's result equals EncodeFrame's for
// it alone.
uint64_t EncodeFrames( const EncodeSlot* slots, uint32_t B, TraceContext& tc, const BvhObject& bvh, const DeviceScene* scene, Wavefront& wf,
                       hipStream_t s );

} // namespace rg
