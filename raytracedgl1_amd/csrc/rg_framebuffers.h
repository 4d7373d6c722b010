// Per-pixel planes of the hot path (the images of Source/Generated/
// GenerateShaderCommon.py:802-887), stored in HBM as one linear array per
// plane, row-major, components interleaved (FP32 storage mode: 4 bytes per
// component).  Planes with a previous-frame copy are double-buffered; Swap()
// exchanges the copies at the start of a frame like FrameIndexToFBIndex
// (Source/Framebuffers.cpp:35-47).
#pragma once

#include "rg_internal.h"

namespace rg
{

enum PlaneLayout : uint8_t
{
    LAYOUT_CHECKERBOARD = 0, // W x H, odd pixels in the right half
    LAYOUT_REGULAR      = 1, // W x H
    LAYOUT_STRATA       = 2, // max(1,(W+1)/3) x max(1,(H+1)/3)
};

struct PlaneDesc
{
    PlaneLayout layout;
    uint8_t     comps; // 4-byte components per pixel
    bool        prev;
};

const PlaneDesc& GetPlaneDesc( int plane );

// Planes stored as IEEE half in RG_EXT_STORAGE_FP16 mode: the temporal and
// a-trous buffers, ACCUM_HISTORY_LENGTH .. INDIR_PONG plus
// ATROUS_FILTERED_VARIANCE.
__host__ __device__ constexpr bool IsHalfPlane( int plane )
{
    return ( plane >= RG_EXT_PLANE_ACCUM_HISTORY_LENGTH && plane <= RG_EXT_PLANE_INDIR_PONG ) ||
           plane == RG_EXT_PLANE_ATROUS_FILTERED_VARIANCE;
}

// Planes a frame reads back from the frame before: the double-buffered ones,
// DIFF_COLOR_HISTORY (the a-trous history read by the next temporal pass),
// DIS_GRADIENT_HISTORY and GRADIENT_PREV_PIX (read by the next frame's
// gradient samples and GI reuse), plus the FINAL output.  All other planes
// are written and read within one frame.
bool IsPersistentPlane( int plane );

// Device pointers of every plane, passed to kernels by value.
struct FbPtrs
{
    void*    cur[ RG_EXT_PLANE_COUNT ];
    void*    prev[ RG_EXT_PLANE_COUNT ];
    uint32_t half; // RG_EXT_STORAGE_FP16: IsHalfPlane planes hold IEEE half
};

// The within-frame planes at one capacity, shared by framebuffer sets whose
// frames run one after another on one stream (render sessions of a server).
struct TransientPlanes
{
    uint32_t     capW = 0, capH = 0, storage = RG_EXT_STORAGE_FP32;
    DeviceBuffer planes[ RG_EXT_PLANE_COUNT ];

    void Create( uint32_t maxW, uint32_t maxH, uint32_t storageMode );
};

struct FramebufferSet
{
    uint32_t         width = 0, height = 0, gw = 0, gh = 0;
    uint32_t         storage = RG_EXT_STORAGE_FP32;
    DeviceBuffer     planes[ RG_EXT_PLANE_COUNT ][ 2 ];
    int              parity = 0;
    TransientPlanes* shared = nullptr; // within-frame planes live there when set
    uint32_t         atrousTileW = 64; // a-trous iterations >= 1: workgroup tile width (x 256/width rows)

    // With `transient`, only the persistent planes are allocated here.
    void   Create( uint32_t w, uint32_t h, uint32_t storageMode = RG_EXT_STORAGE_FP32, TransientPlanes* transient = nullptr );
    // Device bytes Create allocates for a w x h set.
    static size_t BytesFor( uint32_t w, uint32_t h, uint32_t storageMode, bool persistentOnly );
    bool   Half() const { return storage == RG_EXT_STORAGE_FP16; }
    size_t PlaneBytes( int plane ) const;
    void*  Cur( int plane );
    void*  Prev( int plane );
    void   Swap() { parity ^= 1; }
    FbPtrs Ptrs();
    size_t TotalBytes() const;
};

} // namespace rg

// Handle type of RTGL1_ext.h's RgExtFramebuffers.
struct RgExtFramebuffers_T
{
    rg::FramebufferSet fb;
};
