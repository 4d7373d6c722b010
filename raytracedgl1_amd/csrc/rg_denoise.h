// Host launchers of the A-SVGF passes (rg_denoise.hip).
#pragma once

#include "rg_framebuffers.h"

namespace rg
{
void DenoiseGradientAtrous( FramebufferSet& fb, const RgExtFrameParams& p, uint32_t iteration, hipStream_t s );
void DenoiseTemporal( FramebufferSet& fb, const RgExtFrameParams& p, hipStream_t s );
void DenoiseAntiFirefly( FramebufferSet& fb, const RgExtFrameParams& p, hipStream_t s );
// The variance and a-trous passes read the denoise guide (decoded normals
// and depths) that the temporal pass writes; on their own (writeGuide) they
// write it first from NORMAL and DEPTH_WORLD.
void DenoiseVariance( FramebufferSet& fb, const RgExtFrameParams& p, hipStream_t s, bool writeGuide = true );
void DenoiseAtrous( FramebufferSet& fb, const RgExtFrameParams& p, uint32_t iteration, hipStream_t s, bool writeGuide = true );
void Denoise( FramebufferSet& fb, const RgExtFrameParams& p, hipStream_t s );
// Workgroup tile widths of the a-trous iterations >= 1 (256 threads each).
constexpr uint32_t ATROUS_TILE_WIDTHS[ 4 ] = { 16, 32, 64, 128 };
void SetAtrousTile( FramebufferSet& fb, uint32_t tileWidth );
// Times a-trous iteration 1 on fb's current planes with every tile width
// (best of 3 after a warm-up, on stream s, which it synchronises) and keeps
// the fastest in fb.  Writes iteration 1's output planes.  msOut (optional)
// receives the four times in ATROUS_TILE_WIDTHS order.
uint32_t AutotuneAtrous( FramebufferSet& fb, const RgExtFrameParams& p, hipStream_t s, float* msOut );
} // namespace rg
