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
} // namespace rg
