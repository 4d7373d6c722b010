// Host launchers of the A-SVGF passes (rg_denoise.hip).
#pragma once

#include "rg_framebuffers.h"

namespace rg
{
void DenoiseGradientAtrous( FramebufferSet& fb, const RgExtFrameParams& p, uint32_t iteration, hipStream_t s );
void DenoiseTemporal( FramebufferSet& fb, const RgExtFrameParams& p, hipStream_t s );
void DenoiseAntiFirefly( FramebufferSet& fb, const RgExtFrameParams& p, hipStream_t s );
void DenoiseVariance( FramebufferSet& fb, const RgExtFrameParams& p, hipStream_t s );
void DenoiseAtrous( FramebufferSet& fb, const RgExtFrameParams& p, uint32_t iteration, hipStream_t s );
void Denoise( FramebufferSet& fb, const RgExtFrameParams& p, hipStream_t s );
} // namespace rg
