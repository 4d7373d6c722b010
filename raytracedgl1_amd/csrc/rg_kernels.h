// Host-side launchers of the hot-path kernels.
#pragma once

#include "rg_internal.h"

namespace rg
{

struct FramebufferSet;

// LPM 709 -> 709 control values (the LpmSetup control block's used entries)
struct LpmConsts
{
    float saturation[ 3 ];
    float contrast;
    float toneScaleBias[ 2 ];
    float lumaT[ 3 ];
    float rcpLumaT[ 3 ];
    float crosstalk[ 3 ];
};

// rg_tonemap.hip
void LuminanceHistogram( const float* d_lum, uint32_t count, float minLog, float maxLog, uint32_t* d_hist, hipStream_t stream );
void LuminanceAverage( uint32_t* d_hist, uint32_t pixelCount, float minLog, float maxLog, float timeDelta, bool stop, float* d_avg, hipStream_t stream );
LpmConsts MakeLpmConsts( const float saturation[ 3 ], const float crosstalk[ 3 ] );
// PRE_FINAL -> FINAL: checkerboard resolve, auto exposure from *d_avgLum, LPM tone map
// emissionMaxScreenColor weights SCREEN_EMISSION into the tone map's input.
void ComposeFinal( FramebufferSet& fb, const LpmConsts& k, float emissionMaxScreenColor, const float* d_avgLum, hipStream_t stream );
// FINAL (pixels float RGBA) -> RGBA8, sRGB transfer, alpha 255: the frame a
// server reads back.
void PackFinalRgba8( const float* d_final, uint32_t* d_out, uint32_t pixels, hipStream_t stream );

} // namespace rg
