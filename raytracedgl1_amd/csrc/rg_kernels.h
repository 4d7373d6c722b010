// Host-side launchers of the hot-path kernels.
#pragma once

#include "rg_internal.h"

namespace rg
{

// rg_tonemap.hip
void LuminanceHistogram( const float* d_lum, uint32_t count, float minLog, float maxLog, uint32_t* d_hist, hipStream_t stream );
void LuminanceAverage( uint32_t* d_hist, uint32_t pixelCount, float minLog, float maxLog, float timeDelta, bool stop, float* d_avg, hipStream_t stream );

} // namespace rg
