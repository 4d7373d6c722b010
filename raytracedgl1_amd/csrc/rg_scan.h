#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace rg
{
// Scratch needed by ExclusiveScan for n elements.
size_t ScanScratchBytes( uint32_t n );

// d_out[i] = sum(d_in[0..i-1]); *d_total = sum of all (d_total may be null).
// d_in may equal d_out.
void ExclusiveScan( const uint32_t* d_in, uint32_t n, uint32_t* d_out, uint32_t* d_total, uint32_t* d_scratch, hipStream_t stream );
} // namespace rg
