// Filter -> project on a resident chunk: values of the rows set in a condition
// bitmap (buff_decode_condition, buff_simd.rs:1528-1712).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "buff_common.h"

namespace buffgpu {

constexpr uint32_t kCondTilesPerWg = 64;   // tiles one count / gather workgroup owns

// Device scratch the decode needs, and where in it the total row count lands (u64).
uint64_t cond_scratch_bytes(uint64_t ntiles);
uint64_t cond_total_offset(uint64_t ntiles);
// Writes the condition rows' values (at most `cap`) to out in row order; the
// row count is left at scratch + cond_total_offset(ntiles). compat: the
// reference's tail placement when F >= 8 and F % 8 != 0.
// The condition rows of n decoded doubles `vals`, in row order (same scratch as
// launch_cond_decode for ceil(n/1024) tiles; the row count at cond_total_offset).
hipError_t launch_cond_values(const double* vals, uint64_t n, const uint64_t* cond, double* out, uint64_t cap,
                              uint8_t* scratch, hipStream_t s);
hipError_t launch_cond_decode(const ChunkView& cv, double inv_scale, bool compat, const uint64_t* cond, double* out,
                              uint64_t cap, uint8_t* scratch, hipStream_t s);

}  // namespace buffgpu
