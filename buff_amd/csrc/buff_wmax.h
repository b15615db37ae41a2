// Windowed max of a resident chunk (buff_max_range, buff_simd.rs:5096-5334).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "buff_common.h"

namespace buffgpu {

// Scratch of one windowed max over rows [s, e) in windows of W rows: the window
// maxima (nwin u64) and two ping-pong buffers of (window, max) pairs.
struct WmaxLayout {
  uint64_t nwin = 0, t_first = 0, ntiles = 0;
  uint64_t wmax = 0, k[2] = {0, 0}, v[2] = {0, 0}, bytes = 0;
};
WmaxLayout wmax_layout(uint64_t s, uint64_t e, uint64_t W);

// Every window's max stored offset into scratch (wmax region).
hipError_t launch_wmax(const ChunkView& cv, uint64_t s, uint64_t e, uint64_t W, const WmaxLayout& L,
                       uint8_t* scratch, hipStream_t st);
// ceil(n/64) u64 words: the rows of [s, e) equal to their window's max (want =
// false: all zero).
hipError_t launch_wmax_argmax(const ChunkView& cv, uint64_t s, uint64_t e, uint64_t W, bool want,
                              const uint8_t* scratch, const WmaxLayout& L, uint64_t* out, hipStream_t st);
// The first k maxima as (x + base) * inv_scale.
hipError_t launch_wmax_values(const uint8_t* scratch, const WmaxLayout& L, uint64_t k, int64_t base, double inv_scale,
                              double* vals, hipStream_t st);

}  // namespace buffgpu
