// Windowed max (buff_max_range, buff_simd.rs:5096-5334): the max stored offset
// of every window of W rows of [s, e), and the rows holding their window's max.
//
// The reference walks plane 0 of [s, e) keeping, per window, the rows tied on
// the running max byte, then narrows them plane by plane; the result is each
// window's max offset. Here every tile of [s, e) decodes its rows' offsets once
// (planes + tail), parks them in LDS and reduces every window segment it holds:
//   k_wmax_tiles   windows that lie inside one tile are final; every tile also
//                  emits (window, max) for its first and last window segment
//                  (k_wmax_tiles_wide: windows of a tile or more, in registers);
//   k_wmax_reduce  the emitted pairs are sorted by window: a reduce-by-key per
//                  block of 256 pairs writes the windows that end inside the
//                  block and passes on the two boundary runs, until one block
//                  remains (three levels for 134M rows at any window size);
//   k_wmax_argmax  each row of [s, e) against its window's max -> bitmap;
//   k_wmax_values  max offsets -> the reference's f64 (x + base) / 2^dlen.
#include <algorithm>

#include "buff_device.h"
#include "buff_launch.h"
#include "buff_wmax.h"

namespace buffgpu {

// Offsets of the lane's 16 rows of tile c: NP planes, then r tail bits.
template <int NP>
__device__ __forceinline__ void lane_offsets(const ChunkView& cv, uint64_t c, uint32_t lane, uint64_t (&off)[16]) {
  RowBytes<NP> rb;
  load_row_bytes<NP>(cv, c, lane, rb);
#pragma unroll
  for (uint32_t b = 0; b < 16; ++b) off[b] = row_offset<NP>(rb, b, cv.r);
}

// One wave per tile of [s, e): tiles t_first + blockIdx.x.
template <int NP>
__global__ __launch_bounds__(64) void k_wmax_tiles(ChunkView cv, uint64_t s, uint64_t e, uint64_t W, uint64_t t_first,
                                                   unsigned long long* __restrict__ wmax, uint32_t* __restrict__ ekey,
                                                   unsigned long long* __restrict__ eval) {
  __shared__ uint64_t so[kLdsRow * 64];
  const uint32_t lane = __lane_id();
  const uint64_t c = t_first + blockIdx.x;
  const uint64_t r0 = c * kTileRows;
  uint64_t off[16];
  lane_offsets<NP>(cv, c, lane, off);
#pragma unroll
  for (uint32_t b = 0; b < 16; ++b) so[kLdsRow * lane + b] = off[b];
  wave_lds_sync();
  // this tile's rows of [s, e): tile-local [a, b); its windows fw .. lw
  const uint64_t a = (s > r0 ? s : r0) - r0;
  const uint64_t b = (e < r0 + kTileRows ? e : r0 + kTileRows) - r0;
  const uint64_t fw = (r0 + a - s) / W, lw = (r0 + b - 1 - s) / W;
  const uint64_t nwin = lw - fw + 1;
  const uint32_t e0 = 2 * blockIdx.x, e1 = 2 * blockIdx.x + 1;
  if (W >= 64) {
    // few windows: the wave reduces each segment together
    for (uint64_t w = fw; w <= lw; ++w) {
      const uint64_t ws = s + w * W, we = ws + W < e ? ws + W : e;
      const uint64_t x0 = (ws > r0 + a ? ws : r0 + a) - r0, x1 = (we < r0 + b ? we : r0 + b) - r0;
      uint64_t m = 0;
      for (uint64_t q = x0 + lane; q < x1; q += 64) m = max(m, so[lds_slot((uint32_t)q)]);
      m = wave_max_u64(m);
      if (lane == 0) {
        if (ws >= r0 && we <= r0 + kTileRows) wmax[w] = m;   // the whole window is in this tile
        if (w == fw) {
          ekey[e0] = (uint32_t)w;
          eval[e0] = m;
        }
        if (w == lw) {
          ekey[e1] = (uint32_t)w;
          eval[e1] = m;
        }
      }
    }
  } else {
    // many short windows: one lane per window
    for (uint64_t wi = lane; wi < nwin; wi += 64) {
      const uint64_t w = fw + wi;
      const uint64_t ws = s + w * W, we = ws + W < e ? ws + W : e;
      const uint64_t x0 = (ws > r0 + a ? ws : r0 + a) - r0, x1 = (we < r0 + b ? we : r0 + b) - r0;
      uint64_t m = 0;
      for (uint64_t q = x0; q < x1; ++q) m = max(m, so[lds_slot((uint32_t)q)]);
      if (ws >= r0 && we <= r0 + kTileRows) wmax[w] = m;
      if (wi == 0) {
        ekey[e0] = (uint32_t)w;
        eval[e0] = m;
      }
      if (wi == nwin - 1) {
        ekey[e1] = (uint32_t)w;
        eval[e1] = m;
      }
    }
  }
}

// Windows of at least one tile (W >= 1024): a tile holds at most two window
// segments, so each lane keeps its rows' max for both and the wave reduces
// them; no LDS. One wave per tile, four per workgroup.
template <int NP>
__global__ __launch_bounds__(256) void k_wmax_tiles_wide(ChunkView cv, uint64_t s, uint64_t e, uint64_t W,
                                                         uint64_t t_first, uint64_t ntiles,
                                                         unsigned long long* __restrict__ wmax,
                                                         uint32_t* __restrict__ ekey,
                                                         unsigned long long* __restrict__ eval) {
  const uint32_t lane = __lane_id();
  const uint64_t k = wave_global();
  if (k >= ntiles) return;
  const uint64_t c = t_first + k;
  const uint64_t r0 = c * kTileRows, lr0 = r0 + 16ull * lane;
  uint64_t off[16];
  lane_offsets<NP>(cv, c, lane, off);
  const uint64_t a = (s > r0 ? s : r0), b = (e < r0 + kTileRows ? e : r0 + kTileRows);   // rows [a, b)
  const uint64_t fw = (a - s) / W, lw = (b - 1 - s) / W;
  const uint64_t split = s + (fw + 1) * W;   // first row of window fw + 1
  uint64_t m0 = 0, m1 = 0;
#pragma unroll
  for (uint32_t j = 0; j < 16; ++j) {
    const uint64_t row = lr0 + j;
    if (row < a || row >= b) continue;
    if (row < split) m0 = max(m0, off[j]);
    else m1 = max(m1, off[j]);
  }
  m0 = wave_max_u64(m0);
  m1 = wave_max_u64(m1);
  if (lane == 0) {
    const uint64_t ws0 = s + fw * W, we0 = ws0 + W < e ? ws0 + W : e;
    if (ws0 >= r0 && we0 <= r0 + kTileRows) wmax[fw] = m0;
    ekey[2 * k] = (uint32_t)fw;
    eval[2 * k] = m0;
    ekey[2 * k + 1] = (uint32_t)lw;
    eval[2 * k + 1] = lw == fw ? m0 : m1;
    if (lw != fw) {
      const uint64_t ws1 = s + lw * W, we1 = ws1 + W < e ? ws1 + W : e;
      if (ws1 >= r0 && we1 <= r0 + kTileRows) wmax[lw] = m1;
    }
  }
}

// Reduce-by-key over pairs sorted by window (one block of 256 pairs per
// workgroup). A run of equal windows that touches neither end of its block is
// complete: its max is written. The runs at the block's ends go to the next
// level as its pairs 2b and 2b+1 (the same run twice when it fills the block).
// last: one block holds every pair, so every run is complete.
__global__ __launch_bounds__(256) void k_wmax_reduce(const uint32_t* __restrict__ kin,
                                                     const unsigned long long* __restrict__ vin, uint64_t nin,
                                                     uint32_t* __restrict__ kout, unsigned long long* __restrict__ vout,
                                                     unsigned long long* __restrict__ wmax, uint32_t last) {
  __shared__ uint32_t sk[256];
  __shared__ unsigned long long sv[256];
  const uint32_t t = threadIdx.x;
  const uint64_t i = (uint64_t)blockIdx.x * 256 + t;
  const bool valid = i < nin;
  const uint32_t key = valid ? kin[i] : 0xFFFFFFFFu;
  sk[t] = key;
  sv[t] = valid ? vin[i] : 0ull;
  __syncthreads();
  // segmented inclusive max: keys are sorted, so sk[t-o] == key means the
  // whole span t-o .. t belongs to the run
  for (uint32_t o = 1; o < 256; o <<= 1) {
    const unsigned long long v = (t >= o && sk[t - o] == key) ? sv[t - o] : 0ull;
    __syncthreads();
    sv[t] = sv[t] > v ? sv[t] : v;
    __syncthreads();
  }
  if (!valid) return;
  const bool at_end = t == 255 || i == nin - 1;
  const bool run_end = at_end || sk[t + 1] != key;
  if (!run_end) return;
  const bool at_start = sk[0] == key;
  const unsigned long long m = sv[t];
  if (last || (!at_start && !at_end)) wmax[key] = m;
  if (!last) {
    if (at_start) {
      kout[2 * blockIdx.x] = key;
      vout[2 * blockIdx.x] = m;
    }
    if (at_end) {
      kout[2 * blockIdx.x + 1] = key;
      vout[2 * blockIdx.x + 1] = m;
    }
  }
}

// Every tile of the chunk: the rows of [s, e) equal to their window's max
// (want = 0: none), as the lane's 16 bits of the bitmap.
template <int NP>
__global__ __launch_bounds__(256) void k_wmax_argmax(ChunkView cv, uint64_t s, uint64_t e, uint64_t W, uint32_t want,
                                                     const unsigned long long* __restrict__ wmax,
                                                     uint16_t* __restrict__ out16, uint64_t nout16) {
  const uint32_t lane = __lane_id();
  {
    const uint64_t c = wave_global();   // one tile per wave
    if (c >= cv.ntiles) return;
    const uint64_t idx = c * 64 + lane;
    const uint64_t r0 = c * kTileRows, lr0 = r0 + 16ull * lane;
    uint32_t bits = 0;
    const bool tile_in = want && r0 < e && r0 + kTileRows > s;
    if (tile_in) {
      uint64_t off[16];
      lane_offsets<NP>(cv, c, lane, off);
      const bool lane_in = lr0 < e && lr0 + 16 > s;
      if (lane_in) {
        const uint64_t first = lr0 > s ? lr0 : s;
        uint64_t w = (first - s) / W;
        uint64_t next = s + (w + 1) * W;   // first row of the following window
        unsigned long long m = wmax[w];
#pragma unroll
        for (uint32_t b = 0; b < 16; ++b) {
          const uint64_t row = lr0 + b;
          if (row < s || row >= e) continue;
          if (row >= next) {
            w = (row - s) / W;
            next = s + (w + 1) * W;
            m = wmax[w];
          }
          bits |= (uint32_t)(off[b] == m) << b;
        }
      }
    }
    if (idx < nout16) out16[idx] = (uint16_t)bits;
  }
}

__global__ void k_wmax_values(const unsigned long long* __restrict__ wmax, uint64_t k, int64_t base, double inv_scale,
                              double* __restrict__ vals) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x)
    vals[i] = (double)(int64_t)((uint64_t)base + wmax[i]) * inv_scale;
}

// ---------------------------------------------------------------------------

WmaxLayout wmax_layout(uint64_t s, uint64_t e, uint64_t W) {
  WmaxLayout L;
  const uint64_t rows = e - s;
  L.nwin = rows ? (rows + W - 1) / W : 0;
  L.t_first = rows ? s / kTileRows : 0;
  L.ntiles = rows ? (e - 1) / kTileRows - L.t_first + 1 : 0;
  uint64_t at = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t here = at;
    at += (bytes + 255) & ~255ull;
    return here;
  };
  L.wmax = take(8 * (L.nwin ? L.nwin : 1));
  const uint64_t np = 2 * (L.ntiles ? L.ntiles : 1);
  L.k[0] = take(4 * np);
  L.v[0] = take(8 * np);
  const uint64_t np2 = 2 * ((np + 255) / 256);
  L.k[1] = take(4 * np2);
  L.v[1] = take(8 * np2);
  L.bytes = at;
  return L;
}

template <int NP>
static hipError_t launch_wmax_np(const ChunkView& cv, uint64_t s, uint64_t e, uint64_t W, const WmaxLayout& L,
                                 uint8_t* scratch, hipStream_t st) {
  unsigned long long* wmax = reinterpret_cast<unsigned long long*>(scratch + L.wmax);
  uint32_t* kb[2] = {reinterpret_cast<uint32_t*>(scratch + L.k[0]), reinterpret_cast<uint32_t*>(scratch + L.k[1])};
  unsigned long long* vb[2] = {reinterpret_cast<unsigned long long*>(scratch + L.v[0]),
                               reinterpret_cast<unsigned long long*>(scratch + L.v[1])};
  if (W >= kTileRows)
    launch_k(k_wmax_tiles_wide<NP>, dim3((uint32_t)((L.ntiles + 3) / 4)), dim3(256), 0, st, cv, s, e, W, L.t_first,
             L.ntiles, wmax, kb[0], vb[0]);
  else
    launch_k(k_wmax_tiles<NP>, dim3((uint32_t)L.ntiles), dim3(64), 0, st, cv, s, e, W, L.t_first, wmax, kb[0], vb[0]);
  uint64_t nin = 2 * L.ntiles;
  uint32_t cur = 0;
  while (true) {
    const uint64_t nb = (nin + 255) / 256;
    const uint32_t last = nb == 1 ? 1u : 0u;
    launch_k(k_wmax_reduce, dim3((uint32_t)nb), dim3(256), 0, st, static_cast<const uint32_t*>(kb[cur]),
             static_cast<const unsigned long long*>(vb[cur]), nin, kb[cur ^ 1], vb[cur ^ 1], wmax, last);
    if (last) break;
    nin = 2 * nb;
    cur ^= 1;
  }
  return hipGetLastError();
}

hipError_t launch_wmax(const ChunkView& cv, uint64_t s, uint64_t e, uint64_t W, const WmaxLayout& L,
                       uint8_t* scratch, hipStream_t st) {
  if (L.ntiles == 0) return hipSuccess;
  switch (cv.P) {
    case 0: return launch_wmax_np<0>(cv, s, e, W, L, scratch, st);
    case 1: return launch_wmax_np<1>(cv, s, e, W, L, scratch, st);
    case 2: return launch_wmax_np<2>(cv, s, e, W, L, scratch, st);
    case 3: return launch_wmax_np<3>(cv, s, e, W, L, scratch, st);
    case 4: return launch_wmax_np<4>(cv, s, e, W, L, scratch, st);
    case 5: return launch_wmax_np<5>(cv, s, e, W, L, scratch, st);
    case 6: return launch_wmax_np<6>(cv, s, e, W, L, scratch, st);
    case 7: return launch_wmax_np<7>(cv, s, e, W, L, scratch, st);
    default: return launch_wmax_np<8>(cv, s, e, W, L, scratch, st);
  }
}

hipError_t launch_wmax_argmax(const ChunkView& cv, uint64_t s, uint64_t e, uint64_t W, bool want,
                              const uint8_t* scratch, const WmaxLayout& L, uint64_t* out, hipStream_t st) {
  if (cv.ntiles == 0) return hipSuccess;
  const unsigned long long* wmax = reinterpret_cast<const unsigned long long*>(scratch + L.wmax);
  uint16_t* o16 = reinterpret_cast<uint16_t*>(out);
  const uint64_t nout16 = 4 * ((cv.n + 63) / 64);
  const uint32_t wg = (uint32_t)((cv.ntiles + 3) / 4);
  const uint32_t w = want && L.nwin ? 1u : 0u;
#define BUFF_ARG(NP) launch_k(k_wmax_argmax<NP>, dim3(wg), dim3(256), 0, st, cv, s, e, W, w, wmax, o16, nout16)
  switch (cv.P) {
    case 0: BUFF_ARG(0); break;
    case 1: BUFF_ARG(1); break;
    case 2: BUFF_ARG(2); break;
    case 3: BUFF_ARG(3); break;
    case 4: BUFF_ARG(4); break;
    case 5: BUFF_ARG(5); break;
    case 6: BUFF_ARG(6); break;
    case 7: BUFF_ARG(7); break;
    default: BUFF_ARG(8); break;
  }
#undef BUFF_ARG
  return hipGetLastError();
}

hipError_t launch_wmax_values(const uint8_t* scratch, const WmaxLayout& L, uint64_t k, int64_t base, double inv_scale,
                              double* vals, hipStream_t st) {
  if (k == 0) return hipSuccess;
  const unsigned long long* wmax = reinterpret_cast<const unsigned long long*>(scratch + L.wmax);
  const uint32_t wg = (uint32_t)std::min<uint64_t>((k + 255) / 256, 1024);
  launch_k(k_wmax_values, dim3(wg), dim3(256), 0, st, wmax, k, base, inv_scale, vals);
  return hipGetLastError();
}

}  // namespace buffgpu
