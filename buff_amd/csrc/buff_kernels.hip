// HIP kernels of the BUFF hot path for gfx950 (wave64). Layout: buff_common.h.
//
// Every kernel is HBM-bound integer/byte work; nothing here is a contraction, so
// no MFMA. The designs share three ideas:
//   * one wave owns one 1024-row tile per step; a lane reads its 16 bytes of a
//     plane with one dwordx4 load (the whole wave: 1 KiB contiguous),
//   * per-row predicates become bitmap words with __ballot (v_cmp_*_sdwa picks
//     the byte straight out of the loaded dword), and
//   * uniform 64-bit masks leave the scalar registers through v_writelane into
//     vector registers and are written with vector stores only.
#include <hip/hip_runtime.h>

#include "buff_common.h"
#include "buff_kernels.h"

namespace buffgpu {

// ----------------------------------------------------------------------------
// wave helpers

__device__ __forceinline__ uint32_t writelane(uint32_t old, uint32_t val, uint32_t lane) {
  // GFX9 reads at most one SGPR per VALU op: the lane select goes through M0.
  // `val` is wave-uniform; readfirstlane pins it to an SGPR for the asm operand.
  const uint32_t sv = __builtin_amdgcn_readfirstlane(val);
  asm volatile("v_writelane_b32 %0, %1, m0" : "+v"(old) : "s"(sv), "{m0}"(lane));
  return old;
}

__device__ __forceinline__ uint32_t byte_of(const uint4& w, uint32_t g) {
  const uint32_t d = g < 4 ? w.x : g < 8 ? w.y : g < 12 ? w.z : w.w;
  return (d >> (8 * (g & 3))) & 0xFFu;
}

__device__ __forceinline__ int64_t wave_min_i64(int64_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = min(v, (int64_t)__shfl_xor((long long)v, m, 64));
  return v;
}
__device__ __forceinline__ int64_t wave_max_i64(int64_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = max(v, (int64_t)__shfl_xor((long long)v, m, 64));
  return v;
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    uint64_t o = (uint64_t)__shfl_xor((unsigned long long)v, m, 64);
    v = o > v ? o : v;
  }
  return v;
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += (uint64_t)__shfl_xor((unsigned long long)v, m, 64);
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, m, 64));
  return v;
}
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Stores 16 uniform u64 words (128 bytes) at dst[0..15] through lanes 0..31,
// keeping only words with index < limit.
__device__ __forceinline__ void store_words16(uint64_t* dst, const uint64_t (&w)[16], uint32_t limit) {
  const uint32_t lane = __lane_id();
  uint32_t v = 0;
#pragma unroll
  for (uint32_t g = 0; g < 16; ++g) {
    v = writelane(v, (uint32_t)w[g], 2 * g);
    v = writelane(v, (uint32_t)(w[g] >> 32), 2 * g + 1);
  }
  if (lane < 32 && (lane >> 1) < limit) reinterpret_cast<uint32_t*>(dst)[lane] = v;
}

// Turns per-lane offsets of one tile into the tile's 16*r tail bit-plane words
// (word g*r + j = bit j of rows 64g + lane) and stores them at dst.
__device__ __forceinline__ void store_tail_bits(uint64_t* dst, const uint64_t (&off)[16], uint32_t r) {
  const uint32_t lane = __lane_id();
  uint32_t v0 = 0, v1 = 0, v2 = 0, v3 = 0;   // lane L holds dwords 4L..4L+3
#pragma unroll
  for (uint32_t g = 0; g < 16; ++g) {
    for (uint32_t j = 0; j < r; ++j) {
      uint64_t m = __ballot((off[g] >> j) & 1ull);
      uint32_t d = 2 * (g * r + j);          // dword index of the low half
      uint32_t lo = (uint32_t)m, hi = (uint32_t)(m >> 32);
      // d is even: the pair lands in components (0,1) or (2,3) of lane d/4
      if ((d & 3) == 0) { v0 = writelane(v0, lo, d >> 2); v1 = writelane(v1, hi, d >> 2); }
      else { v2 = writelane(v2, lo, d >> 2); v3 = writelane(v3, hi, d >> 2); }
    }
  }
  if (lane < 8 * r) reinterpret_cast<uint4*>(dst)[lane] = make_uint4(v0, v1, v2, v3);
}

// ----------------------------------------------------------------------------
// K1: fixed-point conversion + min / max reduction (encode pass 1).
// split_double.rs:321-330. Workgroup w owns the contiguous tile range
// [w*tpw, (w+1)*tpw), so its partial min is a segment min the prefix pass reuses.

__global__ __launch_bounds__(256) void k_minmax(const double* __restrict__ x, uint64_t n, uint64_t ntiles,
                                                uint64_t tpw, int32_t pexp, uint32_t dlen,
                                                int64_t* __restrict__ partial, MinMaxOut* __restrict__ out) {
  const uint32_t lane = __lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  const uint64_t t0 = (uint64_t)blockIdx.x * tpw;
  const uint64_t t1 = min(t0 + tpw, ntiles);
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  for (uint64_t c = t0 + wave; c < t1; c += 4) {
    const double* xt = x + c * kTileRows;
    const bool full = (c + 1) * kTileRows <= n;
    uint64_t v[16];
#pragma unroll
    for (uint32_t g = 0; g < 16; ++g) {
      const uint64_t row = c * kTileRows + 64 * g + lane;
      v[g] = (full || row < n) ? __double_as_longlong(xt[64 * g + lane]) : 0;
    }
#pragma unroll
    for (uint32_t g = 0; g < 16; ++g) {
      const uint64_t row = c * kTileRows + 64 * g + lane;
      const int64_t f = fetch_fixed_aligned_bits(v[g], pexp, dlen);
      if (full || row < n) {
        mn = min(mn, f);
        if (row != 0) mx = max(mx, f);
      }
    }
  }
  mn = wave_min_i64(mn);
  mx = wave_max_i64(mx);
  __shared__ int64_t smn[4], smx[4];
  if (lane == 0) { smn[wave] = mn; smx[wave] = mx; }
  __syncthreads();
  __shared__ bool last;
  if (threadIdx.x == 0) {
    mn = min(min(smn[0], smn[1]), min(smn[2], smn[3]));
    mx = max(max(smx[0], smx[1]), max(smx[2], smx[3]));
    partial[blockIdx.x] = mn;
    partial[gridDim.x + blockIdx.x] = mx;
    __threadfence();
    last = atomicAdd(&out->done, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  // Last workgroup folds the partials (threadfence reduction).
  __threadfence();
  int64_t a = INT64_MAX, b = INT64_MIN;
  for (uint32_t i = threadIdx.x; i < gridDim.x; i += blockDim.x) {
    a = min(a, __atomic_load_n(&partial[i], __ATOMIC_RELAXED));
    b = max(b, __atomic_load_n(&partial[gridDim.x + i], __ATOMIC_RELAXED));
  }
  a = wave_min_i64(a);
  b = wave_max_i64(b);
  if (lane == 0) { smn[wave] = a; smx[wave] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    out->min_all = min(min(smn[0], smn[1]), min(smn[2], smn[3]));
    out->max_rest = max(max(smx[0], smx[1]), max(smx[2], smx[3]));
    out->first = n ? fetch_fixed_aligned_bits(__double_as_longlong(x[0]), pexp, dlen) : 0;
    out->done = 0;
  }
}

// K1b: exact max of the reference's `if fixed<min {..} else if fixed>max {..}`
// loop (split_double.rs:321-330) for the rare case where row 0 holds the unique
// maximum: max over rows i with fixed_i >= min(fixed_0..fixed_{i-1})
// (min of the empty prefix = i64::MAX). One wave per workgroup walks its
// segment in row order with a wave prefix-min scan.
__global__ __launch_bounds__(64) void k_minmax_prefix(const double* __restrict__ x, uint64_t n, uint64_t ntiles,
                                                      uint64_t tpw, int32_t pexp, uint32_t dlen,
                                                      const int64_t* __restrict__ partial,
                                                      int64_t* __restrict__ partial2, MinMaxOut* __restrict__ out) {
  const uint32_t lane = __lane_id();
  int64_t carry = INT64_MAX;
  for (uint32_t i = lane; i < blockIdx.x; i += 64) carry = min(carry, partial[i]);
  carry = wave_min_i64(carry);
  const uint64_t t0 = (uint64_t)blockIdx.x * tpw;
  const uint64_t t1 = min(t0 + tpw, ntiles);
  int64_t qmax = INT64_MIN;
  for (uint64_t c = t0; c < t1; ++c) {
    for (uint32_t g = 0; g < 16; ++g) {
      const uint64_t row = c * kTileRows + 64 * g + lane;
      const bool valid = row < n;
      int64_t f = valid ? fetch_fixed_aligned_bits(__double_as_longlong(x[row]), pexp, dlen) : INT64_MAX;
      int64_t s = f;   // inclusive prefix min over lanes
#pragma unroll
      for (uint32_t o = 1; o < 64; o <<= 1) {
        int64_t u = (int64_t)__shfl_up((long long)s, o, 64);
        if (lane >= o) s = min(s, u);
      }
      int64_t e = (int64_t)__shfl_up((long long)s, 1, 64);
      e = lane == 0 ? carry : min(carry, e);
      if (valid && f >= e) qmax = max(qmax, f);
      carry = min(carry, (int64_t)__shfl((long long)s, 63, 64));
    }
  }
  qmax = wave_max_i64(qmax);
  __shared__ bool last;
  if (lane == 0) {
    partial2[blockIdx.x] = qmax;
    __threadfence();
    last = atomicAdd(&out->done, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  int64_t b = INT64_MIN;
  for (uint32_t i = lane; i < gridDim.x; i += 64) b = max(b, __atomic_load_n(&partial2[i], __ATOMIC_RELAXED));
  b = wave_max_i64(b);
  if (lane == 0) {
    out->qmax = b;
    out->done = 0;
  }
}

// ----------------------------------------------------------------------------
// K2: encode pass 2 — offsets, byte planes and tail bit-planes.
// split_double.rs:363-414: plane p byte = (off >> (F-8(p+1))) as u8, tail = low r bits.

__global__ __launch_bounds__(256) void k_encode(const double* __restrict__ x, uint64_t n, uint64_t ntiles,
                                                int32_t pexp, uint32_t dlen, int64_t base, uint32_t F, uint32_t P,
                                                uint32_t r, uint8_t* __restrict__ planes, uint64_t pitch,
                                                uint64_t* __restrict__ tail) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t c = wave; c < ntiles; c += nwaves) {
    const double* xt = x + c * kTileRows;
    const bool full = (c + 1) * kTileRows <= n;
    uint64_t off[16];
#pragma unroll
    for (uint32_t g = 0; g < 16; ++g) {
      const uint64_t row = c * kTileRows + 64 * g + lane;
      off[g] = (full || row < n) ? __double_as_longlong(xt[64 * g + lane]) : 0;
    }
#pragma unroll
    for (uint32_t g = 0; g < 16; ++g) {
      const uint64_t row = c * kTileRows + 64 * g + lane;
      const int64_t f = fetch_fixed_aligned_bits(off[g], pexp, dlen);
      off[g] = (full || row < n) ? (uint64_t)f - (uint64_t)base : 0;
    }
    for (uint32_t p = 0; p < P; ++p) {
      const uint32_t sh = F - 8 * (p + 1);
      uint32_t d[4];
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) {
        d[q] = (uint32_t)((off[4 * q] >> sh) & 0xFF) | ((uint32_t)((off[4 * q + 1] >> sh) & 0xFF) << 8) |
               ((uint32_t)((off[4 * q + 2] >> sh) & 0xFF) << 16) | ((uint32_t)((off[4 * q + 3] >> sh) & 0xFF) << 24);
      }
      reinterpret_cast<uint4*>(planes + p * pitch + c * kTileRows)[lane] = make_uint4(d[0], d[1], d[2], d[3]);
    }
    if (r) store_tail_bits(tail + c * 16 * r, off, r);
  }
}

// ----------------------------------------------------------------------------
// K3/K4: range (x > pred) and equality filters with early termination.
// buff_simd.rs:4507-4679 / 4925-5093. Plane 0 is read for every tile; plane k
// (and then the tail) only for tiles that still hold candidate rows, i.e. rows
// whose higher bytes all equal the target's. Bitmap words come from ballots.

template <int MODE>
__global__ __launch_bounds__(256) void k_filter(ChunkView cv, FilterTarget tg, uint64_t* __restrict__ out,
                                                uint64_t nwords, unsigned long long* __restrict__ stats) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t c = wave; c < cv.ntiles; c += nwaves) {
    uint64_t gt[16], eq[16];
    uint64_t any = 0;
    if (cv.P > 0) {
      const uint4 w = reinterpret_cast<const uint4*>(cv.planes + c * kTileRows)[lane];
      const uint32_t d0 = tg.digit[0];
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) {
        const uint32_t b = byte_of(w, g);
        if (MODE == 0) gt[g] = __ballot(b > d0);
        eq[g] = __ballot(b == d0) & group_valid_mask(cv.n, c, g);
        any |= eq[g];
      }
    } else {
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) {
        gt[g] = 0;
        eq[g] = group_valid_mask(cv.n, c, g);
        any |= eq[g];
      }
    }
    for (uint32_t p = 1; p < cv.P && any; ++p) {
      if (stats) {
        uint64_t k = 0;
#pragma unroll
        for (uint32_t g = 0; g < 16; ++g) k += __popcll(eq[g]);
        if (lane == 0) atomicAdd(&stats[p], (unsigned long long)k);
      }
      const uint4 w = reinterpret_cast<const uint4*>(cv.planes + p * cv.pitch + c * kTileRows)[lane];
      const uint32_t dp = tg.digit[p];
      any = 0;
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) {
        const uint32_t b = byte_of(w, g);
        if (MODE == 0) gt[g] |= eq[g] & __ballot(b > dp);
        eq[g] &= __ballot(b == dp);
        any |= eq[g];
      }
    }
    if (cv.r && any) {
      const uint64_t* tw = cv.tail + c * 16 * cv.r;
      if (stats) {
        uint64_t k = 0;
#pragma unroll
        for (uint32_t g = 0; g < 16; ++g) k += __popcll(eq[g]);
        if (lane == 0) atomicAdd(&stats[cv.P], (unsigned long long)k);
      }
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) {
        if (!eq[g]) continue;
        for (int j = (int)cv.r - 1; j >= 0; --j) {
          const uint64_t bits = tw[g * cv.r + j];
          if ((tg.tail >> j) & 1) {
            eq[g] &= bits;
          } else {
            if (MODE == 0) gt[g] |= eq[g] & bits;
            eq[g] &= ~bits;
          }
        }
      }
    }
    uint64_t res[16];
#pragma unroll
    for (uint32_t g = 0; g < 16; ++g) res[g] = MODE == 0 ? (gt[g] & group_valid_mask(cv.n, c, g)) : eq[g];
    const uint64_t w0 = c * 16;
    store_words16(out + w0, res, nwords > w0 ? (uint32_t)min<uint64_t>(16, nwords - w0) : 0u);
  }
}

// Bitmap of all rows (x > pred when pred lies below the column base) or none.
__global__ void k_fill_bitmap(uint64_t* __restrict__ out, uint64_t nwords, uint64_t n, int ones) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t v = 0;
    if (ones) {
      const uint64_t r0 = i * 64;
      v = r0 + 64 <= n ? ~0ull : (r0 < n ? (1ull << (n - r0)) - 1 : 0);
    }
    out[i] = v;
  }
}

// ----------------------------------------------------------------------------
// K5: per-sub-column byte sums (split_double.rs:2158-2204), exact in u64.
// Bytes are summed with v_sad_u8; the tail from bit-plane popcounts.

__global__ __launch_bounds__(256) void k_sum(ChunkView cv, unsigned long long* __restrict__ sums) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  uint32_t acc[kMaxPlanes] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t tacc = 0;
  const uint32_t twords = 16 * cv.r;
  uint64_t wide[kMaxPlanes] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t iter = 0;
  for (uint64_t c = wave; c < cv.ntiles; c += nwaves) {
#pragma unroll
    for (uint32_t p = 0; p < kMaxPlanes; ++p) {
      if (p < cv.P) {
        const uint4 w = reinterpret_cast<const uint4*>(cv.planes + p * cv.pitch + c * kTileRows)[lane];
        uint32_t a = __builtin_amdgcn_sad_u8(w.x, 0u, acc[p]);
        a = __builtin_amdgcn_sad_u8(w.y, 0u, a);
        a = __builtin_amdgcn_sad_u8(w.z, 0u, a);
        acc[p] = __builtin_amdgcn_sad_u8(w.w, 0u, a);
      }
    }
    for (uint32_t i = lane; i < twords; i += 64) tacc += (uint64_t)__popcll(cv.tail[c * twords + i]) << (i % cv.r);
    // a lane adds at most 16*255 per tile: fold into u64 well before u32 overflow
    if (++iter == (1u << 16)) {
#pragma unroll
      for (uint32_t p = 0; p < kMaxPlanes; ++p) { wide[p] += acc[p]; acc[p] = 0; }
      iter = 0;
    }
  }
#pragma unroll
  for (uint32_t p = 0; p < kMaxPlanes; ++p) {
    if (p < cv.P) {
      const uint64_t s = wave_sum_u64(wide[p] + acc[p]);
      if (lane == 0 && s) atomicAdd(&sums[p], (unsigned long long)s);
    }
  }
  if (cv.r) {
    const uint64_t s = wave_sum_u64(tacc);
    if (lane == 0 && s) atomicAdd(&sums[cv.P], (unsigned long long)s);
  }
}

// ----------------------------------------------------------------------------
// K6: max stored offset (split_double.rs:3919-4105). The reference narrows
// candidates byte by byte; its result is the largest stored offset. Each wave
// keeps its running best and only reconstructs whole offsets of a tile whose
// top byte can reach it, so most tiles cost one plane read.

__device__ __forceinline__ uint64_t tail_value(const uint64_t* tw, uint32_t r, uint32_t g, uint32_t lane) {
  uint64_t v = 0;
  for (uint32_t j = 0; j < r; ++j) v |= ((tw[g * r + j] >> lane) & 1ull) << j;
  return v;
}

__global__ __launch_bounds__(256) void k_max(ChunkView cv, unsigned long long* __restrict__ best_out) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  uint64_t best = 0;
  bool have = false;
  const uint32_t top_shift = cv.F >= 8 ? cv.F - 8 : 0;
  for (uint64_t c = wave; c < cv.ntiles; c += nwaves) {
    uint4 w0 = make_uint4(0, 0, 0, 0);
    if (cv.P > 0) {
      w0 = reinterpret_cast<const uint4*>(cv.planes + c * kTileRows)[lane];
      uint32_t mb = 0;
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) mb = max(mb, byte_of(w0, g));
      mb = wave_max_u32(mb);
      if (have && (uint64_t)mb < (best >> top_shift)) continue;
    }
    uint64_t off[16];
#pragma unroll
    for (uint32_t g = 0; g < 16; ++g) off[g] = cv.P > 0 ? (uint64_t)byte_of(w0, g) : 0;
    for (uint32_t p = 1; p < cv.P; ++p) {
      const uint4 w = reinterpret_cast<const uint4*>(cv.planes + p * cv.pitch + c * kTileRows)[lane];
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) off[g] = (off[g] << 8) | byte_of(w, g);
    }
    uint64_t m = 0;
    const uint64_t* tw = cv.tail + c * 16 * cv.r;
#pragma unroll
    for (uint32_t g = 0; g < 16; ++g) {
      uint64_t v = off[g];
      if (cv.r) v = (v << cv.r) | tail_value(tw, cv.r, g, lane);
      m = v > m ? v : m;   // rows >= n are zero and never exceed a real row
    }
    m = uniform_u64(wave_max_u64(m));
    best = m > best ? m : best;
    have = true;
  }
  if (have && lane == 0) atomicMax(best_out, (unsigned long long)best);
}

// ----------------------------------------------------------------------------
// K7: materialization. Full (split_double.rs:1243-1350), semantic reduced
// precision (top R bits kept, re-scaled), and the reference's reduced-precision
// decode (split_double.rs:1612-1893), which places the top R bits unshifted and,
// when R % 8 != 0, reads (R % 8)-bit fields at stride R % 8 from the packed tail.

// Bit `pos` of the packed reference tail (stream bit i*r + j <-> row i, bit j).
__device__ __forceinline__ uint32_t packed_tail_bit(const ChunkView& cv, uint64_t pos) {
  const uint64_t row = pos / cv.r;
  const uint32_t j = (uint32_t)(pos % cv.r);
  if (row >= cv.n) return 0;
  const uint64_t c = row / kTileRows;
  const uint32_t g = (uint32_t)((row % kTileRows) / 64), t = (uint32_t)(row % 64);
  return (uint32_t)((cv.tail[c * 16 * cv.r + g * cv.r + j] >> t) & 1ull);
}

__global__ __launch_bounds__(256) void k_materialize(ChunkView cv, MatParams mp, double* __restrict__ out) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t c = wave; c < cv.ntiles; c += nwaves) {
    uint64_t off[16];
#pragma unroll
    for (uint32_t g = 0; g < 16; ++g) off[g] = 0;
    // whole planes read (top `mp.planes` bytes)
    for (uint32_t p = 0; p < mp.planes; ++p) {
      const uint4 w = reinterpret_cast<const uint4*>(cv.planes + p * cv.pitch + c * kTileRows)[lane];
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) off[g] = (off[g] << 8) | byte_of(w, g);
    }
    if (mp.mode == kMatFull || mp.mode == kMatSemantic) {
      if (mp.tail_bits) {
        const uint64_t* tw = cv.tail + c * 16 * cv.r;
#pragma unroll
        for (uint32_t g = 0; g < 16; ++g) {
          // top mp.tail_bits of the r tail bits
          uint64_t v = 0;
          for (uint32_t k = 0; k < mp.tail_bits; ++k)
            v = (v << 1) | ((tw[g * cv.r + (cv.r - 1 - k)] >> lane) & 1ull);
          off[g] = (off[g] << mp.tail_bits) | v;
        }
      }
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) off[g] = (off[g] >> mp.drop) << mp.shift_back;
    } else {  // kMatRefPrecision: field of s bits at stride s in the packed tail
      if (mp.tail_bits) {
#pragma unroll
        for (uint32_t g = 0; g < 16; ++g) {
          const uint64_t row = c * kTileRows + 64 * g + lane;
          uint64_t v = 0;
          if (row < cv.n) {
            for (uint32_t q = 0; q < mp.tail_bits; ++q)
              v |= (uint64_t)packed_tail_bit(cv, row * mp.tail_bits + q) << q;
          }
          off[g] = (off[g] << mp.tail_bits) | v;
        }
      }
    }
    double* ot = out + c * kTileRows;
    const bool full = (c + 1) * kTileRows <= cv.n;
#pragma unroll
    for (uint32_t g = 0; g < 16; ++g) {
      const uint64_t row = c * kTileRows + 64 * g + lane;
      const double v = (double)(int64_t)((uint64_t)cv.base + off[g]) * mp.inv_scale;
      if (full || row < cv.n) ot[64 * g + lane] = v;
    }
  }
}

// ----------------------------------------------------------------------------
// Stream <-> resident conversion (split_double.rs:349-414 byte layout).

// body = stream + 20. Plane p at body + p*n, tail at body + P*n.
__global__ __launch_bounds__(256) void k_serialize(ChunkView cv, uint8_t flipmask, uint8_t* __restrict__ body,
                                                   uint64_t tail_bytes) {
  const uint32_t lane = __lane_id();
  const uint32_t wslot = threadIdx.x >> 6;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  __shared__ uint64_t tw_s[4][16 * 8];
  for (uint64_t c = wave; c < cv.ntiles; c += nwaves) {
    for (uint32_t p = 0; p < cv.P; ++p) {
      const uint4 w = reinterpret_cast<const uint4*>(cv.planes + p * cv.pitch + c * kTileRows)[lane];
      uint8_t* dst = body + p * cv.n;
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) {
        const uint64_t row = c * kTileRows + 64 * g + lane;
        if (row < cv.n) dst[row] = (uint8_t)byte_of(w, g) ^ flipmask;
      }
    }
    if (cv.r) {
      const uint32_t nw = 16 * cv.r;
      for (uint32_t i = lane; i < nw; i += 64) tw_s[wslot][i] = cv.tail[c * nw + i];
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      uint8_t* tdst = body + cv.P * cv.n;
      const uint32_t tile_bytes = 128 * cv.r;   // 1024 rows * r bits
      for (uint32_t m = lane; m < tile_bytes; m += 64) {
        const uint64_t gb = c * tile_bytes + m;
        if (gb >= tail_bytes) break;
        uint32_t byte = 0;
        for (uint32_t q = 0; q < 8; ++q) {
          const uint32_t bit = 8 * m + q;
          const uint32_t row = bit / cv.r, j = bit % cv.r;
          const uint32_t g = row / 64, t = row % 64;
          byte |= (uint32_t)((tw_s[wslot][g * cv.r + j] >> t) & 1ull) << q;
        }
        tdst[gb] = (uint8_t)byte;
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
  }
}

__global__ __launch_bounds__(256) void k_load(ChunkView cv, uint8_t flipmask, const uint8_t* __restrict__ body,
                                              uint64_t tail_bytes) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t c = wave; c < cv.ntiles; c += nwaves) {
    for (uint32_t p = 0; p < cv.P; ++p) {
      const uint8_t* src = body + p * cv.n;
      uint32_t d[4] = {0, 0, 0, 0};
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) {
        const uint64_t row = c * kTileRows + 64 * g + lane;
        const uint32_t b = row < cv.n ? (uint32_t)(src[row] ^ flipmask) : 0u;
        d[g >> 2] |= b << (8 * (g & 3));
      }
      reinterpret_cast<uint4*>(cv.planes + p * cv.pitch + c * kTileRows)[lane] = make_uint4(d[0], d[1], d[2], d[3]);
    }
    if (cv.r) {
      const uint8_t* tsrc = body + cv.P * cv.n;
      uint64_t off[16];
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) {
        const uint64_t row = c * kTileRows + 64 * g + lane;
        uint64_t v = 0;
        if (row < cv.n) {
          const uint64_t bit = row * cv.r;
          const uint64_t b0 = bit >> 3;
          uint32_t two = tsrc[b0];
          if (b0 + 1 < tail_bytes) two |= (uint32_t)tsrc[b0 + 1] << 8;
          v = (two >> (bit & 7)) & ((1u << cv.r) - 1u);
        }
        off[g] = v;
      }
      store_tail_bits(cv.tail + c * 16 * cv.r, off, cv.r);
    }
  }
}

// ----------------------------------------------------------------------------
// Host-side launchers (declared in buff_kernels.h).

static inline uint32_t grid_for(uint64_t ntiles, uint32_t waves_per_wg, uint32_t max_wg) {
  uint64_t wgs = (ntiles + waves_per_wg - 1) / waves_per_wg;
  if (wgs < 1) wgs = 1;
  if (wgs > max_wg) wgs = max_wg;
  return (uint32_t)wgs;
}

hipError_t launch_minmax(const double* x, uint64_t n, int32_t pexp, uint32_t dlen, int64_t* partial,
                         int64_t* partial2, MinMaxOut* out, uint32_t nwg, bool prefix, hipStream_t s) {
  const uint64_t ntiles = (n + kTileRows - 1) / kTileRows;
  const uint64_t tpw = (ntiles + nwg - 1) / nwg;
  if (!prefix) {
    hipLaunchKernelGGL(k_minmax, dim3(nwg), dim3(256), 0, s, x, n, ntiles, tpw, pexp, dlen, partial, out);
  } else {
    hipLaunchKernelGGL(k_minmax_prefix, dim3(nwg), dim3(64), 0, s, x, n, ntiles, tpw, pexp, dlen,
                       (const int64_t*)partial, partial2, out);
  }
  return hipGetLastError();
}

hipError_t launch_encode(const double* x, const ChunkView& cv, int32_t pexp, hipStream_t s, const LaunchShape& ls) {
  const uint32_t wg = grid_for(cv.ntiles, ls.waves_per_wg, ls.max_wg);
  hipLaunchKernelGGL(k_encode, dim3(wg), dim3(64 * ls.waves_per_wg), 0, s, x, cv.n, cv.ntiles, pexp, cv.dlen,
                     cv.base, cv.F, cv.P, cv.r, cv.planes, cv.pitch, cv.tail);
  return hipGetLastError();
}

hipError_t launch_filter(const ChunkView& cv, int op, const FilterTarget& tg, uint64_t* out, uint64_t nwords,
                         unsigned long long* stats, hipStream_t s, const LaunchShape& ls) {
  const uint32_t wg = grid_for(cv.ntiles, ls.waves_per_wg, ls.max_wg);
  if (op == 0)
    hipLaunchKernelGGL(k_filter<0>, dim3(wg), dim3(64 * ls.waves_per_wg), 0, s, cv, tg, out, nwords, stats);
  else
    hipLaunchKernelGGL(k_filter<1>, dim3(wg), dim3(64 * ls.waves_per_wg), 0, s, cv, tg, out, nwords, stats);
  return hipGetLastError();
}

hipError_t launch_fill_bitmap(uint64_t* out, uint64_t nwords, uint64_t n, int ones, hipStream_t s) {
  uint32_t wg = (uint32_t)std::min<uint64_t>((nwords + 255) / 256, 4096);
  if (wg == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_bitmap, dim3(wg), dim3(256), 0, s, out, nwords, n, ones);
  return hipGetLastError();
}

hipError_t launch_sum(const ChunkView& cv, unsigned long long* sums, hipStream_t s, const LaunchShape& ls) {
  const uint32_t wg = grid_for(cv.ntiles, ls.waves_per_wg, ls.max_wg);
  hipLaunchKernelGGL(k_sum, dim3(wg), dim3(64 * ls.waves_per_wg), 0, s, cv, sums);
  return hipGetLastError();
}

hipError_t launch_max(const ChunkView& cv, unsigned long long* best, hipStream_t s, const LaunchShape& ls) {
  const uint32_t wg = grid_for(cv.ntiles, ls.waves_per_wg, ls.max_wg);
  hipLaunchKernelGGL(k_max, dim3(wg), dim3(64 * ls.waves_per_wg), 0, s, cv, best);
  return hipGetLastError();
}

hipError_t launch_materialize(const ChunkView& cv, const MatParams& mp, double* out, hipStream_t s,
                              const LaunchShape& ls) {
  const uint32_t wg = grid_for(cv.ntiles, ls.waves_per_wg, ls.max_wg);
  hipLaunchKernelGGL(k_materialize, dim3(wg), dim3(64 * ls.waves_per_wg), 0, s, cv, mp, out);
  return hipGetLastError();
}

hipError_t launch_serialize(const ChunkView& cv, uint8_t flipmask, uint8_t* body, uint64_t tail_bytes,
                            hipStream_t s) {
  const uint32_t wg = grid_for(cv.ntiles, 4, 2048);
  hipLaunchKernelGGL(k_serialize, dim3(wg), dim3(256), 0, s, cv, flipmask, body, tail_bytes);
  return hipGetLastError();
}

hipError_t launch_load(const ChunkView& cv, uint8_t flipmask, const uint8_t* body, uint64_t tail_bytes,
                       hipStream_t s) {
  const uint32_t wg = grid_for(cv.ntiles, 4, 2048);
  hipLaunchKernelGGL(k_load, dim3(wg), dim3(256), 0, s, cv, flipmask, body, tail_bytes);
  return hipGetLastError();
}

}  // namespace buffgpu
