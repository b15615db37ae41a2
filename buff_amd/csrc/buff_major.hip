// HIP kernels of the buff-major layout (buff_major.h) for gfx950 (wave64).
//
// A wave owns a 1024-row tile and lane t its rows 16t .. 16t+15, as in the plain
// kernels. A sparse column's bytes for a lane are the majority byte replicated,
// with the outlier bytes patched in: the lane's 16-bit slice of the outlier
// bitmap says which rows, the tile's prefix count plus an exclusive scan of the
// lanes' popcounts says where their bytes start in the compacted array. The
// dense remainder is read exactly like a plain chunk. Everything stays in the
// integer domain and is HBM/latency bound; there is no MFMA work here.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "buff_common.h"
#include "buff_device.h"
#include "buff_major.h"

namespace buffgpu {

__device__ __forceinline__ uint32_t wave_excl_scan_u32(uint32_t v, uint32_t lane) {
  uint32_t s = v;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t u = (uint32_t)__shfl_up((int)s, o, 64);
    if (lane >= o) s += u;
  }
  return s - v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += (uint32_t)__shfl_xor((int)v, m, 64);
  return v;
}

// Replaces byte b (row 16*lane + b: dword b%4, byte b/4) of a lane's 16 bytes.
__device__ __forceinline__ void set_byte(uint32_t& d0, uint32_t& d1, uint32_t& d2, uint32_t& d3, uint32_t b,
                                         uint32_t v) {
  const uint32_t sh = 8 * (b >> 2);
  const uint32_t keep = ~(0xFFu << sh), val = v << sh;
  const uint32_t i = b & 3;
  d0 = i == 0 ? (d0 & keep) | val : d0;
  d1 = i == 1 ? (d1 & keep) | val : d1;
  d2 = i == 2 ? (d2 & keep) | val : d2;
  d3 = i == 3 ? (d3 & keep) | val : d3;
}

// The lane's 16 bytes of a sparse column in tile c: the majority byte, with the
// outlier bytes of the rows in `want` (a subset of m16) patched in. Every lane of
// the wave calls it together (the exclusive scan).
__device__ __forceinline__ uint4 sparse_bytes(const SparseDev& sd, uint64_t c, uint32_t lane, uint32_t m16,
                                              uint32_t want) {
  const uint32_t excl = wave_excl_scan_u32(__popc(m16), lane);
  const uint32_t M = 0x01010101u * sd.major;
  uint32_t d0 = M, d1 = M, d2 = M, d3 = M;
  if (want) {
    const uint8_t* src = sd.outl + (uint64_t)sd.prefix[c] + excl;
    uint32_t bits = want;
    while (bits) {
      const uint32_t b = (uint32_t)__builtin_ctz(bits);
      bits &= bits - 1;
      set_byte(d0, d1, d2, d3, b, src[__popc(m16 & ((1u << b) - 1u))]);
    }
  }
  return make_uint4(d0, d1, d2, d3);
}

// max over the rows in `mask` of (byte + 1); 0 when the mask is empty
__device__ __forceinline__ uint32_t masked_max_plus1(const uint4& w, uint32_t mask) {
  uint32_t v = 0;
#pragma unroll
  for (uint32_t b = 0; b < 16; ++b) v = ((mask >> b) & 1u) ? max(v, byte_at(w, b) + 1u) : v;
  return v;
}

// ----------------------------------------------------------------------------
// Encode (buff_simd.rs:451-555). Pass 1: dense planes / tail of the low F - 8S
// bits, outlier bitmaps, per-tile outlier counts. Pass 2: outlier bytes at
// prefix[c] + lane offset. Both passes convert the float64 input the same way.

template <int PASS>
__global__ __launch_bounds__(256) void k_major_encode(const double* __restrict__ x, uint64_t n, uint64_t ntiles,
                                                      int32_t pexp, uint32_t dlen, int64_t base, ChunkView dense,
                                                      MajorEnc me) {
  __shared__ uint64_t s_off[4][64 * kLdsRow];
  const uint32_t lane = __lane_id();
  uint64_t* so = s_off[threadIdx.x >> 6];
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t c = wave; c < ntiles; c += nwaves) {
    uint64_t v[16];
    const bool full = tile_full(n, c);
    load_tile_f64(x, n, c, lane, full, v);
#pragma unroll
    for (uint32_t g = 0; g < 16; ++g) {
      const int64_t f = fetch_fixed_aligned_bits(v[g], pexp, dlen);
      so[lds_slot(64 * g + lane)] = (full || c * kTileRows + 64 * g + lane < n) ? (uint64_t)f - (uint64_t)base : 0;
    }
    wave_lds_sync();
    uint64_t off[16];
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) off[b] = so[kLdsRow * lane + b];
    wave_lds_sync();
    const uint32_t valid = lane_valid16(n, c, lane);
    if (PASS == 1) {
      for (uint32_t p = 0; p < dense.P; ++p) {
        const uint32_t sh = dense.F - 8 * (p + 1);
        uint32_t by[16];
#pragma unroll
        for (uint32_t b = 0; b < 16; ++b) by[b] = (uint32_t)((off[b] >> sh) & 0xFF);
        reinterpret_cast<uint4*>(dense.planes + p * dense.pitch + c * kTileRows)[lane] = pack_lane(by);
      }
      for (uint32_t j = 0; j < dense.r; ++j) {
        uint32_t h = 0;
#pragma unroll
        for (uint32_t b = 0; b < 16; ++b) h |= (uint32_t)((off[b] >> j) & 1ull) << b;
        dense.tail[(c * dense.r + j) * 64 + lane] = (uint16_t)h;
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < kMaxSparse; ++j) {
      if (j >= me.S) break;
      const uint32_t sh = me.shift[j], mj = me.major[j];
      uint32_t m = 0;
#pragma unroll
      for (uint32_t b = 0; b < 16; ++b) m |= (uint32_t)(((off[b] >> sh) & 0xFF) != mj) << b;
      m &= valid;
      if (PASS == 1) {
        me.mask16[j][c * 64 + lane] = (uint16_t)m;
        const uint32_t cnt = wave_sum_u32(__popc(m));
        if (lane == 0) me.counts[j][c] = cnt;
      } else {
        const uint32_t excl = wave_excl_scan_u32(__popc(m), lane);
        uint8_t* dst = me.outl[j] + (uint64_t)me.prefix[j][c] + excl;
        uint32_t k = 0;
#pragma unroll
        for (uint32_t b = 0; b < 16; ++b) {
          if ((m >> b) & 1u) {
            dst[k] = (uint8_t)(off[b] >> sh);
            ++k;
          }
        }
      }
    }
  }
}

// Per-tile outlier counts from the bitmaps (stream load path).
__global__ __launch_bounds__(256) void k_major_counts(MajorEnc me, uint64_t ntiles) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t c = wave; c < ntiles; c += nwaves) {
#pragma unroll
    for (uint32_t j = 0; j < kMaxSparse; ++j) {
      if (j >= me.S) break;
      const uint32_t cnt = wave_sum_u32(__popc((uint32_t)me.mask16[j][c * 64 + lane]));
      if (lane == 0) me.counts[j][c] = cnt;
    }
  }
}

// Exclusive scan of one column's tile counts (one workgroup per column), its
// total and the row of its last outlier.
__global__ __launch_bounds__(256) void k_major_scan(MajorEnc me, uint64_t ntiles,
                                                    unsigned long long* __restrict__ totals) {
  const uint32_t j = blockIdx.x;
  const uint32_t tid = threadIdx.x;
  const uint32_t* cnt = me.counts[j];
  uint32_t* pre = const_cast<uint32_t*>(me.prefix[j]);
  const uint64_t per = (ntiles + 255) / 256;
  const uint64_t t0 = min((uint64_t)tid * per, ntiles), t1 = min(t0 + per, ntiles);
  uint64_t local = 0;
  int64_t lastc = -1;
  for (uint64_t c = t0; c < t1; ++c) {
    const uint32_t k = cnt[c];
    local += k;
    if (k) lastc = (int64_t)c;
  }
  __shared__ uint64_t s_sum[256];
  __shared__ int64_t s_last[256];
  s_sum[tid] = local;
  s_last[tid] = lastc;
  __syncthreads();
  for (uint32_t o = 1; o < 256; o <<= 1) {
    const uint64_t a = tid >= o ? s_sum[tid - o] : 0;
    const int64_t l = tid >= o ? s_last[tid - o] : -1;
    __syncthreads();
    s_sum[tid] += a;
    s_last[tid] = max(s_last[tid], l);
    __syncthreads();
  }
  uint64_t run = s_sum[tid] - local;
  for (uint64_t c = t0; c < t1; ++c) {
    pre[c] = (uint32_t)run;
    run += cnt[c];
  }
  const uint64_t total = s_sum[255];
  const int64_t lc = s_last[255];
  if (tid == 0) pre[ntiles] = (uint32_t)total;
  if (tid < 64) {
    uint64_t row = 0;
    bool have = false;
    if (lc >= 0) {
      const uint32_t m = me.mask16[j][(uint64_t)lc * 64 + tid];
      if (m) {
        have = true;
        row = (uint64_t)lc * kTileRows + 16ull * tid + (31u - (uint32_t)__clz(m));
      }
    }
    const uint64_t best = wave_max_u64(have ? row + 1 : 0);
    if (tid == 0) {
      totals[2 * j] = total;
      totals[2 * j + 1] = best ? best - 1 : ~0ull;
    }
  }
}

// ----------------------------------------------------------------------------
// Filters: buff_range_filter_majority (buff_simd.rs:863-1277) and
// buff_equal_filter_majority (1916-2292). Per row, with cand the rows equal to
// the target so far and res the rows known above it:
//   column 0:  res = byte > t0, cand = byte == t0;
//   column j>0 (range): if t_j == 0 and major_j == 0: res |= cand & outlier, cand = {};
//              otherwise res |= outlier & byte > t_j (every outlier, not only
//              candidates), res |= cand & !outlier when major_j > t_j,
//              cand &= byte == t_j;
//   dense planes and tail as in the plain filter.
// SEM drops the two column>0 rules (plain digit-wise compare).

template <int MODE, int SEM>
__global__ __launch_bounds__(256) void k_major_filter(MajorView mv, MajorTarget tg, uint64_t* __restrict__ out,
                                                      uint64_t nwords) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const ChunkView& dv = mv.dense;
  uint16_t* out16 = reinterpret_cast<uint16_t*>(out);
  const uint64_t nout16 = 4 * nwords;
  for (uint64_t c = wave; c < dv.ntiles; c += nwaves) {
    const uint32_t valid = lane_valid16(dv.n, c, lane);
    uint32_t res = 0, cand = valid;
#pragma unroll
    for (uint32_t j = 0; j < kMaxSparse; ++j) {
      if (j >= mv.S) break;
      const SparseDev& sd = mv.col[j];
      const uint32_t m16 = sd.mask16[c * 64 + lane];
      const uint32_t want = (MODE == 0 && !SEM) ? m16 : (m16 & cand);
      const uint4 w = sparse_bytes(sd, c, lane, m16, want);
      const uint32_t t = tg.digit[j];
      const Digit dg = make_digit(t);
      if (MODE == 1) {
        cand &= mask_eq(w, dg);
        continue;
      }
      const uint32_t gt = mask_gt(w, dg);
      if (SEM || j == 0) {
        res |= cand & gt;
        cand &= mask_eq(w, dg);
      } else if (t == 0 && sd.major == 0) {
        res |= cand & m16;
        cand = 0;
      } else {
        res |= m16 & gt;
        if (sd.major > t) res |= cand & ~m16;
        cand &= mask_eq(w, dg);
      }
    }
    for (uint32_t p = 0; p < dv.P; ++p) {
      if (!__ballot(cand != 0)) break;
      const uint4 w = load_plane(dv, p, c, cand ? lane : 0u);
      const Digit dg = make_digit(tg.digit[mv.S + p]);
      if (MODE == 0) res |= cand & mask_gt(w, dg);
      cand &= mask_eq(w, dg);
    }
    if (dv.r && __ballot(cand != 0)) refine_tail<MODE>(dv, c, tg.tail, lane, res, cand);
    const uint32_t r = MODE == 0 ? (res & valid) : cand;
    const uint64_t idx = c * 64 + lane;
    if (idx < nout16) out16[idx] = (uint16_t)r;
  }
}

// ----------------------------------------------------------------------------
// Decode: buff_decode_majority (buff_simd.rs:633-861). Column 0 places its byte;
// a later sparse column ORs outlier bytes at its shift and the majority byte at
// twice its shift, and only for rows up to its last outlier (SEM: every row gets
// its byte once). Values are staged in LDS and written coalesced.

template <int SEM>
__global__ __launch_bounds__(256) void k_major_decode(MajorView mv, double inv_scale, double* __restrict__ out) {
  __shared__ double s_val[4][64 * kLdsRow];
  const uint32_t lane = __lane_id();
  double* sv = s_val[threadIdx.x >> 6];
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const ChunkView& dv = mv.dense;
  for (uint64_t c = wave; c < dv.ntiles; c += nwaves) {
    uint64_t fx[16];
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) fx[b] = 0;
    const uint64_t row0 = c * kTileRows + 16ull * lane;
#pragma unroll
    for (uint32_t j = 0; j < kMaxSparse; ++j) {
      if (j >= mv.S) break;
      const SparseDev& sd = mv.col[j];
      const uint32_t m16 = sd.mask16[c * 64 + lane];
      const uint4 w = sparse_bytes(sd, c, lane, m16, m16);
      const uint32_t sh = sd.shift;
      if (SEM || j == 0) {
#pragma unroll
        for (uint32_t b = 0; b < 16; ++b) fx[b] |= (uint64_t)byte_at(w, b) << sh;
      } else {
        const uint64_t twice = ((uint64_t)sd.major << sh) << sh;
#pragma unroll
        for (uint32_t b = 0; b < 16; ++b) {
          const uint64_t v = ((m16 >> b) & 1u) ? ((uint64_t)byte_at(w, b) << sh) : twice;
          fx[b] |= row0 + b <= sd.last ? v : 0ull;
        }
      }
    }
    for (uint32_t p = 0; p < dv.P; ++p) {
      const uint4 w = load_plane(dv, p, c, lane);
      const uint32_t sh = dv.F - 8 * (p + 1);
#pragma unroll
      for (uint32_t b = 0; b < 16; ++b) fx[b] |= (uint64_t)byte_at(w, b) << sh;
    }
    for (uint32_t j = 0; j < dv.r; ++j) {
      const uint32_t h = load_tail(dv, j, c, lane);
#pragma unroll
      for (uint32_t b = 0; b < 16; ++b) fx[b] |= (uint64_t)((h >> b) & 1u) << j;
    }
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) sv[kLdsRow * lane + b] = (double)(int64_t)((uint64_t)dv.base + fx[b]) * inv_scale;
    wave_lds_sync();
    double* ot = out + c * kTileRows + lane;
    if (tile_full(dv.n, c)) {
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) ot[64 * g] = sv[lds_slot(64 * g + lane)];
    } else {
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g)
        if (c * kTileRows + 64 * g + lane < dv.n) ot[64 * g] = sv[lds_slot(64 * g + lane)];
    }
    wave_lds_sync();
  }
}

// ----------------------------------------------------------------------------
// Sum of the outlier bytes of every sparse column (buff_simd.rs:2343-2351 adds
// them to major * (n - outliers)). Grid (blocks, S); one partial per block.

struct OutlierCounts {
  uint64_t n[kMaxSparse];
};

__global__ __launch_bounds__(256) void k_major_outsum(MajorView mv, OutlierCounts oc,
                                                      unsigned long long* __restrict__ rows) {
  const uint32_t j = blockIdx.y;
  const uint8_t* p = mv.col[j].outl;
  const uint64_t nb = oc.n[j];
  const uint64_t n16 = nb / 16;
  uint64_t acc = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint4 w = reinterpret_cast<const uint4*>(p)[i];
    uint32_t a = __builtin_amdgcn_sad_u8(w.x, 0u, 0u);
    a = __builtin_amdgcn_sad_u8(w.y, 0u, a);
    a = __builtin_amdgcn_sad_u8(w.z, 0u, a);
    acc += __builtin_amdgcn_sad_u8(w.w, 0u, a);
  }
  if (blockIdx.x == 0 && threadIdx.x < 16 && n16 * 16 + threadIdx.x < nb) acc += p[n16 * 16 + threadIdx.x];
  acc = wave_sum_u64(acc);
  __shared__ uint64_t part[4];
  if (__lane_id() == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) rows[(uint64_t)j * gridDim.x + blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

__global__ __launch_bounds__(64) void k_major_outsum_final(const unsigned long long* __restrict__ rows,
                                                           uint32_t nblk, unsigned long long* __restrict__ sums) {
  const uint32_t j = blockIdx.x;
  uint64_t acc = 0;
  for (uint32_t i = __lane_id(); i < nblk; i += 64) acc += rows[(uint64_t)j * nblk + i];
  acc = wave_sum_u64(acc);
  if (__lane_id() == 0) sums[j] = acc;
}

// ----------------------------------------------------------------------------
// Max: buff_max_majority (buff_simd.rs:2444-2861), one stage per byte column,
// then the tail, then the count of rows tying the max. Stage k first narrows the
// candidates of stage k-1 to the rows whose byte k-1 equals the byte chosen
// there, then reduces (byte + 1) over the candidates of column k; the final
// kernel turns the reduction into the chosen byte:
//   column 0:    max(major, max outlier byte);
//   column j>0:  no candidate is an outlier -> major, not ORed into the max;
//                otherwise max(major, max candidate outlier byte);
//   dense / tail: max candidate byte (0 when there is none).
// SEM: a sparse column chooses the max byte over all candidates.

enum : uint32_t { kStSparse = 0, kStDense = 1, kStTail = 2, kStCount = 3 };

__device__ __forceinline__ uint32_t tail_value(const ChunkView& dv, uint64_t c, uint32_t lane, uint32_t b,
                                               const uint32_t (&h)[8]) {
  uint32_t v = 0;
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j) v |= j < dv.r ? ((h[j] >> b) & 1u) << j : 0u;
  return v;
}

template <int KIND, int SEM>
__global__ __launch_bounds__(256) void k_major_max_stage(MajorView mv, uint32_t k, const uint16_t* __restrict__ cand_in,
                                                         uint16_t* __restrict__ cand_out,
                                                         const MajorMaxState* __restrict__ st,
                                                         uint32_t* __restrict__ partial) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const ChunkView& dv = mv.dense;
  const uint32_t ncol = mv.S + dv.P;
  uint32_t acc = 0;
  for (uint64_t c = wave; c < dv.ntiles; c += nwaves) {
    uint32_t cand = cand_in ? (uint32_t)cand_in[c * 64 + lane] : lane_valid16(dv.n, c, lane);
    uint32_t h[8];
    const bool need_tail = (KIND == kStTail || KIND == kStCount) && __ballot(cand != 0);
    if (need_tail) {
#pragma unroll
      for (uint32_t j = 0; j < 8; ++j) h[j] = load_tail(dv, min(j, dv.r - 1), c, cand ? lane : 0u);
    }
    if (k > 0) {
      const uint32_t prev = k - 1;
      const uint32_t chosen = st->chosen[prev];
      if (prev < mv.S) {
        const SparseDev& sd = mv.col[prev];
        const uint32_t m16 = sd.mask16[c * 64 + lane];
        const uint4 w = sparse_bytes(sd, c, lane, m16, m16 & cand);
        cand &= mask_eq(w, make_digit(chosen));
      } else if (prev < ncol) {
        if (__ballot(cand != 0)) {
          const uint4 w = load_plane(dv, prev - mv.S, c, cand ? lane : 0u);
          cand &= mask_eq(w, make_digit(chosen));
        }
      } else if (need_tail) {   // the tail narrows the count stage
        uint32_t eq = 0;
#pragma unroll
        for (uint32_t b = 0; b < 16; ++b) eq |= (uint32_t)(tail_value(dv, c, lane, b, h) == chosen) << b;
        cand &= eq;
      }
      if (KIND != kStCount) cand_out[c * 64 + lane] = (uint16_t)cand;
    }
    if (KIND == kStSparse) {
      const SparseDev& sd = mv.col[k];
      const uint32_t m16 = sd.mask16[c * 64 + lane];
      const uint32_t co = cand & m16;
      const uint4 w = sparse_bytes(sd, c, lane, m16, co);
      acc = max(acc, masked_max_plus1(w, co));
      if (SEM && (cand & ~m16)) acc = max(acc, sd.major + 1u);
    } else if (KIND == kStDense) {
      if (__ballot(cand != 0)) {
        const uint4 w = load_plane(dv, k - mv.S, c, cand ? lane : 0u);
        acc = max(acc, masked_max_plus1(w, cand));
      }
    } else if (KIND == kStTail) {
      if (need_tail) {
#pragma unroll
        for (uint32_t b = 0; b < 16; ++b)
          acc = ((cand >> b) & 1u) ? max(acc, tail_value(dv, c, lane, b, h) + 1u) : acc;
      }
    } else {
      acc += __popc(cand);
    }
  }
  acc = KIND == kStCount ? wave_sum_u32(acc) : wave_max_u32(acc);
  __shared__ uint32_t part[4];
  if (lane == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    partial[blockIdx.x] = KIND == kStCount ? part[0] + part[1] + part[2] + part[3]
                                           : max(max(part[0], part[1]), max(part[2], part[3]));
}

__global__ __launch_bounds__(256) void k_major_max_final(const uint32_t* __restrict__ partial, uint32_t nwg,
                                                         uint32_t kind, uint32_t k, uint32_t major, uint32_t shift,
                                                         uint32_t semantic, MajorMaxState* __restrict__ st) {
  __shared__ uint32_t part[4];
  uint32_t v = 0;
  for (uint32_t i = threadIdx.x; i < nwg; i += blockDim.x) v = kind == kStCount ? v + partial[i] : max(v, partial[i]);
  v = kind == kStCount ? wave_sum_u32(v) : wave_max_u32(v);
  if (__lane_id() == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x != 0) return;
  v = kind == kStCount ? part[0] + part[1] + part[2] + part[3] : max(max(part[0], part[1]), max(part[2], part[3]));
  if (kind == kStCount) {
    st->argmax = v;
    return;
  }
  uint32_t chosen;
  bool merge = true;
  if (kind == kStSparse && !semantic) {
    if (v == 0) {
      chosen = major;
      merge = k == 0;
    } else {
      chosen = max(major, v - 1);
    }
  } else {
    chosen = v ? v - 1 : 0;
  }
  st->chosen[k] = chosen;
  if (merge) st->acc |= (unsigned long long)chosen << shift;
}

// ----------------------------------------------------------------------------
// launchers

static inline uint32_t tiles_grid(uint64_t ntiles, uint32_t max_wg) {
  uint64_t wgs = (ntiles + 3) / 4;
  if (wgs < 1) wgs = 1;
  if (wgs > max_wg) wgs = max_wg;
  return (uint32_t)wgs;
}

hipError_t launch_major_encode(int pass, const double* x, uint64_t n, int32_t pexp, uint32_t dlen, int64_t base,
                               const ChunkView& dense, const MajorEnc& me, hipStream_t s) {
  const uint64_t ntiles = (n + kTileRows - 1) / kTileRows;
  const uint32_t grid = tiles_grid(ntiles, 4096);
  if (pass == 1)
    hipLaunchKernelGGL(k_major_encode<1>, dim3(grid), dim3(256), 0, s, x, n, ntiles, pexp, dlen, base, dense, me);
  else
    hipLaunchKernelGGL(k_major_encode<2>, dim3(grid), dim3(256), 0, s, x, n, ntiles, pexp, dlen, base, dense, me);
  return hipGetLastError();
}

hipError_t launch_major_counts(const MajorEnc& me, uint64_t ntiles, hipStream_t s) {
  hipLaunchKernelGGL(k_major_counts, dim3(tiles_grid(ntiles, 4096)), dim3(256), 0, s, me, ntiles);
  return hipGetLastError();
}

hipError_t launch_major_scan(const MajorEnc& me, uint64_t ntiles, unsigned long long* totals, hipStream_t s) {
  hipLaunchKernelGGL(k_major_scan, dim3(me.S), dim3(256), 0, s, me, ntiles, totals);
  return hipGetLastError();
}

hipError_t launch_major_filter(const MajorView& mv, int op, bool semantic, const MajorTarget& tg, uint64_t* out,
                               uint64_t nwords, hipStream_t s, const LaunchShape& ls) {
  const uint32_t grid = tiles_grid(mv.dense.ntiles, ls.max_wg);
  if (op == 0) {
    if (semantic)
      hipLaunchKernelGGL((k_major_filter<0, 1>), dim3(grid), dim3(256), 0, s, mv, tg, out, nwords);
    else
      hipLaunchKernelGGL((k_major_filter<0, 0>), dim3(grid), dim3(256), 0, s, mv, tg, out, nwords);
  } else {
    if (semantic)
      hipLaunchKernelGGL((k_major_filter<1, 1>), dim3(grid), dim3(256), 0, s, mv, tg, out, nwords);
    else
      hipLaunchKernelGGL((k_major_filter<1, 0>), dim3(grid), dim3(256), 0, s, mv, tg, out, nwords);
  }
  return hipGetLastError();
}

hipError_t launch_major_decode(const MajorView& mv, bool semantic, double inv_scale, double* out, hipStream_t s,
                               const LaunchShape& ls) {
  const uint32_t grid = tiles_grid(mv.dense.ntiles, ls.max_wg);
  if (semantic)
    hipLaunchKernelGGL(k_major_decode<1>, dim3(grid), dim3(256), 0, s, mv, inv_scale, out);
  else
    hipLaunchKernelGGL(k_major_decode<0>, dim3(grid), dim3(256), 0, s, mv, inv_scale, out);
  return hipGetLastError();
}

static constexpr uint32_t kOutsumBlocks = 64;

uint64_t major_sum_scratch_bytes(uint32_t S) { return (uint64_t)S * kOutsumBlocks * 8; }

hipError_t launch_major_outlier_sums(const MajorView& mv, const uint64_t* n_out, unsigned long long* sums,
                                     unsigned long long* scratch, hipStream_t s) {
  if (mv.S == 0) return hipSuccess;
  OutlierCounts oc;
  for (uint32_t j = 0; j < kMaxSparse; ++j) oc.n[j] = j < mv.S ? n_out[j] : 0;
  hipLaunchKernelGGL(k_major_outsum, dim3(kOutsumBlocks, mv.S), dim3(256), 0, s, mv, oc, scratch);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_major_outsum_final, dim3(mv.S), dim3(64), 0, s, scratch, kOutsumBlocks, sums);
  return hipGetLastError();
}

uint64_t major_max_scratch_bytes(uint64_t ntiles) { return 2 * ntiles * 128 + 4096 * 4 + 256; }

hipError_t launch_major_max(const MajorView& mv, MajorMaxState* st, uint8_t* scratch, hipStream_t s,
                            const LaunchShape& ls, bool semantic) {
  const ChunkView& dv = mv.dense;
  const uint64_t ntiles = dv.ntiles;
  uint16_t* cand[2] = {reinterpret_cast<uint16_t*>(scratch), reinterpret_cast<uint16_t*>(scratch + ntiles * 128)};
  uint32_t* partial = reinterpret_cast<uint32_t*>(scratch + 2 * ntiles * 128);
  const uint32_t grid = tiles_grid(ntiles, std::min<uint32_t>(ls.max_wg, 4096));
  hipError_t e = launch_fill_bytes(st, sizeof(MajorMaxState), 0, s);
  if (e != hipSuccess) return e;
  const uint32_t ncol = mv.S + dv.P;
  const uint32_t nstage = ncol + (dv.r ? 2 : 0);
  for (uint32_t k = 0; k < nstage; ++k) {
    const uint32_t kind = k < mv.S ? kStSparse : k < ncol ? kStDense : k == ncol ? kStTail : kStCount;
    const uint16_t* in = k >= 2 ? cand[(k - 1) & 1] : nullptr;
    uint16_t* outc = cand[k & 1];
#define BUFF_MAX_STAGE(KIND, SEM)                                                                            \
  hipLaunchKernelGGL((k_major_max_stage<KIND, SEM>), dim3(grid), dim3(256), 0, s, mv, k, in, outc, st, partial)
    if (kind == kStSparse) {
      if (semantic) BUFF_MAX_STAGE(kStSparse, 1);
      else BUFF_MAX_STAGE(kStSparse, 0);
    } else if (kind == kStDense) {
      BUFF_MAX_STAGE(kStDense, 0);
    } else if (kind == kStTail) {
      BUFF_MAX_STAGE(kStTail, 0);
    } else {
      BUFF_MAX_STAGE(kStCount, 0);
    }
#undef BUFF_MAX_STAGE
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t major = k < mv.S ? mv.col[k].major : 0;
    const uint32_t shift = k < mv.S ? mv.col[k].shift : k < ncol ? dv.F - 8 * (k - mv.S + 1) : 0;
    hipLaunchKernelGGL(k_major_max_final, dim3(1), dim3(256), 0, s, partial, grid, kind, k, major, shift,
                       semantic ? 1u : 0u, st);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace buffgpu
