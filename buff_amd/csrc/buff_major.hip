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
#include "buff_launch.h"
#include "buff_major.h"
#include "buff_major_device.h"

namespace buffgpu {

// ----------------------------------------------------------------------------
// Encode (buff_simd.rs:451-555). Pass 1 reads the float64 column once and, per
// sparse column, stores each lane's outlier bytes packed into its own 16-byte
// slot (one aligned store) next to the outlier bitmap; per-workgroup outlier
// totals come from per-lane popcounts. The one-workgroup scan makes the totals
// bases (and finds the last outlier row); pass 2 walks each workgroup's tiles in
// order and moves the packed bytes to their place in the compacted array.

__device__ __forceinline__ void pack_outliers(const uint64_t (&off)[16], uint32_t sh, uint32_t m, uint64_t& lo,
                                              uint64_t& hi) {
  lo = 0;
  hi = 0;
  uint32_t k = 0;
#pragma unroll
  for (uint32_t b = 0; b < 16; ++b) {
    const uint64_t v = (off[b] >> sh) & 0xFF;
    if ((m >> b) & 1u) {
      if (k < 8) lo |= v << (8 * k);
      else hi |= v << (8 * (k - 8));
      ++k;
    }
  }
}

__global__ __launch_bounds__(256) void k_major_encode(const double* __restrict__ x, uint64_t n, uint64_t ntiles,
                                                      int32_t pexp, uint32_t dlen, int64_t base, ChunkView dense,
                                                      MajorEnc me) {
  __shared__ uint64_t s_off[4][64 * kLdsRow];
  __shared__ uint32_t s_cnt[kMaxSparse][4];
  const uint32_t lane = __lane_id();
  const uint32_t wslot = wave_in_wg();
  uint64_t* so = s_off[wslot];
  const uint64_t t0 = (uint64_t)blockIdx.x * me.tpw;
  const uint64_t t1 = min(t0 + me.tpw, ntiles);
  uint32_t lcnt[kMaxSparse];
#pragma unroll
  for (uint32_t j = 0; j < kMaxSparse; ++j) lcnt[j] = 0;
  for (uint64_t c = t0 + wslot; c < t1; c += 4) {
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
    store_tile_planes(off, dense.P, dense.r, dense.planes, dense.pitch, dense.tail, c, lane);
#pragma unroll 1
    for (uint32_t j = 0; j < me.S; ++j) {
      const uint32_t sh = me.shift[j], mj = me.major[j];
      uint32_t m = 0;
#pragma unroll
      for (uint32_t b = 0; b < 16; ++b) m |= (uint32_t)(((off[b] >> sh) & 0xFF) != mj) << b;
      m &= valid;
      me.mask16[j][c * 64 + lane] = (uint16_t)m;
      const uint32_t tile_cnt = wave_total_u32(__popc(m));
      if (lane == 0) me.counts[j][c] = tile_cnt;
      if (m) {
        uint64_t lo, hi;
        pack_outliers(off, sh, m, lo, hi);
        reinterpret_cast<uint4*>(me.stage[j] + c * kTileRows)[lane] =
            make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
      }
      const uint32_t pc = __popc(m);
#pragma unroll
      for (uint32_t jj = 0; jj < kMaxSparse; ++jj) lcnt[jj] += jj == j ? pc : 0u;   // registers, not scratch
    }
  }
#pragma unroll
  for (uint32_t j = 0; j < kMaxSparse; ++j) {
    if (j >= me.S) break;
    const uint32_t w = wave_sum_u32(lcnt[j]);
    if (lane == 0) s_cnt[j][wslot] = w;
  }
  __syncthreads();
  if (threadIdx.x < me.S) {
    const uint32_t j = threadIdx.x;
    me.wgsum[j][blockIdx.x] = s_cnt[j][0] + s_cnt[j][1] + s_cnt[j][2] + s_cnt[j][3];
  }
}

// One workgroup: per column, the workgroup totals become exclusive bases, and
// the last outlier row is found in the last workgroup range holding outliers.
__global__ __launch_bounds__(256) void k_major_scan(MajorEnc me, uint64_t ntiles,
                                                    unsigned long long* __restrict__ totals) {
  __shared__ uint32_t s_sum[256];
  __shared__ int32_t s_lastwg[256];
  __shared__ unsigned long long s_row[256];
  const uint32_t tid = threadIdx.x;
  const uint32_t per = (me.nwg + 255) / 256;
  for (uint32_t j = 0; j < me.S; ++j) {
    uint32_t* ws = me.wgsum[j];
    const uint32_t a0 = min(tid * per, me.nwg), a1 = min(a0 + per, me.nwg);
    uint32_t local = 0;
    int32_t lastwg = -1;
    for (uint32_t i = a0; i < a1; ++i) {
      const uint32_t k = ws[i];
      local += k;
      if (k) lastwg = (int32_t)i;
    }
    s_sum[tid] = local;
    s_lastwg[tid] = lastwg;
    __syncthreads();
    for (uint32_t o = 1; o < 256; o <<= 1) {
      const uint32_t a = tid >= o ? s_sum[tid - o] : 0u;
      __syncthreads();
      s_sum[tid] += a;
      __syncthreads();
    }
    for (uint32_t o = 128; o > 0; o >>= 1) {
      if (tid < o) s_lastwg[tid] = max(s_lastwg[tid], s_lastwg[tid + o]);
      __syncthreads();
    }
    uint32_t run = s_sum[tid] - local;
    for (uint32_t i = a0; i < a1; ++i) {
      const uint32_t k = ws[i];
      ws[i] = run;
      run += k;
    }
    // last outlier row: highest set bit in the masks of that workgroup's tiles
    const int32_t lw = s_lastwg[0];
    unsigned long long best = 0;   // row + 1
    if (lw >= 0) {
      const uint64_t w0 = (uint64_t)lw * me.tpw * 64;
      const uint64_t w1 = min(((uint64_t)lw + 1) * me.tpw, ntiles) * 64;
      for (uint64_t w = w0 + tid; w < w1; w += 256) {
        const uint32_t m = me.mask16[j][w];
        if (m) best = max(best, (unsigned long long)(16 * w + (32u - (uint32_t)__clz(m))));
      }
    }
    s_row[tid] = best;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
      if (tid < o) s_row[tid] = max(s_row[tid], s_row[tid + o]);
      __syncthreads();
    }
    if (tid == 0) {
      const uint32_t total = s_sum[255];
      me.prefix[j][ntiles] = total;
      totals[2 * j] = total;
      totals[2 * j + 1] = s_row[0] ? s_row[0] - 1 : ~0ull;
    }
    __syncthreads();
  }
}

// Pass 2a: workgroup w scans the counts of its tiles (256 at a time) from its base.
__global__ __launch_bounds__(256) void k_major_tileprefix(MajorEnc me, uint64_t ntiles) {
  __shared__ uint32_t s_sum[256];
  const uint32_t tid = threadIdx.x;
  const uint64_t t0 = (uint64_t)blockIdx.x * me.tpw;
  const uint64_t t1 = min(t0 + me.tpw, ntiles);
  for (uint32_t j = 0; j < me.S; ++j) {
    uint32_t carry = me.wgsum[j][blockIdx.x];
    for (uint64_t g = t0; g < t1; g += 256) {
      const uint64_t c = g + tid;
      const uint32_t cnt = c < t1 ? me.counts[j][c] : 0u;
      s_sum[tid] = cnt;
      __syncthreads();
      for (uint32_t o = 1; o < 256; o <<= 1) {
        const uint32_t a = tid >= o ? s_sum[tid - o] : 0u;
        __syncthreads();
        s_sum[tid] += a;
        __syncthreads();
      }
      if (c < t1) me.prefix[j][c] = carry + s_sum[tid] - cnt;
      carry += s_sum[255];
      __syncthreads();
    }
  }
}

// Pass 2b: every tile's staged outlier bytes to prefix[c] + lane offset.
__global__ __launch_bounds__(256) void k_major_copy(MajorEnc me, uint64_t ntiles) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  for (uint64_t c = wave; c < ntiles; c += nwaves) {
#pragma unroll 1
    for (uint32_t j = 0; j < me.S; ++j) {
      const uint32_t m = me.mask16[j][c * 64 + lane];
      const uint32_t pre = me.prefix[j][c];
      const uint32_t pc = __popc(m);
      const uint32_t excl = wave_excl_scan_u32(pc, lane);
      if (m) {
        const uint4 w = reinterpret_cast<const uint4*>(me.stage[j] + c * kTileRows)[lane];
        const uint64_t lo = (uint64_t)w.x | ((uint64_t)w.y << 32), hi = (uint64_t)w.z | ((uint64_t)w.w << 32);
        uint8_t* dst = me.outl[j] + pre + excl;
        for (uint32_t k = 0; k < pc; ++k) dst[k] = (uint8_t)(((k < 8 ? lo : hi) >> (8 * (k & 7))) & 0xFF);
      }
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

// U tiles per wave iteration, every load of a phase for all of them in flight:
// masks / tile prefixes / first dense plane, then the outlier windows, then the
// candidates' later planes and tail. S (sparse columns) is a template argument so
// the per-column state stays in registers.
template <int MODE, int S, int U, bool PF>
__global__ __launch_bounds__(256) void k_major_filter(MajorView mv, MajorTarget tg, uint32_t sem,
                                                      uint64_t* __restrict__ out, uint64_t nwords) {
  __shared__ unsigned long long s_dep[4][64];
  unsigned long long* dep = s_dep[wave_in_wg()];
  const uint32_t lane = __lane_id();
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  const ChunkView& dv = mv.dense;
  uint16_t* out16 = reinterpret_cast<uint16_t*>(out);
  const uint64_t nout16 = 4 * nwords;
  const uint64_t ngroups = (dv.ntiles + U - 1) / U;
  for (uint64_t grp = wave; grp < ngroups; grp += nwaves) {
    const uint64_t c0 = grp * U;
    uint32_t res[U];
    major_filter_group<MODE, S, U, PF>(mv, tg, sem != 0, c0, lane, dep, res);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t c = c0 + u;
      const uint64_t idx = c * 64 + lane;
      if (c < dv.ntiles && idx < nout16) out16[idx] = (uint16_t)res[u];
    }
  }
}

// Row positions of one sparse column: lane t of tile c writes the rows of its
// outliers (16t + b for the set bits b of its bitmap word) from its exclusive
// position on.
__global__ __launch_bounds__(256) void k_major_positions(const uint16_t* __restrict__ mask16,
                                                         const uint32_t* __restrict__ prefix, uint64_t ntiles,
                                                         uint16_t* __restrict__ pos) {
  const uint32_t lane = __lane_id();
  const uint64_t nwaves = waves_in_grid();
  for (uint64_t c = wave_global(); c < ntiles; c += nwaves) {
    uint32_t m = mask16[c * 64 + lane];
    const uint32_t excl = wave_excl_scan_u32(__popc(m), lane);
    uint64_t at = (uint64_t)prefix[c] + excl;
    while (m) {
      const uint32_t b = (uint32_t)__builtin_ctz(m);
      m &= m - 1u;
      pos[at++] = (uint16_t)(16 * lane + b);
    }
  }
}

// ----------------------------------------------------------------------------
// Decode: buff_decode_majority (buff_simd.rs:633-861), the offsets built by
// major_tile_offsets (SEM: the true offsets, truncated to their top F - drop
// bits for a reduced-precision read). Values are staged in LDS and written
// coalesced.

template <int SEM, int S>
__global__ __launch_bounds__(256) void k_major_decode(MajorView mv, double inv_scale, uint32_t drop,
                                                      double* __restrict__ out) {
  __shared__ double s_val[4][64 * kLdsRow];
  const uint32_t lane = __lane_id();
  double* sv = s_val[wave_in_wg()];
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  const ChunkView& dv = mv.dense;
  for (uint64_t c = wave; c < dv.ntiles; c += nwaves) {
    uint64_t v[16];
    major_tile_offsets<SEM, S>(mv, c, lane, v);
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) {
      // SEM 1: the top F - drop bits in place; SEM 2: the top F - drop bits as an
      // (F - drop)-bit number (the residue format's reduced-precision read)
      const uint64_t o = SEM == 1 ? (v[b] >> drop) << drop : SEM == 2 ? v[b] >> drop : v[b];
      sv[kLdsRow * lane + b] = (double)(int64_t)((uint64_t)dv.base + o) * inv_scale;
    }
    wave_lds_sync();
    double* ot = out + c * kTileRows + lane;
    if (tile_full(dv.n, c)) {
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) store_nt(ot + 64 * g, sv[lds_slot(64 * g + lane)]);
    } else {
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g)
        if (c * kTileRows + 64 * g + lane < dv.n) store_nt(ot + 64 * g, sv[lds_slot(64 * g + lane)]);
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

// Stages after the first also keep one flag byte per tile (any candidate
// left): a stage skips the tiles whose input flag is clear without reading
// their candidates or column bytes, so once the chosen bytes single out a few
// rows (an outlier holding the max), later stages touch only those tiles.
template <int KIND, int SEM>
__global__ __launch_bounds__(256) void k_major_max_stage(MajorView mv, uint32_t k, const uint16_t* __restrict__ cand_in,
                                                         uint16_t* __restrict__ cand_out,
                                                         const uint8_t* __restrict__ flag_in,
                                                         uint8_t* __restrict__ flag_out,
                                                         const MajorMaxState* __restrict__ st,
                                                         uint32_t* __restrict__ partial) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  const ChunkView& dv = mv.dense;
  const uint32_t ncol = mv.S + dv.P;
  uint32_t acc = 0;
  for (uint64_t c = wave; c < dv.ntiles; c += nwaves) {
    if (flag_in && !flag_in[c]) {
      if (flag_out && lane == 0) flag_out[c] = 0;
      continue;
    }
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
        uint32_t gt_unused, eq;
        sparse_cmp(lane_outliers(sd, c, lane, m16), m16, m16 & cand, sd.major, chosen, gt_unused, eq);
        cand &= eq;
      } else if (prev < ncol) {
        if (__ballot(cand != 0)) {
          const uint4 w = load_plane(dv, prev - mv.S, c, refine_src(cand, lane));
          cand &= mask_eq(w, make_digit(chosen));
        }
      } else if (need_tail) {   // the tail narrows the count stage
        uint32_t eq = 0;
#pragma unroll
        for (uint32_t b = 0; b < 16; ++b) eq |= (uint32_t)(tail_value(dv, c, lane, b, h) == chosen) << b;
        cand &= eq;
      }
      if (KIND != kStCount) {
        cand_out[c * 64 + lane] = (uint16_t)cand;
        const bool any = __ballot(cand != 0) != 0;
        if (flag_out && lane == 0) flag_out[c] = any ? 1 : 0;
      }
    }
    if (KIND == kStSparse) {
      const SparseDev& sd = mv.col[k];
      const uint32_t m16 = sd.mask16[c * 64 + lane];
      const uint32_t co = cand & m16;
      acc = max(acc, sparse_max_plus1(lane_outliers(sd, c, lane, m16), m16, co));
      if (SEM && (cand & ~m16)) acc = max(acc, sd.major + 1u);
    } else if (KIND == kStDense) {
      if (__ballot(cand != 0)) {
        const uint4 w = load_plane(dv, k - mv.S, c, refine_src(cand, lane));
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
// Max by a row list. Column 0's chosen byte depends only on its outlier bytes
// (and whether any row is not an outlier), so it is a reduction over the
// compacted bytes. When it is not the majority byte, the candidates of every
// later stage are among the rows whose column-0 outlier byte equals it: outliers
// hold the extremes, so these are few. One pass over the tiles collects each
// such row's byte in every column and its tail; one workgroup then runs the
// stage rules of k_major_max_final on that list.

struct MaxRow {
  uint8_t byte[kMaxPlanes];   // sparse columns, then dense planes (S + P <= 8)
  uint8_t tail;
  uint8_t outlier;            // bit j: an outlier of sparse column j
  uint16_t pad;
  uint32_t row;
};
static_assert(sizeof(MaxRow) == 16, "MaxRow is one dwordx4");

// Column 0's chosen byte (the rule of k_major_max_final for stage 0).
__device__ __forceinline__ uint32_t list_chosen0(uint32_t top1, uint32_t major, bool semantic, bool has_major) {
  if (!semantic) return top1 ? max(major, top1 - 1) : major;
  const uint32_t v = max(top1, has_major ? major + 1 : 0u);
  return v ? v - 1 : 0u;
}

// The collect pass appends rows to kListRegions regions (a counter each, 128
// bytes apart, and kMaxListRows / kListRegions rows): hit rows turn up in many
// waves at once, and same-address device atomics serialize.
constexpr uint32_t kListRegions = 64;
constexpr uint32_t kListRegionRows = kMaxListRows / kListRegions;
constexpr uint32_t kListCounterStride = 32;   // u32s: 128 bytes
static_assert(kListRegionRows * kListRegions == kMaxListRows, "list regions");

// Per tile, column 0's largest outlier byte (0 for a tile without outliers:
// never column 0's chosen byte on the list path, which is then above the
// majority byte, or every row is an outlier) into tilemax; per workgroup the
// max over its tiles, plus 1 (0: no outlier), into partial[blockIdx.x]. A
// tile's outlier bytes are contiguous in the compacted array; lane i reads byte
// i (and i + 64, ...). Workgroup 0 also clears the collect pass's counters.
constexpr int kTopTiles = 16;   // tiles per wave step, their loads in flight together

__global__ __launch_bounds__(256) void k_major_max_tiles(SparseDev c0, uint64_t ntiles,
                                                         uint8_t* __restrict__ tilemax,
                                                         uint32_t* __restrict__ partial,
                                                         uint32_t* __restrict__ counters) {
  constexpr int U = kTopTiles;
  __shared__ uint32_t part[4];
  if (blockIdx.x == 0 && threadIdx.x < kListRegions) counters[threadIdx.x * kListCounterStride] = 0;
  const uint32_t lane = __lane_id();
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  uint32_t best = 0;
  const uint64_t ngroups = (ntiles + U - 1) / U;
  for (uint64_t grp = wave; grp < ngroups; grp += nwaves) {
    const uint64_t cb = grp * U;
    uint32_t p[U + 1];
#pragma unroll
    for (int u = 0; u <= U; ++u) p[u] = c0.prefix[min(cb + u, ntiles)];   // prefix[ntiles] = outliers
    uint32_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = lane < p[u + 1] - p[u] ? (uint32_t)c0.outl[p[u] + lane] + 1u : 0u;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      for (uint32_t i = lane + 64; i < p[u + 1] - p[u]; i += 64) v[u] = max(v[u], (uint32_t)c0.outl[p[u] + i] + 1u);
      const uint32_t m = wave_max_u32_dpp(v[u]);
      best = max(best, m);
      if (lane == (uint32_t)u && cb + u < ntiles) tilemax[cb + u] = (uint8_t)(m ? m - 1 : 0);
    }
  }
  if (lane == 0) part[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = max(max(part[0], part[1]), max(part[2], part[3]));
}

// max over the top pass's partials (at most kTopBlocks; lane i reads 16 of them
// as dwordx4s), every lane of the wave
__device__ __forceinline__ uint32_t list_top1(const uint32_t* __restrict__ partial, uint32_t nparts) {
  const uint32_t lane = __lane_id();
  uint32_t v = 0;
#pragma unroll
  for (uint32_t q = 0; q < kTopBlocks / 256; ++q) {
    const uint32_t i = 4 * (lane + 64 * q);
    if (i < nparts) {
      const uint4 w = reinterpret_cast<const uint4*>(partial)[lane + 64 * q];
      v = max(v, max(max(w.x, i + 1 < nparts ? w.y : 0u), max(i + 2 < nparts ? w.z : 0u, i + 3 < nparts ? w.w : 0u)));
    }
  }
  return wave_max_u32_dpp(v);
}

// Tiles per wave of the collect pass (lane u < kCollectTiles reads tilemax of one).
constexpr uint32_t kCollectTiles = 16;

// The rows of the tiles whose column-0 max is the chosen byte: per hit tile,
// every column's masks, the dense planes and the tail in one round of loads,
// the outlier windows in a second.
template <int S>
__global__ __launch_bounds__(256) void k_major_max_collect(MajorView mv, const uint8_t* __restrict__ tilemax,
                                                           const uint32_t* __restrict__ partial, uint32_t nparts,
                                                           uint32_t* __restrict__ counters, MaxRow* __restrict__ rows,
                                                           uint32_t semantic, uint32_t has_major) {
  constexpr int NPM = (int)kMaxPlanes - S > 0 ? (int)kMaxPlanes - S : 1;
  const uint32_t lane = __lane_id();
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  const ChunkView& dv = mv.dense;
  const uint32_t chosen0 = list_chosen0(list_top1(partial, nparts), mv.col[0].major, semantic != 0, has_major != 0);
  if (chosen0 == mv.col[0].major) return;   // the majority rows are candidates: staged max
  const uint32_t region = (uint32_t)(wave % kListRegions);
  uint32_t* counter = counters + region * kListCounterStride;
  MaxRow* rrows = rows + region * kListRegionRows;
  const uint64_t ngroups = (dv.ntiles + kCollectTiles - 1) / kCollectTiles;
  for (uint64_t grp = wave; grp < ngroups; grp += nwaves) {
    const uint64_t cb = grp * kCollectTiles;
    const bool mine = lane < kCollectTiles && cb + lane < dv.ntiles;
    uint64_t hit_tiles = __ballot(mine && tilemax[cb + lane] == chosen0);
    while (hit_tiles) {   // rare: the tiles holding the chosen byte, one at a time
      const uint64_t c = cb + (uint32_t)__builtin_ctzll(hit_tiles);
      hit_tiles &= hit_tiles - 1;
      uint32_t m[S], pre[S];
#pragma unroll
      for (int j = 0; j < S; ++j) {
        m[j] = mv.col[j].mask16[c * 64 + lane];
        pre[j] = mv.col[j].prefix[c];
      }
      uint4 wp[NPM];
#pragma unroll
      for (int p = 0; p < NPM; ++p) wp[p] = (uint32_t)p < dv.P ? load_plane(dv, p, c, lane) : make_uint4(0, 0, 0, 0);
      uint32_t h[8];
#pragma unroll
      for (uint32_t j = 0; j < 8; ++j) h[j] = j < dv.r ? load_tail(dv, j, c, lane) : 0u;
      Packed16 pk[S];
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const uint32_t excl = wave_excl_scan_u32(__popc(m[j]), lane);
        pk[j] = m[j] ? load_packed(mv.col[j].outl, (uint64_t)pre[j] + excl) : Packed16{0, 0, 0, 0};
      }
      uint32_t gt_unused, hit;
      sparse_cmp(pk[0], m[0], m[0], mv.col[0].major, chosen0, gt_unused, hit);   // chosen0 != major: outliers only
      uint4 ws[S];
#pragma unroll
      for (int j = 1; j < S; ++j) ws[j] = patch_outliers(pk[j], m[j], m[j] & hit, mv.col[j].major);
      uint32_t left = hit;
      while (left) {
        const uint32_t b = (uint32_t)__builtin_ctz(left);
        left &= left - 1u;
        MaxRow rec;
#pragma unroll
        for (uint32_t j = 0; j < kMaxPlanes; ++j) rec.byte[j] = 0;
        rec.byte[0] = (uint8_t)chosen0;
        rec.outlier = 1;
#pragma unroll
        for (int j = 1; j < S; ++j) {
          rec.byte[j] = (uint8_t)byte_at(ws[j], b);
          rec.outlier |= (uint8_t)(((m[j] >> b) & 1u) << j);
        }
#pragma unroll
        for (int p = 0; p < NPM; ++p)
          if (S + p < (int)kMaxPlanes && (uint32_t)p < dv.P) rec.byte[S + p] = (uint8_t)byte_at(wp[p], b);
        uint32_t t = 0;
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) t |= j < dv.r ? ((h[j] >> b) & 1u) << j : 0u;
        rec.tail = (uint8_t)t;
        rec.pad = 0;
        rec.row = (uint32_t)(c * kTileRows + 16ull * lane + b);
        const uint32_t idx = atomicAdd(counter, 1u);
        if (idx < kListRegionRows) rrows[idx] = rec;
      }
    }
  }
}

// One workgroup of 1024 threads; every thread holds up to kListPerThread rows.
constexpr uint32_t kListPerThread = kMaxListRows / 1024;

// Max over the workgroup; slots alternate between calls, so one barrier a call
// suffices (a wave writes a slot again only after every wave has passed the
// next call's barrier, i.e. has finished reading it).
__device__ __forceinline__ uint32_t block_max_1024(uint32_t v, uint32_t (&red)[2][16], uint32_t& parity) {
  v = wave_max_u32_dpp(v);
  uint32_t* slot = red[parity];
  parity ^= 1u;
  if (__lane_id() == 0) slot[threadIdx.x >> 6] = v;
  __syncthreads();
  uint32_t r = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) r = max(r, slot[i]);
  return r;
}

// A list row as loaded: x, y = the column bytes (byte k of the pair is column k),
// z = tail | outlier bits << 8.
__device__ __forceinline__ uint32_t row_byte(const uint4& r, uint32_t k) {
  return ((k < 4 ? r.x : r.y) >> (8 * (k & 3))) & 0xFFu;
}

// Writes every field of st the host reads: acc, argmax, count, fallback and the
// chosen bytes (st is not cleared before the list path).
__global__ __launch_bounds__(1024) void k_major_max_finish(MajorView mv, const uint32_t* __restrict__ partial,
                                                           uint32_t nparts, MajorMaxState* __restrict__ st,
                                                           const uint32_t* __restrict__ counters,
                                                           const MaxRow* __restrict__ rows, uint32_t semantic,
                                                           uint32_t has_major) {
  __shared__ uint32_t red[2][16];
  __shared__ uint32_t s_cnt[kListRegions];
  uint32_t parity = 0;
  const ChunkView& dv = mv.dense;
  const uint32_t S = mv.S, ncol = mv.S + dv.P;
  const uint32_t top1 = list_top1(partial, nparts);
  const uint32_t chosen0 = list_chosen0(top1, mv.col[0].major, semantic != 0, has_major != 0);
  // region counts: a region past its rows (or no row at all) means the staged max
  const uint32_t rc = threadIdx.x < kListRegions ? counters[threadIdx.x * kListCounterStride] : 0u;
  if (threadIdx.x < kListRegions) s_cnt[threadIdx.x] = rc;
  const uint32_t total = wave_total_u32(min(rc, kListRegionRows + 1));   // the first wave holds every region
  const bool over = __syncthreads_or(rc > kListRegionRows) != 0;
  const bool none = __syncthreads_or(threadIdx.x == 0 && total == 0) != 0;
  if (threadIdx.x == 0) {
    st->top1 = top1;
    st->count = total;
    st->argmax = 0;
    st->fallback = chosen0 == mv.col[0].major || over || none ? 1u : 0u;
  }
  if (chosen0 == mv.col[0].major || over || none) return;
  const uint4* rv = reinterpret_cast<const uint4*>(rows);
  uint4 r[kListPerThread];
  uint32_t active = 0;
#pragma unroll
  for (uint32_t q = 0; q < kListPerThread; ++q) {
    const uint32_t slot = threadIdx.x + 1024 * q;   // region slot / kListRegionRows, row slot % kListRegionRows
    const bool valid = slot % kListRegionRows < s_cnt[slot / kListRegionRows];
    r[q] = valid ? rv[slot] : make_uint4(0, 0, 0, 0);
    active |= valid ? 1u << q : 0u;
  }
  unsigned long long acc = (unsigned long long)chosen0 << mv.col[0].shift;
  if (threadIdx.x == 0) st->chosen[0] = chosen0;
  for (uint32_t k = 1; k < ncol; ++k) {
    const bool sparse = k < S;
    uint32_t v = 0;
#pragma unroll
    for (uint32_t q = 0; q < kListPerThread; ++q) {
      const bool outl = (r[q].z >> (8 + k)) & 1u;
      if (((active >> q) & 1u) && (!sparse || semantic || outl)) v = max(v, row_byte(r[q], k) + 1u);
    }
    v = block_max_1024(v, red, parity);
    uint32_t chosen;
    bool merge = true;
    if (sparse && !semantic) {
      chosen = v ? max(mv.col[k].major, v - 1) : mv.col[k].major;
      merge = v != 0;
    } else {
      chosen = v ? v - 1 : 0u;
    }
    const uint32_t shift = sparse ? mv.col[k].shift : dv.F - 8 * (k - S + 1);
    if (merge) acc |= (unsigned long long)chosen << shift;
    if (threadIdx.x == 0) st->chosen[k] = chosen;
#pragma unroll
    for (uint32_t q = 0; q < kListPerThread; ++q)
      if (row_byte(r[q], k) != chosen) active &= ~(1u << q);
  }
  if (dv.r) {
    uint32_t v = 0;
#pragma unroll
    for (uint32_t q = 0; q < kListPerThread; ++q)
      if ((active >> q) & 1u) v = max(v, (r[q].z & 0xFFu) + 1u);
    v = block_max_1024(v, red, parity);
    const uint32_t chosen = v ? v - 1 : 0u;
    acc |= chosen;
    if (threadIdx.x == 0) st->chosen[ncol] = chosen;
#pragma unroll
    for (uint32_t q = 0; q < kListPerThread; ++q)
      if ((r[q].z & 0xFFu) != chosen) active &= ~(1u << q);
    // rows left: the count stage's argmax
    const uint32_t nleft = wave_total_u32((uint32_t)__popc(active));
    uint32_t* slot = red[parity];
    if (__lane_id() == 0) slot[threadIdx.x >> 6] = nleft;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t tot = 0;
      for (int i = 0; i < 16; ++i) tot += slot[i];
      st->argmax = tot;
    }
  }
  if (threadIdx.x == 0) st->acc = acc;
}

// ----------------------------------------------------------------------------
// launchers

static inline uint32_t tiles_grid(uint64_t ntiles, uint32_t max_wg) {
  uint64_t wgs = (ntiles + 3) / 4;
  if (wgs < 1) wgs = 1;
  if (wgs > max_wg) wgs = max_wg;
  return (uint32_t)wgs;
}

void major_enc_geometry(uint64_t ntiles, MajorEnc* me) {
  uint64_t nwg = (ntiles + 3) / 4;
  if (nwg > 4096) nwg = 4096;
  if (nwg < 1) nwg = 1;
  me->tpw = (ntiles + nwg - 1) / nwg;
  me->nwg = (uint32_t)((ntiles + me->tpw - 1) / me->tpw);
  if (me->nwg < 1) me->nwg = 1;
}

hipError_t launch_major_encode(const double* x, uint64_t n, int32_t pexp, uint32_t dlen, int64_t base,
                               const ChunkView& dense, const MajorEnc& me, hipStream_t s) {
  const uint64_t ntiles = (n + kTileRows - 1) / kTileRows;
  launch_k(k_major_encode, dim3(me.nwg), dim3(256), 0, s, x, n, ntiles, pexp, dlen, base, dense, me);
  return hipGetLastError();
}

hipError_t launch_major_scan(const MajorEnc& me, uint64_t ntiles, unsigned long long* totals, hipStream_t s) {
  launch_k(k_major_scan, dim3(1), dim3(256), 0, s, me, ntiles, totals);
  return hipGetLastError();
}

hipError_t launch_major_place(const MajorEnc& me, uint64_t ntiles, hipStream_t s) {
  launch_k(k_major_tileprefix, dim3(me.nwg), dim3(256), 0, s, me, ntiles);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  launch_k(k_major_copy, dim3(tiles_grid(ntiles, 4096)), dim3(256), 0, s, me, ntiles);
  return hipGetLastError();
}

hipError_t launch_major_positions(const uint16_t* mask16, const uint32_t* prefix, uint64_t ntiles, uint16_t* pos,
                                  hipStream_t s) {
  if (ntiles == 0) return hipSuccess;
  launch_k(k_major_positions, dim3(tiles_grid(ntiles, 4096)), dim3(256), 0, s, mask16, prefix, ntiles, pos);
  return hipGetLastError();
}

// Workgroups of 256 threads of kernel k that fit on the device at once (0 if
// unknown): a grid no larger than that runs in one round, with no second round
// of stragglers.
template <typename K>
static uint32_t resident_wgs(K k) {
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0) == hipSuccess && hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && per_cu > 0 && cus > 0)
    return (uint32_t)(per_cu * cus);
  return 0;
}

template <int MODE, int S, bool PF>
static void launch_filter_pf(const MajorView& mv, bool semantic, const MajorTarget& tg, uint64_t* out,
                             uint64_t nwords, hipStream_t s, uint32_t max_wg) {
  constexpr int U = 2;
  const uint64_t groups = (mv.dense.ntiles + U - 1) / U;
  static const uint32_t fit = resident_wgs(k_major_filter<MODE, S, U, PF>);   // once per kernel
  const uint32_t grid = tiles_grid(groups, fit ? std::min(max_wg, fit) : max_wg);
  launch_k((k_major_filter<MODE, S, U, PF>), dim3(grid), dim3(256), 0, s, mv, tg, semantic ? 1u : 0u, out, nwords);
}

template <int MODE, int S>
static void launch_filter_s(const MajorView& mv, bool semantic, const MajorTarget& tg, uint64_t* out,
                            uint64_t nwords, hipStream_t s, uint32_t max_wg) {
  // the target's sparse bytes all majority bytes: most rows reach the dense planes
  bool pf = mv.dense.P >= 2;
  for (uint32_t j = 0; j < mv.S; ++j) pf = pf && tg.digit[j] == mv.col[j].major;
  if (pf)
    launch_filter_pf<MODE, S, true>(mv, semantic, tg, out, nwords, s, max_wg);
  else
    launch_filter_pf<MODE, S, false>(mv, semantic, tg, out, nwords, s, max_wg);
}

template <int MODE>
static hipError_t launch_filter_mode(const MajorView& mv, bool semantic, const MajorTarget& tg, uint64_t* out,
                                     uint64_t nwords, hipStream_t s, uint32_t max_wg) {
  switch (mv.S) {
    case 1: launch_filter_s<MODE, 1>(mv, semantic, tg, out, nwords, s, max_wg); break;
    case 2: launch_filter_s<MODE, 2>(mv, semantic, tg, out, nwords, s, max_wg); break;
    case 3: launch_filter_s<MODE, 3>(mv, semantic, tg, out, nwords, s, max_wg); break;
    case 4: launch_filter_s<MODE, 4>(mv, semantic, tg, out, nwords, s, max_wg); break;
    case 5: launch_filter_s<MODE, 5>(mv, semantic, tg, out, nwords, s, max_wg); break;
    case 6: launch_filter_s<MODE, 6>(mv, semantic, tg, out, nwords, s, max_wg); break;
    case 7: launch_filter_s<MODE, 7>(mv, semantic, tg, out, nwords, s, max_wg); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_major_filter(const MajorView& mv, int op, bool semantic, const MajorTarget& tg, uint64_t* out,
                               uint64_t nwords, hipStream_t s, const LaunchShape& ls) {
  const uint32_t cap = ls.max_wg ? ls.max_wg : 2048;
  return op == 0 ? launch_filter_mode<0>(mv, semantic, tg, out, nwords, s, cap)
                 : launch_filter_mode<1>(mv, semantic, tg, out, nwords, s, cap);
}

template <int SEM>
static hipError_t launch_decode_sem(const MajorView& mv, double inv_scale, uint32_t drop, double* out, uint32_t grid,
                                    hipStream_t s) {
  switch (mv.S) {
    case 0: launch_k((k_major_decode<SEM, 0>), dim3(grid), dim3(256), 0, s, mv, inv_scale, drop, out); break;
    case 1: launch_k((k_major_decode<SEM, 1>), dim3(grid), dim3(256), 0, s, mv, inv_scale, drop, out); break;
    case 2: launch_k((k_major_decode<SEM, 2>), dim3(grid), dim3(256), 0, s, mv, inv_scale, drop, out); break;
    case 3: launch_k((k_major_decode<SEM, 3>), dim3(grid), dim3(256), 0, s, mv, inv_scale, drop, out); break;
    case 4: launch_k((k_major_decode<SEM, 4>), dim3(grid), dim3(256), 0, s, mv, inv_scale, drop, out); break;
    case 5: launch_k((k_major_decode<SEM, 5>), dim3(grid), dim3(256), 0, s, mv, inv_scale, drop, out); break;
    case 6: launch_k((k_major_decode<SEM, 6>), dim3(grid), dim3(256), 0, s, mv, inv_scale, drop, out); break;
    case 7: launch_k((k_major_decode<SEM, 7>), dim3(grid), dim3(256), 0, s, mv, inv_scale, drop, out); break;
    case 8: launch_k((k_major_decode<SEM, 8>), dim3(grid), dim3(256), 0, s, mv, inv_scale, drop, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_major_decode(const MajorView& mv, int mode, uint32_t drop, double inv_scale, double* out,
                               hipStream_t s, const LaunchShape& ls) {
  if (drop && mode == 0) return hipErrorInvalidValue;
  const uint32_t grid = tiles_grid(mv.dense.ntiles, ls.max_wg ? ls.max_wg : 2048);
  return mode == 1 ? launch_decode_sem<1>(mv, inv_scale, drop, out, grid, s)
       : mode == 2 ? launch_decode_sem<2>(mv, inv_scale, drop, out, grid, s)
                   : launch_decode_sem<0>(mv, inv_scale, 0, out, grid, s);
}

static constexpr uint32_t kOutsumBlocks = 64;

uint64_t major_sum_scratch_bytes(uint32_t S) { return (uint64_t)S * kOutsumBlocks * 8; }

hipError_t launch_major_outlier_sums(const MajorView& mv, const uint64_t* n_out, unsigned long long* sums,
                                     unsigned long long* scratch, hipStream_t s) {
  if (mv.S == 0) return hipSuccess;
  OutlierCounts oc;
  for (uint32_t j = 0; j < kMaxSparse; ++j) oc.n[j] = j < mv.S ? n_out[j] : 0;
  launch_k(k_major_outsum, dim3(kOutsumBlocks, mv.S), dim3(256), 0, s, mv, oc, scratch);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  launch_k(k_major_outsum_final, dim3(mv.S), dim3(64), 0, s, scratch, kOutsumBlocks, sums);
  return hipGetLastError();
}

hipError_t launch_major_max_list(const MajorView& mv, uint64_t n_out0, MajorMaxState* st, uint8_t* scratch,
                                 hipStream_t s, bool semantic) {
  const ChunkView& dv = mv.dense;
  if (mv.S < 1 || mv.S > kMaxSparse || dv.n == 0) return hipErrorInvalidValue;
  const uint32_t has_major = n_out0 < dv.n ? 1u : 0u;
  const uint32_t sem = semantic ? 1u : 0u;
  uint32_t* counters = reinterpret_cast<uint32_t*>(scratch);
  uint32_t* partial = reinterpret_cast<uint32_t*>(scratch + kListCounterBytes);
  MaxRow* rows = reinterpret_cast<MaxRow*>(scratch + kListCounterBytes + kTopBlocks * 4);
  uint8_t* tilemax = scratch + kMaxListScratch;
  const uint64_t groups = (dv.ntiles + kTopTiles - 1) / kTopTiles;   // one wave each
  const uint32_t g1 = (uint32_t)std::min<uint64_t>(std::max<uint64_t>((groups + 3) / 4, 1), kTopBlocks);
  launch_k(k_major_max_tiles, dim3(g1), dim3(256), 0, s, mv.col[0], dv.ntiles, tilemax, partial, counters);
  const uint64_t cgroups = (dv.ntiles + kCollectTiles - 1) / kCollectTiles;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(std::max<uint64_t>((cgroups + 3) / 4, 1), 4096);
#define BUFF_COLLECT(SS)                                                                               \
  launch_k((k_major_max_collect<SS>), dim3(grid), dim3(256), 0, s, mv, tilemax, partial, g1, counters, rows, sem, \
           has_major)
  switch (mv.S) {
    case 1: BUFF_COLLECT(1); break;
    case 2: BUFF_COLLECT(2); break;
    case 3: BUFF_COLLECT(3); break;
    case 4: BUFF_COLLECT(4); break;
    case 5: BUFF_COLLECT(5); break;
    case 6: BUFF_COLLECT(6); break;
    case 7: BUFF_COLLECT(7); break;
    default: BUFF_COLLECT(8); break;
  }
#undef BUFF_COLLECT
  launch_k(k_major_max_finish, dim3(1), dim3(1024), 0, s, mv, partial, g1, st, counters, rows, sem, has_major);
  return hipGetLastError();
}

uint64_t major_max_scratch_bytes(uint64_t ntiles) { return 2 * ntiles * 128 + 4096 * 4 + 2 * ntiles + 512; }

hipError_t launch_major_max(const MajorView& mv, MajorMaxState* st, uint8_t* scratch, hipStream_t s,
                            const LaunchShape& ls, bool semantic) {
  const ChunkView& dv = mv.dense;
  const uint64_t ntiles = dv.ntiles;
  uint16_t* cand[2] = {reinterpret_cast<uint16_t*>(scratch), reinterpret_cast<uint16_t*>(scratch + ntiles * 128)};
  uint32_t* partial = reinterpret_cast<uint32_t*>(scratch + 2 * ntiles * 128);
  uint8_t* flags[2] = {scratch + 2 * ntiles * 128 + 4096 * 4, scratch + 2 * ntiles * 128 + 4096 * 4 + ntiles};
  const uint32_t grid = tiles_grid(ntiles, std::min<uint32_t>(ls.max_wg ? ls.max_wg : 2048, 4096));
  hipError_t e = launch_fill_bytes(st, sizeof(MajorMaxState), 0, s);
  if (e != hipSuccess) return e;
  const uint32_t ncol = mv.S + dv.P;
  const uint32_t nstage = ncol + (dv.r ? 2 : 0);
  for (uint32_t k = 0; k < nstage; ++k) {
    const uint32_t kind = k < mv.S ? kStSparse : k < ncol ? kStDense : k == ncol ? kStTail : kStCount;
    const uint16_t* in = k >= 2 ? cand[(k - 1) & 1] : nullptr;
    uint16_t* outc = cand[k & 1];
    const uint8_t* fin = k >= 2 ? flags[(k - 1) & 1] : nullptr;
    uint8_t* fout = k >= 1 && kind != kStCount ? flags[k & 1] : nullptr;
#define BUFF_MAX_STAGE(KIND, SEM)                                                                            \
  launch_k((k_major_max_stage<KIND, SEM>), dim3(grid), dim3(256), 0, s, mv, k, in, outc, fin, fout, st, partial)
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
    launch_k(k_major_max_final, dim3(1), dim3(256), 0, s, partial, grid, kind, k, major, shift,
                       semantic ? 1u : 0u, st);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace buffgpu
