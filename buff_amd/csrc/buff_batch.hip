// Batched kernels (buff_batch.h) for gfx950: persistent workgroups take items
// (owner, tile range) from a list in device memory; the per-tile work is that of
// the single-chunk kernels (buff_kernels.hip) on the owner's chunk view.
#include <hip/hip_runtime.h>

#include "buff_batch.h"
#include "buff_common.h"
#include "buff_device.h"

namespace buffgpu {

// ----------------------------------------------------------------------------
// Batch encode, pass 1: per item min over all rows and max over the rows other
// than the chunk's row 0 (split_double.rs:321-330 keeps row 0 out of the max).

__global__ __launch_bounds__(256) void k_bminmax(const BEncChunk* __restrict__ chunks, const BItem* __restrict__ items,
                                                 uint32_t nitems, int64_t* __restrict__ part) {
  __shared__ int64_t smn[4], smx[4];
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  for (uint32_t i = blockIdx.x; i < nitems; i += gridDim.x) {
    const BItem it = items[i];
    const BEncChunk& ck = chunks[it.owner];
    const uint64_t n = ck.view.n;
    int64_t mn = INT64_MAX, mx = INT64_MIN;
    for (uint64_t c = it.c0 + wave; c < it.c1; c += 4) {
      uint64_t v[16];
      const bool full = tile_full(n, c);
      load_tile_f64(ck.x, n, c, lane, full, v);
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) {
        const uint64_t row = c * kTileRows + 64 * g + lane;
        const int64_t f = fetch_fixed_aligned_bits(v[g], ck.pexp, ck.view.dlen);
        if (full || row < n) {
          mn = min(mn, f);
          if (row != 0) mx = max(mx, f);
        }
      }
    }
    mn = wave_min_i64(mn);
    mx = wave_max_i64(mx);
    if (lane == 0) {
      smn[wave] = mn;
      smx[wave] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      part[2 * (uint64_t)i] = min(min(smn[0], smn[1]), min(smn[2], smn[3]));
      part[2 * (uint64_t)i + 1] = max(max(smx[0], smx[1]), max(smx[2], smx[3]));
    }
    __syncthreads();
  }
}

// One wave per chunk folds its items.
__global__ __launch_bounds__(256) void k_bminmax_final(const BEncChunk* __restrict__ chunks, uint32_t nchunks,
                                                       const int64_t* __restrict__ part, BMinMax* __restrict__ out) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t k = wave; k < nchunks; k += nwaves) {
    const BEncChunk& ck = chunks[k];
    int64_t mn = INT64_MAX, mx = INT64_MIN;
    for (uint32_t i = lane; i < ck.nitems; i += 64) {
      mn = min(mn, part[2 * (uint64_t)(ck.item0 + i)]);
      mx = max(mx, part[2 * (uint64_t)(ck.item0 + i) + 1]);
    }
    mn = wave_min_i64(mn);
    mx = wave_max_i64(mx);
    if (lane == 0) {
      BMinMax r;
      r.min_all = mn;
      r.max_rest = mx;
      r.first = ck.view.n ? fetch_fixed_aligned_bits(__double_as_longlong(ck.x[0]), ck.pexp, ck.view.dlen) : 0;
      r.pad = 0;
      out[k] = r;
    }
  }
}

// Batch encode, pass 2: offsets, byte planes and tail bit-planes of every chunk
// (the k_encode tile body on the item's chunk).
__global__ __launch_bounds__(256) void k_bencode(const BEncChunk* __restrict__ chunks, const BItem* __restrict__ items,
                                                 uint32_t nitems) {
  __shared__ uint64_t s_off[4][64 * kLdsRow];
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  uint64_t* so = s_off[wave];
  for (uint32_t i = blockIdx.x; i < nitems; i += gridDim.x) {
    const BItem it = items[i];
    const BEncChunk& ck = chunks[it.owner];
    const ChunkView& cv = ck.view;
    for (uint64_t c = it.c0 + wave; c < it.c1; c += 4) {
      uint64_t v[16];
      const bool full = tile_full(cv.n, c);
      load_tile_f64(ck.x, cv.n, c, lane, full, v);
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) {
        const int64_t f = fetch_fixed_aligned_bits(v[g], ck.pexp, cv.dlen);
        so[lds_slot(64 * g + lane)] =
            (full || c * kTileRows + 64 * g + lane < cv.n) ? (uint64_t)f - (uint64_t)cv.base : 0;
      }
      wave_lds_sync();
      uint64_t off[16];
#pragma unroll
      for (uint32_t b = 0; b < 16; ++b) off[b] = so[kLdsRow * lane + b];
      wave_lds_sync();
      for (uint32_t p = 0; p < cv.P; ++p) {
        const uint32_t sh = cv.F - 8 * (p + 1);
        uint32_t by[16];
#pragma unroll
        for (uint32_t b = 0; b < 16; ++b) by[b] = (uint32_t)((off[b] >> sh) & 0xFF);
        reinterpret_cast<uint4*>(cv.planes + p * cv.pitch + c * kTileRows)[lane] = pack_lane(by);
      }
      for (uint32_t j = 0; j < cv.r; ++j) {
        uint32_t h = 0;
#pragma unroll
        for (uint32_t b = 0; b < 16; ++b) h |= (uint32_t)((off[b] >> j) & 1ull) << b;
        cv.tail[(c * cv.r + j) * 64 + lane] = (uint16_t)h;
      }
    }
  }
}

hipError_t launch_batch_minmax(const BEncChunk* chunks, uint32_t nchunks, const BItem* items, uint32_t nitems,
                               int64_t* item_part, BMinMax* out, uint32_t grid, hipStream_t s) {
  if (nitems) hipLaunchKernelGGL(k_bminmax, dim3(grid), dim3(256), 0, s, chunks, items, nitems, item_part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bminmax_final, dim3((nchunks + 3) / 4 ? (nchunks + 3) / 4 : 1), dim3(256), 0, s, chunks,
                     nchunks, item_part, out);
  return hipGetLastError();
}

hipError_t launch_batch_encode(const BEncChunk* chunks, const BItem* items, uint32_t nitems, uint32_t grid,
                               hipStream_t s) {
  if (nitems) hipLaunchKernelGGL(k_bencode, dim3(grid), dim3(256), 0, s, chunks, items, nitems);
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Batch queries. Kernel 1 (k_bq_tiles) does the per-tile work of every item:
// filters write the tile bitmaps and count them, sums add the sub-column bytes,
// maxima record every tile's top byte, materializations write values. Kernel 2
// folds items per query; kernel 3 refines the maxima on the tiles that hold the
// top byte; kernel 4 folds those.

template <int MODE>
__device__ __forceinline__ uint32_t bq_filter_tile(const ChunkView& cv, const FilterTarget& tg, uint64_t c,
                                                   uint32_t lane, uint32_t first_planes) {
  const uint32_t valid = lane_valid16(cv.n, c, lane);
  uint32_t gt = 0, eq = valid;
  if (cv.P == 0) {
    uint32_t g2 = 0, e2 = valid;
    refine_tail<MODE>(cv, c, tg.tail, lane, g2, e2);
    return MODE == 0 ? (g2 & valid) : e2;
  }
  const uint4 w0 = load_plane(cv, 0, c, lane);
  uint4 w1 = make_uint4(0, 0, 0, 0);
  const bool two = first_planes > 1 && cv.P > 1;
  if (two) w1 = load_plane(cv, 1, c, lane);
  const Digit d0 = make_digit(tg.digit[0]);
  if (MODE == 0) gt = mask_gt(w0, d0);
  eq &= mask_eq(w0, d0);
  for (uint32_t p = 1; p < cv.P; ++p) {
    if (!__ballot(eq != 0)) break;
    const uint4 w = (p == 1 && two) ? w1 : load_plane(cv, p, c, eq ? lane : 0u);
    const Digit dp = make_digit(tg.digit[p]);
    if (MODE == 0) gt |= eq & mask_gt(w, dp);
    eq &= mask_eq(w, dp);
  }
  if (cv.r && __ballot(eq != 0)) refine_tail<MODE>(cv, c, tg.tail, lane, gt, eq);
  return MODE == 0 ? (gt & valid) : eq;
}

__device__ __forceinline__ uint64_t tile_offset_max(const ChunkView& cv, uint64_t c, uint32_t lane) {
  uint64_t off[16];
#pragma unroll
  for (uint32_t b = 0; b < 16; ++b) off[b] = 0;
  for (uint32_t p = 0; p < cv.P; ++p) {
    const uint4 w = load_plane(cv, p, c, lane);
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) off[b] = (off[b] << 8) | byte_at(w, b);
  }
  for (uint32_t j0 = cv.r; j0 > 0; --j0) {
    const uint32_t h = load_tail(cv, j0 - 1, c, lane);
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) off[b] = (off[b] << 1) | ((h >> b) & 1u);
  }
  uint64_t m = 0;
#pragma unroll
  for (uint32_t b = 0; b < 16; ++b) m = off[b] > m ? off[b] : m;   // rows >= n are zero
  return wave_max_u64(m);
}

__device__ __forceinline__ uint32_t tile_top_byte(const ChunkView& cv, uint64_t c, uint32_t lane) {
  if (cv.P == 0) return (uint32_t)tile_offset_max(cv, c, lane);   // F < 8: the whole value
  const uint4 w = load_plane(cv, 0, c, lane);
  const uint32_t mb = max(max(max_byte4(w.x), max_byte4(w.y)), max(max_byte4(w.z), max_byte4(w.w)));
  return wave_max_u32(mb);
}

__global__ __launch_bounds__(256) void k_bq_tiles(const BatchBuffers* __restrict__ bbp, BatchTuning tu) {
  __shared__ double s_val[4][64 * kLdsRow];
  __shared__ unsigned long long s_acc[4][kMaxPlanes + 1];
  const BatchBuffers bb = *bbp;
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  const uint32_t nitems = bb.hdr->nitems;
  for (uint32_t i = blockIdx.x; i < nitems; i += gridDim.x) {
    const BItem it = bb.items[i];
    const BQuery& q = bb.q[it.owner];
    const ChunkView& cv = bb.views[q.chunk];
    const uint32_t op = q.op;
    unsigned long long acc[kMaxPlanes + 1];
#pragma unroll
    for (uint32_t k = 0; k <= kMaxPlanes; ++k) acc[k] = 0;
    if (op == kBqGt || op == kBqEq) {
      uint16_t* out16 = reinterpret_cast<uint16_t*>(q.out);
      for (uint64_t c = it.c0 + wave; c < it.c1; c += 4) {
        uint32_t r;
        if (q.fill == 1) r = 0;
        else if (q.fill == 2) r = lane_valid16(cv.n, c, lane);
        else r = op == kBqGt ? bq_filter_tile<0>(cv, q.tg, c, lane, tu.first_planes)
                             : bq_filter_tile<1>(cv, q.tg, c, lane, tu.first_planes);
        if (c * 64 + lane < 4 * ((cv.n + 63) / 64)) out16[c * 64 + lane] = (uint16_t)r;
        acc[0] += __popc(r);
      }
    } else if (op == kBqSum) {
      for (uint64_t c = it.c0 + wave; c < it.c1; c += 4) {
        for (uint32_t p = 0; p < cv.P; ++p) {
          const uint4 w = load_plane(cv, p, c, lane);
          uint32_t a = __builtin_amdgcn_sad_u8(w.x, 0u, 0u);
          a = __builtin_amdgcn_sad_u8(w.y, 0u, a);
          a = __builtin_amdgcn_sad_u8(w.z, 0u, a);
          a = __builtin_amdgcn_sad_u8(w.w, 0u, a);
#pragma unroll
          for (uint32_t k = 0; k < kMaxPlanes; ++k) acc[k] += k == p ? a : 0u;
        }
        uint64_t t = 0;
        for (uint32_t j = 0; j < cv.r; ++j) t += (uint64_t)__popc(load_tail(cv, j, c, lane)) << j;
        acc[kMaxPlanes] += t;
      }
    } else if (op == kBqMax) {
      uint32_t best = 0;
      for (uint64_t c = it.c0 + wave; c < it.c1; c += 4) {
        const uint32_t tb = tile_top_byte(cv, c, lane);
        if (lane == 0) bb.tile_top[q.top + c] = (uint8_t)tb;
        best = max(best, tb);
      }
      acc[0] = best;
    } else {   // materialize
      double* sv = s_val[wave];
      const MatParams& mp = q.mp;
      double* outv = reinterpret_cast<double*>(q.out);
      for (uint64_t c = it.c0 + wave; c < it.c1; c += 4) {
        if (mp.mode == kMatRefPrecision) {
          // the reference's reduced-precision read: fields at stride tail_bits of the packed tail
          for (uint32_t g = 0; g < 16; ++g) {
            const uint64_t row = c * kTileRows + 64 * g + lane;
            if (row >= cv.n) continue;
            const uint32_t t = (uint32_t)((row % kTileRows) / 16), b = (uint32_t)(row % 16);
            uint64_t v = 0;
            for (uint32_t p = 0; p < mp.planes; ++p)
              v = (v << 8) | cv.planes[p * cv.pitch + c * kTileRows + 16 * t + 4 * (b & 3) + (b >> 2)];
            if (mp.tail_bits) {
              uint64_t tv = 0;
              for (uint32_t qb = 0; qb < mp.tail_bits; ++qb) {
                const uint64_t pos = row * mp.tail_bits + qb;
                const uint64_t rr = pos / cv.r;
                const uint32_t jj = (uint32_t)(pos % cv.r);
                uint32_t bit = 0;
                if (rr < cv.n) {
                  const uint64_t cc = rr / kTileRows;
                  const uint32_t tt = (uint32_t)((rr % kTileRows) / 16), bb2 = (uint32_t)(rr % 16);
                  bit = (cv.tail[(cc * cv.r + jj) * 64 + tt] >> bb2) & 1u;
                }
                tv |= (uint64_t)bit << qb;
              }
              v = (v << mp.tail_bits) | tv;
            }
            outv[row] = (double)(int64_t)((uint64_t)cv.base + v) * mp.inv_scale;
          }
          continue;
        }
        uint64_t off[16];
#pragma unroll
        for (uint32_t b = 0; b < 16; ++b) off[b] = 0;
        for (uint32_t p = 0; p < mp.planes; ++p) {
          const uint4 w = load_plane(cv, p, c, lane);
#pragma unroll
          for (uint32_t b = 0; b < 16; ++b) off[b] = (off[b] << 8) | byte_at(w, b);
        }
        const uint32_t lo = cv.r - mp.tail_bits;
        for (uint32_t k = mp.tail_bits; k > 0; --k) {
          const uint32_t h = load_tail(cv, lo + k - 1, c, lane);
#pragma unroll
          for (uint32_t b = 0; b < 16; ++b) off[b] = (off[b] << 1) | ((h >> b) & 1u);
        }
#pragma unroll
        for (uint32_t b = 0; b < 16; ++b) {
          const uint64_t o = (off[b] >> mp.drop) << mp.shift_back;
          sv[kLdsRow * lane + b] = (double)(int64_t)((uint64_t)cv.base + o) * mp.inv_scale;
        }
        wave_lds_sync();
        double* ot = outv + c * kTileRows + lane;
#pragma unroll
        for (uint32_t g = 0; g < 16; ++g)
          if (c * kTileRows + 64 * g + lane < cv.n) ot[64 * g] = sv[lds_slot(64 * g + lane)];
        wave_lds_sync();
      }
    }
    // fold the item's accumulators over the four waves
#pragma unroll
    for (uint32_t k = 0; k <= kMaxPlanes; ++k) {
      const unsigned long long v = op == kBqMax ? wave_max_u64(acc[k]) : wave_sum_u64(acc[k]);
      if (lane == 0) s_acc[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x <= kMaxPlanes) {
      const uint32_t k = threadIdx.x;
      unsigned long long v;
      if (op == kBqMax)
        v = max(max(s_acc[0][k], s_acc[1][k]), max(s_acc[2][k], s_acc[3][k]));
      else
        v = s_acc[0][k] + s_acc[1][k] + s_acc[2][k] + s_acc[3][k];
      bb.item_acc[(uint64_t)i * (kMaxPlanes + 1) + k] = v;
    }
    __syncthreads();
  }
}

// One wave per query: fold its items (filters: count, sums: sub-column sums,
// maxima: the top byte into v[1]).
__global__ __launch_bounds__(256) void k_bq_reduce(const BatchBuffers* __restrict__ bbp) {
  const BatchBuffers bb = *bbp;
  const uint32_t lane = __lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint32_t nq = bb.hdr->nq;
  for (uint64_t qi = wave; qi < nq; qi += nwaves) {
    const BQuery& q = bb.q[qi];
    unsigned long long v[kMaxPlanes + 1];
#pragma unroll
    for (uint32_t k = 0; k <= kMaxPlanes; ++k) v[k] = 0;
    for (uint32_t i = lane; i < q.nitems; i += 64) {
      const unsigned long long* a = bb.item_acc + (uint64_t)(q.item0 + i) * (kMaxPlanes + 1);
#pragma unroll
      for (uint32_t k = 0; k <= kMaxPlanes; ++k) v[k] = q.op == kBqMax ? max(v[k], a[k]) : v[k] + a[k];
    }
#pragma unroll
    for (uint32_t k = 0; k <= kMaxPlanes; ++k) v[k] = q.op == kBqMax ? wave_max_u64(v[k]) : wave_sum_u64(v[k]);
    if (lane == 0) {
      BResult r;
#pragma unroll
      for (uint32_t k = 0; k <= kMaxPlanes; ++k) r.v[k] = v[k];
      if (q.op == kBqMax) {
        r.v[1] = v[0];   // top byte; v[0] is filled by the refine fold
        r.v[0] = 0;
      }
      bb.res[qi] = r;
    }
  }
}

// Maxima: tiles whose top byte equals the query's top byte give the max offset.
__global__ __launch_bounds__(256) void k_bq_max_refine(const BatchBuffers* __restrict__ bbp) {
  __shared__ unsigned long long s_m[4];
  const BatchBuffers bb = *bbp;
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  const uint32_t nitems = bb.hdr->nitems;
  for (uint32_t i = blockIdx.x; i < nitems; i += gridDim.x) {
    const BItem it = bb.items[i];
    const BQuery& q = bb.q[it.owner];
    if (q.op != kBqMax) continue;   // uniform over the workgroup
    const ChunkView& cv = bb.views[q.chunk];
    const uint32_t top = (uint32_t)bb.res[it.owner].v[1];
    unsigned long long best = 0;    // max offset + 1
    for (uint64_t c = it.c0 + wave; c < it.c1; c += 4) {
      if (bb.tile_top[q.top + c] != top) continue;
      const uint64_t m = tile_offset_max(cv, c, lane);
      best = max(best, (unsigned long long)m + 1);
    }
    if (lane == 0) s_m[wave] = best;
    __syncthreads();
    if (threadIdx.x == 0)
      bb.item_acc[(uint64_t)i * (kMaxPlanes + 1) + 1] = max(max(s_m[0], s_m[1]), max(s_m[2], s_m[3]));
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_bq_max_final(const BatchBuffers* __restrict__ bbp) {
  const BatchBuffers bb = *bbp;
  const uint32_t lane = __lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint32_t nq = bb.hdr->nq;
  for (uint64_t qi = wave; qi < nq; qi += nwaves) {
    const BQuery& q = bb.q[qi];
    if (q.op != kBqMax) continue;
    unsigned long long m = 0;
    for (uint32_t i = lane; i < q.nitems; i += 64)
      m = max(m, bb.item_acc[(uint64_t)(q.item0 + i) * (kMaxPlanes + 1) + 1]);
    m = wave_max_u64(m);
    if (lane == 0) bb.res[qi].v[0] = m ? m - 1 : 0;
  }
}

hipError_t launch_batch_queries(const BatchBuffers* bb, const BatchTuning& tu, hipStream_t s) {
  hipLaunchKernelGGL(k_bq_tiles, dim3(tu.grid), dim3(256), 0, s, bb, tu);
  hipLaunchKernelGGL(k_bq_reduce, dim3(256), dim3(256), 0, s, bb);
  hipLaunchKernelGGL(k_bq_max_refine, dim3(tu.grid), dim3(256), 0, s, bb);
  hipLaunchKernelGGL(k_bq_max_final, dim3(256), dim3(256), 0, s, bb);
  return hipGetLastError();
}

}  // namespace buffgpu
