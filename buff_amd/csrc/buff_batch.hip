// Batched kernels (buff_batch.h) for gfx950: persistent workgroups take items
// (owner, tile range) from a list in device memory; the per-tile work is that of
// the single-chunk kernels (buff_kernels.hip) on the owner's chunk view.
#include <hip/hip_runtime.h>

#include "buff_batch.h"
#include "buff_major_device.h"
#include "buff_common.h"
#include "buff_device.h"

namespace buffgpu {

// ----------------------------------------------------------------------------
// Batch encode, pass 1: per item min over all rows and max over the rows other
// than the chunk's row 0 (split_double.rs:321-330 keeps row 0 out of the max).

__global__ __launch_bounds__(256) void k_bminmax(const BEncChunk* __restrict__ chunks, const BItem* __restrict__ items,
                                                 uint32_t nitems, int64_t* __restrict__ part) {
  __shared__ int64_t smn[4], smx[4];
  const uint32_t lane = __lane_id(), wave = wave_in_wg();
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
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
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
  const uint32_t lane = __lane_id(), wave = wave_in_wg();
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
      store_tile_planes(off, cv.P, cv.r, cv.planes, cv.pitch, cv.tail, c, lane);
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
// Batch queries. One wave works on one item at a time (items of different
// queries interleave freely; the item loops hold no workgroup barrier):
//   k_bq_fetch  copies the header and query descriptors from pinned host memory;
//   k_bq_items  writes the work items from the per-query item ranges;
//   k_bq_scan   items [0, nscan): filters write tile bitmaps and count them,
//               sums add the sub-column bytes and tail bits, maxima record every
//               tile's top byte;
//   k_bq_reduce one wave per query folds its items' partials;
//   k_bq_max_refine / k_bq_max_final: the tiles holding a max query's top byte
//               give its max offset;
//   k_bq_mat    items [nscan, nitems): materializations, staged through LDS for
//               coalesced f64 stores (kept apart so the scan kernel holds no LDS
//               and runs at full occupancy);
//   k_bq_publish copies the results to pinned host memory.
// Per-tile loads are issued together: a filter reads the top byte plane, then
// every remaining plane and tail bit-plane of the tied rows at once; sums and
// materializations read all their planes before combining them.

struct WaveSlot {
  uint32_t lane;
  uint32_t local;    // wave index inside the workgroup
  uint32_t first;    // this wave's first item
  uint32_t stride;   // waves in the grid
};
__device__ __forceinline__ WaveSlot wave_slot() {
  WaveSlot w;
  w.lane = __lane_id();
  w.local = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t per_wg = blockDim.x >> 6;
  w.first = blockIdx.x * per_wg + w.local;
  w.stride = gridDim.x * per_wg;
  return w;
}

constexpr int kPlanesPerRound = 4;   // plane loads in flight per wave and round (register budget)

// Planes P0 .. NP-1 and the tail bit-planes of tile c, for the rows still in
// `eq` (gt gains the rows above the target), in rounds whose loads are all in
// flight together; the tail bit-planes travel with the last round. Lanes
// without a live row load the first live lane's bytes instead, which costs no
// extra HBM lines.
template <int MODE, int P0, int NP>
__device__ __forceinline__ void filter_rest(const ChunkView& cv, const FilterTarget& tg, uint64_t c, uint32_t lane,
                                            uint32_t& gt, uint32_t& eq) {
  // plane 1 alone (rows that tie on the top byte often differ on the next),
  // then rounds of up to four planes
  constexpr int P1 = P0 == 1 ? (NP > 2 ? 2 : NP) : (NP - P0 > kPlanesPerRound ? P0 + kPlanesPerRound : NP);
  constexpr bool last = P1 == NP;
  constexpr int K = P1 > P0 ? P1 - P0 : 1;
  const uint32_t src = refine_src(eq, lane);
  const uint32_t r = last ? cv.r : 0u;
  uint4 w[K];
#pragma unroll
  for (int p = P0; p < P1; ++p) w[p - P0] = load_plane(cv, p, c, src);
  uint32_t h[7];
#pragma unroll
  for (uint32_t j = 0; j < 7; ++j) h[j] = j < r ? load_tail(cv, j, c, src) : 0u;
#pragma unroll
  for (int p = P0; p < P1; ++p) {
    const Digit dp = make_digit(tg.digit[p]);
    if (MODE == 0) gt |= eq & mask_gt(w[p - P0], dp);
    eq &= mask_eq(w[p - P0], dp);
  }
  if constexpr (!last) {
    if (__ballot(eq != 0)) filter_rest<MODE, P1, NP>(cv, tg, c, lane, gt, eq);
  } else {
#pragma unroll
    for (int j = 6; j >= 0; --j) {
      if ((uint32_t)j >= r) continue;
      if ((tg.tail >> j) & 1) {
        eq &= h[j];
      } else {
        if (MODE == 0) gt |= eq & h[j];
        eq &= ~h[j];
      }
    }
  }
}

template <int MODE, int P0>
__device__ __forceinline__ void filter_rest_np(const ChunkView& cv, const FilterTarget& tg, uint64_t c, uint32_t lane,
                                               uint32_t& gt, uint32_t& eq) {
  switch (cv.P) {
    case 0: filter_rest<MODE, 0, 0>(cv, tg, c, lane, gt, eq); break;
    case 1: filter_rest<MODE, P0, 1>(cv, tg, c, lane, gt, eq); break;
    case 2: filter_rest<MODE, P0, 2>(cv, tg, c, lane, gt, eq); break;
    case 3: filter_rest<MODE, P0, 3>(cv, tg, c, lane, gt, eq); break;
    case 4: filter_rest<MODE, P0, 4>(cv, tg, c, lane, gt, eq); break;
    case 5: filter_rest<MODE, P0, 5>(cv, tg, c, lane, gt, eq); break;
    case 6: filter_rest<MODE, P0, 6>(cv, tg, c, lane, gt, eq); break;
    case 7: filter_rest<MODE, P0, 7>(cv, tg, c, lane, gt, eq); break;
    default: filter_rest<MODE, P0, 8>(cv, tg, c, lane, gt, eq); break;
  }
}

// One filter item: tiles [c0, c1). Returns the rows set. With the top plane
// first, two tiles' top planes are in flight together and each tile with tied
// rows then reads the rest; otherwise (every plane in one pass, tail-only
// chunks, or a host-decided fill) one tile at a time.
template <int MODE>
__device__ __forceinline__ void filter_store(const ChunkView& cv, BUFF_GLOBAL uint16_t* out16, uint64_t c, uint32_t lane,
                                             uint32_t res, uint32_t& count) {
  const uint64_t idx = c * 64 + lane;
  if (idx < 4 * ((cv.n + 63) / 64)) out16[idx] = (uint16_t)res;
  count += __popc(res);
}

template <int MODE>
__device__ __forceinline__ uint32_t filter_item(const ChunkView& cv, const BQuery& q, const BItem& it, uint32_t lane,
                                                bool all_first) {
  BUFF_GLOBAL uint16_t* out16 = gmem_w(reinterpret_cast<uint16_t*>(q.out));
  uint32_t count = 0;
  if (q.fill || cv.P == 0 || all_first) {
#pragma unroll 1
    for (uint64_t c = it.c0; c < it.c1; ++c) {
      const uint32_t valid = lane_valid16(cv.n, c, lane);
      uint32_t gt = 0, eq = valid;
      if (q.fill) gt = eq = q.fill == 2 ? valid : 0u;
      else filter_rest_np<MODE, 0>(cv, q.tg, c, lane, gt, eq);
      filter_store<MODE>(cv, out16, c, lane, MODE == 0 ? (gt & valid) : eq, count);
    }
    return count;
  }
  const Digit d0 = make_digit(q.tg.digit[0]);
  const bool rest = cv.P > 1 || cv.r;
#pragma unroll 1
  for (uint64_t c = it.c0; c < it.c1; c += 2) {
    const bool two = c + 1 < it.c1;
    const uint64_t c1 = two ? c + 1 : c;
    const uint4 w0 = load_plane(cv, 0, c, lane);
    const uint4 w1 = load_plane(cv, 0, c1, lane);
    const uint32_t v0 = lane_valid16(cv.n, c, lane), v1 = lane_valid16(cv.n, c1, lane);
    uint32_t g0 = MODE == 0 ? mask_gt(w0, d0) : 0u, e0 = v0 & mask_eq(w0, d0);
    uint32_t g1 = MODE == 0 ? mask_gt(w1, d0) : 0u, e1 = v1 & mask_eq(w1, d0);
    if (rest && __ballot(e0 != 0)) filter_rest_np<MODE, 1>(cv, q.tg, c, lane, g0, e0);
    filter_store<MODE>(cv, out16, c, lane, MODE == 0 ? (g0 & v0) : e0, count);
    if (two) {
      if (rest && __ballot(e1 != 0)) filter_rest_np<MODE, 1>(cv, q.tg, c1, lane, g1, e1);
      filter_store<MODE>(cv, out16, c1, lane, MODE == 0 ? (g1 & v1) : e1, count);
    }
  }
  return count;
}

// Sum item: per-lane byte sums of each plane and weighted tail-bit counts
// (u32 per lane: at most kMaxTilesPerItem * 16 rows * 255 per item and plane).
// Planes are read in rounds of up to four loads in flight.
template <int P0, int NP>
__device__ __forceinline__ void sum_planes(const ChunkView& cv, uint64_t c, uint32_t lane,
                                           uint32_t (&acc)[kMaxPlanes + 1]) {
  constexpr int P1 = NP - P0 > kPlanesPerRound ? P0 + kPlanesPerRound : NP;
  constexpr bool last = P1 == NP;
  constexpr int K = P1 > P0 ? P1 - P0 : 1;
  const uint32_t r = last ? cv.r : 0u;
  uint4 w[K];
#pragma unroll
  for (int p = P0; p < P1; ++p) w[p - P0] = load_plane(cv, p, c, lane);
  uint32_t h[7];
#pragma unroll
  for (uint32_t j = 0; j < 7; ++j) h[j] = j < r ? load_tail(cv, j, c, lane) : 0u;
#pragma unroll
  for (int p = P0; p < P1; ++p) {
    uint32_t a = __builtin_amdgcn_sad_u8(w[p - P0].x, 0u, 0u);
    a = __builtin_amdgcn_sad_u8(w[p - P0].y, 0u, a);
    a = __builtin_amdgcn_sad_u8(w[p - P0].z, 0u, a);
    acc[p] = __builtin_amdgcn_sad_u8(w[p - P0].w, 0u, a + acc[p]);
  }
#pragma unroll
  for (uint32_t j = 0; j < 7; ++j) acc[kMaxPlanes] += (uint32_t)__popc(h[j]) << j;
  if constexpr (!last) sum_planes<P1, NP>(cv, c, lane, acc);
}

template <int NP>
__device__ __forceinline__ void sum_item(const ChunkView& cv, const BItem& it, uint32_t lane,
                                         uint32_t (&acc)[kMaxPlanes + 1]) {
  #pragma unroll 1
  for (uint64_t c = it.c0; c < it.c1; ++c) sum_planes<0, NP>(cv, c, lane, acc);
}

// Max stored offset among the lane's 16 rows (rows >= n are zero): the max
// byte plane by plane among the rows tying so far, then bit by bit on the tail.
__device__ __forceinline__ uint64_t lane_offset_max(const ChunkView& cv, uint64_t c, uint32_t lane) {
  uint32_t cand = 0xFFFFu;
  uint64_t m = 0;
#pragma unroll 1
  for (uint32_t p = 0; p < cv.P; ++p) {
    const uint4 w = load_plane(cv, p, c, lane);
    const uint32_t mb = masked_max_plus1(w, cand) - 1u;   // cand is never empty
    cand &= mask_eq(w, make_digit(mb));
    m = (m << 8) | mb;
  }
#pragma unroll 1
  for (uint32_t j0 = cv.r; j0 > 0; --j0) {
    const uint32_t h = load_tail(cv, j0 - 1, c, lane) & cand;
    m <<= 1;
    if (h) {
      m |= 1u;
      cand = h;
    }
  }
  return m;
}

// F < 8: the lane's max value from the tail bit-planes alone.
__device__ __forceinline__ uint32_t lane_tail_max(const ChunkView& cv, uint64_t c, uint32_t lane) {
  uint32_t cand = 0xFFFFu, m = 0;
#pragma unroll 1
  for (uint32_t j0 = cv.r; j0 > 0; --j0) {
    const uint32_t h = load_tail(cv, j0 - 1, c, lane) & cand;
    m <<= 1;
    if (h) {
      m |= 1u;
      cand = h;
    }
  }
  return m;
}

__device__ __forceinline__ uint64_t tile_offset_max(const ChunkView& cv, uint64_t c, uint32_t lane) {
  return wave_max_u64(lane_offset_max(cv, c, lane));
}

__device__ __forceinline__ uint32_t lane_max_byte(const uint4& w) {
  return max(max(max_byte4(w.x), max_byte4(w.y)), max(max_byte4(w.z), max_byte4(w.w)));
}

// Max item: every tile's top byte (the whole value when F < 8), four tiles in flight.
__device__ __forceinline__ uint32_t max_top_item(const ChunkView& cv, const BQuery& q, const BItem& it,
                                                 uint8_t* tile_top, uint32_t lane) {
  uint32_t best = 0;
  if (cv.P == 0) {
    #pragma unroll 1
    for (uint64_t c = it.c0; c < it.c1; ++c) {
      const uint32_t tb = wave_max_u32(lane_tail_max(cv, c, lane));
      if (lane == 0) tile_top[q.top + c] = (uint8_t)tb;
      best = max(best, tb);
    }
    return best;
  }
  #pragma unroll 1
  for (uint64_t c = it.c0; c < it.c1; c += 4) {
    uint4 w[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) w[u] = load_plane(cv, 0, min(c + u, it.c1 - 1), lane);
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const uint32_t tb = wave_max_u32(lane_max_byte(w[u]));
      if (c + u < it.c1) {
        if (lane == 0) tile_top[q.top + c + u] = (uint8_t)tb;
        best = max(best, tb);
      }
    }
  }
  return best;
}

__global__ __launch_bounds__(256) void k_bq_scan(const BatchBuffers* __restrict__ bbp, BatchTuning tu) {
  const BatchBuffers bb = cload(bbp);
  const WaveSlot ws = wave_slot();
  const uint32_t lane = ws.lane;
  const uint32_t nscan = cmem(bb.hdr)->nscan;
  const bool all_first = tu.first_planes >= kMaxPlanes;
  #pragma unroll 1
  for (uint32_t i = ws.first; i < nscan; i += ws.stride) {
    const BItem it = cload(bb.items, i);
    const BQuery q = cload(bb.q, it.owner);
    if (q.sparse) continue;   // k_bq_major
    const ChunkView cv = cload(bb.views, it.chunk);
    const uint32_t op = q.op;
    unsigned long long* dst = bb.item_acc + (uint64_t)i * (kMaxPlanes + 1);
    if (op == kBqGt || op == kBqEq) {
      const uint32_t cnt = op == kBqGt ? filter_item<0>(cv, q, it, lane, all_first)
                                       : filter_item<1>(cv, q, it, lane, all_first);
      const uint32_t tot = wave_total_u32(cnt);
      if (lane == 0) dst[0] = tot;
    } else if (op == kBqSum) {
      uint32_t acc[kMaxPlanes + 1];
#pragma unroll
      for (uint32_t k = 0; k <= kMaxPlanes; ++k) acc[k] = 0;
      switch (cv.P) {
        case 0: sum_item<0>(cv, it, lane, acc); break;
        case 1: sum_item<1>(cv, it, lane, acc); break;
        case 2: sum_item<2>(cv, it, lane, acc); break;
        case 3: sum_item<3>(cv, it, lane, acc); break;
        case 4: sum_item<4>(cv, it, lane, acc); break;
        case 5: sum_item<5>(cv, it, lane, acc); break;
        case 6: sum_item<6>(cv, it, lane, acc); break;
        case 7: sum_item<7>(cv, it, lane, acc); break;
        default: sum_item<8>(cv, it, lane, acc); break;
      }
#pragma unroll
      for (uint32_t k = 0; k <= kMaxPlanes; ++k) {
        const bool used = k < cv.P || k == kMaxPlanes;
        const uint32_t tot = used ? wave_total_u32(acc[k]) : 0u;
        if (lane == 0) dst[k] = tot;
      }
    } else if (op == kBqMax) {
      const uint32_t best = max_top_item(cv, q, it, bb.tile_top, lane);
      if (lane == 0) dst[0] = best;
    }
  }
}

// Copies the batch's header and query descriptors from pinned host memory to
// device memory (the host only writes them and launches the graph: no runtime
// copy, which could queue behind another slot's transfers).
__global__ __launch_bounds__(256) void k_bq_fetch(const BatchBuffers* __restrict__ bbp) {
  const BatchBuffers bb = cload(bbp);
  const BUFF_GLOBAL uint32_t* hin = gmem(reinterpret_cast<const uint32_t*>(bb.host_in));
  BUFF_GLOBAL uint32_t* hdr = gmem_w(reinterpret_cast<uint32_t*>(const_cast<BHeader*>(bb.hdr)));
  const uint32_t nq = hin[0];
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, nthreads = (uint64_t)gridDim.x * blockDim.x;
  if (tid < sizeof(BHeader) / 4) hdr[tid] = hin[tid];
  const uint64_t nwords = (uint64_t)nq * (sizeof(BQuery) / 4);
  const BUFF_GLOBAL uint32_t* src = hin + kHostQueries / 4;
  BUFF_GLOBAL uint32_t* dst = gmem_w(reinterpret_cast<uint32_t*>(const_cast<BQuery*>(bb.q)));
  for (uint64_t i = tid; i < nwords; i += nthreads) dst[i] = src[i];
}

// Writes the batch's results to pinned host memory, visible to the host once
// the batch's completion event has fired.
__global__ __launch_bounds__(256) void k_bq_publish(const BatchBuffers* __restrict__ bbp) {
  const BatchBuffers bb = cload(bbp);
  const uint32_t nq = cmem(bb.hdr)->nq;
  const uint64_t nwords = (uint64_t)nq * (sizeof(BResult) / 8);
  const BUFF_GLOBAL uint64_t* src = gmem(reinterpret_cast<const uint64_t*>(bb.res));
  BUFF_GLOBAL uint64_t* dst = gmem_w(reinterpret_cast<uint64_t*>(bb.host_res));
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, nthreads = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = tid; i < nwords; i += nthreads) dst[i] = src[i];
  __threadfence_system();
}

// One workgroup per query: writes the query's work items (tile ranges of
// tiles_per_item tiles over its chunk) at the range the host assigned it.
__global__ __launch_bounds__(256) void k_bq_items(const BatchBuffers* __restrict__ bbp) {
  const BatchBuffers bb = cload(bbp);
  const uint32_t nq = cmem(bb.hdr)->nq;
  const uint64_t T = cmem(bb.hdr)->tiles_per_item;
  for (uint32_t qi = blockIdx.x; qi < nq; qi += gridDim.x) {
    const BQuery q = cload(bb.q, qi);
    if (!q.nitems) continue;
    const uint64_t ntiles = cmem(&bb.views[q.chunk].ntiles)[0];
    for (uint32_t j = threadIdx.x; j < q.nitems; j += blockDim.x) {
      const uint64_t c0 = (uint64_t)j * T;
      BUFF_GLOBAL uint64_t* w = gmem_w(reinterpret_cast<uint64_t*>(bb.items + q.item0 + j));
      w[0] = (uint64_t)qi | ((uint64_t)q.chunk << 32);   // owner, chunk
      w[1] = c0;
      w[2] = c0 + T < ntiles ? c0 + T : ntiles;
    }
  }
}

// One workgroup per query folds its items' partials (filters: count, sums:
// sub-column sums, maxima: the top byte into v[1]); each thread keeps four
// partial rows in flight. Materializations have nothing to fold.
__global__ __launch_bounds__(256) void k_bq_reduce(const BatchBuffers* __restrict__ bbp) {
  __shared__ unsigned long long s_part[4][kMaxPlanes + 1];
  const BatchBuffers bb = cload(bbp);
  const uint32_t lane = __lane_id(), wave = wave_in_wg();
  const uint32_t nq = cmem(bb.hdr)->nq;
  for (uint32_t qi = blockIdx.x; qi < nq; qi += gridDim.x) {
    const BQuery q = cload(bb.q, qi);
    const bool is_max = q.op == kBqMax;
    const uint32_t nk = q.op == kBqSum ? kMaxPlanes + 1 : (q.op == kBqMat ? 0u : 1u);
    unsigned long long v[kMaxPlanes + 1];
#pragma unroll
    for (uint32_t k = 0; k <= kMaxPlanes; ++k) v[k] = 0;
    const unsigned long long* acc = bb.item_acc + (uint64_t)q.item0 * (kMaxPlanes + 1);
    for (uint32_t i0 = threadIdx.x; i0 < q.nitems; i0 += 4 * blockDim.x) {
#pragma unroll
      for (uint32_t k = 0; k <= kMaxPlanes; ++k) {
        if (k >= nk) continue;
        unsigned long long a[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
          const uint32_t i = i0 + u * blockDim.x;
          a[u] = i < q.nitems ? gmem(acc)[(uint64_t)i * (kMaxPlanes + 1) + k] : 0ull;
        }
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) v[k] = is_max ? max(v[k], a[u]) : v[k] + a[u];
      }
    }
#pragma unroll
    for (uint32_t k = 0; k <= kMaxPlanes; ++k) {
      if (k >= nk) continue;
      v[k] = is_max ? wave_max_u64(v[k]) : wave_sum_u64(v[k]);
      if (lane == 0) s_part[wave][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      BResult r;
#pragma unroll
      for (uint32_t k = 0; k <= kMaxPlanes; ++k) {
        unsigned long long t = 0;
        if (k < nk) {
          t = s_part[0][k];
          for (uint32_t w = 1; w < (blockDim.x >> 6); ++w) t = is_max ? max(t, s_part[w][k]) : t + s_part[w][k];
        }
        r.v[k] = t;
      }
      if (is_max) {
        r.v[1] = r.v[0];   // top byte; v[0] is filled by the refine fold
        r.v[0] = 0;
      }
      bb.res[qi] = r;
    }
    __syncthreads();
  }
}

// Maxima: tiles whose top byte equals the query's top byte give the max offset.
__global__ __launch_bounds__(256) void k_bq_max_refine(const BatchBuffers* __restrict__ bbp) {
  const BatchBuffers bb = cload(bbp);
  const WaveSlot ws = wave_slot();
  const uint32_t lane = ws.lane;
  const uint32_t nscan = cmem(bb.hdr)->nscan;
  #pragma unroll 1
  for (uint32_t i = ws.first; i < nscan; i += ws.stride) {
    const BItem it = cload(bb.items, i);
    const BQuery q = cload(bb.q, it.owner);
    if (q.op != kBqMax) continue;
    const ChunkView cv = cload(bb.views, it.chunk);
    const uint32_t top = (uint32_t)cmem(bb.res[it.owner].v)[1];   // written by k_bq_reduce
    unsigned long long best = 0;    // max offset + 1
    #pragma unroll 1
    for (uint64_t c = it.c0; c < it.c1; ++c) {
      if (cmem(bb.tile_top)[q.top + c] != top) continue;   // written by k_bq_scan
      const uint64_t m = tile_offset_max(cv, c, lane);
      best = max(best, (unsigned long long)m + 1);
    }
    if (lane == 0) bb.item_acc[(uint64_t)i * (kMaxPlanes + 1) + 1] = best;
  }
}

__global__ __launch_bounds__(256) void k_bq_max_final(const BatchBuffers* __restrict__ bbp) {
  const BatchBuffers bb = cload(bbp);
  const WaveSlot ws = wave_slot();
  const uint32_t lane = ws.lane;
  const uint32_t nq = cmem(bb.hdr)->nq;
  for (uint32_t qi = ws.first; qi < nq; qi += ws.stride) {
    const BQuery q = cload(bb.q, qi);
    if (q.op != kBqMax) continue;
    unsigned long long m = 0;
    for (uint32_t i = lane; i < q.nitems; i += 64)
      m = max(m, bb.item_acc[(uint64_t)(q.item0 + i) * (kMaxPlanes + 1) + 1]);
    m = wave_max_u64(m);
    if (lane == 0) bb.res[qi].v[0] = m ? m - 1 : 0;
  }
}

// Materialize one tile from the top NP planes and mp.tail_bits tail bits below
// them (every load in flight before the values are built).
template <int NP>
__device__ __forceinline__ void mat_tile(const ChunkView& cv, const MatParams& mp, uint64_t c, uint32_t lane,
                                         double* sv, double* outv) {
  uint4 w[NP > 0 ? NP : 1];
#pragma unroll
  for (int p = 0; p < NP; ++p) w[p] = load_plane(cv, p, c, lane);
  const uint32_t tb = mp.tail_bits, lo = cv.r - tb;
  uint32_t h[7];   // h[k]: tail bit-plane lo + k
#pragma unroll
  for (uint32_t k = 0; k < 7; ++k) h[k] = k < tb ? load_tail(cv, lo + k, c, lane) : 0u;
  uint64_t off[16];
#pragma unroll
  for (uint32_t b = 0; b < 16; ++b) off[b] = 0;
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) off[b] = (off[b] << 8) | byte_at(w[p], b);
#pragma unroll
  for (int k = 6; k >= 0; --k) {
    if ((uint32_t)k >= tb) continue;
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) off[b] = (off[b] << 1) | ((h[k] >> b) & 1u);
  }
#pragma unroll
  for (uint32_t b = 0; b < 16; ++b) {
    const uint64_t o = (off[b] >> mp.drop) << mp.shift_back;
    sv[kLdsRow * lane + b] = (double)(int64_t)((uint64_t)cv.base + o) * mp.inv_scale;
  }
  wave_lds_sync();
  double* ot = outv + c * kTileRows + lane;
  if (tile_full(cv.n, c)) {
#pragma unroll
    for (uint32_t g = 0; g < 16; ++g) store_nt(ot + 64 * g, sv[lds_slot(64 * g + lane)]);
  } else {
#pragma unroll
    for (uint32_t g = 0; g < 16; ++g)
      if (c * kTileRows + 64 * g + lane < cv.n) store_nt(ot + 64 * g, sv[lds_slot(64 * g + lane)]);
  }
  wave_lds_sync();
}

// The reference's reduced-precision read: fields of width tail_bits at stride
// tail_bits of the packed reference tail (rows one by one).
__device__ void mat_tile_ref(const ChunkView& cv, const MatParams& mp, uint64_t c, uint32_t lane, double* outv) {
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
}

__global__ __launch_bounds__(256) void k_bq_mat(const BatchBuffers* __restrict__ bbp) {
  __shared__ double s_val[4][64 * kLdsRow];
  const BatchBuffers bb = cload(bbp);
  const WaveSlot ws = wave_slot();
  const uint32_t lane = ws.lane;
  const uint32_t nscan = cmem(bb.hdr)->nscan, nitems = cmem(bb.hdr)->nitems;
  double* sv = s_val[ws.local];
  #pragma unroll 1
  for (uint32_t i = nscan + ws.first; i < nitems; i += ws.stride) {
    const BItem it = cload(bb.items, i);
    const BQuery q = cload(bb.q, it.owner);
    if (q.sparse) continue;   // k_bq_major_mat
    const ChunkView cv = cload(bb.views, it.chunk);
    const MatParams& mp = q.mp;
    double* outv = reinterpret_cast<double*>(q.out);
    #pragma unroll 1
    for (uint64_t c = it.c0; c < it.c1; ++c) {
      if (mp.mode == kMatRefPrecision) {
        mat_tile_ref(cv, mp, c, lane, outv);
        continue;
      }
      switch (mp.planes) {
        case 0: mat_tile<0>(cv, mp, c, lane, sv, outv); break;
        case 1: mat_tile<1>(cv, mp, c, lane, sv, outv); break;
        case 2: mat_tile<2>(cv, mp, c, lane, sv, outv); break;
        case 3: mat_tile<3>(cv, mp, c, lane, sv, outv); break;
        case 4: mat_tile<4>(cv, mp, c, lane, sv, outv); break;
        case 5: mat_tile<5>(cv, mp, c, lane, sv, outv); break;
        case 6: mat_tile<6>(cv, mp, c, lane, sv, outv); break;
        case 7: mat_tile<7>(cv, mp, c, lane, sv, outv); break;
        default: mat_tile<8>(cv, mp, c, lane, sv, outv); break;
      }
    }
  }
}

// ---- queries on buff-major chunks with sparse columns -------------------------
// Filters run the single-chunk filter's tile function (major_filter_group:
// the reference's rules, or with BUFF_SEMANTIC the intended answer); sums add
// up the dense planes and tail like a plain chunk and each sparse column's
// outlier bytes (the host adds major * (n - outliers) and folds the columns);
// materializations build the offsets with major_tile_offsets.

template <int MODE, int S>
__device__ __forceinline__ uint32_t major_filter_item(const MajorPrefix<S>& mv, const BQuery& q, const BItem& it,
                                                      uint32_t lane, bool sem, unsigned long long* dep) {
  BUFF_GLOBAL uint16_t* out16 = gmem_w(reinterpret_cast<uint16_t*>(q.out));
  MajorTarget tg;
#pragma unroll
  for (uint32_t k = 0; k < kMaxPlanes; ++k) tg.digit[k] = q.tg.digit[k];
  tg.tail = q.tg.tail;
  uint32_t count = 0;
  // two tiles in flight per step (the second may lie past the item: computed, not stored)
  #pragma unroll 1
  for (uint64_t c = it.c0; c < it.c1; c += 2) {
    uint32_t res[2];
    major_filter_group<MODE, S, 2>(mv, tg, sem, c, lane, dep, res);
    filter_store<MODE>(mv.dense, out16, c, lane, res[0], count);
    if (c + 1 < it.c1) filter_store<MODE>(mv.dense, out16, c + 1, lane, res[1], count);
  }
  return count;
}

// Per-lane outlier byte sums of every sparse column over tiles [c0, c1): their
// outlier bytes are contiguous, [prefix[c0], prefix[c1]) of the compacted array,
// read as aligned dwords with the bytes outside the range masked off (the
// column's buffer is padded past its last byte).
template <int S>
__device__ __forceinline__ void major_outlier_sums(const MajorPrefix<S>& mv, const BItem& it, uint32_t lane,
                                                   uint32_t (&osum)[S]) {
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const uint32_t a = mv.col[j].prefix[it.c0], b = mv.col[j].prefix[it.c1];
    const uint32_t* w = reinterpret_cast<const uint32_t*>(mv.col[j].outl);
    for (uint32_t k = (a >> 2) + lane; 4 * k < b; k += 64) {
      uint32_t v = w[k];
      if (4 * k < a) v &= 0xFFFFFFFFu << (8 * (a - 4 * k));        // bytes before a
      if (4 * k + 4 > b) v &= 0xFFFFFFFFu >> (8 * (4 * k + 4 - b));  // bytes from b on
      osum[j] = __builtin_amdgcn_sad_u8(v, 0u, osum[j]);
    }
  }
}

// One scan item of a query on a buff-major chunk with S sparse columns (only
// the view's first S columns are read into registers).
template <int S>
__device__ __forceinline__ void major_scan_item(const BatchBuffers& bb, const BQuery& q, const BItem& it, uint32_t i,
                                                uint32_t lane, unsigned long long* dep) {
  const MajorPrefix<S> mv = cload(reinterpret_cast<const MajorPrefix<S>*>(bb.majors + it.chunk));
  const bool sem = (q.sparse & kBqSparseSem) != 0;
  unsigned long long* dst = bb.item_acc + (uint64_t)i * (kMaxPlanes + 1);
  if (q.op == kBqGt || q.op == kBqEq) {
    const uint32_t cnt = q.op == kBqGt ? major_filter_item<0, S>(mv, q, it, lane, sem, dep)
                                       : major_filter_item<1, S>(mv, q, it, lane, sem, dep);
    const uint32_t tot = wave_total_u32(cnt);
    if (lane == 0) dst[0] = tot;
  } else if (q.op == kBqSum) {
    const ChunkView& cv = mv.dense;
    uint32_t acc[kMaxPlanes + 1];
#pragma unroll
    for (uint32_t k = 0; k <= kMaxPlanes; ++k) acc[k] = 0;
    switch (cv.P) {
      case 0: sum_item<0>(cv, it, lane, acc); break;
      case 1: sum_item<1>(cv, it, lane, acc); break;
      case 2: sum_item<2>(cv, it, lane, acc); break;
      case 3: sum_item<3>(cv, it, lane, acc); break;
      case 4: sum_item<4>(cv, it, lane, acc); break;
      case 5: sum_item<5>(cv, it, lane, acc); break;
      default: sum_item<6>(cv, it, lane, acc); break;
    }
    uint32_t osum[S];
#pragma unroll
    for (int j = 0; j < S; ++j) osum[j] = 0;
    major_outlier_sums<S>(mv, it, lane, osum);
    // slots: dense planes 0..P-1, outlier sums P..P+S-1, tail kMaxPlanes
#pragma unroll
    for (uint32_t k = 0; k < kMaxPlanes; ++k) {
      uint32_t v = k < cv.P ? acc[k] : 0u;
#pragma unroll
      for (int j = 0; j < S; ++j) v = k == cv.P + j ? osum[j] : v;
      const uint32_t tot = wave_total_u32(v);
      if (lane == 0) dst[k] = tot;
    }
    const uint32_t tail = wave_total_u32(acc[kMaxPlanes]);
    if (lane == 0) dst[kMaxPlanes] = tail;
  }
}

__global__ __launch_bounds__(256) void k_bq_major(const BatchBuffers* __restrict__ bbp) {
  __shared__ unsigned long long s_dep[4][64];
  const BatchBuffers bb = cload(bbp);
  const WaveSlot ws = wave_slot();
  const uint32_t lane = ws.lane;
  const uint32_t nscan = cmem(bb.hdr)->nscan;
  #pragma unroll 1
  for (uint32_t i = ws.first; i < nscan; i += ws.stride) {
    const BItem it = cload(bb.items, i);
    const BQuery q = cload(bb.q, it.owner);
    switch (q.sparse & 0xFFu) {
      case 0: break;
      case 1: major_scan_item<1>(bb, q, it, i, lane, s_dep[threadIdx.x >> 6]); break;
      case 2: major_scan_item<2>(bb, q, it, i, lane, s_dep[threadIdx.x >> 6]); break;
      case 3: major_scan_item<3>(bb, q, it, i, lane, s_dep[threadIdx.x >> 6]); break;
      case 4: major_scan_item<4>(bb, q, it, i, lane, s_dep[threadIdx.x >> 6]); break;
      case 5: major_scan_item<5>(bb, q, it, i, lane, s_dep[threadIdx.x >> 6]); break;
      case 6: major_scan_item<6>(bb, q, it, i, lane, s_dep[threadIdx.x >> 6]); break;
      default: major_scan_item<7>(bb, q, it, i, lane, s_dep[threadIdx.x >> 6]); break;
    }
  }
}

template <int SEM, int S>
__device__ __forceinline__ void major_mat_item(const BatchBuffers& bb, const BQuery& q, const BItem& it, uint32_t lane,
                                               double* sv) {
  const MajorPrefix<S> mv = cload(reinterpret_cast<const MajorPrefix<S>*>(bb.majors + it.chunk));
  const ChunkView& dv = mv.dense;
  double* outv = reinterpret_cast<double*>(q.out);
  const uint32_t drop = SEM ? q.mp.drop : 0u;
  #pragma unroll 1
  for (uint64_t c = it.c0; c < it.c1; ++c) {
    uint64_t v[16];
    major_tile_offsets<SEM, S>(mv, c, lane, v);
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) {
      const uint64_t o = (v[b] >> drop) << drop;
      sv[kLdsRow * lane + b] = (double)(int64_t)((uint64_t)dv.base + o) * q.mp.inv_scale;
    }
    wave_lds_sync();
    double* ot = outv + c * kTileRows + lane;
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

template <int SEM>
__device__ __forceinline__ void major_mat_item_s(const BatchBuffers& bb, const BQuery& q, const BItem& it,
                                                 uint32_t lane, double* sv) {
  switch (q.sparse & 0xFFu) {
    case 1: major_mat_item<SEM, 1>(bb, q, it, lane, sv); break;
    case 2: major_mat_item<SEM, 2>(bb, q, it, lane, sv); break;
    case 3: major_mat_item<SEM, 3>(bb, q, it, lane, sv); break;
    case 4: major_mat_item<SEM, 4>(bb, q, it, lane, sv); break;
    case 5: major_mat_item<SEM, 5>(bb, q, it, lane, sv); break;
    case 6: major_mat_item<SEM, 6>(bb, q, it, lane, sv); break;
    default: major_mat_item<SEM, 7>(bb, q, it, lane, sv); break;
  }
}

__global__ __launch_bounds__(256) void k_bq_major_mat(const BatchBuffers* __restrict__ bbp) {
  __shared__ double s_val[4][64 * kLdsRow];
  const BatchBuffers bb = cload(bbp);
  const WaveSlot ws = wave_slot();
  const uint32_t lane = ws.lane;
  const uint32_t nscan = cmem(bb.hdr)->nscan, nitems = cmem(bb.hdr)->nitems;
  double* sv = s_val[ws.local];
  #pragma unroll 1
  for (uint32_t i = nscan + ws.first; i < nitems; i += ws.stride) {
    const BItem it = cload(bb.items, i);
    const BQuery q = cload(bb.q, it.owner);
    if (!q.sparse) continue;
    if (q.sparse & kBqSparseSem) major_mat_item_s<1>(bb, q, it, lane, sv);
    else major_mat_item_s<0>(bb, q, it, lane, sv);
  }
}

hipError_t launch_batch_queries(const BatchBuffers* bb, const BatchTuning& tu, uint32_t parts, hipStream_t s) {
  hipLaunchKernelGGL(k_bq_fetch, dim3(64), dim3(256), 0, s, bb);
  hipLaunchKernelGGL(k_bq_items, dim3(1024), dim3(256), 0, s, bb);
  if (parts & kBqPartScan) {
    hipLaunchKernelGGL(k_bq_scan, dim3(tu.grid), dim3(256), 0, s, bb, tu);
    if (parts & kBqPartMajor) hipLaunchKernelGGL(k_bq_major, dim3(tu.grid), dim3(256), 0, s, bb);
    hipLaunchKernelGGL(k_bq_reduce, dim3(1024), dim3(256), 0, s, bb);
  }
  if (parts & kBqPartMax) {
    hipLaunchKernelGGL(k_bq_max_refine, dim3(tu.grid), dim3(256), 0, s, bb);
    hipLaunchKernelGGL(k_bq_max_final, dim3(256), dim3(256), 0, s, bb);
  }
  // four resident workgroups per CU fit the LDS staging of k_bq_mat
  if (parts & kBqPartMat) hipLaunchKernelGGL(k_bq_mat, dim3(tu.grid < 1024 ? tu.grid : 1024), dim3(256), 0, s, bb);
  if (parts & kBqPartMajorMat)
    hipLaunchKernelGGL(k_bq_major_mat, dim3(tu.grid < 1024 ? tu.grid : 1024), dim3(256), 0, s, bb);
  hipLaunchKernelGGL(k_bq_publish, dim3(16), dim3(256), 0, s, bb);
  return hipGetLastError();
}

}  // namespace buffgpu
