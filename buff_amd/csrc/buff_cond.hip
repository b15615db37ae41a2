// Filter -> project: the values of the rows set in a condition bitmap,
// compacted in row order (buff_decode_condition, buff_simd.rs:1528-1712).
//
// The condition has the layout every filter here writes: ceil(n/64) u64 words,
// bit i = row i, so the lane owning rows 16t..16t+15 of tile c reads u16 word
// c*64 + t. Four launches:
//   k_cond_count   per-tile and per-workgroup counts of condition rows and of
//                  tiles holding any (workgroup w owns 64 consecutive tiles);
//   k_cond_scan    one workgroup: exclusive prefixes of the per-workgroup
//                  counts and the totals;
//   k_cond_index   the tiles holding condition rows, compacted in tile order,
//                  each with its first output row;
//   k_cond_gather  per compacted tile, the waves of the grid in turn: the
//                  planes and tail bits of the
//                  lanes that hold such rows (idle lanes load a live lane's
//                  bytes), the values built per row, compacted through LDS and
//                  written with contiguous stores at the tile's output offset.
#include <algorithm>

#include "buff_cond.h"
#include "buff_device.h"
#include "buff_launch.h"

namespace buffgpu {

// The lane's condition rows of tile c. Lanes past the last row read word 0
// instead (the bitmap holds ceil(n/64) words, fewer than a full last tile).
__device__ __forceinline__ uint32_t cond_lane(const uint16_t* cond, uint64_t n, uint64_t c, uint32_t lane) {
  const uint32_t valid = lane_valid16(n, c, lane);
  const uint64_t idx = valid ? c * 64 + lane : 0;
  return (uint32_t)cond[idx] & valid;
}

// Sum of v over each row of 16 lanes, in the row's last lane.
__device__ __forceinline__ uint32_t row16_sum(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);   // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);   // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);   // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);   // row_shr:8
  return v;
}

// Workgroup w owns tiles [64w, 64w + 64): 1024 bitmap words, four per thread
// (coalesced, all in flight together). A tile's 16 words sit in one row of 16
// lanes, whose last lane writes the tile's count.
__global__ __launch_bounds__(256) void k_cond_count(const uint64_t* __restrict__ cond, uint64_t n, uint64_t ntiles,
                                                    uint16_t* __restrict__ tile_cnt, uint32_t* __restrict__ wg_rows,
                                                    uint32_t* __restrict__ wg_tiles) {
  const uint32_t t = threadIdx.x, lane = __lane_id(), wave = wave_in_wg();
  const uint64_t nw = (n + 63) / 64;
  const uint64_t w0 = (uint64_t)blockIdx.x * kCondTilesPerWg * 16;
  uint32_t pc[4];
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j) {
    const uint64_t i = w0 + t + 256 * j;
    uint64_t v = cond[i < nw ? i : 0];
    if (i >= nw) v = 0;
    else if (i == nw - 1 && (n & 63)) v &= (1ull << (n & 63)) - 1;
    pc[j] = (uint32_t)__popcll(v);
  }
  uint32_t rows = 0, tiles = 0;
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j) {
    const uint32_t tc = row16_sum(pc[j]);
    const uint64_t c = (uint64_t)blockIdx.x * kCondTilesPerWg + 16 * j + t / 16;
    if ((t & 15) == 15 && c < ntiles) {
      tile_cnt[c] = (uint16_t)tc;   // at most 1024
      rows += tc;
      tiles += tc ? 1u : 0u;
    }
  }
  rows = wave_sum_u32(rows);
  tiles = wave_sum_u32(tiles);
  __shared__ uint32_t part[8];
  if (lane == 0) {
    part[wave] = rows;
    part[4 + wave] = tiles;
  }
  __syncthreads();
  if (t == 0) {
    wg_rows[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
    wg_tiles[blockIdx.x] = part[4] + part[5] + part[6] + part[7];
  }
}

// One workgroup: exclusive prefixes of the per-workgroup row and tile counts
// (their totals land at index nwg), 4096 entries per round.
__global__ __launch_bounds__(256) void k_cond_scan(const uint32_t* __restrict__ wg_rows,
                                                   const uint32_t* __restrict__ wg_tiles, uint32_t nwg,
                                                   unsigned long long* __restrict__ off_rows,
                                                   uint32_t* __restrict__ off_tiles) {
  __shared__ unsigned long long sr[256];
  __shared__ uint32_t st[256];
  const uint32_t t = threadIdx.x;
  unsigned long long carry_r = 0;
  uint32_t carry_t = 0;
  for (uint32_t base = 0; base < nwg; base += 4096) {
    uint32_t vr[16], vt[16];
    unsigned long long lr = 0;
    uint32_t lt = 0;
#pragma unroll
    for (uint32_t j = 0; j < 16; ++j) {
      const uint32_t i = base + t * 16 + j;
      vr[j] = i < nwg ? wg_rows[i] : 0u;
      vt[j] = i < nwg ? wg_tiles[i] : 0u;
      lr += vr[j];
      lt += vt[j];
    }
    sr[t] = lr;
    st[t] = lt;
    __syncthreads();
    for (uint32_t o = 1; o < 256; o <<= 1) {
      const unsigned long long a = t >= o ? sr[t - o] : 0ull;
      const uint32_t b = t >= o ? st[t - o] : 0u;
      __syncthreads();
      sr[t] += a;
      st[t] += b;
      __syncthreads();
    }
    unsigned long long rr = carry_r + sr[t] - lr;
    uint32_t rt = carry_t + st[t] - lt;
#pragma unroll
    for (uint32_t j = 0; j < 16; ++j) {
      const uint32_t i = base + t * 16 + j;
      if (i < nwg) {
        off_rows[i] = rr;
        off_tiles[i] = rt;
      }
      rr += vr[j];
      rt += vt[j];
    }
    carry_r += sr[255];
    carry_t += st[255];
    __syncthreads();
  }
  if (t == 0) {
    off_rows[nwg] = carry_r;
    off_tiles[nwg] = carry_t;
  }
}

// One wave per workgroup range (one tile per lane): the tiles holding
// condition rows, compacted in tile order, each with its first output row.
__global__ __launch_bounds__(64) void k_cond_index(const uint16_t* __restrict__ tile_cnt, uint64_t ntiles,
                                                   const unsigned long long* __restrict__ off_rows,
                                                   const uint32_t* __restrict__ off_tiles,
                                                   uint32_t* __restrict__ list,
                                                   unsigned long long* __restrict__ row_off) {
  const uint32_t lane = __lane_id();
  const uint64_t c = (uint64_t)blockIdx.x * kCondTilesPerWg + lane;
  const uint32_t cnt = c < ntiles ? (uint32_t)tile_cnt[c] : 0u;
  const uint32_t rp = wave_excl_scan_u32(cnt, lane);
  const uint32_t tp = wave_excl_scan_u32(cnt ? 1u : 0u, lane);
  if (cnt) {
    const uint32_t k = off_tiles[blockIdx.x] + tp;
    list[k] = (uint32_t)c;
    row_off[k] = off_rows[blockIdx.x] + rp;
  }
}

// Offset of row b (lane's rows) from NP planes and r tail bits. compat: when
// F >= 8 and r > 0 the tail field is OR-ed in shifted left by r, as the
// reference does (buff_simd.rs:1693, 1701); otherwise in place.
template <int NP>
__device__ __forceinline__ uint64_t cond_offset(const RowBytes<NP>& rb, uint32_t r, uint32_t b, bool compat) {
  const uint64_t X = row_bytes_x<NP>(rb, b);
  if (NP == 0 || !compat) return X >> (8u - r);
  const uint64_t v = X >> 8, t = (X & 0xFFu) >> (8u - r);
  return (v | t) << r;
}

// Waves take the compacted tiles in turn, so clustered condition rows spread
// over the whole grid.
template <int NP>
__global__ __launch_bounds__(64) void k_cond_gather(ChunkView cv, double inv_scale, uint32_t compat,
                                                     const uint16_t* __restrict__ cond,
                                                     const uint32_t* __restrict__ ntiles_live,
                                                     const uint32_t* __restrict__ list,
                                                     const unsigned long long* __restrict__ row_off,
                                                     double* __restrict__ out, uint64_t cap) {
  __shared__ double sv[kTileRows];
  const uint32_t lane = __lane_id();
  const uint32_t total = *cmem(ntiles_live);
  for (uint64_t i = wave_global(); i < total; i += waves_in_grid()) {
    const uint64_t c = cmem(list)[i];
    const unsigned long long o = cmem(row_off)[i];
    const uint32_t live = cond_lane(cond, cv.n, c, lane);
    const uint32_t src = refine_src(live, lane);
    RowBytes<NP> rb;
    load_row_bytes<NP>(cv, c, src, rb);
    const uint32_t cntl = (uint32_t)__popc(live);
    uint32_t pos = wave_excl_scan_u32(cntl, lane);
    const uint32_t tot = wave_total_u32(cntl);
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) {
      if ((live >> b) & 1u) {
        const uint64_t off = cond_offset<NP>(rb, cv.r, b, compat != 0);
        sv[pos++] = (double)(int64_t)((uint64_t)cv.base + off) * inv_scale;
      }
    }
    wave_lds_sync();
    for (uint32_t k = lane; k < tot; k += 64)
      if (o + k < cap) __builtin_nontemporal_store(sv[k], gmem_w(out) + o + k);
    wave_lds_sync();
  }
}

// Scratch layout (each region 256-byte aligned): off_rows (nwg + 1 u64),
// row_off (ntiles u64), off_tiles (nwg + 1 u32), wg_rows, wg_tiles (nwg u32
// each), list (ntiles u32), tile_cnt (ntiles u16).
// The condition rows of an already decoded column `vals` (n doubles), in row
// order: the same tile list and row offsets as k_cond_gather.
__global__ __launch_bounds__(64) void k_cond_gather_values(const double* __restrict__ vals, uint64_t n,
                                                            const uint16_t* __restrict__ cond,
                                                            const uint32_t* __restrict__ ntiles_live,
                                                            const uint32_t* __restrict__ list,
                                                            const unsigned long long* __restrict__ row_off,
                                                            double* __restrict__ out, uint64_t cap) {
  __shared__ double sv[kTileRows];
  const uint32_t lane = __lane_id();
  const uint32_t total = *cmem(ntiles_live);
  for (uint64_t i = wave_global(); i < total; i += waves_in_grid()) {
    const uint64_t c = cmem(list)[i];
    const unsigned long long o = cmem(row_off)[i];
    const uint32_t live = cond_lane(cond, n, c, lane);
    const uint32_t cntl = (uint32_t)__popc(live);
    uint32_t pos = wave_excl_scan_u32(cntl, lane);
    const uint32_t tot = wave_total_u32(cntl);
    const uint64_t r0 = c * kTileRows + 16ull * lane;
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b)
      if ((live >> b) & 1u) sv[pos++] = vals[r0 + b];
    wave_lds_sync();
    for (uint32_t k = lane; k < tot; k += 64)
      if (o + k < cap) __builtin_nontemporal_store(sv[k], gmem_w(out) + o + k);
    wave_lds_sync();
  }
}

struct CondLayout {
  uint32_t nwg;
  uint64_t off_rows, row_off, off_tiles, wg_rows, wg_tiles, list, tile_cnt, bytes;
};

static CondLayout cond_layout(uint64_t ntiles) {
  CondLayout L;
  L.nwg = (uint32_t)((ntiles + kCondTilesPerWg - 1) / kCondTilesPerWg);
  if (L.nwg < 1) L.nwg = 1;
  uint64_t at = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t here = at;
    at += (bytes + 255) & ~255ull;
    return here;
  };
  L.off_rows = take(8ull * (L.nwg + 1));
  L.row_off = take(8ull * ntiles);
  L.off_tiles = take(4ull * (L.nwg + 1));
  L.wg_rows = take(4ull * L.nwg);
  L.wg_tiles = take(4ull * L.nwg);
  L.list = take(4ull * ntiles);
  L.tile_cnt = take(2ull * ntiles);
  L.bytes = at;
  return L;
}

uint64_t cond_scratch_bytes(uint64_t ntiles) { return cond_layout(ntiles).bytes; }

uint64_t cond_total_offset(uint64_t ntiles) {
  const CondLayout L = cond_layout(ntiles);
  return L.off_rows + 8ull * L.nwg;
}

hipError_t launch_cond_values(const double* vals, uint64_t n, const uint64_t* cond, double* out, uint64_t cap,
                              uint8_t* scratch, hipStream_t s) {
  const uint64_t ntiles = (n + kTileRows - 1) / kTileRows;
  if (ntiles == 0) return hipSuccess;
  const CondLayout L = cond_layout(ntiles);
  unsigned long long* off_rows = reinterpret_cast<unsigned long long*>(scratch + L.off_rows);
  unsigned long long* row_off = reinterpret_cast<unsigned long long*>(scratch + L.row_off);
  uint32_t* off_tiles = reinterpret_cast<uint32_t*>(scratch + L.off_tiles);
  uint32_t* wg_rows = reinterpret_cast<uint32_t*>(scratch + L.wg_rows);
  uint32_t* wg_tiles = reinterpret_cast<uint32_t*>(scratch + L.wg_tiles);
  uint32_t* list = reinterpret_cast<uint32_t*>(scratch + L.list);
  uint16_t* tile_cnt = reinterpret_cast<uint16_t*>(scratch + L.tile_cnt);
  launch_k(k_cond_count, dim3(L.nwg), dim3(256), 0, s, cond, n, ntiles, tile_cnt, wg_rows, wg_tiles);
  launch_k(k_cond_scan, dim3(1), dim3(256), 0, s, static_cast<const uint32_t*>(wg_rows),
           static_cast<const uint32_t*>(wg_tiles), L.nwg, off_rows, off_tiles);
  launch_k(k_cond_index, dim3(L.nwg), dim3(64), 0, s, static_cast<const uint16_t*>(tile_cnt), ntiles,
           static_cast<const unsigned long long*>(off_rows), static_cast<const uint32_t*>(off_tiles), list, row_off);
  const uint32_t grid = (uint32_t)std::min<uint64_t>(ntiles, 5120);
  launch_k(k_cond_gather_values, dim3(grid), dim3(64), 0, s, vals, n, reinterpret_cast<const uint16_t*>(cond),
           static_cast<const uint32_t*>(off_tiles + L.nwg), static_cast<const uint32_t*>(list),
           static_cast<const unsigned long long*>(row_off), out, cap);
  return hipGetLastError();
}

hipError_t launch_cond_decode(const ChunkView& cv, double inv_scale, bool compat, const uint64_t* cond, double* out,
                              uint64_t cap, uint8_t* scratch, hipStream_t s) {
  if (cv.ntiles == 0) return hipSuccess;
  const CondLayout L = cond_layout(cv.ntiles);
  unsigned long long* off_rows = reinterpret_cast<unsigned long long*>(scratch + L.off_rows);
  unsigned long long* row_off = reinterpret_cast<unsigned long long*>(scratch + L.row_off);
  uint32_t* off_tiles = reinterpret_cast<uint32_t*>(scratch + L.off_tiles);
  uint32_t* wg_rows = reinterpret_cast<uint32_t*>(scratch + L.wg_rows);
  uint32_t* wg_tiles = reinterpret_cast<uint32_t*>(scratch + L.wg_tiles);
  uint32_t* list = reinterpret_cast<uint32_t*>(scratch + L.list);
  uint16_t* tile_cnt = reinterpret_cast<uint16_t*>(scratch + L.tile_cnt);
  const uint16_t* c16 = reinterpret_cast<const uint16_t*>(cond);
  launch_k(k_cond_count, dim3(L.nwg), dim3(256), 0, s, cond, cv.n, cv.ntiles, tile_cnt, wg_rows, wg_tiles);
  launch_k(k_cond_scan, dim3(1), dim3(256), 0, s, static_cast<const uint32_t*>(wg_rows),
           static_cast<const uint32_t*>(wg_tiles), L.nwg, off_rows, off_tiles);
  launch_k(k_cond_index, dim3(L.nwg), dim3(64), 0, s, static_cast<const uint16_t*>(tile_cnt), cv.ntiles,
           static_cast<const unsigned long long*>(off_rows), static_cast<const uint32_t*>(off_tiles), list, row_off);
  const uint32_t cm = compat ? 1u : 0u;
  const uint32_t* nlive = off_tiles + L.nwg;
  const uint32_t* lst = list;
  const unsigned long long* ro = row_off;
  // one wave per workgroup, 8 KB of LDS staging each: 20 waves fit per CU
  const uint32_t grid = (uint32_t)std::min<uint64_t>(cv.ntiles, 5120);
#define BUFF_GATHER(NP) \
  launch_k(k_cond_gather<NP>, dim3(grid), dim3(64), 0, s, cv, inv_scale, cm, c16, nlive, lst, ro, out, cap)
  switch (cv.P) {
    case 0: BUFF_GATHER(0); break;
    case 1: BUFF_GATHER(1); break;
    case 2: BUFF_GATHER(2); break;
    case 3: BUFF_GATHER(3); break;
    case 4: BUFF_GATHER(4); break;
    case 5: BUFF_GATHER(5); break;
    case 6: BUFF_GATHER(6); break;
    case 7: BUFF_GATHER(7); break;
    default: BUFF_GATHER(8); break;
  }
#undef BUFF_GATHER
  return hipGetLastError();
}

}  // namespace buffgpu
