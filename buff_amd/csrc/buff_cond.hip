// Filter -> project: the values of the rows set in a condition bitmap,
// compacted in row order (buff_decode_condition, buff_simd.rs:1528-1712).
//
// The condition has the layout every filter here writes: ceil(n/64) u64 words,
// bit i = row i, so the lane owning rows 16t..16t+15 of tile c reads u16 word
// c*64 + t. Three launches:
//   k_cond_count   per-workgroup counts of condition rows (workgroup w owns a
//                  contiguous range of tiles);
//   k_cond_scan    one workgroup: exclusive prefix of those counts (u64) and the
//                  total;
//   k_cond_gather  per tile with condition rows: the planes and tail bits of the
//                  lanes that hold such rows (idle lanes load a live lane's
//                  bytes), the values built per row, compacted through LDS and
//                  written with contiguous stores at the tile's output offset.
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

__global__ __launch_bounds__(256) void k_cond_count(const uint16_t* __restrict__ cond, uint64_t n, uint64_t ntiles,
                                                    uint64_t tpw, uint32_t* __restrict__ wg_tot) {
  const uint32_t lane = __lane_id(), wave = wave_in_wg();
  const uint64_t t0 = (uint64_t)blockIdx.x * tpw;
  const uint64_t t1 = min(t0 + tpw, ntiles);
  uint32_t cnt = 0;
  for (uint64_t c = t0 + wave; c < t1; c += 4) cnt += (uint32_t)__popc(cond_lane(cond, n, c, lane));
  cnt = wave_sum_u32(cnt);
  __shared__ uint32_t part[4];
  if (lane == 0) part[wave] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) wg_tot[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Exclusive prefix of nwg counts into wg_off[0..nwg), the total in wg_off[nwg].
__global__ __launch_bounds__(256) void k_cond_scan(const uint32_t* __restrict__ wg_tot, uint32_t nwg,
                                                   unsigned long long* __restrict__ wg_off) {
  __shared__ unsigned long long s[256];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (nwg + 255) / 256;
  const uint32_t i0 = min(t * per, nwg), i1 = min(i0 + per, nwg);
  unsigned long long local = 0;
  for (uint32_t i = i0; i < i1; ++i) local += wg_tot[i];
  s[t] = local;
  __syncthreads();
  for (uint32_t o = 1; o < 256; o <<= 1) {
    const unsigned long long v = t >= o ? s[t - o] : 0ull;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  unsigned long long run = s[t] - local;
  for (uint32_t i = i0; i < i1; ++i) {
    wg_off[i] = run;
    run += wg_tot[i];
  }
  if (t == 255) wg_off[nwg] = s[255];
}

// Offset of row b (lane's rows) from NP planes and r tail bits. compat: when
// F >= 8 and r > 0 the tail field is OR-ed in shifted left by r, as the
// reference does (buff_simd.rs:1693, 1701); otherwise in place.
template <int NP>
__device__ __forceinline__ uint64_t cond_offset(const uint4 (&w)[NP > 0 ? NP : 1], const uint32_t (&h)[8], uint32_t r,
                                                uint32_t b, bool compat) {
  uint64_t v = 0;
#pragma unroll
  for (int p = 0; p < NP; ++p) v = (v << 8) | byte_at(w[p], b);
  uint32_t t = 0;
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j) t |= ((h[j] >> b) & 1u) << j;   // h[j] = 0 for j >= r
  if (NP == 0) return t;
  return compat ? (v | t) << r : (v << r) | t;
}

template <int NP>
__global__ __launch_bounds__(256) void k_cond_gather(ChunkView cv, double inv_scale, uint32_t compat,
                                                     const uint16_t* __restrict__ cond, uint64_t tpw,
                                                     const unsigned long long* __restrict__ wg_off,
                                                     double* __restrict__ out, uint64_t cap) {
  __shared__ uint32_t s_cnt[kCondMaxTilesPerWg];
  __shared__ double s_val[4][kTileRows];
  const uint32_t lane = __lane_id(), wave = wave_in_wg();
  const uint64_t t0 = (uint64_t)blockIdx.x * tpw;
  const uint64_t t1 = min(t0 + tpw, cv.ntiles);
  const uint32_t nt = (uint32_t)(t1 > t0 ? t1 - t0 : 0);
  // per-tile counts of this workgroup's range, then their exclusive prefix
  for (uint32_t k = wave; k < nt; k += 4) {
    const uint32_t tot = wave_total_u32((uint32_t)__popc(cond_lane(cond, cv.n, t0 + k, lane)));
    if (lane == 0) s_cnt[k] = tot;
  }
  __syncthreads();
  if (wave == 0) {
    const uint32_t per = (nt + 63) / 64;
    const uint32_t k0 = min(lane * per, nt), k1 = min(k0 + per, nt);
    uint32_t local = 0;
    for (uint32_t k = k0; k < k1; ++k) local += s_cnt[k];
    uint32_t run = wave_excl_scan_u32(local, lane);
    for (uint32_t k = k0; k < k1; ++k) {
      const uint32_t v = s_cnt[k];
      s_cnt[k] = run;
      run += v;
    }
  }
  __syncthreads();
  const unsigned long long base_out = wg_off[blockIdx.x];
  double* sv = s_val[wave];
  for (uint32_t k = wave; k < nt; k += 4) {
    const uint64_t c = t0 + k;
    const uint32_t live = cond_lane(cond, cv.n, c, lane);
    if (!__ballot(live != 0)) continue;
    const uint32_t src = refine_src(live, lane);
    uint4 w[NP > 0 ? NP : 1];
#pragma unroll
    for (int p = 0; p < NP; ++p) w[p] = load_plane(cv, p, c, src);
    uint32_t h[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) h[j] = j < cv.r ? load_tail(cv, j, c, src) : 0u;
    const uint32_t lp = wave_excl_scan_u32((uint32_t)__popc(live), lane);
    const uint32_t tot = wave_total_u32((uint32_t)__popc(live));
    uint32_t pos = lp;
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) {
      if ((live >> b) & 1u) {
        const uint64_t off = cond_offset<NP>(w, h, cv.r, b, compat != 0);
        sv[pos++] = (double)(int64_t)((uint64_t)cv.base + off) * inv_scale;
      }
    }
    wave_lds_sync();
    const unsigned long long o = base_out + s_cnt[k];
    for (uint32_t i = lane; i < tot; i += 64)
      if (o + i < cap) __builtin_nontemporal_store(sv[i], gmem_w(out) + o + i);
    wave_lds_sync();
  }
}

void cond_geometry(uint64_t ntiles, uint32_t* nwg, uint64_t* tpw) {
  uint64_t w = (ntiles + 15) / 16;
  if (w > 4096) w = 4096;
  const uint64_t need = (ntiles + kCondMaxTilesPerWg - 1) / kCondMaxTilesPerWg;
  if (w < need) w = need;
  if (w < 1) w = 1;
  *tpw = (ntiles + w - 1) / w;
  if (*tpw < 1) *tpw = 1;
  *nwg = (uint32_t)((ntiles + *tpw - 1) / *tpw);
  if (*nwg < 1) *nwg = 1;
}

uint64_t cond_scratch_bytes(uint64_t ntiles) {
  uint32_t nwg;
  uint64_t tpw;
  cond_geometry(ntiles, &nwg, &tpw);
  return 8ull * (nwg + 1) + 4ull * nwg + 256;
}

uint64_t cond_total_offset(uint64_t ntiles) {
  uint32_t nwg;
  uint64_t tpw;
  cond_geometry(ntiles, &nwg, &tpw);
  return 8ull * nwg;
}

hipError_t launch_cond_decode(const ChunkView& cv, double inv_scale, bool compat, const uint64_t* cond, double* out,
                              uint64_t cap, uint8_t* scratch, hipStream_t s) {
  if (cv.ntiles == 0) return hipSuccess;
  uint32_t nwg;
  uint64_t tpw;
  cond_geometry(cv.ntiles, &nwg, &tpw);
  unsigned long long* wg_off = reinterpret_cast<unsigned long long*>(scratch);   // nwg + 1 entries
  uint32_t* wg_tot = reinterpret_cast<uint32_t*>(scratch + 8ull * (nwg + 1));
  const uint16_t* c16 = reinterpret_cast<const uint16_t*>(cond);
  launch_k(k_cond_count, dim3(nwg), dim3(256), 0, s, c16, cv.n, cv.ntiles, tpw, wg_tot);
  launch_k(k_cond_scan, dim3(1), dim3(256), 0, s, static_cast<const uint32_t*>(wg_tot), nwg, wg_off);
  const unsigned long long* off = wg_off;
  const uint32_t cm = compat ? 1u : 0u;
#define BUFF_GATHER(NP) launch_k(k_cond_gather<NP>, dim3(nwg), dim3(256), 0, s, cv, inv_scale, cm, c16, tpw, off, out, cap)
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
