// HIP kernels of the BUFF hot path for gfx950 (wave64). Layout: buff_common.h.
//
// Every kernel is HBM-bound integer/byte work; nothing here is a contraction, so
// no MFMA. The designs share three ideas:
//   * a wave owns 1024-row tiles; a lane reads its 16 bytes of a plane with one
//     dwordx4 load (the wave: 1 KiB contiguous), and a wave issues the loads of
//     several tiles before it consumes any, so enough bytes are in flight;
//   * a lane owns 16 consecutive rows, so byte compares done four at a time in
//     a dword (SWAR) gather into the lane's 16-bit slice of the result bitmap,
//     written with one 2-byte store per lane (a wave: 128 B contiguous);
//   * float64 traffic (encode input, materialize output) stays coalesced: lanes
//     read/write rows 64g + t and LDS moves each row to the lane that owns it.
#include <hip/hip_runtime.h>

#include "buff_common.h"
#include "buff_device.h"
#include "buff_launch.h"
#include "buff_kernels.h"

namespace buffgpu {

LaunchEvents& launch_events() {
  static thread_local LaunchEvents le;
  return le;
}

// ----------------------------------------------------------------------------
// K1: fixed-point conversion + min / max reduction (encode pass 1).
// split_double.rs:321-330. Workgroup w owns the contiguous tile range
// [w*tpw, (w+1)*tpw), so its partial min is a segment min the prefix pass reuses.
// max is taken over rows i >= 1 (row 0 only ever updates the min). Row order
// does not matter here, so lanes read rows 64g + t (coalesced).

__global__ __launch_bounds__(256) void k_minmax(const double* __restrict__ x, uint64_t n, uint64_t ntiles,
                                                uint64_t tpw, int32_t pexp, uint32_t dlen,
                                                int64_t* __restrict__ partial, int64_t* __restrict__ set,
                                                int64_t* __restrict__ next) {
  const uint32_t lane = __lane_id();
  const uint32_t wave = wave_in_wg();
  const uint64_t t0 = (uint64_t)blockIdx.x * tpw;
  const uint64_t t1 = min(t0 + tpw, ntiles);
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  for (uint64_t c = t0 + wave; c < t1; c += 4) {
    uint64_t v[16];
    if (tile_full(n, c) && c != 0) {
      load_tile_f64(x, n, c, lane, true, v);
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) {
        const int64_t f = fetch_fixed_aligned_bits(v[g], pexp, dlen);
        mn = min(mn, f);
        mx = max(mx, f);
      }
    } else {
      load_tile_f64(x, n, c, lane, false, v);
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) {
        const uint64_t row = c * kTileRows + 64 * g + lane;
        const int64_t f = fetch_fixed_aligned_bits(v[g], pexp, dlen);
        if (row < n) {
          mn = min(mn, f);
          if (row != 0) mx = max(mx, f);
        }
      }
    }
  }
  mn = wave_min_i64(mn);
  mx = wave_max_i64(mx);
  __shared__ int64_t smn[4], smx[4];
  if (lane == 0) { smn[wave] = mn; smx[wave] = mx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int64_t a = min(min(smn[0], smn[1]), min(smn[2], smn[3]));
    const int64_t b = max(max(smx[0], smx[1]), max(smx[2], smx[3]));
    partial[blockIdx.x] = a;   // kept for the prefix-min pass
    partial[gridDim.x + blockIdx.x] = b;
    if (set) {
      const uint32_t slot = blockIdx.x % kMinMaxSlots;
      if (a != INT64_MAX) atomicMin((long long*)&set[slot], (long long)a);
      if (b != INT64_MIN) atomicMax((long long*)&set[kMinMaxSlots + slot], (long long)b);
      if (blockIdx.x == 0) set[2 * kMinMaxSlots] = n ? fetch_fixed_aligned_bits(__double_as_longlong(x[0]), pexp, dlen) : 0;
    }
  }
  if (set && blockIdx.x == 0)
    for (uint32_t i = threadIdx.x; i < kMinMaxSetWords; i += blockDim.x)
      next[i] = i < kMinMaxSlots ? INT64_MAX : (i < 2 * kMinMaxSlots ? INT64_MIN : 0);
}

// Sets of slots to their empty values (mins INT64_MAX, maxes INT64_MIN).
__global__ __launch_bounds__(256) void k_minmax_slots_init(int64_t* __restrict__ sets, uint32_t nsets) {
  for (uint32_t i = threadIdx.x; i < nsets * kMinMaxSetWords; i += blockDim.x) {
    const uint32_t k = i % kMinMaxSetWords;
    sets[i] = k < kMinMaxSlots ? INT64_MAX : (k < 2 * kMinMaxSlots ? INT64_MIN : 0);
  }
}

// Folds the k_minmax partials (one workgroup).
__global__ __launch_bounds__(256) void k_minmax_final(const double* __restrict__ x, uint64_t n, int32_t pexp,
                                                      uint32_t dlen, const int64_t* __restrict__ partial,
                                                      uint32_t nwg, MinMaxOut* __restrict__ out) {
  const uint32_t lane = __lane_id(), wave = wave_in_wg();
  __shared__ int64_t smn[4], smx[4];
  int64_t a = INT64_MAX, b = INT64_MIN;
  for (uint32_t i = threadIdx.x; i < nwg; i += blockDim.x) {
    a = min(a, partial[i]);
    b = max(b, partial[nwg + i]);
  }
  a = wave_min_i64(a);
  b = wave_max_i64(b);
  if (lane == 0) { smn[wave] = a; smx[wave] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    out->min_all = min(min(smn[0], smn[1]), min(smn[2], smn[3]));
    out->max_rest = max(max(smx[0], smx[1]), max(smx[2], smx[3]));
    out->first = n ? fetch_fixed_aligned_bits(__double_as_longlong(x[0]), pexp, dlen) : 0;
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
                                                      int64_t* __restrict__ partial2) {
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
  if (lane == 0) partial2[blockIdx.x] = qmax;
}

__global__ __launch_bounds__(64) void k_minmax_prefix_final(const int64_t* __restrict__ partial2, uint32_t nwg,
                                                            MinMaxOut* __restrict__ out) {
  int64_t b = INT64_MIN;
  for (uint32_t i = __lane_id(); i < nwg; i += 64) b = max(b, partial2[i]);
  b = wave_max_i64(b);
  if (__lane_id() == 0) out->qmax = b;
}

// ----------------------------------------------------------------------------
// K2: encode pass 2 — offsets, byte planes and tail bit-planes.
// split_double.rs:363-414: plane p byte = (off >> (F-8(p+1))) as u8, tail = low r bits.
// Lanes read rows 64g + t (coalesced), park the offsets in LDS in row order
// (one pad word per 16 rows against bank conflicts) and read back their own 16
// consecutive rows.

__global__ __launch_bounds__(256) void k_encode(const double* __restrict__ x, uint64_t n, uint64_t ntiles,
                                                int32_t pexp, uint32_t dlen, int64_t base, uint32_t F, uint32_t P,
                                                uint32_t r, uint8_t* __restrict__ planes, uint64_t pitch,
                                                uint16_t* __restrict__ tail) {
  __shared__ uint64_t s_off[4][64 * kLdsRow];
  const uint32_t lane = __lane_id();
  const uint32_t wslot = wave_in_wg();
  uint64_t* so = s_off[wslot];
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  for (uint64_t c = wave; c < ntiles; c += nwaves) {
    uint64_t v[16];
    const bool full = tile_full(n, c);
    load_tile_f64(x, n, c, lane, full, v);
    if (full) {
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g)
        so[lds_slot(64 * g + lane)] = (uint64_t)fetch_fixed_aligned_bits(v[g], pexp, dlen) - (uint64_t)base;
    } else {
#pragma unroll
      for (uint32_t g = 0; g < 16; ++g) {
        const uint64_t o = (uint64_t)fetch_fixed_aligned_bits(v[g], pexp, dlen) - (uint64_t)base;
        so[lds_slot(64 * g + lane)] = c * kTileRows + 64 * g + lane < n ? o : 0;
      }
    }
    wave_lds_sync();
    uint64_t off[16];
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) off[b] = so[kLdsRow * lane + b];
    wave_lds_sync();
    store_tile_planes(off, P, r, planes, pitch, tail, c, lane);
  }
}

// ----------------------------------------------------------------------------
// K3/K4: range (x > pred) and equality filters with early termination.
// buff_simd.rs:4507-4679 / 4925-5093.
//
// Per tile and lane: the top sub-column's 16 bytes against the target byte,
// four rows per dword op (SWAR), give the lane's 16-row masks of rows above
// (`gt`) and tied with (`eq`) the target. Tiles where some row ties are refined
// together, one sub-column at a time with all their loads in flight; a lane
// without ties reads the bytes of the tile's first lane with ties instead of its
// own, so a refined tile costs only the lines that hold ties. The tail compares
// bit-plane by bit-plane on the 16-row masks. The result mask is the lane's
// 16 bits of the tile bitmap.

// PF: the next group's top plane is loaded before this group is compared and
// refined, so the refinement's dependent loads overlap it (columns that tie on
// the top byte in most tiles).
// COND: only the rows set in `cond` (the bitmap layout of `out`) take part; a
// group of tiles without any such row is written as zeros without reading its
// planes.
template <int MODE, int U, bool PF, bool COND = false>
__global__ __launch_bounds__(256) void k_filter(ChunkView cv, FilterTarget tg, uint64_t* __restrict__ out,
                                                uint64_t nwords, unsigned long long* __restrict__ stats,
                                                const uint16_t* __restrict__ cond) {
  static_assert(!(COND && PF), "conditional filters run without the prefetch");
  const uint32_t lane = __lane_id();
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  const uint64_t ngroups = (cv.ntiles + U - 1) / U;
  const bool has_planes = cv.P > 0;
  const Digit d0 = make_digit(tg.digit[0]);
  uint16_t* out16 = reinterpret_cast<uint16_t*>(out);
  const uint64_t nout16 = 4 * nwords;
  uint4 wn[U];
  if (PF && has_planes && wave < ngroups) {
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) wn[u] = load_plane(cv, 0, min(wave * U + u, cv.ntiles - 1), lane);
  }
  for (uint64_t grp = wave; grp < ngroups; grp += nwaves) {
    const uint64_t c0 = grp * U;
    uint32_t cm[U];
    if (COND) {
      bool anyc = false;
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint64_t idx = (c0 + u) * 64 + lane;
        const bool in = c0 + u < cv.ntiles && idx < nout16;
        cm[u] = (uint32_t)cond[in ? idx : 0] & (in ? 0xFFFFu : 0u);   // the address stays inside the bitmap
        anyc |= cm[u] != 0;
      }
      if (!__ballot(anyc)) {
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
          const uint64_t idx = (c0 + u) * 64 + lane;
          if (c0 + u < cv.ntiles && idx < nout16) __builtin_nontemporal_store((uint16_t)0, gmem_w(out16) + idx);
        }
        continue;
      }
    }
    uint4 w[U];
    if (has_planes) {
      if (PF) {
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) w[u] = wn[u];
        const uint64_t gn = grp + nwaves;
        if (gn < ngroups) {
#pragma unroll
          for (uint32_t u = 0; u < U; ++u) wn[u] = load_plane(cv, 0, min(gn * U + u, cv.ntiles - 1), lane);
        }
      } else {
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) w[u] = load_plane(cv, 0, min(c0 + u, cv.ntiles - 1), lane);
      }
    }
    uint32_t gtm[U], eqm[U];
    bool any = false;
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint32_t valid = lane_valid16(cv.n, c0 + u, lane);
      const uint32_t live = COND ? valid & cm[u] : valid;
      if (has_planes) {
        gtm[u] = MODE == 0 ? mask_gt(w[u], d0) & live : 0u;
        eqm[u] = mask_eq(w[u], d0) & live;
      } else {
        gtm[u] = 0;
        eqm[u] = live;
      }
      any |= eqm[u] != 0;
    }
    if (__ballot(any)) {
      for (uint32_t p = 1; p < cv.P; ++p) {
        bool live = false;
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) live |= eqm[u] != 0;
        if (!__ballot(live)) break;
        if (stats) {
          uint32_t k = 0;
#pragma unroll
          for (uint32_t u = 0; u < U; ++u) k += __popc(eqm[u]);
          const uint64_t tot = wave_sum_u64(k);
          if (lane == 0) atomicAdd(&stats[p], (unsigned long long)tot);
        }
        uint4 wp[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) wp[u] = load_plane(cv, p, min(c0 + u, cv.ntiles - 1), refine_src(eqm[u], lane));
        const Digit dp = make_digit(tg.digit[p]);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
          if (MODE == 0) gtm[u] |= eqm[u] & mask_gt(wp[u], dp);
          eqm[u] &= mask_eq(wp[u], dp);
        }
      }
      if (cv.r) {
        bool live = false;
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) live |= eqm[u] != 0;
        if (__ballot(live)) {
          if (stats) {
            uint32_t k = 0;
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) k += __popc(eqm[u]);
            const uint64_t tot = wave_sum_u64(k);
            if (lane == 0) atomicAdd(&stats[cv.P], (unsigned long long)tot);
          }
#pragma unroll
          for (uint32_t u = 0; u < U; ++u)
            if (c0 + u < cv.ntiles) refine_tail<MODE>(cv, c0 + u, tg.tail, lane, gtm[u], eqm[u]);
        }
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint64_t c = c0 + u;
      const uint64_t idx = c * 64 + lane;
      const uint32_t res = (MODE == 0 ? gtm[u] & lane_valid16(cv.n, c, lane) : eqm[u]);
      if (c < cv.ntiles && idx < nout16) __builtin_nontemporal_store((uint16_t)res, gmem_w(out16) + idx);
    }
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

// Fills n16 16-byte words, a workgroup writing 16 KiB contiguous per step: the
// device's streaming write rate (faster than a grid-strided fill: 0.73 of peak
// against 0.68 at 1 GiB).
__global__ __launch_bounds__(256) void k_fill_bytes(uint4* __restrict__ out, uint64_t n16, uint32_t v) {
  const uint4 w = make_uint4(v, v, v, v);
  constexpr uint64_t kPer = 4 * 256;
  for (uint64_t b = (uint64_t)blockIdx.x * kPer; b < n16; b += (uint64_t)gridDim.x * kPer) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint64_t i = b + k * 256 + threadIdx.x;
      if (i < n16) out[i] = w;
    }
  }
}

// Copies n16 16-byte words with non-temporal loads and stores: the device's
// streaming read + write rate, the ceiling of the kernels that read and write
// (encode, materialize). A workgroup moves 16 KiB contiguous per step, four
// loads per thread in flight (faster than a grid-strided copy: 0.74 of peak
// against 0.63).
__global__ __launch_bounds__(256) void k_copy_bytes(const u32x4_t* __restrict__ in, u32x4_t* __restrict__ out,
                                                    uint64_t n16) {
  constexpr uint64_t kPer = 4 * 256;
  for (uint64_t b = (uint64_t)blockIdx.x * kPer; b < n16; b += (uint64_t)gridDim.x * kPer) {
    u32x4_t v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint64_t i = b + k * 256 + threadIdx.x;
      if (i < n16) v[k] = __builtin_nontemporal_load(gmem(in) + i);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint64_t i = b + k * 256 + threadIdx.x;
      if (i < n16) __builtin_nontemporal_store(v[k], gmem_w(out) + i);
    }
  }
}

// Reads n16 16-byte words with non-temporal loads, one 1 KiB row of 64 lanes
// per wave and four in flight, in the wave-strided order of the plane kernels:
// the device's streaming read rate for a plane-sized read.
__global__ __launch_bounds__(128) void k_read_bytes(const u32x4_t* __restrict__ in, uint64_t n16,
                                                    uint32_t* __restrict__ sink) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = wave_global(), nwaves = waves_in_grid();
  const uint64_t nrows = (n16 + 63) / 64, ngroups = (nrows + 3) / 4;
  uint32_t acc = 0;
  for (uint64_t g = wave; g < ngroups; g += nwaves) {
    u32x4_t w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint64_t i = (g * 4 + u) * 64 + lane;
      w[u] = i < n16 ? __builtin_nontemporal_load(gmem(in) + i) : u32x4_t{0, 0, 0, 0};
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) acc ^= w[u].x ^ w[u].y ^ w[u].z ^ w[u].w;
  }
  if (acc == 0x9E3779B9u) sink[0] = acc;   // keeps the loads; practically never taken
}

// Reads `bytes` (a multiple of 16) and keeps nothing: evicts the memory-side
// cache with clean lines before a cold measurement.
__global__ void k_touch_bytes(const uint4* __restrict__ in, uint64_t n16, uint32_t* __restrict__ sink) {
  uint32_t acc = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint4 w = in[i];
    acc ^= w.x ^ w.y ^ w.z ^ w.w;
  }
  if (acc == 0x9E3779B9u) sink[0] = acc;   // keeps the loads; practically never taken
}

// ----------------------------------------------------------------------------
// K5: per-sub-column byte sums (split_double.rs:2158-2204), exact in u64.
// Bytes are summed with v_sad_u8, tail bit-planes with popcounts weighted by
// their bit position. P (planes) and R (tail bits) are template parameters and
// every load of a step is unconditional (the last group clamps its tile index
// and skips the duplicate), so all of them are in flight together. Each
// workgroup writes one row of partial sums, which k_sum_final folds; or, with
// `next` given (one launch), adds its row into one of kSumSlots slot rows of
// `rows` (u64 atomics, workgroup w into slot w % kSumSlots, laid out
// [k][slot]) for launch_publish to fold, and workgroup 0 clears `next`, the set the
// following call adds into.

template <int P, int R, int U>
__global__ __launch_bounds__(256) void k_sum(ChunkView cv, unsigned long long* __restrict__ rows,
                                             unsigned long long* __restrict__ next) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  const uint64_t ngroups = (cv.ntiles + U - 1) / U;
  constexpr int PP = P > 0 ? P : 1;
  constexpr int RR = R > 0 ? R : 1;
  uint32_t acc[PP];
  uint64_t wide[PP];
#pragma unroll
  for (int p = 0; p < PP; ++p) { acc[p] = 0; wide[p] = 0; }
  uint32_t tacc[RR];   // per-bit popcounts
#pragma unroll
  for (int j = 0; j < RR; ++j) tacc[j] = 0;
  uint32_t iter = 0;
  for (uint64_t grp = wave; grp < ngroups; grp += nwaves) {
    const uint64_t c0 = grp * U;
    uint4 w[U][PP];
    uint32_t h[U][RR];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t c = min(c0 + u, cv.ntiles - 1);
#pragma unroll
      for (int p = 0; p < P; ++p) w[u][p] = load_plane(cv, p, c, lane);
#pragma unroll
      for (int j = 0; j < R; ++j) h[u][j] = cv.tail[(c * R + j) * 64 + lane];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (c0 + u >= cv.ntiles) break;   // clamped duplicate of the last tile
#pragma unroll
      for (int p = 0; p < P; ++p) {
        uint32_t a = __builtin_amdgcn_sad_u8(w[u][p].x, 0u, acc[p]);
        a = __builtin_amdgcn_sad_u8(w[u][p].y, 0u, a);
        a = __builtin_amdgcn_sad_u8(w[u][p].z, 0u, a);
        acc[p] = __builtin_amdgcn_sad_u8(w[u][p].w, 0u, a);
      }
#pragma unroll
      for (int j = 0; j < R; ++j) tacc[j] += __popc(h[u][j]);
    }
    // a lane adds at most 16*255 per tile and plane: fold into u64 long before u32 overflow
    if (++iter == (1u << 14)) {
#pragma unroll
      for (int p = 0; p < P; ++p) { wide[p] += acc[p]; acc[p] = 0; }
      iter = 0;
    }
  }
  // wave sums -> workgroup row in LDS -> one partial row per workgroup
  // (a last-workgroup fold inside this kernel would need a device-scope
  // release per workgroup, which writes back the XCD's L2; a second tiny
  // launch is cheaper)
  __shared__ uint64_t wrow[16][kMaxPlanes + 1];
  const uint32_t wslot = wave_in_wg();
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const uint64_t s = wave_sum_u64(wide[p] + acc[p]);
    if (lane == 0) wrow[wslot][p] = s;
  }
  uint64_t t = 0;
#pragma unroll
  for (int j = 0; j < R; ++j) t += (uint64_t)tacc[j] << j;
  t = wave_sum_u64(t);
  if (lane == 0) wrow[wslot][P] = t;
  __syncthreads();
  if (threadIdx.x <= (uint32_t)P) {
    uint64_t v = 0;
    for (uint32_t k = 0; k < (blockDim.x >> 6); ++k) v += wrow[k][threadIdx.x];
    if (!next)
      rows[(uint64_t)blockIdx.x * (kMaxPlanes + 1) + threadIdx.x] = v;
    else if (v)
      atomicAdd(&rows[threadIdx.x * kSumSlots + blockIdx.x % kSumSlots], (unsigned long long)v);
  }
  if (next && blockIdx.x == 0)
    for (uint32_t i = threadIdx.x; i < kSumSetWords; i += blockDim.x) next[i] = 0ull;
}

// Sums the per-workgroup rows of k_sum (one workgroup; each thread keeps its
// rows' loads in flight together).
template <int W>
__global__ __launch_bounds__(256) void k_sum_final(const unsigned long long* __restrict__ rows, uint32_t nrows,
                                                   unsigned long long* __restrict__ out) {
  __shared__ unsigned long long part[256 / 64][kMaxPlanes + 1];
  const uint32_t lane = __lane_id(), wslot = wave_in_wg();
  uint64_t acc[W];
#pragma unroll
  for (int k = 0; k < W; ++k) acc[k] = 0;
  for (uint32_t i0 = threadIdx.x; i0 < nrows; i0 += 8 * blockDim.x) {
    uint64_t v[8][W];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t i = i0 + u * blockDim.x;
#pragma unroll
      for (int k = 0; k < W; ++k) v[u][k] = i < nrows ? rows[(uint64_t)i * (kMaxPlanes + 1) + k] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int k = 0; k < W; ++k) acc[k] += v[u][k];
  }
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const uint64_t v = wave_sum_u64(acc[k]);
    if (lane == 0) part[wslot][k] = v;
  }
  __syncthreads();
  if (threadIdx.x <= kMaxPlanes)
    out[threadIdx.x] = threadIdx.x < (uint32_t)W ? part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
                                                       part[3][threadIdx.x]
                                                 : 0ull;
}

// ----------------------------------------------------------------------------
// K5b: the tail part of the reduced-precision sum
// (byte_residue_sum_with_precision, split_double.rs:2611-2622) when it reads
// q < r bits per row. The reference then reads n q-bit fields from the start of
// the packed r-bit tail: field i holds stream tail bits [i*q, i*q + q), i.e. tail
// bit b of row i (stream bit i*r + b) lands in field bit (i*r + b) mod q when
// i*r + b < n*q. out[0] accumulates the sum of those fields; out[1] the sum of
// every row's top q tail bits (tail >> (r - q)), the exact truncated value.
// One lane per 16 rows of a tile (its r bit-plane words), grid-stride over
// tiles; one atomic add per workgroup and value.
__global__ __launch_bounds__(256) void k_tail_fields(ChunkView cv, uint32_t q, unsigned long long* __restrict__ out) {
  const uint32_t lane = __lane_id();
  const uint32_t r = cv.r;
  const uint64_t limit = cv.n * q;   // stream bits the reference reads
  uint64_t ref = 0, sem = 0;
  for (uint64_t c = wave_global(); c < cv.ntiles; c += waves_in_grid()) {
    uint32_t h[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) h[j] = j < (int)r ? cv.tail[(c * r + j) * 64 + lane] : 0u;
    for (uint32_t j = r - q; j < r; ++j) sem += (uint64_t)__popc(h[j]) << (j - (r - q));
    const uint64_t row0 = c * kTileRows + 16ull * lane;
    uint64_t pos = row0 * r;             // stream bit of (row0, bit 0)
    uint32_t m = (uint32_t)(pos % q);    // its field bit
    for (int k = 0; k < 16; ++k) {
      for (uint32_t b = 0; b < r; ++b) {
        if (pos < limit && ((h[b] >> k) & 1u)) ref += 1ull << m;
        ++pos;
        m = (m + 1 == q) ? 0u : m + 1;
      }
    }
  }
  __shared__ uint64_t part[2][4];
  ref = wave_sum_u64(ref);
  sem = wave_sum_u64(sem);
  if (lane == 0) {
    part[0][wave_in_wg()] = ref;
    part[1][wave_in_wg()] = sem;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    uint64_t v = 0;
    for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) v += part[threadIdx.x][w];
    if (v) atomicAdd(out + threadIdx.x, (unsigned long long)v);
  }
}

hipError_t launch_tail_fields(const ChunkView& cv, uint32_t q, unsigned long long* out, hipStream_t s) {
  if (q == 0 || q >= cv.r || cv.r > 7) return hipErrorInvalidValue;
  const uint32_t wg = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((cv.ntiles + 3) / 4, 2048));
  launch_k(k_tail_fields, dim3(wg), dim3(256), 0, s, cv, q, out);
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// K6: max stored offset (split_double.rs:3919-4105). The reference narrows the
// candidates byte by byte; its result is the largest stored offset, whose top
// byte is the largest top-plane byte. Pass 1 streams the top plane once and
// records each tile's largest byte and folds each workgroup's into one of 64
// slots (atomicMax, workgroup w into slot w % 64: same-address device atomics
// serialize); pass 2 takes the max of the slots and rebuilds whole offsets only
// in tiles whose largest byte equals it. A slot holds (tag << 8) | byte, where
// the tag grows by one per call: the values a call writes exceed every value
// of earlier calls, so the slots are never cleared between calls. Pass 1 also
// clears *best before pass 2 folds into it.


constexpr uint32_t kMaxSlots = 64;

template <int U>
__global__ __launch_bounds__(256) void k_max_top(ChunkView cv, uint8_t* __restrict__ tile_max,
                                                 unsigned int* __restrict__ slots, uint32_t tag,
                                                 unsigned long long* __restrict__ best_out) {
  const uint32_t lane = __lane_id();
  if (blockIdx.x == 0 && threadIdx.x == 0) *best_out = 0ull;
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  const uint64_t ngroups = (cv.ntiles + U - 1) / U;
  uint32_t best = 0;
  for (uint64_t grp = wave; grp < ngroups; grp += nwaves) {
    const uint64_t c0 = grp * U;
    uint4 w[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) w[u] = load_plane(cv, 0, min(c0 + u, cv.ntiles - 1), lane);
    uint32_t mine = 0;   // lane u holds tile c0+u's max after the loop
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      uint32_t mb = max(max(max_byte4(w[u].x), max_byte4(w[u].y)), max(max_byte4(w[u].z), max_byte4(w[u].w)));
      mb = wave_max_u32(mb);
      best = max(best, mb);
      if (lane == u) mine = mb;
    }
    if (lane < U && c0 + lane < cv.ntiles) tile_max[c0 + lane] = (uint8_t)mine;
  }
  __shared__ uint32_t wbest[16];
  if (lane == 0) wbest[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t b = 0;
    for (uint32_t k = 0; k < (blockDim.x >> 6); ++k) b = max(b, wbest[k]);
    atomicMax(&slots[blockIdx.x % kMaxSlots], (tag << 8) | b);
  }
}

__global__ __launch_bounds__(256) void k_max_refine(ChunkView cv, const uint8_t* __restrict__ tile_max,
                                                    const unsigned int* __restrict__ slots, uint32_t tag,
                                                    unsigned long long* __restrict__ best_out) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  const bool has_planes = cv.P > 0;
  // 64 tiles per probe, one byte per lane; the first probe's bytes are loaded
  // together with the slots
  uint64_t base = wave * 64;
  uint32_t tm = (has_planes && base + lane < cv.ntiles) ? tile_max[base + lane] : 0u;
  uint32_t t = 0;
  if (has_planes) {
    const uint32_t m = wave_max_u32_dpp(slots[lane]);   // one slot per lane
    t = (m >> 8) == tag ? (m & 0xFFu) : 0u;             // every workgroup of pass 1 wrote this call's tag
  }
  uint64_t best = 0;
  bool have = false;
  for (; base < cv.ntiles; base += nwaves * 64) {
    const uint64_t c = base + lane;
    const bool hit = c < cv.ntiles && (!has_planes || tm == t);
    const uint64_t nb = base + nwaves * 64;
    tm = (has_planes && nb + lane < cv.ntiles) ? tile_max[nb + lane] : 0u;   // the next probe's bytes
    uint64_t hits = __ballot(hit);
    while (hits) {
      const uint32_t k = (uint32_t)__builtin_ctzll(hits);
      hits &= hits - 1;
      const uint64_t ct = base + k;
      // every plane and tail bit-plane of the tile in flight together
      uint4 wp[kMaxPlanes];
      uint32_t th[7];
#pragma unroll
      for (uint32_t p = 0; p < kMaxPlanes; ++p) wp[p] = p < cv.P ? load_plane(cv, p, ct, lane) : make_uint4(0, 0, 0, 0);
#pragma unroll
      for (uint32_t j = 0; j < 7; ++j) th[j] = j < cv.r ? load_tail(cv, j, ct, lane) : 0u;
      uint64_t off[16];
#pragma unroll
      for (uint32_t b = 0; b < 16; ++b) off[b] = 0;
#pragma unroll
      for (uint32_t p = 0; p < kMaxPlanes; ++p) {
        if (p < cv.P) {
#pragma unroll
          for (uint32_t b = 0; b < 16; ++b) off[b] = (off[b] << 8) | byte_at(wp[p], b);
        }
      }
#pragma unroll
      for (int j = 6; j >= 0; --j) {
        if ((uint32_t)j < cv.r) {
#pragma unroll
          for (uint32_t b = 0; b < 16; ++b) off[b] = (off[b] << 1) | ((th[j] >> b) & 1u);
        }
      }
      uint64_t m = 0;
#pragma unroll
      for (uint32_t b = 0; b < 16; ++b) m = off[b] > m ? off[b] : m;   // rows >= n are zero
      m = uniform_u64(wave_max_u64(m));
      best = m > best ? m : best;
      have = true;
    }
  }
  if (have && lane == 0) atomicMax(best_out, (unsigned long long)best);
}

// One pass (a tuner choice; the two-pass form below is the default): every
// wave streams the top plane of its tiles
// and keeps (top, best), the largest top byte it has seen and the largest
// offset among its tiles whose top byte reached `top`. Such a tile is narrowed
// as the reference narrows its candidates: the rows holding the tile's top
// byte, then those holding the largest next byte among them, ..., then the tail
// bits from the highest (refinement loads redirected to lanes that still hold
// candidates, so they touch only lines with candidate rows), stopping as soon
// as the prefix falls below best's. A wave's best is the exact max of its
// tiles; each workgroup folds its waves' into *best_out with one atomic.
template <int U>
__global__ __launch_bounds__(256) void k_max_fused(ChunkView cv, unsigned long long* __restrict__ best_out) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  const uint64_t ngroups = (cv.ntiles + U - 1) / U;
  const uint32_t P = cv.P, r = cv.r, F = cv.F;
  uint32_t top = 0;
  uint64_t best = 0;
  bool have = false;
  for (uint64_t grp = wave; grp < ngroups; grp += nwaves) {
    const uint64_t c0 = grp * U;
    uint4 w[U];
    if (P > 0) {
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) w[u] = load_plane(cv, 0, min(c0 + u, cv.ntiles - 1), lane);
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint64_t c = c0 + u;
      if (c >= cv.ntiles) break;
      uint32_t mb = 0;
      if (P > 0)
        mb = __builtin_amdgcn_readfirstlane(wave_max_u32(
            max(max(max_byte4(w[u].x), max_byte4(w[u].y)), max(max_byte4(w[u].z), max_byte4(w[u].w)))));
      if (have && mb < top) continue;   // wave-uniform
      const bool first_top = !have || mb > top;
      uint32_t cand = P > 0 ? mask_eq(w[u], make_digit(mb)) : 0xFFFFu;
      uint64_t pre = mb;
      uint32_t bits = P > 0 ? 8u : 0u;
      bool beat = first_top;   // the prefix is already above best's
      bool lost = false;       // ... or below it: the tile cannot hold the max
      for (uint32_t p = 1; p < P && !lost; ++p) {
        const uint4 wp = load_plane(cv, p, c, refine_src(cand, lane));
        const uint32_t pb = __builtin_amdgcn_readfirstlane(wave_max_u32(masked_max_plus1(wp, cand))) - 1u;
        pre = (pre << 8) | pb;
        bits += 8;
        if (!beat) {
          const uint64_t bp = best >> (F - bits);
          lost = pre < bp;
          beat = pre > bp;
        }
        cand &= mask_eq(wp, make_digit(pb));
      }
      if (!lost && r > 0) {
        const uint32_t src = refine_src(cand, lane);
        uint32_t h[8];
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) h[j] = j < r ? load_tail(cv, j, c, src) : 0u;
#pragma unroll
        for (int j = 7; j >= 0; --j) {
          if ((uint32_t)j >= r || lost) continue;
          const uint32_t one = __ballot((cand & h[j]) != 0) != 0 ? 1u : 0u;
          pre = (pre << 1) | one;
          bits += 1;
          if (!beat) {
            const uint64_t bp = best >> (F - bits);
            lost = pre < bp;
            beat = pre > bp;
          }
          cand &= one ? h[j] : ~h[j];
        }
      }
      if (lost) continue;
      if (first_top) {
        top = mb;
        best = pre;
        have = true;
      } else if (pre > best) {
        best = pre;
      }
    }
  }
  __shared__ unsigned long long wb[16];
  if (lane == 0) wb[wave_in_wg()] = best;   // 0 for a wave without tiles: no larger than the max
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long b = 0;
    for (uint32_t k = 0; k < (blockDim.x >> 6); ++k) b = max(b, wb[k]);
    atomicMax(best_out, b);
  }
}

// ----------------------------------------------------------------------------
// K7: materialization. Full (split_double.rs:1243-1350) and semantic reduced
// precision (top R bits kept, re-scaled). A lane builds the values of its 16
// rows, parks them in LDS in row order, and the wave writes rows 64g + t
// (coalesced).

template <int NP, int NT, int U>
__global__ __launch_bounds__(256) void k_materialize(ChunkView cv, MatParams mp, double* __restrict__ out) {
  __shared__ double s_val[4][64 * kLdsRow];
  const uint32_t lane = __lane_id();
  const uint32_t wslot = wave_in_wg();
  double* sv = s_val[wslot];
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  const uint64_t ngroups = (cv.ntiles + U - 1) / U;
  const uint32_t lo = cv.r - NT;   // lowest tail bit kept
  constexpr int NPP = NP > 0 ? NP : 1;
  constexpr int NTT = NT > 0 ? NT : 1;
  for (uint64_t grp = wave; grp < ngroups; grp += nwaves) {
    const uint64_t c0 = grp * U;
    uint4 w[U][NPP];
    uint32_t h[U][NTT];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t c = min(c0 + u, cv.ntiles - 1);
#pragma unroll
      for (int p = 0; p < NP; ++p) w[u][p] = load_plane(cv, p, c, lane);
#pragma unroll
      for (int k = 0; k < NT; ++k) h[u][k] = cv.tail[(c * cv.r + lo + k) * 64 + lane];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t c = c0 + u;
      if (c >= cv.ntiles) break;
#pragma unroll
      for (uint32_t b = 0; b < 16; ++b) {
        uint64_t v = 0;
#pragma unroll
        for (int p = 0; p < NP; ++p) v = (v << 8) | byte_at(w[u][p], b);
#pragma unroll
        for (int k = NT - 1; k >= 0; --k) v = (v << 1) | ((h[u][k] >> b) & 1u);
        const uint64_t o = (v >> mp.drop) << mp.shift_back;
        sv[kLdsRow * lane + b] = (double)(int64_t)((uint64_t)cv.base + o) * mp.inv_scale;
      }
      wave_lds_sync();
      double* ot = out + c * kTileRows + lane;
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
  }
}

// Bit `pos` of the packed reference tail (stream bit i*r + j <-> row i, bit j).
__device__ __forceinline__ uint32_t packed_tail_bit(const ChunkView& cv, uint64_t pos) {
  const uint64_t row = pos / cv.r;
  const uint32_t j = (uint32_t)(pos % cv.r);
  if (row >= cv.n) return 0;
  const uint64_t c = row / kTileRows;
  const uint32_t t = (uint32_t)((row % kTileRows) / 16), b = (uint32_t)(row % 16);
  return (cv.tail[(c * cv.r + j) * 64 + t] >> b) & 1u;
}

// The reference's reduced-precision decode (split_double.rs:1612-1893): the top
// R bits placed unshifted and, when R % 8 != 0, (R % 8)-bit fields read at
// stride R % 8 from the packed tail. Compatibility only; one row per thread.
__global__ __launch_bounds__(256) void k_materialize_ref(ChunkView cv, MatParams mp, double* __restrict__ out) {
  for (uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; row < cv.n;
       row += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t c = row / kTileRows;
    const uint32_t t = (uint32_t)((row % kTileRows) / 16), b = (uint32_t)(row % 16);
    uint64_t v = 0;
    for (uint32_t p = 0; p < mp.planes; ++p) {
      const uint8_t* lanebytes = cv.planes + p * cv.pitch + c * kTileRows + 16 * t;
      v = (v << 8) | lanebytes[4 * (b & 3) + (b >> 2)];
    }
    if (mp.tail_bits) {
      uint64_t tv = 0;
      for (uint32_t q = 0; q < mp.tail_bits; ++q) tv |= (uint64_t)packed_tail_bit(cv, row * mp.tail_bits + q) << q;
      v = (v << mp.tail_bits) | tv;
    }
    out[row] = (double)(int64_t)((uint64_t)cv.base + v) * mp.inv_scale;
  }
}

// ----------------------------------------------------------------------------
// Stream <-> resident conversion (split_double.rs:349-414 byte layout).
// body = stream + 20. Plane p at body + p*n; the tail at body + P*n packs row
// i's r bits at stream bits i*r .. i*r+r-1, so a lane's 16 rows own the 2r
// bytes at 128*r*c + 2*r*t.

__global__ __launch_bounds__(256) void k_serialize(ChunkView cv, uint8_t flipmask, uint8_t* __restrict__ body,
                                                   uint64_t tail_bytes) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  for (uint64_t c = wave; c < cv.ntiles; c += nwaves) {
    const uint64_t r0 = c * kTileRows + 16ull * lane;
    for (uint32_t p = 0; p < cv.P; ++p) {
      const uint4 w = load_plane(cv, p, c, lane);
      uint8_t* dst = body + p * cv.n + r0;
#pragma unroll
      for (uint32_t b = 0; b < 16; ++b)
        if (r0 + b < cv.n) dst[b] = (uint8_t)byte_at(w, b) ^ flipmask;
    }
    if (cv.r) {
      uint32_t h[8];
      for (uint32_t j = 0; j < 8; ++j) h[j] = j < cv.r ? load_tail(cv, j, c, lane) : 0u;
      // 16 values of r bits, LSB-first: bit b*r + j = bit j of row b
      uint8_t* tdst = body + cv.P * cv.n + 128ull * cv.r * c + 2ull * cv.r * lane;
      const uint64_t tbase = 128ull * cv.r * c + 2ull * cv.r * lane;
      for (uint32_t m = 0; m < 2 * cv.r; ++m) {
        if (tbase + m >= tail_bytes) break;
        uint32_t byte = 0;
        for (uint32_t q = 0; q < 8; ++q) {
          const uint32_t bit = 8 * m + q;
          const uint32_t b = bit / cv.r, j = bit % cv.r;
          byte |= ((h[j] >> b) & 1u) << q;
        }
        tdst[m] = (uint8_t)byte;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_load(ChunkView cv, uint8_t flipmask, const uint8_t* __restrict__ body,
                                              uint64_t tail_bytes) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  for (uint64_t c = wave; c < cv.ntiles; c += nwaves) {
    const uint64_t r0 = c * kTileRows + 16ull * lane;
    for (uint32_t p = 0; p < cv.P; ++p) {
      const uint8_t* src = body + p * cv.n + r0;
      uint32_t by[16];
#pragma unroll
      for (uint32_t b = 0; b < 16; ++b) by[b] = r0 + b < cv.n ? (uint32_t)(src[b] ^ flipmask) : 0u;
      store_plane_nt(cv.planes + p * cv.pitch + c * kTileRows, lane, pack_lane(by));
    }
    if (cv.r) {
      const uint8_t* tsrc = body + cv.P * cv.n + 128ull * cv.r * c + 2ull * cv.r * lane;
      const uint64_t tbase = 128ull * cv.r * c + 2ull * cv.r * lane;
      uint32_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (uint32_t m = 0; m < 2 * cv.r; ++m) {
        if (tbase + m >= tail_bytes) break;
        const uint32_t byte = tsrc[m];
        for (uint32_t q = 0; q < 8; ++q) {
          const uint32_t bit = 8 * m + q;
          const uint32_t b = bit / cv.r, j = bit % cv.r;
          if (r0 + b < cv.n) h[j] |= ((byte >> q) & 1u) << b;
        }
      }
      for (uint32_t j = 0; j < cv.r; ++j) cv.tail[(c * cv.r + j) * 64 + lane] = (uint16_t)h[j];
    }
  }
}

// ----------------------------------------------------------------------------
// Byte-slice stream (buff_slice.rs:298-437): the P full planes are the byte_fixed
// planes XOR 0x80 (k_serialize / k_load with flipmask 0x80 and no tail); the r
// tail bits of row i become one more plane, byte i = (tail << (8 - r)) ^ 0x80.

__global__ __launch_bounds__(256) void k_slice_tail_serialize(ChunkView cv, uint8_t* __restrict__ plane) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  for (uint64_t c = wave; c < cv.ntiles; c += nwaves) {
    const uint64_t r0 = c * kTileRows + 16ull * lane;
    uint32_t h[8];
    for (uint32_t j = 0; j < 8; ++j) h[j] = j < cv.r ? load_tail(cv, j, c, lane) : 0u;
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) {
      uint32_t tv = 0;
      for (uint32_t j = 0; j < cv.r; ++j) tv |= ((h[j] >> b) & 1u) << j;
      if (r0 + b < cv.n) plane[r0 + b] = (uint8_t)((tv << (8 - cv.r)) ^ 0x80u);
    }
  }
}

__global__ __launch_bounds__(256) void k_slice_tail_load(ChunkView cv, const uint8_t* __restrict__ plane) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = wave_global();
  const uint64_t nwaves = waves_in_grid();
  for (uint64_t c = wave; c < cv.ntiles; c += nwaves) {
    const uint64_t r0 = c * kTileRows + 16ull * lane;
    uint32_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) {
      const uint32_t tv = r0 + b < cv.n ? ((uint32_t)(plane[r0 + b] ^ 0x80u) >> (8 - cv.r)) : 0u;
      for (uint32_t j = 0; j < 8; ++j) h[j] |= ((tv >> j) & 1u) << b;
    }
    for (uint32_t j = 0; j < cv.r; ++j) cv.tail[(c * cv.r + j) * 64 + lane] = (uint16_t)h[j];
  }
}

// out[w] = (a[w] or all rows, complemented when inv_a) & (b[w] or its
// complement when inv_b; absent: all rows) & (cond[w] if given), restricted to
// rows < cover. Element-wise, so out may alias a, b or cond.
__global__ void k_bitmap_logic(uint64_t* out, const uint64_t* a, const uint64_t* b, const uint64_t* cond,
                               uint64_t nwords, uint64_t cover, int inv_a, int inv_b) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t x = a ? a[i] : ~0ull;
    if (inv_a) x = ~x;
    if (b) x &= inv_b ? ~b[i] : b[i];
    if (cond) x &= cond[i];
    const uint64_t r0 = i * 64;
    x &= r0 + 64 <= cover ? ~0ull : (r0 < cover ? (1ull << (cover - r0)) - 1 : 0ull);
    out[i] = x;
  }
}

// ----------------------------------------------------------------------------
// Host-side launchers (declared in buff_kernels.h).

// Grid cap of a launch shape: its max_wg, else `dflt`.
static inline uint32_t cap_of(const LaunchShape& ls, uint32_t dflt) { return ls.max_wg ? ls.max_wg : dflt; }

static inline uint32_t grid_for(uint64_t units, uint32_t waves_per_wg, uint32_t max_wg) {
  uint64_t wgs = (units + waves_per_wg - 1) / waves_per_wg;
  if (wgs < 1) wgs = 1;
  if (wgs > max_wg) wgs = max_wg;
  return (uint32_t)wgs;
}

uint32_t minmax_grid(uint64_t ntiles) {
  // 1024 workgroups (16 waves per CU) measured faster than the ~1800 that fit
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>({ntiles, 1024, resident_wgs(k_minmax, 256)}));
}

hipError_t launch_minmax(const double* x, uint64_t n, int32_t pexp, uint32_t dlen, int64_t* partial,
                         int64_t* partial2, MinMaxOut* out, uint32_t nwg, bool prefix, hipStream_t s) {
  const uint64_t ntiles = (n + kTileRows - 1) / kTileRows;
  const uint64_t tpw = (ntiles + nwg - 1) / nwg;
  if (!prefix) {
    launch_k(k_minmax, dim3(nwg), dim3(256), 0, s, x, n, ntiles, tpw, pexp, dlen, partial, (int64_t*)nullptr,
             (int64_t*)nullptr);
    launch_k(k_minmax_final, dim3(1), dim3(256), 0, s, x, n, pexp, dlen, (const int64_t*)partial, nwg,
                       out);
  } else {
    launch_k(k_minmax_prefix, dim3(nwg), dim3(64), 0, s, x, n, ntiles, tpw, pexp, dlen,
                       (const int64_t*)partial, partial2);
    launch_k(k_minmax_prefix_final, dim3(1), dim3(64), 0, s, (const int64_t*)partial2, nwg, out);
  }
  return hipGetLastError();
}

hipError_t launch_minmax_slots(const double* x, uint64_t n, int32_t pexp, uint32_t dlen, int64_t* partial,
                               int64_t* set, int64_t* next, uint32_t nwg, hipStream_t s) {
  if (!set || !next || n == 0) return hipErrorInvalidValue;
  const uint64_t ntiles = (n + kTileRows - 1) / kTileRows;
  const uint64_t tpw = (ntiles + nwg - 1) / nwg;
  launch_k(k_minmax, dim3(nwg), dim3(256), 0, s, x, n, ntiles, tpw, pexp, dlen, partial, set, next);
  return hipGetLastError();
}

// One wave: the folded result of the reduction before it on the stream to the
// mailbox (lane k writes word k), then the sequence number (a system-scope
// release after every lane's words).
__global__ __launch_bounds__(64) void k_publish(const uint64_t* __restrict__ src, uint32_t mode, uint32_t nk,
                                                Publish pub) {
  const uint32_t lane = __lane_id();
  uint64_t mine = 0;
  uint32_t nw = nk;
  if (mode == kPublishSumSlots) {
    for (uint32_t k = 0; k < nk; ++k) {
      const uint64_t t = wave_sum_u64(src[k * kSumSlots + lane]);
      if (lane == k) mine = t;
    }
  } else if (mode == kPublishMinMax) {
    const int64_t a = wave_min_i64((int64_t)src[lane]);
    const int64_t b = wave_max_i64((int64_t)src[kMinMaxSlots + lane]);
    mine = lane == 0 ? (uint64_t)a : (lane == 1 ? (uint64_t)b : src[2 * kMinMaxSlots]);
    nw = 3;
  } else if (lane < nk) {
    mine = src[lane];
  }
  if (lane < nw) pub.mbox[1 + lane] = mine;
  __threadfence_system();
  __syncthreads();
  if (lane == 0) __hip_atomic_store(pub.mbox, pub.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

hipError_t launch_publish(const void* src, uint32_t mode, uint32_t nk, const Publish& pub, hipStream_t s) {
  if (!src || !pub.mbox || mode > kPublishCopy || nk >= kPublishWords ||
      (mode == kPublishSumSlots && (nk == 0 || nk > kMaxPlanes + 1)))
    return hipErrorInvalidValue;
  launch_k(k_publish, dim3(1), dim3(64), 0, s, static_cast<const uint64_t*>(src), mode, nk, pub);
  return hipGetLastError();
}

hipError_t launch_minmax_slots_init(int64_t* sets, uint32_t nsets, hipStream_t s) {
  launch_k(k_minmax_slots_init, dim3(1), dim3(256), 0, s, sets, nsets);
  return hipGetLastError();
}

hipError_t launch_encode(const double* x, const ChunkView& cv, int32_t pexp, hipStream_t s, const LaunchShape& ls) {
  // k_encode's LDS holds 4 waves
  const uint32_t wg = grid_for(cv.ntiles, 4, cap_of(ls, 2048));
  launch_k(k_encode, dim3(wg), dim3(256), 0, s, x, cv.n, cv.ntiles, pexp, cv.dlen, cv.base, cv.F, cv.P,
                     cv.r, cv.planes, cv.pitch, cv.tail);
  return hipGetLastError();
}

template <int MODE, int U, bool PF>
static void launch_filter_u(const ChunkView& cv, const FilterTarget& tg, uint64_t* out, uint64_t nwords,
                            unsigned long long* stats, hipStream_t s, const LaunchShape& ls) {
  const uint32_t block = 64 * ls.waves_per_wg;
  const uint32_t wg =
      grid_for((cv.ntiles + U - 1) / U, ls.waves_per_wg, cap_of(ls, resident_wgs(k_filter<MODE, U, PF>, block)));
  launch_k((k_filter<MODE, U, PF>), dim3(wg), dim3(block), 0, s, cv, tg, out, nwords, stats,
           static_cast<const uint16_t*>(nullptr));
}

template <int MODE, bool PF>
static void launch_filter_mp(const ChunkView& cv, const FilterTarget& tg, uint64_t* out, uint64_t nwords,
                             unsigned long long* stats, hipStream_t s, const LaunchShape& ls) {
  switch (ls.tiles_per_wave) {
    case 2: launch_filter_u<MODE, 2, PF>(cv, tg, out, nwords, stats, s, ls); break;
    case 8: launch_filter_u<MODE, 8, PF>(cv, tg, out, nwords, stats, s, ls); break;
    default: launch_filter_u<MODE, 4, PF>(cv, tg, out, nwords, stats, s, ls); break;
  }
}

template <int MODE>
static void launch_filter_m(const ChunkView& cv, const FilterTarget& tg, uint64_t* out, uint64_t nwords,
                            unsigned long long* stats, hipStream_t s, const LaunchShape& ls) {
  if (ls.prefetch) launch_filter_mp<MODE, true>(cv, tg, out, nwords, stats, s, ls);
  else launch_filter_mp<MODE, false>(cv, tg, out, nwords, stats, s, ls);
}

hipError_t launch_filter(const ChunkView& cv, int op, const FilterTarget& tg, uint64_t* out, uint64_t nwords,
                         unsigned long long* stats, hipStream_t s, const LaunchShape& ls) {
  if (op == 0) launch_filter_m<0>(cv, tg, out, nwords, stats, s, ls);
  else launch_filter_m<1>(cv, tg, out, nwords, stats, s, ls);
  return hipGetLastError();
}

hipError_t launch_filter_cond(const ChunkView& cv, int op, const FilterTarget& tg, const uint64_t* cond,
                              uint64_t* out, uint64_t nwords, hipStream_t s, const LaunchShape& ls) {
  const uint32_t block = 64 * ls.waves_per_wg;
  const uint16_t* c16 = reinterpret_cast<const uint16_t*>(cond);
  if (op == 0) {
    const uint32_t wg = grid_for((cv.ntiles + 3) / 4, ls.waves_per_wg,
                                 cap_of(ls, resident_wgs(k_filter<0, 4, false, true>, block)));
    launch_k((k_filter<0, 4, false, true>), dim3(wg), dim3(block), 0, s, cv, tg, out, nwords,
             static_cast<unsigned long long*>(nullptr), c16);
  } else {
    const uint32_t wg = grid_for((cv.ntiles + 3) / 4, ls.waves_per_wg,
                                 cap_of(ls, resident_wgs(k_filter<1, 4, false, true>, block)));
    launch_k((k_filter<1, 4, false, true>), dim3(wg), dim3(block), 0, s, cv, tg, out, nwords,
             static_cast<unsigned long long*>(nullptr), c16);
  }
  return hipGetLastError();
}

// Rows of the condition (mode 1, masked to the n rows) or none (mode 0).
__global__ void k_cond_fill(uint64_t* __restrict__ out, const uint64_t* __restrict__ cond, uint64_t nwords,
                            uint64_t n, int copy) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t w = copy ? cond[i] : 0ull;
    if (copy && i == nwords - 1 && (n & 63)) w &= (1ull << (n & 63)) - 1;
    out[i] = w;
  }
}

hipError_t launch_cond_fill(uint64_t* out, const uint64_t* cond, uint64_t nwords, uint64_t n, int copy,
                            hipStream_t s) {
  uint32_t wg = (uint32_t)std::min<uint64_t>((nwords + 255) / 256, 4096);
  if (wg == 0) return hipSuccess;
  launch_k(k_cond_fill, dim3(wg), dim3(256), 0, s, out, cond, nwords, n, copy);
  return hipGetLastError();
}

hipError_t launch_fill_bitmap(uint64_t* out, uint64_t nwords, uint64_t n, int ones, hipStream_t s) {
  uint32_t wg = (uint32_t)std::min<uint64_t>((nwords + 255) / 256, 4096);
  if (wg == 0) return hipSuccess;
  launch_k(k_fill_bitmap, dim3(wg), dim3(256), 0, s, out, nwords, n, ones);
  return hipGetLastError();
}

hipError_t launch_fill_bytes(void* out, uint64_t bytes, uint8_t value, hipStream_t s) {
  const uint32_t v = 0x01010101u * value;
  const uint64_t n16 = bytes / 16;
  const uint64_t wg = std::min<uint64_t>(std::max<uint64_t>((n16 + 1023) / 1024, 1), 16384);
  if (n16) launch_k(k_fill_bytes, dim3((uint32_t)wg), dim3(256), 0, s, (uint4*)out, n16, v);
  if (bytes % 16) {
    hipError_t e = hipMemsetAsync((uint8_t*)out + n16 * 16, value, bytes % 16, s);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

hipError_t launch_copy_bytes(void* out, const void* in, uint64_t bytes, hipStream_t s) {
  if (bytes % 16) return hipErrorInvalidValue;
  const uint64_t n16 = bytes / 16;
  const uint64_t wg = std::min<uint64_t>(std::max<uint64_t>((n16 + 1023) / 1024, 1), 16384);
  if (n16) launch_k(k_copy_bytes, dim3((uint32_t)wg), dim3(256), 0, s, (const u32x4_t*)in, (u32x4_t*)out, n16);
  return hipGetLastError();
}

hipError_t launch_read_bytes(const void* in, uint64_t bytes, uint32_t* sink, hipStream_t s) {
  if (bytes % 16) return hipErrorInvalidValue;
  const uint64_t n16 = bytes / 16;
  const uint64_t groups = (n16 / 64 + 4) / 4;
  const uint64_t wg = std::min<uint64_t>(std::max<uint64_t>((groups + 1) / 2, 1), 16384);   // two waves each
  if (n16) launch_k(k_read_bytes, dim3((uint32_t)wg), dim3(128), 0, s, (const u32x4_t*)in, n16, sink);
  return hipGetLastError();
}

hipError_t launch_touch_bytes(const void* in, uint64_t bytes, uint32_t* sink, hipStream_t s) {
  const uint64_t n16 = bytes / 16;
  if (n16) launch_k(k_touch_bytes, dim3(2048), dim3(256), 0, s, (const uint4*)in, n16, sink);
  return hipGetLastError();
}

// next == nullptr: k_sum's rows after sums[0 .. kMaxPlanes], folded by
// k_sum_final into sums; else one launch adding into the slot set `sums`.
template <int P, int R, int U>
static void launch_sum_pru(const ChunkView& cv, unsigned long long* sums, unsigned long long* next, hipStream_t s,
                           const LaunchShape& ls) {
  const uint32_t block = 64 * ls.waves_per_wg;
  const uint32_t wg = std::min<uint32_t>(
      grid_for((cv.ntiles + U - 1) / U, ls.waves_per_wg, cap_of(ls, resident_wgs(k_sum<P, R, U>, block))),
      kMaxSumRows);
  if (next) {
    launch_k((k_sum<P, R, U>), dim3(wg), dim3(block), 0, s, cv, sums, next);
    return;
  }
  unsigned long long* rows = sums + (kMaxPlanes + 1);
  launch_k((k_sum<P, R, U>), dim3(wg), dim3(block), 0, s, cv, rows, (unsigned long long*)nullptr);
  launch_k((k_sum_final<P + 1>), dim3(1), dim3(256), 0, s, (const unsigned long long*)rows, wg, sums);
}

template <int P, int R>
static void launch_sum_pr(const ChunkView& cv, unsigned long long* sums, unsigned long long* next, hipStream_t s,
                          const LaunchShape& ls) {
  switch (ls.tiles_per_wave) {
    case 1: launch_sum_pru<P, R, 1>(cv, sums, next, s, ls); break;
    case 2: launch_sum_pru<P, R, 2>(cv, sums, next, s, ls); break;
    case 4: launch_sum_pru<P, R, 4>(cv, sums, next, s, ls); break;
    default: launch_sum_pru<P, R, (P + R) <= 3 ? 4 : 2>(cv, sums, next, s, ls); break;
  }
}

uint64_t sum_scratch_bytes(const LaunchShape& ls) {
  (void)ls;
  return (uint64_t)(kMaxSumRows + 1) * (kMaxPlanes + 1) * 8;
}

template <int P>
static hipError_t launch_sum_p(const ChunkView& cv, unsigned long long* sums, unsigned long long* next, hipStream_t s,
                               const LaunchShape& ls) {
  switch (cv.r) {
    case 0: launch_sum_pr<P, 0>(cv, sums, next, s, ls); break;
    case 1: launch_sum_pr<P, 1>(cv, sums, next, s, ls); break;
    case 2: launch_sum_pr<P, 2>(cv, sums, next, s, ls); break;
    case 3: launch_sum_pr<P, 3>(cv, sums, next, s, ls); break;
    case 4: launch_sum_pr<P, 4>(cv, sums, next, s, ls); break;
    case 5: launch_sum_pr<P, 5>(cv, sums, next, s, ls); break;
    case 6: launch_sum_pr<P, 6>(cv, sums, next, s, ls); break;
    case 7: launch_sum_pr<P, 7>(cv, sums, next, s, ls); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

static hipError_t launch_sum_any(const ChunkView& cv, unsigned long long* sums, unsigned long long* next,
                                 hipStream_t s, const LaunchShape& ls) {
  switch (cv.P) {
    case 0: return launch_sum_p<0>(cv, sums, next, s, ls);
    case 1: return launch_sum_p<1>(cv, sums, next, s, ls);
    case 2: return launch_sum_p<2>(cv, sums, next, s, ls);
    case 3: return launch_sum_p<3>(cv, sums, next, s, ls);
    case 4: return launch_sum_p<4>(cv, sums, next, s, ls);
    case 5: return launch_sum_p<5>(cv, sums, next, s, ls);
    case 6: return launch_sum_p<6>(cv, sums, next, s, ls);
    case 7: return launch_sum_p<7>(cv, sums, next, s, ls);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_sum(const ChunkView& cv, unsigned long long* sums, hipStream_t s, const LaunchShape& ls) {
  return launch_sum_any(cv, sums, nullptr, s, ls);
}

hipError_t launch_sum_slots(const ChunkView& cv, unsigned long long* set, unsigned long long* next, hipStream_t s,
                            const LaunchShape& ls) {
  if (!set || !next) return hipErrorInvalidValue;
  return launch_sum_any(cv, set, next, s, ls);
}

uint64_t max_scratch_bytes(uint64_t ntiles) { return ((ntiles + 255) / 256) * 256; }

// scratch: per-tile maxima (ntiles bytes, padded); `best` is followed by the
// 64 u32 slots of pass 1 (kMaxBestBytes in all, zeroed by the caller).
template <int U>
static void launch_max_top_u(const ChunkView& cv, uint8_t* scratch, unsigned int* slots, uint32_t tag,
                             unsigned long long* best, hipStream_t s, const LaunchShape& ls) {
  const uint32_t block = 64 * ls.waves_per_wg;
  const uint32_t wg = std::min<uint32_t>(
      grid_for((cv.ntiles + U - 1) / U, ls.waves_per_wg, cap_of(ls, resident_wgs(k_max_top<U>, block))), 8192);
  launch_k((k_max_top<U>), dim3(wg), dim3(block), 0, s, cv, scratch, slots, tag, best);
}

template <int U>
static void launch_max_fused_u(const ChunkView& cv, unsigned long long* best, hipStream_t s, const LaunchShape& ls) {
  const uint32_t block = 64 * ls.waves_per_wg;
  const uint32_t wg =
      grid_for((cv.ntiles + U - 1) / U, ls.waves_per_wg, cap_of(ls, resident_wgs(k_max_fused<U>, block)));
  launch_k((k_max_fused<U>), dim3(wg), dim3(block), 0, s, cv, best);
}

hipError_t launch_max(const ChunkView& cv, unsigned long long* best, uint8_t* scratch, unsigned int* slots,
                      uint32_t tag, hipStream_t s, const LaunchShape& ls) {
  if (ls.prefetch == 1 || cv.P == 0) {   // no pass 1 to clear *best
    const hipError_t e = launch_fill_bytes(best, 16, 0, s);
    if (e != hipSuccess) return e;
  }
  if (ls.prefetch == 1) {   // one pass
    switch (ls.tiles_per_wave) {
      case 4: launch_max_fused_u<4>(cv, best, s, ls); break;
      case 16: launch_max_fused_u<16>(cv, best, s, ls); break;
      default: launch_max_fused_u<8>(cv, best, s, ls); break;
    }
    return hipGetLastError();
  }
  // two passes: every tile's top byte, then the tiles holding the largest
  if (cv.P > 0) {
    switch (ls.tiles_per_wave) {
      case 4: launch_max_top_u<4>(cv, scratch, slots, tag, best, s, ls); break;
      case 16: launch_max_top_u<16>(cv, scratch, slots, tag, best, s, ls); break;
      default: launch_max_top_u<8>(cv, scratch, slots, tag, best, s, ls); break;
    }
  }
  const uint32_t wg2 = std::min(grid_for((cv.ntiles + 63) / 64, 4, 1024), resident_wgs(k_max_refine, 256));
  launch_k(k_max_refine, dim3(wg2), dim3(256), 0, s, cv, (const uint8_t*)scratch, (const unsigned int*)slots, tag,
           best);
  return hipGetLastError();
}

template <int NP, int NT>
static void launch_mat_pt(const ChunkView& cv, const MatParams& mp, double* out, hipStream_t s,
                          const LaunchShape& ls) {
  constexpr int U = (NP + NT) <= 3 ? 2 : 1;
  // k_materialize's LDS holds 4 waves
  const uint32_t wg = grid_for((cv.ntiles + U - 1) / U, 4, cap_of(ls, 2048));
  launch_k((k_materialize<NP, NT, U>), dim3(wg), dim3(256), 0, s, cv, mp, out);
}

template <int NP>
static hipError_t launch_mat_p(const ChunkView& cv, const MatParams& mp, double* out, hipStream_t s,
                               const LaunchShape& ls) {
  switch (mp.tail_bits) {
    case 0: launch_mat_pt<NP, 0>(cv, mp, out, s, ls); break;
    case 1: launch_mat_pt<NP, 1>(cv, mp, out, s, ls); break;
    case 2: launch_mat_pt<NP, 2>(cv, mp, out, s, ls); break;
    case 3: launch_mat_pt<NP, 3>(cv, mp, out, s, ls); break;
    case 4: launch_mat_pt<NP, 4>(cv, mp, out, s, ls); break;
    case 5: launch_mat_pt<NP, 5>(cv, mp, out, s, ls); break;
    case 6: launch_mat_pt<NP, 6>(cv, mp, out, s, ls); break;
    case 7: launch_mat_pt<NP, 7>(cv, mp, out, s, ls); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_materialize(const ChunkView& cv, const MatParams& mp, double* out, hipStream_t s,
                              const LaunchShape& ls) {
  if (mp.mode == kMatRefPrecision) {
    const uint32_t wg = (uint32_t)std::min<uint64_t>((cv.n + 255) / 256, 4096);
    if (wg) launch_k(k_materialize_ref, dim3(wg), dim3(256), 0, s, cv, mp, out);
    return hipGetLastError();
  }
  switch (mp.planes) {
    case 0: return launch_mat_p<0>(cv, mp, out, s, ls);
    case 1: return launch_mat_p<1>(cv, mp, out, s, ls);
    case 2: return launch_mat_p<2>(cv, mp, out, s, ls);
    case 3: return launch_mat_p<3>(cv, mp, out, s, ls);
    case 4: return launch_mat_p<4>(cv, mp, out, s, ls);
    case 5: return launch_mat_p<5>(cv, mp, out, s, ls);
    case 6: return launch_mat_p<6>(cv, mp, out, s, ls);
    case 7: return launch_mat_p<7>(cv, mp, out, s, ls);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_serialize(const ChunkView& cv, uint8_t flipmask, uint8_t* body, uint64_t tail_bytes,
                            hipStream_t s) {
  const uint32_t wg = grid_for(cv.ntiles, 4, 2048);
  launch_k(k_serialize, dim3(wg), dim3(256), 0, s, cv, flipmask, body, tail_bytes);
  return hipGetLastError();
}

hipError_t launch_load(const ChunkView& cv, uint8_t flipmask, const uint8_t* body, uint64_t tail_bytes,
                       hipStream_t s) {
  const uint32_t wg = grid_for(cv.ntiles, 4, 2048);
  launch_k(k_load, dim3(wg), dim3(256), 0, s, cv, flipmask, body, tail_bytes);
  return hipGetLastError();
}

hipError_t launch_slice_tail_serialize(const ChunkView& cv, uint8_t* plane, hipStream_t s) {
  const uint32_t wg = grid_for(cv.ntiles, 4, 2048);
  launch_k(k_slice_tail_serialize, dim3(wg), dim3(256), 0, s, cv, plane);
  return hipGetLastError();
}

hipError_t launch_slice_tail_load(const ChunkView& cv, const uint8_t* plane, hipStream_t s) {
  const uint32_t wg = grid_for(cv.ntiles, 4, 2048);
  launch_k(k_slice_tail_load, dim3(wg), dim3(256), 0, s, cv, plane);
  return hipGetLastError();
}

hipError_t launch_bitmap_logic(uint64_t* out, const uint64_t* a, const uint64_t* b, const uint64_t* cond,
                               uint64_t nwords, uint64_t cover, int inv_a, int inv_b, hipStream_t s) {
  const uint32_t wg = (uint32_t)std::min<uint64_t>((nwords + 255) / 256, 2048);
  if (wg) launch_k(k_bitmap_logic, dim3(wg), dim3(256), 0, s, out, a, b, cond, nwords, cover, inv_a, inv_b);
  return hipGetLastError();
}

}  // namespace buffgpu
