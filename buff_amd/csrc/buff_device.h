// Device-side helpers shared by the kernel files (wave reductions, tile
// access in the resident layout of buff_common.h, SWAR byte compares).
#pragma once
#include <hip/hip_runtime.h>

#include "buff_common.h"

namespace buffgpu {

// ----------------------------------------------------------------------------
// wave helpers

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
// Exclusive prefix sum over the 64 lanes with DPP: Hillis-Steele inside each
// 16-lane row (row_shr 1, 2, 4, 8), then the row totals carried by row_bcast15
// (into rows 1 and 3) and row_bcast31 (into rows 2 and 3). Lanes whose DPP
// source lies outside the row read 0.
__device__ __forceinline__ uint32_t wave_excl_scan_u32(uint32_t v, uint32_t lane) {
  (void)lane;
  uint32_t s = v;
  s += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)s, 0x111, 0xf, 0xf, false);   // row_shr:1
  s += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)s, 0x112, 0xf, 0xf, false);   // row_shr:2
  s += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)s, 0x114, 0xf, 0xf, false);   // row_shr:4
  s += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)s, 0x118, 0xf, 0xf, false);   // row_shr:8
  s += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)s, 0x142, 0xa, 0xf, false);   // row_bcast:15
  s += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)s, 0x143, 0xc, 0xf, false);   // row_bcast:31
  return s - v;
}

// Max over the 64 lanes with the same DPP steps (every lane gets it).
__device__ __forceinline__ uint32_t wave_max_u32_dpp(uint32_t v) {
  uint32_t s = v;
  s = max(s, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)s, 0x111, 0xf, 0xf, false));   // row_shr:1
  s = max(s, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)s, 0x112, 0xf, 0xf, false));   // row_shr:2
  s = max(s, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)s, 0x114, 0xf, 0xf, false));   // row_shr:4
  s = max(s, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)s, 0x118, 0xf, 0xf, false));   // row_shr:8
  s = max(s, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)s, 0x142, 0xa, 0xf, false));   // row_bcast:15
  s = max(s, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)s, 0x143, 0xc, 0xf, false));   // row_bcast:31
  return (uint32_t)__builtin_amdgcn_readlane((int)s, 63);
}

__device__ __forceinline__ uint32_t wave_total_u32(uint32_t v) {
  const uint32_t incl = wave_excl_scan_u32(v, 0) + v;
  return (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += (uint32_t)__shfl_xor((int)v, m, 64);
  return v;
}

// Wave indices as the compiler can see they are wave-uniform (so tile and
// plane addresses derived from them stay in scalar registers and loads use a
// scalar base with a per-lane offset).
__device__ __forceinline__ uint32_t wave_in_wg() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }
__device__ __forceinline__ uint64_t wave_global() {
  return (uint64_t)blockIdx.x * (blockDim.x >> 6) + wave_in_wg();
}
__device__ __forceinline__ uint64_t waves_in_grid() { return (uint64_t)gridDim.x * (blockDim.x >> 6); }

__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
// Orders this wave's LDS writes before its later LDS reads (and vice versa).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ----------------------------------------------------------------------------
// Tile access

__device__ __forceinline__ bool tile_full(uint64_t n, uint64_t c) { return (c + 1) * kTileRows <= n; }

// Address-space views of generic pointers. Chunk data lives in global memory;
// descriptors that no kernel writes while it reads them (work lists, chunk
// views, query descriptors) are read as constant memory, so uniform reads of
// them become scalar loads and their fields stay in scalar registers.
#define BUFF_GLOBAL __attribute__((address_space(1)))
#define BUFF_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const BUFF_GLOBAL T* gmem(const T* p) {
  return (const BUFF_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ BUFF_GLOBAL T* gmem_w(T* p) {
  return (BUFF_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ const BUFF_CONST T* cmem(const T* p) {
  return (const BUFF_CONST T*)p;
}
// Element i of a descriptor array read as constant memory (word by word: class
// types cannot be copied through an address-space pointer).
template <typename T>
__device__ __forceinline__ T cload(const T* base, uint64_t i = 0) {
  static_assert(sizeof(T) % 4 == 0, "descriptor of whole words");
  T out;
  const BUFF_CONST uint32_t* src = (const BUFF_CONST uint32_t*)(base + i);
  uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
  for (uint32_t k = 0; k < sizeof(T) / 4; ++k) dst[k] = src[k];
  return out;
}

typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

// Plane bytes are streamed once per operation: loads are non-temporal, so they
// do not displace other lines from the caches.
__device__ __forceinline__ uint4 load_plane(const ChunkView& cv, uint32_t p, uint64_t c, uint32_t lane) {
  const u32x4_t v =
      __builtin_nontemporal_load(gmem(reinterpret_cast<const u32x4_t*>(cv.planes + p * cv.pitch + c * kTileRows)) + lane);
  return make_uint4(v.x, v.y, v.z, v.w);
}

// Streaming stores (written once, read by a later operation).
__device__ __forceinline__ void store_nt(double* p, double v) { __builtin_nontemporal_store(v, gmem_w(p)); }
__device__ __forceinline__ void store_plane_nt(uint8_t* tile, uint32_t lane, const uint4& w) {
  const u32x4_t v = {w.x, w.y, w.z, w.w};
  __builtin_nontemporal_store(v, gmem_w(reinterpret_cast<u32x4_t*>(tile)) + lane);
}

// The lane whose bytes a lane loads in a refinement round: its own when it
// still has a candidate row, else the first lane that has one (a line the
// round reads anyway, so idle lanes add no HBM traffic).
__device__ __forceinline__ uint32_t refine_src(uint32_t live, uint32_t lane) {
  const uint64_t b = __ballot(live != 0);
  return live ? lane : (b ? (uint32_t)__builtin_ctzll(b) : 0u);
}

// bit-plane j of tile c, lane's 16 rows (bit b <-> row 16*lane + b)
__device__ __forceinline__ uint32_t load_tail(const ChunkView& cv, uint32_t j, uint64_t c, uint32_t lane) {
  return gmem(cv.tail)[(c * cv.r + j) * 64 + lane];
}

// byte of row 16*lane + b inside the lane's 16 bytes (dword b%4, byte b/4)
__device__ __forceinline__ uint32_t byte_at(const uint4& w, uint32_t b) {
  const uint32_t i = b & 3, k = b >> 2;
  const uint32_t d = i == 0 ? w.x : i == 1 ? w.y : i == 2 ? w.z : w.w;
  return (d >> (8 * k)) & 0xFFu;
}

// 16 per-row bytes (row offsets b = 0..15) -> the lane's uint4
__device__ __forceinline__ uint4 pack_lane(const uint32_t (&by)[16]) {
  uint32_t d[4];
#pragma unroll
  for (uint32_t i = 0; i < 4; ++i) d[i] = by[i] | (by[4 + i] << 8) | (by[8 + i] << 16) | (by[12 + i] << 24);
  return make_uint4(d[0], d[1], d[2], d[3]);
}

// Byte k of each of 16 words, gathered into the lane's plane layout (dword i
// byte k' = word 4k' + i): two byte permutes and a shift-or per dword.
__device__ __forceinline__ uint4 gather_byte16(const uint32_t (&W)[16], uint32_t k) {
  const uint32_t sel = k | ((k + 4) << 8) | 0x0C0C0000u;   // bytes 0,1 <- byte k of the two sources; 2,3 <- 0
  uint32_t d[4];
#pragma unroll
  for (uint32_t i = 0; i < 4; ++i) {
    const uint32_t lo = __builtin_amdgcn_perm(W[4 + i], W[i], sel);
    const uint32_t hi = __builtin_amdgcn_perm(W[12 + i], W[8 + i], sel);
    d[i] = lo | (hi << 16);
  }
  return make_uint4(d[0], d[1], d[2], d[3]);
}

// Byte planes and tail bit-planes of tile c from the lane's 16 offsets (row
// 16*lane + b in off[b]; bits above the low 8P + r are ignored). Plane p holds
// bits [r + 8(P-1-p), +8) of an offset, which is byte P-1-p of (off >> r): one
// shift by r byte-aligns every plane, and each plane is a byte gather. Tail
// bit j of the 16 rows is gathered from the low bytes (already in the plane
// layout) by one mask per dword and a nibble merge.
__device__ __forceinline__ void store_tile_planes(const uint64_t (&off)[16], uint32_t P, uint32_t r,
                                                  uint8_t* planes, uint64_t pitch, uint16_t* tail, uint64_t c,
                                                  uint32_t lane) {
  uint32_t lo[16], hi[16], low8[16];
#pragma unroll
  for (uint32_t b = 0; b < 16; ++b) {
    const uint64_t a = off[b] >> r;
    lo[b] = (uint32_t)a;
    hi[b] = (uint32_t)(a >> 32);
    low8[b] = (uint32_t)off[b];
  }
  uint8_t* tile = planes + c * kTileRows;
  uint32_t p = 0;
#pragma unroll 1
  for (; p + 4 < P; ++p) store_plane_nt(tile + p * pitch, lane, gather_byte16(hi, P - 5 - p));
#pragma unroll 1
  for (; p < P; ++p) store_plane_nt(tile + p * pitch, lane, gather_byte16(lo, P - 1 - p));
  if (r == 0) return;
  const uint4 L = gather_byte16(low8, 0);
#pragma unroll 1
  for (uint32_t j = 0; j < r; ++j) {
    // bit 8k + i of G: bit j of row 4k + i
    uint32_t G = (L.x >> j) & 0x01010101u;
    G |= ((L.y >> j) & 0x01010101u) << 1;
    G |= ((L.z >> j) & 0x01010101u) << 2;
    G |= ((L.w >> j) & 0x01010101u) << 3;
    const uint32_t G2 = G | (G >> 4);   // byte 0: rows 0-7, byte 2: rows 8-15
    tail[(c * r + j) * 64 + lane] = (uint16_t)__builtin_amdgcn_perm(G2, G2, 0x0C0C0200u);
  }
}

// The lane's 16 rows of tile c as whole offsets (planes then r tail bits),
// assembled with byte permutes: the tail bit-planes are first spread into the
// plane layout as one byte per row holding the bits at its top (bit j of the
// tail at bit 8 - r + j), so row b's NP plane bytes and that byte form one
// integer X = planes * 2^8 + tail * 2^(8 - r), and the offset is X >> (8 - r).
// `src`: the lane whose bytes are loaded (a lane's own, or another live lane's).
template <int NP>
struct RowBytes {
  uint4 w[NP > 0 ? NP : 1];
  uint32_t T[4];   // tail byte of row 4k + i in byte k of T[i]
};

template <int NP>
__device__ __forceinline__ void load_row_bytes(const ChunkView& cv, uint64_t c, uint32_t src, RowBytes<NP>& rb) {
#pragma unroll
  for (int p = 0; p < NP; ++p) rb.w[p] = load_plane(cv, p, c, src);
  const uint32_t r = cv.r;
  uint32_t h[7];
#pragma unroll
  for (uint32_t j = 0; j < 7; ++j) h[j] = j < r ? load_tail(cv, j, c, src) : 0u;
  rb.T[0] = rb.T[1] = rb.T[2] = rb.T[3] = 0u;
#pragma unroll
  for (uint32_t j = 0; j < 7; ++j) {
    if (j >= r) break;
    const uint32_t x = h[j];
    // nibble k of the 16 row bits -> byte k
    const uint32_t y = (x & 0xFu) | ((x << 4) & 0xF00u) | ((x << 8) & 0xF0000u) | ((x << 12) & 0xF000000u);
    const uint32_t sh = 8u - r + j;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) rb.T[i] |= ((y >> i) & 0x01010101u) << sh;
  }
}

// X of row b (see above): byte 0 the tail byte, bytes 1.. the planes from the last.
template <int NP>
__device__ __forceinline__ uint64_t row_bytes_x(const RowBytes<NP>& rb, uint32_t b) {
  const uint32_t i = b & 3, k = b >> 2;
  const uint32_t sel = k | ((k + 4) << 8) | 0x0C0C0000u;
  auto byte_src = [&](int m) -> uint32_t {   // source dword of byte m of X
    if (m == 0) return i == 0 ? rb.T[0] : i == 1 ? rb.T[1] : i == 2 ? rb.T[2] : rb.T[3];
    const int p = NP - m;
    if (p < 0) return 0u;
    const uint4& q = rb.w[p < NP ? p : 0];
    return i == 0 ? q.x : i == 1 ? q.y : i == 2 ? q.z : q.w;
  };
  uint64_t X = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (2 * q > NP) break;
    X |= (uint64_t)__builtin_amdgcn_perm(byte_src(2 * q + 1), byte_src(2 * q), sel) << (16 * q);
  }
  return X;
}

template <int NP>
__device__ __forceinline__ uint64_t row_offset(const RowBytes<NP>& rb, uint32_t b, uint32_t r) {
  return row_bytes_x<NP>(rb, b) >> (8u - r);
}

// valid-row mask of the lane's 16 rows
__device__ __forceinline__ uint32_t lane_valid16(uint64_t n, uint64_t c, uint32_t lane) {
  const uint64_t r0 = c * kTileRows + 16ull * lane;
  if (r0 + 16 <= n) return 0xFFFFu;
  if (r0 >= n) return 0u;
  return (1u << (uint32_t)(n - r0)) - 1u;
}

// ---- SWAR byte compares (4 rows per dword) --------------------------------------
// bit 7 of every byte of the result holds the per-byte predicate.

// unsigned x > d for each byte, given D1 = (d+1) replicated and D1lo = D1 & 0x7f7f7f7f
// (x >= d+1: the low 7 bits compare without borrows across bytes, the high bits
// decide where they differ).
__device__ __forceinline__ uint32_t swar_gt(uint32_t x, uint32_t D1, uint32_t D1lo) {
  const uint32_t diff = (x | 0x80808080u) - D1lo;
  const uint32_t e = x ^ D1;
  return ((e & x) | (~e & diff)) & 0x80808080u;
}
// x == d for each byte, D = d replicated
__device__ __forceinline__ uint32_t swar_eq(uint32_t x, uint32_t D) {
  const uint32_t t = x ^ D;
  return ~(((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t) & 0x80808080u;
}
// bit 7 of byte k of gi -> bit 4k + i: the lane's 16-row mask
__device__ __forceinline__ uint32_t gather16(uint32_t g0, uint32_t g1, uint32_t g2, uint32_t g3) {
  const uint32_t G = (g0 >> 7) | (g1 >> 6) | (g2 >> 5) | (g3 >> 4);
  const uint32_t G2 = G | (G >> 4);
  return (G2 & 0xFFu) | ((G2 >> 8) & 0xFF00u);
}

struct Digit {  // one target byte, prepared for the SWAR compares
  uint32_t D, D1, D1lo;
  bool gt_possible;  // d < 255
};
__device__ __forceinline__ Digit make_digit(uint32_t d) {
  Digit r;
  r.D = 0x01010101u * d;
  r.gt_possible = d < 255;
  r.D1 = 0x01010101u * ((d + 1) & 0xFF);
  r.D1lo = r.D1 & 0x7F7F7F7Fu;
  return r;
}
__device__ __forceinline__ uint32_t mask_gt(const uint4& w, const Digit& dg) {
  if (!dg.gt_possible) return 0u;
  return gather16(swar_gt(w.x, dg.D1, dg.D1lo), swar_gt(w.y, dg.D1, dg.D1lo), swar_gt(w.z, dg.D1, dg.D1lo),
                  swar_gt(w.w, dg.D1, dg.D1lo));
}
__device__ __forceinline__ uint32_t mask_eq(const uint4& w, const Digit& dg) {
  return gather16(swar_eq(w.x, dg.D), swar_eq(w.y, dg.D), swar_eq(w.z, dg.D), swar_eq(w.w, dg.D));
}

__device__ __forceinline__ void load_tile_f64(const double* __restrict__ x, uint64_t n, uint64_t c, uint32_t lane,
                                              bool full, uint64_t (&v)[16]) {
  const BUFF_GLOBAL double* xt = gmem(x) + c * kTileRows + lane;
  if (full) {
#pragma unroll
    for (uint32_t g = 0; g < 16; ++g) v[g] = __double_as_longlong(__builtin_nontemporal_load(xt + 64 * g));
  } else {
#pragma unroll
    for (uint32_t g = 0; g < 16; ++g) {
      const uint64_t row = c * kTileRows + 64 * g + lane;
      v[g] = row < n ? __double_as_longlong(__builtin_nontemporal_load(xt + 64 * g)) : 0;
    }
  }
}

// LDS staging of a tile's 1024 u64/f64 values in row order (one pad slot per 16 rows).
constexpr uint32_t kLdsRow = 17;  // u64 slots per 16 rows (+1 pad)

__device__ __forceinline__ uint32_t lds_slot(uint32_t row_in_tile) { return row_in_tile + (row_in_tile >> 4); }

// Bit-sliced refinement of the candidate rows on the tail bit-planes:
// eqm keeps rows equal to the target so far, gtm (MODE 0) gains rows above it.
template <int MODE>
__device__ __forceinline__ void refine_tail(const ChunkView& cv, uint64_t c, uint64_t dt, uint32_t lane, uint32_t& gtm,
                                            uint32_t& eqm) {
  uint32_t h[8];
  const uint32_t src = refine_src(eqm, lane);
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j) h[j] = j < cv.r ? load_tail(cv, j, c, src) : 0u;   // only the r bit-planes
#pragma unroll
  for (int j = 7; j >= 0; --j) {
    if ((uint32_t)j >= cv.r) continue;
    if ((dt >> j) & 1) {
      eqm &= h[j];
    } else {
      if (MODE == 0) gtm |= eqm & h[j];
      eqm &= ~h[j];
    }
  }
}

__device__ __forceinline__ uint32_t max_byte4(uint32_t d) {
  return max(max(d & 0xFFu, (d >> 8) & 0xFFu), max((d >> 16) & 0xFFu, d >> 24));
}

// max over the rows in `mask` of (byte + 1); 0 when the mask is empty
__device__ __forceinline__ uint32_t masked_max_plus1(const uint4& w, uint32_t mask) {
  uint32_t v = 0;
#pragma unroll
  for (uint32_t b = 0; b < 16; ++b) v = ((mask >> b) & 1u) ? max(v, byte_at(w, b) + 1u) : v;
  return v;
}

}  // namespace buffgpu
