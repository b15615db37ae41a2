// Device helpers of the buff-major layout shared by its kernels
// (buff_major.hip) and the batched query kernels (buff_batch.hip).
//
// A sparse column's bytes for a lane are the majority byte replicated, with the
// outlier bytes patched in: the lane's 16-bit slice of the outlier bitmap says
// which rows, the tile's prefix count plus an exclusive scan of the lanes'
// popcounts says where their bytes start in the compacted array.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "buff_common.h"
#include "buff_device.h"
#include "buff_major.h"

namespace buffgpu {

// The lane's packed outlier bytes (k-th outlier of the lane = byte k) as four
// dwords, from a 20-byte window of aligned loads at its compacted position.
struct Packed16 {
  uint32_t d0, d1, d2, d3;
};
__device__ __forceinline__ Packed16 load_packed(const uint8_t* outl, uint64_t pos) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(outl + (pos & ~3ull));
  const uint32_t sh = (uint32_t)(pos & 3);
  const uint32_t a = w[0], b = w[1], c = w[2], d = w[3], e = w[4];
  Packed16 p;
  p.d0 = __builtin_amdgcn_alignbyte(b, a, sh);
  p.d1 = __builtin_amdgcn_alignbyte(c, b, sh);
  p.d2 = __builtin_amdgcn_alignbyte(d, c, sh);
  p.d3 = __builtin_amdgcn_alignbyte(e, d, sh);
  return p;
}
// Row masks (bit b <-> row 16*lane + b) of one sparse column against target byte
// t: majority rows all at once, outlier rows from the packed bytes. The lane's
// k-th outlier (k-th set bit of m16, ascending rows) is byte k of its packed
// window; k runs with a wave-uniform bound in an unrolled loop, so every step
// reads a fixed register and byte (no per-lane select of the byte). Only the
// rows in `want` keep their results.
__device__ __forceinline__ void sparse_cmp(const Packed16& pk, uint32_t m16, uint32_t want, uint32_t major, uint32_t t,
                                           uint32_t& gt, uint32_t& eq) {
  const uint32_t nm = ~m16 & 0xFFFFu;
  const uint32_t d[4] = {pk.d0, pk.d1, pk.d2, pk.d3};
  uint32_t go = 0, eo = 0, rem = m16;
#pragma unroll
  for (uint32_t k = 0; k < 16; ++k) {
    if (!__ballot((rem & want) != 0)) break;
    const uint32_t pos = (uint32_t)__builtin_ctz(rem | 0x10000u);   // 16 once the lane has no outlier left
    rem &= rem - 1u;
    const uint32_t v = (d[k >> 2] >> (8 * (k & 3))) & 0xFFu;
    go |= (uint32_t)(v > t) << pos;
    eo |= (uint32_t)(v == t) << pos;
  }
  gt = (major > t ? nm : 0u) | (go & want);
  eq = (major == t ? nm : 0u) | (eo & want);
}

// Replaces byte b (row 16*lane + b: dword b%4, byte b/4) of a lane's 16 bytes
// when `on` (written with selects: every lane runs the same instructions).
__device__ __forceinline__ void set_byte(uint32_t& d0, uint32_t& d1, uint32_t& d2, uint32_t& d3, uint32_t b,
                                         uint32_t v, bool on) {
  const uint32_t sh = 8 * ((b >> 2) & 3);
  const uint32_t keep = ~(0xFFu << sh), val = v << sh;
  const uint32_t i = b & 3;
  d0 = on && i == 0 ? (d0 & keep) | val : d0;
  d1 = on && i == 1 ? (d1 & keep) | val : d1;
  d2 = on && i == 2 ? (d2 & keep) | val : d2;
  d3 = on && i == 3 ? (d3 & keep) | val : d3;
}

// The lane's 16 bytes of a sparse column: the majority byte, with the outlier
// bytes of the rows in `want` (a subset of m16) patched in. The outliers are
// walked by rank as in sparse_cmp.
__device__ __forceinline__ uint4 patch_outliers(const Packed16& pk, uint32_t m16, uint32_t want, uint32_t major) {
  const uint32_t M = 0x01010101u * major;
  uint32_t o0 = M, o1 = M, o2 = M, o3 = M;
  const uint32_t d[4] = {pk.d0, pk.d1, pk.d2, pk.d3};
  uint32_t rem = m16;
#pragma unroll
  for (uint32_t k = 0; k < 16; ++k) {
    if (!__ballot((rem & want) != 0)) break;
    const uint32_t pos = (uint32_t)__builtin_ctz(rem | 0x10000u);
    rem &= rem - 1u;
    const uint32_t v = (d[k >> 2] >> (8 * (k & 3))) & 0xFFu;
    set_byte(o0, o1, o2, o3, pos, v, (want >> pos) & 1u);
  }
  return make_uint4(o0, o1, o2, o3);
}

// max over the outlier rows in `want` of (byte + 1); 0 when there is none.
__device__ __forceinline__ uint32_t sparse_max_plus1(const Packed16& pk, uint32_t m16, uint32_t want) {
  const uint32_t d[4] = {pk.d0, pk.d1, pk.d2, pk.d3};
  uint32_t rem = m16, best = 0;
#pragma unroll
  for (uint32_t k = 0; k < 16; ++k) {
    if (!__ballot((rem & want) != 0)) break;
    const uint32_t pos = (uint32_t)__builtin_ctz(rem | 0x10000u);
    rem &= rem - 1u;
    const uint32_t v = (d[k >> 2] >> (8 * (k & 3))) & 0xFFu;
    best = (want >> pos) & 1u ? max(best, v + 1u) : best;
  }
  return best;
}

// The outlier window of column sd for the lane in tile c (every lane of the
// wave calls it together: the exclusive scan).
__device__ __forceinline__ Packed16 lane_outliers(const SparseDev& sd, uint64_t c, uint32_t lane, uint32_t m16) {
  const uint32_t excl = wave_excl_scan_u32(__popc(m16), lane);
  return m16 ? load_packed(sd.outl, (uint64_t)sd.prefix[c] + excl) : Packed16{0, 0, 0, 0};
}

// Masks and outlier windows of every sparse column of tile c, with all their
// loads issued before any is consumed (the columns are independent).
struct SparseLanes {
  uint32_t m[kMaxSparse];
  Packed16 pk[kMaxSparse];
};

__device__ __forceinline__ void load_sparse(const MajorView& mv, uint64_t c, uint32_t lane, SparseLanes& sl) {
  uint32_t pre[kMaxSparse];
#pragma unroll
  for (uint32_t j = 0; j < kMaxSparse; ++j) {
    sl.m[j] = j < mv.S ? (uint32_t)mv.col[j].mask16[c * 64 + lane] : 0u;
    pre[j] = j < mv.S ? mv.col[j].prefix[c] : 0u;
  }
#pragma unroll
  for (uint32_t j = 0; j < kMaxSparse; ++j) {
    sl.pk[j] = Packed16{0, 0, 0, 0};
    if (j < mv.S) {
      const uint32_t excl = wave_excl_scan_u32(__popc(sl.m[j]), lane);
      if (sl.m[j]) sl.pk[j] = load_packed(mv.col[j].outl, (uint64_t)pre[j] + excl);
    }
  }
}

// One sparse column's outliers of tile c against target byte t, deposited in
// row order. Outlier k of the tile (k = 0..cnt-1, in row order) goes to lane
// k % 64, which ORs (1 << b) | (byte > t) << (16 + b) | (byte == t) << (32 + b)
// into dep[L] for its row 16L + b. Lane L then reads its word: the outlier rows
// of its 16 (bits 0-15), those above t (16-31) and those equal (32-47). The
// first outlier of each lane arrives preloaded (row0 / byte0, has0); further
// ones (a tile with more than 64) are loaded here.
__device__ __forceinline__ uint64_t deposit_outliers(unsigned long long* dep, const SparseDev& sd, uint32_t p0,
                                                     uint32_t cnt, uint32_t lane, uint32_t row0, uint32_t byte0,
                                                     bool has0, uint32_t t) {
  dep[lane] = 0;
  wave_lds_sync();
  auto put = [&](uint32_t row, uint32_t byte) {
    const uint32_t one = 1u << (row & 15u);
    const uint32_t lo = one | (byte > t ? one << 16 : 0u), hi = byte == t ? one : 0u;
    atomicOr(&dep[row >> 4], ((unsigned long long)hi << 32) | lo);
  };
  if (has0) put(row0, byte0);
  for (uint32_t k = lane + 64; k < cnt; k += 64) put(sd.pos[p0 + k], sd.outl[p0 + k]);
  wave_lds_sync();
  const uint64_t w = dep[lane];
  wave_lds_sync();   // every lane has read its word before the next column clears it
  return w;
}

// Tiles [c0, c0 + U) of a buff-major filter (buff_simd.rs:863-1277 range,
// 1916-2292 equality; with sem the intended answer on the true offsets):
// res[u] is the lane's 16-row result mask of tile c0 + u (tiles past the
// chunk's last are computed empty and are not meant to be stored). dep: the
// wave's 64 LDS words. Every load of the first round (tile prefixes, the first
// dense plane, each lane's first outlier row and byte of every column) is in
// flight together.
//
// PF (the target's sparse bytes are the majority bytes, so most rows stay
// candidates into the dense planes): the second dense plane and a tail of at
// most kPfTail bits are loaded in the first round as well, in full, instead of
// for the candidates' lines in later dependent rounds.
constexpr uint32_t kPfTail = 2;

template <int MODE, int S, int U, bool PF = false, typename MV>
__device__ __forceinline__ void major_filter_group(const MV& mv, const MajorTarget& tg, bool sem, uint64_t c0,
                                                   uint32_t lane, unsigned long long* dep, uint32_t (&res)[U]) {
  const ChunkView& dv = mv.dense;
  const bool all_outliers = MODE == 0 && !sem;   // the range rules read every outlier byte
  uint32_t p0[U][S], cnt[U][S], row0[U][S], byte0[U][S];
  uint4 w0[U], w1[U];
  uint32_t hp[U][kPfTail];
  const bool pf_tail = PF && dv.r <= kPfTail;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const bool in = c0 + u < dv.ntiles;
    const uint64_t c = min(c0 + u, dv.ntiles - 1);
#pragma unroll
    for (int j = 0; j < S; ++j) {
      p0[u][j] = mv.col[j].prefix[c];
      cnt[u][j] = in ? mv.col[j].prefix[c + 1] - p0[u][j] : 0u;
    }
    w0[u] = dv.P ? load_plane(dv, 0, c, lane) : make_uint4(0, 0, 0, 0);
    w1[u] = PF && dv.P > 1 ? load_plane(dv, 1, c, lane) : make_uint4(0, 0, 0, 0);
#pragma unroll
    for (uint32_t j = 0; j < kPfTail; ++j) hp[u][j] = pf_tail && j < dv.r ? load_tail(dv, j, c, lane) : 0u;
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const bool has = lane < cnt[u][j];
      row0[u][j] = has ? mv.col[j].pos[p0[u][j] + lane] : 0u;
      byte0[u][j] = has ? mv.col[j].outl[p0[u][j] + lane] : 0u;
    }
  }
  uint32_t cand[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    res[u] = 0;
    cand[u] = lane_valid16(dv.n, c0 + u, lane);
#pragma unroll
    for (int j = 0; j < S; ++j) {
      if (!all_outliers && !__ballot(cand[u] != 0)) break;   // nothing left for this tile to narrow or add
      const uint32_t major = mv.col[j].major;
      const uint32_t t = tg.digit[j];
      const uint64_t w = deposit_outliers(dep, mv.col[j], p0[u][j], cnt[u][j], lane, row0[u][j], byte0[u][j],
                                          lane < cnt[u][j], t);
      const uint32_t m16 = (uint32_t)w & 0xFFFFu;
      const uint32_t nm = ~m16 & 0xFFFFu;
      const uint32_t want = all_outliers ? m16 : (m16 & cand[u]);
      const uint32_t gt = (major > t ? nm : 0u) | ((uint32_t)(w >> 16) & want);
      const uint32_t eq = (major == t ? nm : 0u) | ((uint32_t)(w >> 32) & want);
      if (MODE == 1) {
        cand[u] &= eq;
        continue;
      }
      if (sem || j == 0) {
        res[u] |= cand[u] & gt;
        cand[u] &= eq;
      } else if (t == 0 && major == 0) {
        res[u] |= cand[u] & m16;
        cand[u] = 0;
      } else {
        res[u] |= m16 & gt;
        if (major > t) res[u] |= cand[u] & ~m16;
        cand[u] &= eq;
      }
    }
  }
  for (uint32_t p = 0; p < dv.P; ++p) {
    bool live = false;
#pragma unroll
    for (int u = 0; u < U; ++u) live |= cand[u] != 0;
    if (!__ballot(live)) break;
    uint4 wp[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      wp[u] = p == 0             ? w0[u]
              : (PF && p == 1) ? w1[u]
                               : load_plane(dv, p, min(c0 + u, dv.ntiles - 1), refine_src(cand[u], lane));
    const Digit dg = make_digit(tg.digit[S + p]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (MODE == 0) res[u] |= cand[u] & mask_gt(wp[u], dg);
      cand[u] &= mask_eq(wp[u], dg);
    }
  }
  if (dv.r) {
    bool live = false;
#pragma unroll
    for (int u = 0; u < U; ++u) live |= cand[u] != 0;
    if (__ballot(live)) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (c0 + u >= dv.ntiles) continue;
        if (pf_tail) {   // the prefetched bit-planes, most significant first
#pragma unroll
          for (int j = (int)kPfTail - 1; j >= 0; --j) {
            if ((uint32_t)j >= dv.r) continue;
            if ((tg.tail >> j) & 1) {
              cand[u] &= hp[u][j];
            } else {
              if (MODE == 0) res[u] |= cand[u] & hp[u][j];
              cand[u] &= ~hp[u][j];
            }
          }
        } else {
          refine_tail<MODE>(dv, c0 + u, tg.tail, lane, res[u], cand[u]);
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < U; ++u) res[u] = MODE == 0 ? (res[u] & lane_valid16(dv.n, c0 + u, lane)) : cand[u];
}

// Sum of the lane's outlier bytes of one sparse column (its k = popc(m16)
// outliers are the first k bytes of its packed window).
__device__ __forceinline__ uint32_t packed_sum(const Packed16& pk, uint32_t m16) {
  const uint32_t k = (uint32_t)__popc(m16);
  auto keep = [k](uint32_t i) -> uint32_t {   // bytes of dword i below k
    return k >= 4 * (i + 1) ? 0xFFFFFFFFu : (k > 4 * i ? (1u << (8 * (k - 4 * i))) - 1u : 0u);
  };
  uint32_t a = __builtin_amdgcn_sad_u8(pk.d0 & keep(0), 0u, 0u);
  a = __builtin_amdgcn_sad_u8(pk.d1 & keep(1), 0u, a);
  a = __builtin_amdgcn_sad_u8(pk.d2 & keep(2), 0u, a);
  return __builtin_amdgcn_sad_u8(pk.d3 & keep(3), 0u, a);
}

// The lane's 16 stored offsets of tile c of a buff-major chunk
// (buff_decode_majority, buff_simd.rs:633-861): column 0 places its byte; a
// later sparse column ORs its outlier bytes at its shift and the majority byte
// at twice its shift, and only for rows up to its last outlier (SEM: every row
// gets its byte once, the true offset). Then the dense planes and tail.
template <int SEM, int S, typename MV>
__device__ __forceinline__ void major_tile_offsets(const MV& mv, uint64_t c, uint32_t lane, uint64_t (&v)[16]) {
  constexpr int SS = S > 0 ? S : 1;
  constexpr int NPM = (int)kMaxPlanes - S > 0 ? (int)kMaxPlanes - S : 1;   // dense planes at most (F <= 64)
  const ChunkView& dv = mv.dense;
  // every load of the tile first: masks and prefixes, dense planes and tail,
  // then the outlier windows (they need the masks' scan)
  uint32_t m[SS], pre[SS];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    m[j] = mv.col[j].mask16[c * 64 + lane];
    pre[j] = mv.col[j].prefix[c];
  }
  uint4 wp[NPM];
#pragma unroll
  for (int p = 0; p < NPM; ++p) wp[p] = (uint32_t)p < dv.P ? load_plane(dv, p, c, lane) : make_uint4(0, 0, 0, 0);
  uint32_t h[7];
#pragma unroll
  for (uint32_t j = 0; j < 7; ++j) h[j] = j < dv.r ? load_tail(dv, j, c, lane) : 0u;
  uint4 ws[SS];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const uint32_t excl = wave_excl_scan_u32(__popc(m[j]), lane);
    const Packed16 pk = m[j] ? load_packed(mv.col[j].outl, (uint64_t)pre[j] + excl) : Packed16{0, 0, 0, 0};
    ws[j] = patch_outliers(pk, m[j], m[j], mv.col[j].major);
  }
  const uint64_t row0 = c * kTileRows + 16ull * lane;
  // one column at a time over the lane's 16 rows, each behind a wave-uniform
  // guard (absent planes and tail bits cost no instructions); the tail bits
  // gather in 32-bit words
#pragma unroll
  for (uint32_t b = 0; b < 16; ++b) v[b] = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const uint32_t sh = mv.col[j].shift;
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) {
      const uint64_t byte = (uint64_t)byte_at(ws[j], b) << sh;
      if (SEM || j == 0) {
        v[b] |= byte;
      } else {
        const uint64_t twice = ((uint64_t)mv.col[j].major << sh) << sh;
        const uint64_t t = ((m[j] >> b) & 1u) ? byte : twice;
        v[b] |= row0 + b <= mv.col[j].last ? t : 0ull;
      }
    }
  }
#pragma unroll
  for (int p = 0; p < NPM; ++p) {
    if ((uint32_t)p >= dv.P) break;
    const uint32_t sh = dv.F - 8 * (p + 1);
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) v[b] |= (uint64_t)byte_at(wp[p], b) << sh;
  }
  uint32_t t[16];
#pragma unroll
  for (uint32_t b = 0; b < 16; ++b) t[b] = 0;
#pragma unroll
  for (uint32_t j = 0; j < 7; ++j) {
    if (j >= dv.r) break;
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) t[b] |= ((h[j] >> b) & 1u) << j;
  }
#pragma unroll
  for (uint32_t b = 0; b < 16; ++b) v[b] |= t[b];
}

// The first S sparse columns of a MajorView (the same layout as its prefix):
// what a kernel reading the view from memory needs in registers for S columns.
template <int S>
struct MajorPrefix {
  ChunkView dense;
  SparseDev col[S];
};
static_assert(sizeof(MajorPrefix<3>) == offsetof(MajorView, col) + 3 * sizeof(SparseDev), "MajorView prefix layout");

}  // namespace buffgpu
