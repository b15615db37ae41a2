// Shared host/device definitions of the resident BUFF layout.
//
// Resident layout of one chunk (n rows, F = P*8 + r bits per row), designed for
// wave64 lanes:
//   * Rows are grouped in tiles of 1024; lane t of a wave owns the 16
//     consecutive rows 16t .. 16t+15 of a tile.
//   * Byte sub-column ("plane") p is `pitch` bytes long; tile c of plane p is the
//     1024 bytes at planes + p*pitch + 1024*c, lane t's 16 bytes at offset 16*t.
//     Inside them, dword i byte k holds row 16t + 4k + i, so the byte-wise
//     compare bits of a dword land on rows i, i+4, i+8, i+12 and four dwords
//     gather into one 16-bit row mask with shifts (see k_filter).
//     A dwordx4 load per lane reads a whole tile of a plane coalesced.
//   * The r tail bits are kept as bit-planes: tile c owns r*64 u16 words at
//     tail + 64*r*c; word j*64 + t holds bit j of rows 16t .. 16t+15 (bit b <->
//     row 16t + b). Bit-plane j of a tile is therefore a 1024-row bitmap in
//     BitVec order, and a filter's 16-bit row mask of lane t is word t of the
//     tile's result bitmap.
//   * Rows >= n are zero in every plane and bit-plane.
// The reference byte stream (split_double.rs:349-414) is produced from and parsed
// into this form by the serialize/load kernels.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define BUFF_HD __host__ __device__ __forceinline__
#else
#define BUFF_HD inline
#endif

namespace buffgpu {

constexpr uint32_t kTileRows = 1024;
constexpr uint32_t kGroups = 16;     // 64-row groups per tile
constexpr uint32_t kMaxPlanes = 8;   // F <= 63 in practice; 8 slots
constexpr uint32_t kStreamHeader = 20;

// Device view of a resident chunk.
struct ChunkView {
  uint8_t* planes;    // plane p at planes + p * pitch
  uint16_t* tail;     // bit-planes, 64*r u16 per tile
  uint64_t pitch;     // bytes per plane = ntiles * 1024
  uint64_t n;         // rows
  uint64_t ntiles;    // ceil(n / 1024)
  int64_t base;       // min fixed value
  uint32_t F;         // stored width
  uint32_t P;         // full byte planes (F >= 8 ? F/8 : 0)
  uint32_t r;         // tail bits (F >= 8 ? F%8 : F)
  uint32_t dlen;
};

// database/src/methods/prec_double.rs:190-205 (fetch_fixed_aligned), with Rust
// release semantics: the shift amount is masked to 6 bits, negation wraps
// (`!(fixed - 1)` == `0 - fixed`). Written without branches.
BUFF_HD int64_t fetch_fixed_aligned_bits(uint64_t bdu, int32_t precision_exp, uint32_t dlen) {
  const int32_t exp = (int32_t)((bdu >> 52) & 0x7FFu) - 1023;
  const uint32_t sh = (uint32_t)(63 - exp - (int32_t)dlen) & 63u;
  const uint64_t mag = ((bdu << 11) | 0x8000000000000000ull) >> sh;
  const uint64_t fixed = (bdu >> 63) ? 0ull - mag : mag;
  return exp >= precision_exp ? (int64_t)fixed : 0;
}

// Per-filter constants computed on the host from the predicate.
struct FilterTarget {
  uint32_t digit[kMaxPlanes];  // target byte per plane: (t >> (F - 8(k+1))) as u8
  uint64_t tail;               // target tail bits (t & (2^r - 1); the full t when P == 0)
};

// Encode-pass reduction results (minmax kernel -> host).
struct MinMaxOut {
  int64_t min_all;    // min over all rows
  int64_t max_rest;   // max over rows i >= 1
  int64_t first;      // fixed value of row 0
  int64_t qmax;       // quirk max from the prefix-min pass
  uint32_t unused0;
  uint32_t pad;
};

}  // namespace buffgpu
