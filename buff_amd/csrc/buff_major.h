// Resident form and kernel launchers of the buff-major (majority / outlier)
// layout (database/src/compress/buff_simd.rs:317-2861).
//
// A buff-major chunk of width F keeps its S top byte columns sparse. Sparse
// column j (byte (off >> (F - 8(j+1))) as u8) stores:
//   * the majority byte chosen by the encoder's vote,
//   * the outlier bitmap (rows whose byte differs) in row order, LSB first:
//     exactly the BitVec<u32> storage the reference serializes (u16 word
//     64c + t holds rows 1024c + 16t .. +15, the 16 rows lane t owns in tile c),
//   * the outlier bytes in row order, compacted,
//   * per tile c the number of outliers in tiles before c (prefix[c]), so a wave
//     finds the position of its lanes' outlier bytes with one exclusive scan,
//   * per outlier its row in its tile (u16, in the order of the bytes): the
//     filters give outlier k of a tile to lane k, which compares its byte and
//     ORs the result bits into the owning lane's LDS word (no per-lane walk of
//     the bitmap).
// The low F - 8S bits are an ordinary resident chunk (buff_common.h).
#pragma once
#include <hip/hip_runtime.h>

#include "buff_common.h"
#include "buff_kernels.h"

namespace buffgpu {

constexpr uint32_t kMaxSparse = 8;

struct SparseDev {
  const uint16_t* mask16;  // ntiles * 64 words
  const uint32_t* prefix;  // ntiles + 1 (prefix[ntiles] = outliers)
  const uint8_t* outl;     // compacted outlier bytes
  const uint16_t* pos;     // row in its tile (0..1023) of each outlier, in the order of outl
  uint64_t last;           // row of the column's last outlier
  uint32_t major;          // majority byte
  uint32_t shift;          // F - 8(j+1)
};

struct MajorView {
  ChunkView dense;               // the low F - 8S bits (base, dlen, n, ntiles set)
  SparseDev col[kMaxSparse];
  uint32_t S;                    // sparse columns
  uint32_t F;                    // full width
};

// Encode passes. Workgroup w owns the contiguous tiles [w*tpw, (w+1)*tpw).
// Pass 1 writes the dense planes / tail, the outlier bitmaps, each lane's
// outlier bytes packed into its 16-byte staging slot, and per-workgroup outlier
// totals. A one-workgroup scan turns the totals into bases; pass 2 walks each
// workgroup's tiles in order, writes the tile prefixes and moves the staged
// bytes to their final place.
struct MajorEnc {
  uint32_t S;
  uint32_t nwg;
  uint64_t tpw;
  uint32_t major[kMaxSparse];
  uint32_t shift[kMaxSparse];
  uint16_t* mask16[kMaxSparse];
  uint32_t* counts[kMaxSparse];          // per tile
  uint8_t* stage[kMaxSparse];            // ntiles * 1024 bytes: 16-byte slot per lane
  uint32_t* wgsum[kMaxSparse];           // per workgroup: total, then exclusive base
  uint32_t* prefix[kMaxSparse];
  uint8_t* outl[kMaxSparse];
};

// Per-query constants: target byte of every byte column (sparse then dense,
// most significant first) and the tail target.
struct MajorTarget {
  uint32_t digit[kMaxPlanes];
  uint64_t tail;
};

// Device state of the staged max: chosen byte per stage and the accumulated
// max offset; argmax = rows that tie the max in the tail stage.
struct MajorMaxState {
  uint32_t chosen[16];
  unsigned long long acc;
  unsigned long long argmax;
  uint32_t top1;       // list path: max outlier byte of column 0, plus 1 (0: none)
  uint32_t count;      // list path: rows that hold column 0's chosen byte (set when taken)
  uint32_t fallback;   // list path not taken: the staged kernels must run
};

// Pass 1 (x read once); me.nwg / me.tpw set by major_enc_geometry().
void major_enc_geometry(uint64_t ntiles, MajorEnc* me);
hipError_t launch_major_encode(const double* x, uint64_t n, int32_t pexp, uint32_t dlen, int64_t base,
                               const ChunkView& dense, const MajorEnc& me, hipStream_t s);
// Workgroup totals -> bases; totals[2j] = outliers, totals[2j+1] = last outlier
// row (UINT64_MAX if none); prefix[j][ntiles] = outliers.
hipError_t launch_major_scan(const MajorEnc& me, uint64_t ntiles, unsigned long long* totals, hipStream_t s);
// Pass 2: tile prefixes (per workgroup range), then the outlier bytes moved to
// their final place (grid-stride over tiles).
hipError_t launch_major_place(const MajorEnc& me, uint64_t ntiles, hipStream_t s);
// pos of one sparse column from its bitmap and tile prefixes.
hipError_t launch_major_positions(const uint16_t* mask16, const uint32_t* prefix, uint64_t ntiles, uint16_t* pos,
                                  hipStream_t s);
hipError_t launch_major_filter(const MajorView& mv, int op, bool semantic, const MajorTarget& tg, uint64_t* out,
                               uint64_t nwords, hipStream_t s, const LaunchShape& ls);
// mode 0: buff_decode_majority's values; 1: the true offsets truncated to their
// top F - drop bits (drop = 0: every bit); 2: the true offsets' top F - drop
// bits as a number (off >> drop), the residue format's reduced-precision read.
hipError_t launch_major_decode(const MajorView& mv, int mode, uint32_t drop, double inv_scale, double* out,
                               hipStream_t s, const LaunchShape& ls);
// sums[j] = sum of the outlier bytes of sparse column j; scratch: major_sum_scratch_bytes().
uint64_t major_sum_scratch_bytes(uint32_t S);
hipError_t launch_major_outlier_sums(const MajorView& mv, const uint64_t* n_out, unsigned long long* sums,
                                     unsigned long long* scratch, hipStream_t s);
// Max by a list of the rows that hold column 0's chosen byte (S >= 1, n >= 1;
// n_out0 = outliers of column 0). The same answer as the staged max; when the
// majority byte is column 0's chosen byte or more than kMaxListRows rows hold
// it, st->fallback is set and the staged max must run. Scratch:
// kMaxListScratch + ntiles bytes.
constexpr uint32_t kMaxListRows = 8192;
constexpr uint64_t kListCounterBytes = 64 * 128;
constexpr uint32_t kTopBlocks = 1024;   // workgroups of the first list pass (one partial max each)
constexpr uint64_t kMaxListScratch = kListCounterBytes + 4ull * kTopBlocks + 16ull * kMaxListRows;
hipError_t launch_major_max_list(const MajorView& mv, uint64_t n_out0, MajorMaxState* st, uint8_t* scratch,
                                 hipStream_t s, bool semantic);
// Staged max over the byte columns and the tail; scratch: major_max_scratch_bytes(ntiles).
uint64_t major_max_scratch_bytes(uint64_t ntiles);
hipError_t launch_major_max(const MajorView& mv, MajorMaxState* st, uint8_t* scratch, hipStream_t s,
                            const LaunchShape& ls, bool semantic);

}  // namespace buffgpu
