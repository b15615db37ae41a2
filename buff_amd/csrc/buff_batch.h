// Batched execution over many resident chunks of uneven length and mixed
// precision: one persistent launch encodes all chunks, and a batch of queries
// (range / equality filters, sums, maxima, materializations) runs as a fixed
// sequence of persistent launches that read their work lists from device memory,
// so the sequence is captured once in a hipGraph and replayed for every batch.
//
// Work is cut into items: (chunk or query, tile range). A workgroup takes one
// item at a time (its waves stride over the item's tiles); item sizes and grid
// sizes are what the autotuner chooses.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "buff_common.h"
#include "buff_kernels.h"
#include "buff_major.h"

namespace buffgpu {

// ---- batch encode -------------------------------------------------------------

struct BEncChunk {
  const double* x;     // first row of the chunk
  ChunkView view;      // planes / tail in the batch pool (n, ntiles, base, F, P, r, dlen)
  int32_t pexp;
  uint32_t item0;      // first item of the chunk
  uint32_t nitems;
  uint32_t pad;
};

struct BItem {
  uint32_t owner;      // chunk (encode) or query (queries)
  uint32_t chunk;      // queries: the query's chunk (its view is read without going through the query)
  uint64_t c0, c1;     // tile range [c0, c1)
};
static_assert(sizeof(BItem) == 24 && offsetof(BItem, chunk) == 4 && offsetof(BItem, c0) == 8, "BItem layout");

// Per-chunk min/max of the `else if` pass (split_double.rs:321-330): min over all
// rows, max over rows >= 1, and the fixed value of row 0.
struct BMinMax {
  int64_t min_all;
  int64_t max_rest;
  int64_t first;
  int64_t pad;
};

hipError_t launch_batch_minmax(const BEncChunk* chunks, uint32_t nchunks, const BItem* items, uint32_t nitems,
                               int64_t* item_part, BMinMax* out, uint32_t grid, hipStream_t s);
hipError_t launch_batch_encode(const BEncChunk* chunks, const BItem* items, uint32_t nitems, uint32_t grid,
                               hipStream_t s);

// ---- batch queries ------------------------------------------------------------

enum : uint32_t { kBqGt = 0, kBqEq = 1, kBqSum = 2, kBqMax = 3, kBqMat = 4 };

struct BQuery {
  uint32_t op;
  uint32_t chunk;        // index into the batch's chunk views
  uint32_t fill;         // filters: 0 = run, 1 = no rows, 2 = all rows (decided on the host)
  uint32_t item0, nitems;
  uint32_t sparse;       // buff-major chunk with sparse columns: S (bits 0-7), BUFF_SEMANTIC (bit 8); else 0
  FilterTarget tg;       // filters (on a buff-major chunk: the target byte of every byte column, sparse first)
  MatParams mp;          // materialize (buff-major: mode kMatSemantic keeps the top F - drop bits)
  uint64_t out;          // filters: device address of the ceil(n/64) bitmap words;
                         // materialize: device address of the n doubles
  uint64_t top;          // max: byte offset of its per-tile top bytes in the max arena
};

// Per-query results the host reads back.
struct BResult {
  unsigned long long v[kMaxPlanes + 1];   // sum: per-sub-column sums; filter: v[0] = count; max: v[0] = offset
};

// Device-side counters of the current batch (written with the descriptors).
// Items [0, nscan) are filters, sums and maxima; [nscan, nitems) materializations.
// The host assigns every query its item range (item0, nitems) and the first
// kernel of the batch writes the items themselves (tiles_per_item tiles each).
struct BHeader {
  uint32_t nq, nitems;
  uint32_t nscan;
  uint32_t tiles_per_item;
};

struct BatchBuffers {
  const uint8_t* host_in;            // pinned host: BHeader at 0, nq BQuery at kHostQueries
  BResult* host_res;                 // pinned host: nq results, written by the last kernel
  const BHeader* hdr;
  const ChunkView* views;
  const MajorView* majors;           // per chunk, read only by queries whose chunk has sparse columns
  const BQuery* q;
  BItem* items;                      // written by the batch's first kernel
  uint8_t* tile_top;                 // max: per-tile top byte
  unsigned long long* item_acc;      // per item: kMaxPlanes + 1 u64
  BResult* res;
};

// `grid` is the persistent grid size and `tiles_per_item` the work item length
// (at most kMaxTilesPerItem, so per-item sums fit 32 bits). Filters read the top
// byte plane of every tile first and the remaining planes only for tiles with
// tied rows (first_planes = 1), or every plane in one pass (first_planes >=
// kMaxPlanes), which saves a dependent round trip on columns where most rows tie.
constexpr uint32_t kMaxTilesPerItem = 4096;
struct BatchTuning {
  uint32_t grid;
  uint32_t first_planes;
  uint32_t tiles_per_item;
  uint32_t pad;
};

constexpr uint64_t kHostQueries = 256;   // byte offset of the queries in host_in

// Kernels a batch needs: the scan (filters, sums, maxima and their fold), the
// max refinement, the materialization, and the same for queries on buff-major
// chunks with sparse columns (kernels of their own, so the plain kernels keep
// their register budget).
enum : uint32_t { kBqPartScan = 1, kBqPartMax = 2, kBqPartMat = 4, kBqPartMajor = 8, kBqPartMajorMat = 16 };
constexpr uint32_t kBqSparseSem = 1u << 8;

// The launch sequence of one batch (captured in a graph by the host). The
// kernels read every buffer address and count from `bb` (device memory), so one
// captured graph serves every batch of the same tuning and parts.
hipError_t launch_batch_queries(const BatchBuffers* bb, const BatchTuning& tu, uint32_t parts, hipStream_t s);

}  // namespace buffgpu
