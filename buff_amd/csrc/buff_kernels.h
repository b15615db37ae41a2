// Launchers of the HIP kernels in buff_kernels.hip (host side).
#pragma once
#include <hip/hip_runtime.h>

#include "buff_common.h"

namespace buffgpu {

// Launch geometry of the tile kernels (chosen per op by the autotuner).
struct LaunchShape {
  uint32_t waves_per_wg;    // 1, 2, 4 (a workgroup is 64*waves_per_wg threads)
  uint32_t max_wg;          // grid cap (0: the workgroups that fit resident for filter, sum and max, else 2048);
                            // waves grid-stride over tiles beyond it
  uint32_t tiles_per_wave;  // tiles a wave loads per step (0: the kernel's default; filter 2/4/8, sum 1/2/4,
                            // max 4/8/16)
  uint32_t prefetch;        // filter: load the next step's top plane before refining this one;
                            // max: 1 = the one-pass form (0 = two passes)
};

enum : uint32_t { kMatFull = 0, kMatSemantic = 1, kMatRefPrecision = 2 };

// Materialization recipe: read the top `planes` byte planes and `tail_bits`
// bits below them, then value = base + ((v >> drop) << shift_back)
// (kMatRefPrecision: tail_bits are read as fields of that width at that stride
// from the packed reference tail, drop = shift_back = 0).
struct MatParams {
  uint32_t mode;
  uint32_t planes;
  uint32_t tail_bits;
  uint32_t drop;
  uint32_t shift_back;
  uint32_t pad;
  double inv_scale;  // exactly 2^-dlen
};

// Hand-over of a reduction's result to the host without a stream wait: a
// one-wave kernel, queued right after the reduction, folds its slots, writes
// the result words to mbox[1 ..] (the device address of mapped, coherent host
// memory) and then stores seq into mbox[0] with a system-scope release; a host
// that sees seq there reads a complete result.
struct Publish {
  unsigned long long* mbox;   // kPublishWords u64
  unsigned long long seq;     // new for every launch, nonzero
};
constexpr uint32_t kPublishWords = 64;
enum : uint32_t {
  kPublishSumSlots = 0,   // nk rows of kSumSlots u64 slots: the row sums
  kPublishMinMax = 1,     // a min/max slot set: min, max over rows >= 1, row 0
  kPublishCopy = 2,       // nk u64 words as they are (src 8-byte aligned)
};
hipError_t launch_publish(const void* src, uint32_t mode, uint32_t nk, const Publish& pub, hipStream_t s);

// Workgroups of the min/max pass over ntiles tiles (all resident at once).
uint32_t minmax_grid(uint64_t ntiles);
hipError_t launch_minmax(const double* x, uint64_t n, int32_t pexp, uint32_t dlen, int64_t* partial,
                         int64_t* partial2, MinMaxOut* out, uint32_t nwg, bool prefix, hipStream_t s);
// One launch (n >= 1): the min over all rows and the max over rows >= 1 folded
// into the slot set `set` (int64: kMinMaxSlots mins, kMinMaxSlots maxes, then
// row 0's fixed value), for launch_publish to fold; `next`, the set the following
// call folds into, is reset; `partial` gets the per-workgroup rows the
// prefix-min pass reads. Sets start reset (launch_minmax_slots_init).
constexpr uint32_t kMinMaxSlots = 64;
constexpr uint32_t kMinMaxSetWords = 2 * kMinMaxSlots + 8;
hipError_t launch_minmax_slots(const double* x, uint64_t n, int32_t pexp, uint32_t dlen, int64_t* partial,
                               int64_t* set, int64_t* next, uint32_t nwg, hipStream_t s);
hipError_t launch_minmax_slots_init(int64_t* sets, uint32_t nsets, hipStream_t s);
hipError_t launch_encode(const double* x, const ChunkView& cv, int32_t pexp, hipStream_t s, const LaunchShape& ls);
hipError_t launch_filter(const ChunkView& cv, int op, const FilterTarget& tg, uint64_t* out, uint64_t nwords,
                         unsigned long long* stats, hipStream_t s, const LaunchShape& ls);
hipError_t launch_fill_bitmap(uint64_t* out, uint64_t nwords, uint64_t n, int ones, hipStream_t s);
// Byte-slice stream: the right-padded last plane <-> the resident tail bit-planes.
hipError_t launch_slice_tail_serialize(const ChunkView& cv, uint8_t* plane, hipStream_t s);
hipError_t launch_slice_tail_load(const ChunkView& cv, const uint8_t* plane, hipStream_t s);
// out = (inv_a ? ~a : a) & (inv_b ? ~b : b) & cond over rows < cover (null a/b/cond: all rows).
hipError_t launch_bitmap_logic(uint64_t* out, const uint64_t* a, const uint64_t* b, const uint64_t* cond,
                               uint64_t nwords, uint64_t cover, int inv_a, int inv_b, hipStream_t s);
// Filter restricted to the rows set in `cond` (ceil(n/64) u64 words, the bitmap layout).
hipError_t launch_filter_cond(const ChunkView& cv, int op, const FilterTarget& tg, const uint64_t* cond,
                              uint64_t* out, uint64_t nwords, hipStream_t s, const LaunchShape& ls);
// out = cond masked to the n rows (copy = 1) or zeros (copy = 0).
hipError_t launch_cond_fill(uint64_t* out, const uint64_t* cond, uint64_t nwords, uint64_t n, int copy,
                            hipStream_t s);
hipError_t launch_fill_bytes(void* out, uint64_t bytes, uint8_t value, hipStream_t s);
// bytes: a multiple of 16; in and out 16-byte aligned and not overlapping
hipError_t launch_copy_bytes(void* out, const void* in, uint64_t bytes, hipStream_t s);
// bytes: a multiple of 16, in 16-byte aligned
hipError_t launch_read_bytes(const void* in, uint64_t bytes, uint32_t* sink, hipStream_t s);
hipError_t launch_touch_bytes(const void* in, uint64_t bytes, uint32_t* sink, hipStream_t s);
// sums: 9 u64 results followed by sum_scratch_bytes(ls) of per-workgroup rows
// (at most kMaxSumRows workgroups).
constexpr uint32_t kMaxSumRows = 8192;
uint64_t sum_scratch_bytes(const LaunchShape& ls);
hipError_t launch_sum(const ChunkView& cv, unsigned long long* sums, hipStream_t s, const LaunchShape& ls);
// One launch: per-plane sums (and the weighted tail sum at k = P) added into
// the slot set `set` (u64 [kMaxPlanes + 1][kSumSlots], to be summed over the
// slots); `next`, the set the following call adds into, is cleared. The first
// set must be zero.
constexpr uint32_t kSumSlots = 64;
constexpr uint32_t kSumSetWords = (kMaxPlanes + 1) * kSumSlots;
hipError_t launch_sum_slots(const ChunkView& cv, unsigned long long* set, unsigned long long* next, hipStream_t s,
                            const LaunchShape& ls);
// Tail part of the reduced-precision sum reading q < r bits per row: out[0] +=
// the sum of the n q-bit fields the reference reads from the packed tail,
// out[1] += the sum of every row's top q tail bits (the caller zeroes out[0..1]).
hipError_t launch_tail_fields(const ChunkView& cv, uint32_t q, unsigned long long* out, hipStream_t s);
// Two passes using `scratch` (max_scratch_bytes(ntiles) bytes) and `slots`
// (kMaxSlotBytes), or one pass (ls.prefetch == 1). The max offset is left in
// best[0] (16 bytes at `best` may be written). A slot holds (tag << 8) | byte:
// `tag` must exceed every tag given with the same slots before, and be below
// kMaxTagLimit; the slots are zeroed once before their first tag.
constexpr uint64_t kMaxSlotBytes = 64 * 4;
constexpr uint32_t kMaxTagLimit = 1u << 24;
uint64_t max_scratch_bytes(uint64_t ntiles);
hipError_t launch_max(const ChunkView& cv, unsigned long long* best, uint8_t* scratch, unsigned int* slots,
                      uint32_t tag, hipStream_t s, const LaunchShape& ls);
hipError_t launch_materialize(const ChunkView& cv, const MatParams& mp, double* out, hipStream_t s,
                              const LaunchShape& ls);
hipError_t launch_serialize(const ChunkView& cv, uint8_t flipmask, uint8_t* body, uint64_t tail_bytes,
                            hipStream_t s);
hipError_t launch_load(const ChunkView& cv, uint8_t flipmask, const uint8_t* body, uint64_t tail_bytes,
                       hipStream_t s);

}  // namespace buffgpu
