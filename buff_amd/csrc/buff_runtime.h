// Host-side runtime objects behind the C ABI: context, chunk, scratch buffers,
// event timing. Internal to libbuffgpu.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "buff_common.h"
#include "buff_kernels.h"
#include "buff_major.h"

namespace buffgpu {

int fail(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return fail(BUFF_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));      \
  } while (0)

#define BUFF_TRY(expr)          \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != BUFF_OK) return rc_; \
  } while (0)

// precision helpers and small transfers (buff_api.hip)
int prec_of_scale(uint64_t scale);
int precision_map(int p);
int32_t precision_exp_of_prec(int prec);
double powi2(int32_t e);
uint32_t width_of_delta(int64_t delta);
uint64_t stream_bytes_of(uint64_t n, uint32_t P, uint32_t r);

inline uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

enum class Op : uint32_t { Encode = 0, FilterGt, FilterEq, Sum, Max, Materialize, kCount };

uint32_t tune_bucket(uint64_t ntiles);
// Launch shapes are kept per (operation, size class, byte planes).
inline uint64_t tune_key(Op op, uint64_t ntiles, uint32_t planes) {
  return ((uint64_t)op << 40) | ((uint64_t)planes << 32) | tune_bucket(ntiles);
}

struct Scratch {
  void* ptr = nullptr;
  uint64_t bytes = 0;
};

struct ProfRec {
  std::string name;
  hipEvent_t a, b;
};

struct Ctx {
  static constexpr uint64_t kPinnedBytes = 4096;

  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t budget = 0;  // 0 = unlimited
  uint64_t held = 0;    // bytes obtained from hipMalloc and not yet returned
  std::mutex mu;
  std::map<void*, uint64_t> live;
  std::multimap<uint64_t, void*> free_blocks;
  void* pinned = nullptr;

  Scratch s_partial, s_minmax, s_stream, s_bitmap, s_sums, s_best, s_values, s_sink, s_major, s_cond, s_condin, s_wmax, s_slice, s_csv, s_csvtab,
      s_csvtmp, s_csvslow, s_condval, s_maxslot, s_sumslot, s_mmslot;
  uint32_t max_tag = 0;        // the last tag given with s_maxslot (launch_max)
  uint32_t sum_set = 0;        // the s_sumslot set the next launch_sum_slots adds into
  uint32_t mm_set = 0;         // the s_mmslot set the next launch_minmax_slots folds into
  bool csv_tab_ready = false;  // s_csvtab holds the table of 5^q
  // result mailbox of the reductions (launch_publish): kPublishWords u64 of
  // mapped, coherent host memory
  uint64_t* mbox = nullptr;
  unsigned long long* mbox_dev = nullptr;
  unsigned long long mbox_seq = 0;

  bool prof_on = false;
  std::vector<hipEvent_t> prof_pool;
  std::vector<ProfRec> prof_pending;
  std::map<std::string, std::pair<uint64_t, double>> prof_stats;

  std::map<uint64_t, LaunchShape> tuned;
  // batched queries: persistent grid, tiles per work item, planes of the first filter pass
  uint32_t bt_grid = 2048, bt_tiles = 64, bt_first = 1;
  bool bt_tuned = false;

  ~Ctx();
  int alloc(uint64_t bytes, void** out);
  int release(void* p);
  int scratch(Scratch& s, uint64_t bytes);
  // max's slots and the next tag (slots zeroed on first use and when tags run out)
  int max_slots(unsigned int** slots, uint32_t* tag);
  // the sum's slot set for this call and the one it clears for the next
  // (both zeroed on first use); sum_slots_used() after the launch went out
  int sum_slots(unsigned long long** set, unsigned long long** next);
  void sum_slots_used() { sum_set ^= 1u; }
  // the min/max pass's slot sets, the same way
  int minmax_slots(int64_t** set, int64_t** next);
  void minmax_slots_used() { mm_set ^= 1u; }
  // the min/max pass over n >= 1 rows of x in one launch, folded on the host
  int minmax_pass(const double* x, uint64_t n, int32_t pexp, uint32_t dlen, MinMaxOut* mm);
  // queues the hand-over of a reduction's result (launch_publish into the
  // mailbox made with the context); published() waits until it has arrived
  // (or the stream failed or ended without it) and copies its nw result words
  int publish(const void* src, uint32_t mode, uint32_t nk, Publish* pub);
  int published(const Publish& pub, uint64_t* words, uint32_t nw);
  void prof_begin(const char* name, hipEvent_t* ev, bool kernels);
  void prof_end(const char* name, hipEvent_t* ev, bool kernels);
  int prof_collect();
  LaunchShape shape(Op op, const ChunkView& v) const;
};

// Times one operation while profiling is on. kernels = true: the interval is
// the execution of the kernels the operation launches through launch_k (first
// start to last end); false: events recorded on the stream around it (used for
// graph launches).
struct ProfScope {
  Ctx* ctx;
  const char* name;
  bool kernels;
  hipEvent_t ev[2];
  ProfScope(Ctx* c, const char* n, bool k = true) : ctx(c), name(n), kernels(k) { ctx->prof_begin(name, ev, kernels); }
  ~ProfScope() { ctx->prof_end(name, ev, kernels); }
};

// One sparse (majority / outlier) byte column of a buff-major chunk.
struct SparseCol {
  uint32_t major = 0;
  uint32_t shift = 0;       // F - 8(j+1)
  uint64_t n_out = 0;       // outlier rows
  uint64_t last = ~0ull;    // row of the last outlier (~0 when there is none)
  uint16_t* mask16 = nullptr;
  uint32_t* prefix = nullptr;
  uint8_t* outl = nullptr;
  uint16_t* pos = nullptr;  // row in tile of each outlier (n_out u16)
};

struct Chunk {
  Ctx* ctx = nullptr;
  uint64_t n = 0;
  uint64_t scale = 0;
  uint32_t flags = 0;  // BUFF_FLIP | BUFF_MAJORITY | BUFF_SLICE
  uint32_t dlen = 0;
  uint32_t ilen = 0;
  uint32_t F = 0;      // full stored width; `view` holds the low F - 8*S bits
  uint32_t S = 0;      // sparse columns (buff-major chunks only)
  ChunkView view{};
  void* mem = nullptr;
  void* mem_sparse = nullptr;  // bitmaps and prefixes of the sparse columns
  void* mem_outl = nullptr;    // compacted outlier bytes
  SparseCol sp[8];
  uint64_t device_bytes = 0;
  bool borrowed = false;       // a view into a batch's pool (freed with the batch)
};

int d2h_small(Ctx* ctx, void* dst, const void* src, uint64_t bytes);
int chunk_alloc(Ctx* ctx, Chunk* ch);
void chunk_set_width(Chunk* ch, uint32_t F);
int default_ctx(Ctx** out);
int quirk_minmax(Ctx* ctx, const double* x, uint64_t n, int32_t pexp, uint32_t dlen, int64_t* mn, int64_t* mx);
bool filter_plan(const Chunk* ch, int op, double pred, uint32_t flags, FilterTarget* tg, int* fill);
double reference_sum_fold(const Chunk* ch, const uint64_t* ps);
int mat_plan(const Chunk* ch, int32_t precision, uint32_t flags, MatParams* out);
// Stored bits a reduced-precision read keeps (split_double.rs:1642-1664, u32 arithmetic).
int precision_bits(const Chunk* ch, int32_t precision, uint32_t* R);
int filter_device(Ctx* ctx, const Chunk* ch, int op, double pred, uint32_t flags, uint64_t* words,
                  unsigned long long* stats);
int sum_op(Ctx* ctx, const Chunk* ch, double* ref_sum, double* exact_sum, uint64_t* plane_sums);
int sum_precision_op(Ctx* ctx, const Chunk* ch, int32_t precision, uint32_t flags, double* ref_sum,
                     double* exact_sum, uint64_t* sums);
int max_op(Ctx* ctx, const Chunk* ch, uint32_t flags, double* value, uint64_t* max_off, uint64_t* argmax);
int materialize_device(Ctx* ctx, const Chunk* ch, int32_t precision, uint32_t flags, double* out);
int filter_cond_device(Ctx* ctx, const Chunk* ch, int op, double pred, uint32_t flags, const uint64_t* cond,
                       uint64_t* words);
int materialize_cond_device(Ctx* ctx, const Chunk* ch, uint32_t flags, const uint64_t* cond, double* out,
                            uint64_t cap, uint64_t* rows);
int max_range_device(Ctx* ctx, const Chunk* ch, uint64_t s, uint64_t e, uint64_t W, uint32_t flags, double* vals,
                     uint64_t cap, uint64_t* nwin, uint64_t* argmax);
int64_t fixed_of_pred(const Chunk* ch, double pred);

// CSV ingest (buff_csv.hip)
int csv_parse_device(Ctx* ctx, const uint8_t* text, uint64_t len, uint8_t delim, uint32_t skip, double* out,
                     uint64_t cap, uint64_t* count);

// byte-slice chunks and the _with_slice filters (buff_slice_api.hip)
uint64_t slice_stream_bytes(uint64_t n, uint32_t F);
int slice_serialize(Ctx* ctx, const Chunk* ch, uint8_t* dst, uint64_t* len);
int slice_load(Ctx* ctx, const uint8_t* s, uint64_t len, uint64_t scale, uint32_t flags, Chunk** out);
int slice_filter_dispatch(Ctx* ctx, const Chunk* ch, int op, double pred, uint32_t flags, bool conditional,
                          const uint64_t* cond, uint64_t* words);
int slice_materialize_check(const Chunk* ch, int32_t precision, uint32_t flags);
int slice_sum(Ctx* ctx, const Chunk* ch, uint32_t flags, double* ref_sum, double* exact_sum, uint64_t* plane_sums);
int slice_max(Ctx* ctx, const Chunk* ch, uint32_t flags, double* value, uint64_t* max_off, uint64_t* argmax);
int slice_max_range_check(const Chunk* ch, uint64_t s, uint64_t e, uint64_t W, uint32_t flags);

// buff-major (buff_major_api.hip)
int major_encode_device(Ctx* ctx, const double* x, uint64_t n, uint64_t scale, uint32_t flags, Chunk** out);
int major_serialize(Ctx* ctx, const Chunk* ch, uint8_t* dst, uint64_t* len);
int major_load(Ctx* ctx, const uint8_t* s, uint64_t len, uint64_t scale, uint32_t flags, Chunk** out);
int major_filter_device(Ctx* ctx, const Chunk* ch, int op, double pred, uint32_t flags, uint64_t* words);
int major_sum(Ctx* ctx, const Chunk* ch, uint32_t flags, double* ref_sum, double* exact_sum, uint64_t* col_sums);
int major_max(Ctx* ctx, const Chunk* ch, uint32_t flags, double* value, uint64_t* max_off, uint64_t* argmax);
int major_materialize_device(Ctx* ctx, const Chunk* ch, int32_t precision, uint32_t flags, double* out,
                             uint64_t* rows);
uint64_t major_decoded_rows(const Chunk* ch, uint32_t flags);
// RAPG residue format (BUFF_RESIDUE chunks: the buff-major resident form, majority bytes 0).
int residue_serialize(Ctx* ctx, const Chunk* ch, uint8_t* dst, uint64_t* len);
int residue_load(Ctx* ctx, const uint8_t* s, uint64_t len, uint64_t scale, uint32_t flags, Chunk** out);
uint64_t residue_decoded_rows(const Chunk* ch, int32_t precision);
int residue_materialize_device(Ctx* ctx, const Chunk* ch, int32_t precision, double* out, uint64_t* rows);
// byte_residue_sum from a residue chunk's outlier byte sums (one per column) and
// the sums of its planes and tail (plane_sums' layout)
int residue_full_sum_fold(const Chunk* ch, const uint64_t* osum, const uint64_t* ps, double* ref_sum,
                          double* exact_sum);
int residue_sum_op(Ctx* ctx, const Chunk* ch, int32_t precision, uint32_t flags, double* ref_sum, double* exact_sum,
                   uint64_t* sums_out);
int major_precision_drop(const Chunk* ch, int32_t precision, uint32_t* drop);
MajorView major_view(const Chunk* ch);
// Filter targets of a buff-major chunk; returns -1 when the filter must run, else
// the fill of every row (1: all rows, 0: none) decided without reading the chunk.
int major_filter_plan(const Chunk* ch, int op, double pred, uint32_t flags, MajorTarget* tg);
void major_sum_fold(const Chunk* ch, const uint64_t* outsum, const uint64_t* dense_ps, double* ref_sum,
                    double* exact_sum, uint64_t* col_sums);
void major_release(Chunk* ch);

}  // namespace buffgpu
