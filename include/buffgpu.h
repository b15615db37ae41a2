/*
 * buffgpu.h — C ABI of the MI355X-native BUFF (decomposed bounded floats) engine.
 *
 * The reference exposes its codec and queries as inherent methods of
 * `SplitBDDoubleCompress` (database/src/compress/split_double.rs:29-39) and of the
 * same struct in database/src/compress/buff_simd.rs. It has no FFI of its own; the
 * entry points below are what a Rust crate keeping that API binds with
 * `extern "C"` (ptr + len for every `Vec<u8>` / `&[f64]`, plain status codes
 * instead of panics). Every `buff_sd_*` function names the reference method it
 * replaces; the `buff_*` chunk functions are the resident (HBM) form those
 * methods are built on.
 *
 * Conventions
 *  - Every function returns a status code (BUFF_OK on success) and never unwinds.
 *  - Host buffers are caller-owned. Device chunks are owned by the library and
 *    released with buff_chunk_free.
 *  - Functions taking a buff_ctx serialize on that context's HIP stream.
 *    Use one context per host thread for concurrency.
 *  - A function that returns a result to the host (a sum, a max, a count,
 *    encode's min/max) receives it in the context's mailbox, a few words of
 *    mapped host memory that a one-wave kernel fills after the work, and
 *    returns once it has arrived: all earlier work on the stream is complete,
 *    only that kernel may still be retiring.
 *  - Bitmaps use the reference's BitVec<u32> layout (bit-vec/src/lib.rs:399-436,
 *    619-627): row i is bit i%32 of u32 word i/32, unused high bits are 0.
 *    On little-endian hosts a bitmap of ceil(n/64) u64 words has the same bytes.
 */
#ifndef BUFFGPU_H
#define BUFFGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define BUFF_OK 0
#define BUFF_E_INVALID 1      /* malformed stream, bad argument */
#define BUFF_E_OOM 2          /* HBM budget or device allocation exhausted */
#define BUFF_E_DEVICE 3       /* HIP runtime error */
#define BUFF_E_REF_PANIC 4    /* an input on which the reference panics */
#define BUFF_E_UNSUPPORTED 5  /* outside what the cited reference branch covers */
#define BUFF_E_SMALL_BUFFER 6 /* caller buffer too small; the needed size is returned */

/* ---- flags --------------------------------------------------------------- */
/* Stream layout of buff_simd256_encode (buff_simd.rs:3116-3250): sub-column bytes
 * XOR 0x80, tail bits unflipped. Only the serialized stream differs; the resident
 * form is the same. */
#define BUFF_FLIP 0x1u
/* Return the mathematically intended answer where the reference has a quirk
 * (equality bitmap left in rb1 when F%8==0, target-byte truncation for
 * predicates above the column maximum, sum without the base term when F%8==0,
 * unscaled reduced-precision decode). Default (0) reproduces the reference. */
#define BUFF_SEMANTIC 0x2u
/* buff_encode_majority layout (buff_simd.rs:317-630): while the top byte of the
 * first 2000 offsets has a Boyer-Moore majority (vote counter > 1200), that byte
 * column is stored sparse as {majority byte, snappy-compressed outlier bitmap,
 * outlier bytes}; the remaining bits are byte planes and a tail as in
 * byte_fixed_encode. Accepted by buff_encode, buff_encode_device and
 * buff_chunk_load; every query dispatches on it. */
#define BUFF_MAJORITY 0x4u
/* Byte-slice layout of buff_slice_encode (buff_slice.rs:298-437): byte_fixed's
 * header and offsets, every sub-column a full byte XOR 0x80 (the last one
 * right-padded; one 0x80 plane when all values are equal). Only the stream
 * differs; the resident form is byte_fixed's. Queries of such a chunk follow
 * BuffSliceCompress (buff_slice.rs:222-1970). Excludes BUFF_FLIP and BUFF_MAJORITY. */
#define BUFF_SLICE 0x8u
/* Byte-slice queries: the scalar driver's variants (run_buff_slice_scalar_encoding_decoding,
 * buff_slice.rs:103-171): the nosimd filters, decode_bs, buff_slice_sum,
 * buff_slice_max instead of the AVX2 / "buff way" ones. */
#define BUFF_SCALAR 0x10u
/* Filters of a byte_fixed chunk: buff_simd_range_filter_with_slice /
 * buff_simd_equal_filter_with_slice (buff_simd.rs:3736-4025). */
#define BUFF_WITH_SLICE 0x20u
/* RAPG residue layout of byte_residue_encode (split_double.rs:624-954): header
 * from the true minimum and maximum; with ilen > 15 a count byte and the
 * leading byte columns whose top byte is nonzero in fewer than 10% of the first
 * 2000 offsets stored as {u32 size, CRoaring portable bitmap of the rows whose
 * byte is nonzero (run-optimized), u32 count, those bytes}, a zero byte, then
 * the remaining planes and tail as in byte_fixed_encode (with ilen <= 15 the
 * byte_fixed layout itself). Accepted by buff_encode, buff_encode_device and
 * buff_chunk_load; such a chunk answers buff_materialize (full and reduced
 * precision: byte_residue_decode[_with_precision]), buff_sum and
 * buff_sum_precision (byte_residue_sum[_with_precision]); filters and maxima
 * (no reference counterpart) only with BUFF_SEMANTIC. Excludes BUFF_FLIP,
 * BUFF_MAJORITY and BUFF_SLICE. */
#define BUFF_RESIDUE 0x40u

/* filter operators */
#define BUFF_OP_GT 0 /* x > pred (buff_range_filter, buff_simd.rs:4507-4679) */
#define BUFF_OP_EQ 1 /* x == pred (buff_equal_filter, buff_simd.rs:4925-5093) */
#define BUFF_OP_LT 2 /* x < pred: buff_slice_range_smaller_filter_condition (buff_slice.rs:1553-1662) */

typedef struct buff_ctx buff_ctx;
typedef struct buff_chunk buff_chunk;

/* Header fields of a chunk (split_double.rs:349-354) plus derived widths. */
typedef struct {
  int64_t base;          /* min fixed value */
  uint32_t n;            /* rows */
  uint32_t ilen;         /* (F - dlen) mod 2^32, as stored */
  uint32_t dlen;         /* PRECISION_MAP[p] */
  uint32_t width;        /* F = ceil(log2(max - min)) */
  uint32_t planes;       /* floor(F/8) full byte sub-columns (0 if F < 8) */
  uint32_t tail_bits;    /* F mod 8 (F itself if F < 8) */
  uint32_t flags;        /* BUFF_FLIP if the chunk serializes to the flipped layout */
  uint64_t scale;        /* the SplitBDDoubleCompress scale it was encoded with */
  uint64_t stream_bytes; /* size of the serialized reference stream */
  uint64_t device_bytes; /* HBM held by the resident form */
  uint32_t sparse_cols;  /* buff-major: sparse byte columns (planes/tail_bits then describe the rest) */
  uint32_t reserved;
  uint64_t decoded_rows; /* values buff_materialize writes at full precision (buff-major: 0 when the
                            stream ends in sparse columns, buff_simd.rs:788-845) */
} buff_chunk_info;

/* ---- context --------------------------------------------------------------- */
int buff_ctx_create(int device, uint64_t hbm_budget_bytes /* 0 = no limit */, buff_ctx** out);
int buff_ctx_destroy(buff_ctx* ctx);
int buff_ctx_synchronize(buff_ctx* ctx);
/* HBM bytes currently held by chunks and scratch of this context. */
int buff_ctx_hbm_used(buff_ctx* ctx, uint64_t* bytes);
/* Message of the last failing call on this thread ("" if none). */
const char* buff_last_error(void);

/* Plain device memory for callers that keep inputs resident (counted against
 * the HBM budget). */
int buff_dev_alloc(buff_ctx* ctx, uint64_t bytes, void** dptr);
int buff_dev_free(buff_ctx* ctx, void* dptr);
int buff_memcpy_h2d(buff_ctx* ctx, void* dst_dev, const void* src_host, uint64_t bytes);
int buff_memcpy_d2h(buff_ctx* ctx, void* dst_host, const void* src_dev, uint64_t bytes);
/* Fills device bytes with `value` on the context's stream (no host sync). */
int buff_dev_fill(buff_ctx* ctx, void* dptr, uint8_t value, uint64_t bytes);
/* Reads device bytes and discards them (no host sync): refills the memory-side
 * cache with clean lines before a cold-cache measurement. */
int buff_dev_touch(buff_ctx* ctx, const void* dptr, uint64_t bytes);
/* Copies device bytes with non-temporal loads and stores on the context's
 * stream (no host sync): the streaming read + write rate that encode and
 * materialize are measured against. bytes, dst and src multiples of 16; the
 * ranges must not overlap. No reference counterpart. */
int buff_dev_copy(buff_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Reads device bytes with non-temporal loads in the plane kernels' order and
 * discards them (no host sync): the streaming read rate that the plane-reading
 * kernels (filters, max, sum) are measured against. bytes and src multiples of
 * 16. No reference counterpart. */
int buff_dev_read(buff_ctx* ctx, const void* src, uint64_t bytes);

/* Per-kernel timing with HIP events recorded on the context's stream.
 * buff_prof_query reports launches and summed durations for one kernel name. */
int buff_prof_enable(buff_ctx* ctx, int on);
int buff_prof_reset(buff_ctx* ctx);
int buff_prof_query(buff_ctx* ctx, const char* kernel, uint64_t* launches, double* total_ms);

/* ---- encode / resident chunks ---------------------------------------------- */
/* byte_fixed_encode (split_double.rs:298-432) into the resident form.
 * x_dev: n doubles in device memory. BUFF_FLIP selects buff_simd256_encode's
 * serialized layout. Returns BUFF_E_REF_PANIC where the reference panics
 * (F == 0 with n >= 1; scale whose precision is not in PRECISION_MAP). */
int buff_encode_device(buff_ctx* ctx, const double* x_dev, uint64_t n, uint64_t scale, uint32_t flags,
                       buff_chunk** out);
/* Same, from host memory (uploads first). */
int buff_encode(buff_ctx* ctx, const double* x_host, uint64_t n, uint64_t scale, uint32_t flags,
                buff_chunk** out);
int buff_chunk_info_get(const buff_chunk* chunk, buff_chunk_info* info);
/* Writes the reference byte stream. dst == NULL: only *len is set. */
int buff_chunk_serialize(buff_ctx* ctx, const buff_chunk* chunk, uint8_t* dst, uint64_t* len);
/* Parses a reference byte stream (BUFF_FLIP: a buff_simd256 stream; BUFF_SLICE:
 * a buff_slice_encode stream) into the resident form. `scale` is the caller's
 * SplitBDDoubleCompress scale. */
int buff_chunk_load(buff_ctx* ctx, const uint8_t* stream, uint64_t len, uint64_t scale, uint32_t flags,
                    buff_chunk** out);
int buff_chunk_free(buff_chunk* chunk);
/* Sparse column j of a buff-major chunk (buff_simd.rs:451-530): its majority
 * byte, outlier rows and the row of its last outlier (UINT64_MAX if none). */
int buff_chunk_sparse_col(const buff_chunk* chunk, uint32_t j, uint32_t* major, uint64_t* outliers,
                          uint64_t* last_row);

/* ---- queries on a resident chunk ------------------------------------------ */
/* Filter into host words: ceil(n/32) u32 words; *count = set bits (the count the
 * reference prints). On a BUFF_SLICE chunk these and the calls below follow
 * BuffSliceCompress: buff_filter is buff_slice_range_filter / _equal_filter
 * (BUFF_SCALAR: the _nosimd forms), buff_filter_cond is
 * buff_slice_range_filter_condition (BUFF_OP_GT) or
 * buff_slice_range_smaller_filter_condition (BUFF_OP_LT), buff_materialize is
 * buff_slice_decode (BUFF_SCALAR: decode_bs), buff_materialize_cond is
 * buff_slice_decode_condition, buff_sum_ex / buff_max_ex are buffsum / buffmax
 * (BUFF_SCALAR: sum / max), buff_max_range is buff_slice_max_range. With
 * BUFF_WITH_SLICE on a byte_fixed chunk, buff_filter is
 * buff_simd_range_filter_with_slice / buff_simd_equal_filter_with_slice. */
int buff_filter(buff_ctx* ctx, const buff_chunk* chunk, int op, double pred, uint32_t flags, uint32_t* words,
                uint64_t nwords, uint64_t* count);
/* Filter into device memory (ceil(n/64) u64 words, 8-byte aligned); no host sync.
 * stats (nullable, device, 9 u64, zeroed by the caller) accumulates rows
 * entering sub-column k >= 1 (index k) and the tail (index planes). */
int buff_filter_device(buff_ctx* ctx, const buff_chunk* chunk, int op, double pred, uint32_t flags,
                       uint64_t* words_dev, uint64_t* stats_dev);
/* ---- filter -> project: queries over the rows of a condition bitmap --------
 * A condition is a bitmap in the layout the filters write (bit i = row i; the
 * _device forms take ceil(n/64) u64 words, the host forms ceil(n/32) u32
 * words), e.g. the result of an earlier filter: predicates chain, and a
 * filter's rows can be materialized. On a buff-major chunk these are
 * buff_range_filter_majority_condition (buff_simd.rs:1279-1526) and
 * buff_major_decode_condition (1714-1914): with sparse columns the reference
 * answers nothing (no rows, no values); there BUFF_SEMANTIC (and the filter's
 * equality form) answer with the chunk's own filter and stored values on the
 * condition rows; without, they equal the byte_fixed forms except
 * that the decode places the tail bits correctly and an empty condition gives
 * no values. F < 8 buff-major chunks are BUFF_E_UNSUPPORTED.
 *
 * buff_range_filter_condition (buff_simd.rs:4683-4921): x > pred over the
 * condition rows (the reference takes their row ids as an iterator); rows
 * outside the condition are 0. Default reproduces its early returns: a
 * predicate below the column base gives no rows, and with F < 8 the target is
 * its low byte. BUFF_SEMANTIC: x > pred on every condition row. op BUFF_OP_EQ
 * has no reference counterpart: buff_filter's equality on the condition rows.
 * *count (host form) = rows set in the result. */
int buff_filter_cond(buff_ctx* ctx, const buff_chunk* chunk, int op, double pred, uint32_t flags,
                     const uint32_t* cond_words, uint32_t* words, uint64_t nwords, uint64_t* count);
int buff_filter_cond_device(buff_ctx* ctx, const buff_chunk* chunk, int op, double pred, uint32_t flags,
                            const uint64_t* cond_dev, uint64_t* words_dev);
/* buff_decode_condition (buff_simd.rs:1528-1712): the values of the condition
 * rows in row order, at most cap of them; *rows = the condition's row count
 * (BUFF_E_SMALL_BUFFER when it exceeds cap). Default reproduces the reference:
 * when F >= 8 and F % 8 != 0 the tail bits are OR-ed in shifted left by the
 * tail width (1693, 1701), and an empty condition is its panic
 * (BUFF_E_REF_PANIC). BUFF_SEMANTIC: the stored values; an empty condition
 * gives 0 rows. The device form reads the row count back (one host sync). */
int buff_materialize_cond(buff_ctx* ctx, const buff_chunk* chunk, uint32_t flags, const uint32_t* cond_words,
                          double* out, uint64_t cap, uint64_t* rows);
int buff_materialize_cond_device(buff_ctx* ctx, const buff_chunk* chunk, uint32_t flags, const uint64_t* cond_dev,
                                 double* out_dev, uint64_t cap, uint64_t* rows);

/* buff_max_range (buff_simd.rs:5096-5334): the max of every window of `window`
 * rows of [s, e), as the reference's f64 (max offset + base) / 2^dlen, in window
 * order (*nwindows of them, at most cap; BUFF_E_SMALL_BUFFER beyond), and
 * (argmax_words, nullable: ceil(n/32) u32 words) the rows holding their
 * window's max. Default reproduces the reference: only full windows are
 * reported; when F > 8 a range the windows do not tile exactly (or an empty
 * one) is its panic (BUFF_E_REF_PANIC); the argmax rows are filled only where
 * its tail stage runs (F < 8, or F > 8 with F % 8 != 0), and for F < 8 they
 * include the rows of a partial last window. BUFF_SEMANTIC: every window,
 * including a partial last one, and the argmax rows in every case. window = 0
 * or a range outside [0, n] is BUFF_E_INVALID. A buff-major chunk is
 * buff_max_majority_range (buff_simd.rs:2863-3114): with sparse columns no
 * window and no argmax row (its log-and-return; BUFF_SEMANTIC there and F < 8
 * are BUFF_E_UNSUPPORTED), without them the byte_fixed result. The device form
 * writes values and ceil(n/64) u64 argmax words to device memory (argmax_dev
 * nullable). */
int buff_max_range(buff_ctx* ctx, const buff_chunk* chunk, uint64_t s, uint64_t e, uint64_t window, uint32_t flags,
                   double* max_values, uint64_t cap, uint64_t* nwindows, uint32_t* argmax_words, uint64_t nwords);
int buff_max_range_device(buff_ctx* ctx, const buff_chunk* chunk, uint64_t s, uint64_t e, uint64_t window,
                          uint32_t flags, double* max_values_dev, uint64_t cap, uint64_t* nwindows,
                          uint64_t* argmax_dev);

/* byte_fixed_sum (split_double.rs:2127-2208). ref_sum reproduces the reference
 * f64 fold bit for bit; exact_sum = (n*base + sum(off)) / 2^dlen rounded once;
 * plane_sums (nullable, 9 u64) are the exact per-sub-column byte sums (tail last). */
int buff_sum(buff_ctx* ctx, const buff_chunk* chunk, double* ref_sum, double* exact_sum, uint64_t* plane_sums);
/* buff_sum with flags. A BUFF_SLICE chunk is buff_slice_buffsum (buff_slice.rs:
 * 702-814; BUFF_SCALAR: buff_slice_sum, 1101-1182, a panic unless 1 <= F <= 32):
 * ref_sum = exact_sum, the f64 running sum's value within its rounding. */
int buff_sum_ex(buff_ctx* ctx, const buff_chunk* chunk, uint32_t flags, double* ref_sum, double* exact_sum,
                uint64_t* plane_sums);
/* buff_sum without allocating: no reference counterpart; for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). The same results as buff_sum on a plain
 * (byte_fixed) chunk; buff-major and byte-slice chunks are BUFF_E_UNSUPPORTED.
 * The partial sums live in a 589,896-byte device array of the library (one per
 * GPU, held until the process ends) that covers every column length (the sum's
 * grid is capped); the call allocates nothing and returns after its work is
 * done. All calls on one GPU share that array: calls that overlap on one GPU
 * race and must be serialized by the caller. */
int buff_sum_noalloc(buff_ctx* ctx, const buff_chunk* chunk, double* ref_sum, double* exact_sum,
                     uint64_t* plane_sums);
/* byte_residue_sum_with_precision (split_double.rs:2409-2629) on a plain chunk
 * (the stream branch taken when ilen <= 15; larger ilen is BUFF_E_UNSUPPORTED,
 * as are buff-major and byte-slice chunks). The read width R (the rule of
 * 2439-2462, shared with the reduced-precision decode) selects R/8 whole
 * sub-columns and, when R % 8 = q > 0, q-bit fields read from the packed tail
 * (reading fields of q < r bits at stride q, as the reference does). ref_sum
 * reproduces the reference f64 fold bit for bit, including the missing base
 * term when R is a positive multiple of 8; exact_sum is the exact sum of the
 * values truncated to their top min(R, F) stored bits, rounded once; sums
 * (nullable, 9 u64) are the u64 sums read, in read order. Reference panics
 * (precision not in PRECISION_MAP, dlen below the requested bits, reads past
 * the stream) are BUFF_E_REF_PANIC. BUFF_SEMANTIC: ref_sum = exact_sum and
 * those panics and the ilen > 15 branch do not apply (more decimals than
 * stored keep every bit). */
int buff_sum_precision(buff_ctx* ctx, const buff_chunk* chunk, int32_t precision, uint32_t flags, double* ref_sum,
                       double* exact_sum, uint64_t* sums);
/* byte_fixed_max (split_double.rs:3919-4105): the printed max value and the max
 * stored offset. */
int buff_max(buff_ctx* ctx, const buff_chunk* chunk, double* max_value, uint64_t* max_off);
/* buff_max with flags (BUFF_SEMANTIC: the true maximum of a buff-major chunk)
 * and, for buff-major chunks, the rows tying the max in the tail stage that
 * buff_max_majority prints (buff_simd.rs:2856-2858; 0 otherwise). */
int buff_max_ex(buff_ctx* ctx, const buff_chunk* chunk, uint32_t flags, double* max_value, uint64_t* max_off,
                uint64_t* argmax_count);
/* byte_fixed_decode (split_double.rs:1243-1350) for precision < 0, else
 * byte_residue_decode_with_precision (split_double.rs:1612-1893) at `precision`
 * decimals. out: n doubles (host). A buff-major chunk decodes as
 * buff_decode_majority (buff_simd.rs:633-861) and writes decoded_rows values
 * (buff_chunk_info); the reference has no reduced-precision decode of it
 * (BUFF_E_UNSUPPORTED), but with BUFF_SEMANTIC a precision >= 0 gives the true
 * values truncated to their top R stored bits (R from the same read rule). */
int buff_materialize(buff_ctx* ctx, const buff_chunk* chunk, int32_t precision, uint32_t flags, double* out);
int buff_materialize_device(buff_ctx* ctx, const buff_chunk* chunk, int32_t precision, uint32_t flags,
                            double* out_dev);

/* ---- launch-shape tuning -----------------------------------------------------
 * Times candidate launch shapes (waves per workgroup, grid cap, tiles a wave
 * loads per step) of the range filter (at `pred`), sum, max and full
 * materialization on this byte_fixed chunk, each run after evicting the
 * memory-side cache, and keeps the fastest per operation for later calls on
 * chunks of the same size class and byte-plane count. choices: 4 entries in
 * that order (may be NULL). No reference counterpart. */
typedef struct {
  uint32_t op;             /* internal operation id */
  uint32_t waves_per_wg;
  uint32_t max_wg;         /* 0: the workgroups that fit resident */
  uint32_t tiles_per_wave; /* 0: the kernel's default */
  uint32_t prefetch;       /* range filter: next step's top plane loaded before refining; max: 1 = the
                              one-pass form (0: the top bytes of every tile, then only the tiles holding
                              the largest) */
  float best_us;           /* kernel time of the chosen shape */
  float default_us;        /* kernel time of the default shape */
} buff_tune_choice;
int buff_tune_chunk(buff_ctx* ctx, const buff_chunk* chunk, double pred, buff_tune_choice* choices);

/* ---- batched execution ------------------------------------------------------
 * Many chunks of uneven length and mixed precision encoded in one persistent
 * launch (plus one min/max launch), and queries over them run as one captured
 * hipGraph per batch. There is no reference counterpart: the reference runs one
 * method call per chunk (the drivers of buff_simd.rs:167-311). Every result equals
 * the single-chunk call's on the same chunk. */
typedef struct buff_batch buff_batch;

/* Encodes nchunks columns laid end to end in device memory: chunk i is rows
 * [row_off[i], row_off[i+1]) of x_dev with scale scales[i] (host arrays). The
 * chunks share one HBM pool owned by the batch. flags: BUFF_FLIP. A chunk on
 * which byte_fixed_encode would panic fails the call (buff_last_error names it). */
int buff_batch_encode_device(buff_ctx* ctx, const double* x_dev, const uint64_t* row_off, const uint64_t* scales,
                             uint32_t nchunks, uint32_t flags, buff_batch** out);
int buff_batch_size(const buff_batch* batch, uint32_t* nchunks);
/* Chunk i as a buff_chunk usable with every single-chunk call; it belongs to the
 * batch (buff_chunk_free refuses it). */
int buff_batch_chunk(buff_batch* batch, uint32_t i, buff_chunk** chunk);
int buff_batch_free(buff_batch* batch);

#define BUFF_Q_GT 0  /* x > pred */
#define BUFF_Q_EQ 1  /* x == pred */
#define BUFF_Q_SUM 2
#define BUFF_Q_MAX 3
#define BUFF_Q_MAT 4 /* materialize at `precision` decimals (-1: full) */

typedef struct {
  uint32_t op;       /* BUFF_Q_* */
  uint32_t chunk;    /* index in the batch */
  double pred;       /* filters */
  int32_t precision; /* BUFF_Q_MAT */
  uint32_t flags;    /* BUFF_SEMANTIC */
} buff_query;

typedef struct {
  int32_t status;      /* what the single-chunk call returns for this query */
  uint32_t reserved;
  uint64_t count;      /* filters: rows set */
  double value;        /* sum: the reference f64 sum; max: the max value */
  double exact;        /* sum: exact sum */
  uint64_t max_off;    /* max: max stored offset */
  uint64_t out_offset; /* buff_batch_query: byte offset of a filter's bitmap / a materialization's
                          values in the batch arena; server requests: their device address */
} buff_query_result;

/* Runs nq queries on one batch: host planning, one graph launch, one result
 * copy. Filter bitmaps (ceil(n/64) u64 words each) and materialized values stay
 * in the batch's device arena (buff_batch_arena gives its base; both pointers are
 * that base) until the next call. */
int buff_batch_query(buff_ctx* ctx, buff_batch* batch, const buff_query* qs, uint32_t nq, buff_query_result* results);
int buff_batch_arena(buff_batch* batch, const uint64_t** bitmaps_dev, const double** values_dev);
/* Times candidate shapes on this query mix and keeps the fastest for the
 * context's later batches: the persistent grid size, the tiles per work item,
 * and the planes a filter reads before it checks for tied rows (1: the top byte
 * plane, then the rest only for tiles with ties; 8: every plane in one pass). */
int buff_batch_tune(buff_ctx* ctx, buff_batch* batch, const buff_query* qs, uint32_t nq, uint32_t* grid,
                    uint32_t* tiles_per_item, uint32_t* first_planes);
/* Sets the shape later batches of this context and servers created afterwards
 * use (as if buff_batch_tune had chosen it): grid 1..65536, tiles_per_item
 * 1..4096, first_planes 1 or 8. Call it while no batch of the context is being
 * planned. */
int buff_batch_set_tuning(buff_ctx* ctx, uint32_t grid, uint32_t tiles_per_item, uint32_t first_planes);

/* ---- serving ------------------------------------------------------------------
 * A request loop over resident chunks with continuous batching: requests that
 * arrive while batches run on the GPU are planned into the next batch as soon as
 * one of `slots` in-flight batch slots is free (each slot has its own stream, so
 * batches overlap on the device) and while the batch has room (max_batch
 * requests). Admission by memory: a filter or materialization submitted
 * without an output address gets its output range in the server's arena
 * (arena_bytes of device memory); requests are admitted in arrival order, and
 * the oldest waits (holding back the ones behind it) until enough of the arena
 * is free. Such a range stays held, and the output at the result's out_offset
 * stays valid, until the client returns it with buff_server_release. A request
 * whose output is larger than the whole arena fails with BUFF_E_OOM. Results
 * equal the single-chunk calls'. No reference counterpart (the reference has
 * no query server). */
typedef struct buff_server buff_server;
typedef struct {
  uint32_t max_batch;   /* requests per batch (0: 4096) */
  uint32_t slots;       /* batches in flight (0: 2) */
  uint64_t arena_bytes; /* the server's output arena for requests without out_dev (0: 1 GiB) */
} buff_server_cfg;

int buff_server_create(buff_ctx* ctx, const buff_server_cfg* cfg, buff_server** out);
/* Makes a resident byte_fixed, buff-major or residue chunk of the same
 * context servable; *id is the chunk index requests name. The caller keeps
 * ownership (free it after the server). Calls from several client threads may
 * overlap; every call on the server is serialized on its lock. On a buff-major
 * chunk with sparse columns, filters, sums and materializations are batched
 * like the others, and so are a residue chunk's sums and full-precision
 * materializations. Requests without a batched kernel (a max on a buff-major
 * chunk with sparse columns; a residue chunk's filters, maxima and
 * reduced-precision materializations) run the single-chunk call
 * (buff_filter_device with the count of its bitmap, buff_max_ex,
 * buff_materialize_device) when their batch completes, on a context the server
 * creates for itself, and answer what that call answers. */
int buff_server_add_chunk(buff_server* server, buff_chunk* chunk, uint32_t* id);
/* Queues one request (buff_query.chunk = server chunk id). out_dev: where a
 * filter writes its ceil(n/64) bitmap words or a materialization its n doubles
 * (NULL: a range of the server arena, held until buff_server_release). */
int buff_server_submit(buff_server* server, const buff_query* q, void* out_dev, uint64_t* ticket);
/* Blocks until the request completes; latency_us = completion - submission.
 * result->out_offset is the device address of a filter's bitmap or a
 * materialization's values (out_dev, or the held arena range); 0 for sums and
 * maxima and for failed requests. The ticket is spent afterwards. */
int buff_server_wait(buff_server* server, uint64_t ticket, buff_query_result* result, double* latency_us);
/* *done = 1 when the request has completed (it can then be waited without blocking). */
int buff_server_poll(buff_server* server, uint64_t ticket, int* done);
/* Returns a completed request's arena output range (its out_offset) to the
 * arena; requests waiting for arena memory are admitted as ranges return. */
int buff_server_release(buff_server* server, const void* out);
/* Arena bytes held now, the most ever held at once, and the arena's size. */
int buff_server_arena_stats(buff_server* server, uint64_t* in_use, uint64_t* peak, uint64_t* capacity);
/* Submits request i at arrival_s[i] seconds after the call (a native open-loop
 * client), then waits for all of them in order, returning each arena output
 * range as soon as its request completes (results[i].out_offset is then 0;
 * give outs[i] to keep an output). outs may be NULL. */
int buff_server_run_trace(buff_server* server, const buff_query* qs, const double* arrival_s, void* const* outs,
                          uint32_t n, buff_query_result* results, double* latency_us);
/* The same open-loop replay without a server: each request, in arrival order,
 * is the plain single-chunk call on the context's stream (buff_filter_device,
 * buff_sum, buff_max, buff_materialize_device; one query per launch), followed
 * by a stream synchronization for a filter or materialization (a sum or max is
 * complete when its call returns); latency_us[i] = completion - arrival. Filters
 * write their bitmap but no count (results[i].count = 0). Outputs of requests
 * without outs[i] go to one context buffer that each call overwrites. chunks:
 * the resident chunks requests name by index. The baseline that continuous
 * batching is measured against. */
int buff_trace_per_call(buff_ctx* ctx, buff_chunk* const* chunks, uint32_t nchunks, const buff_query* qs,
                        const double* arrival_s, void* const* outs, uint32_t n, buff_query_result* results,
                        double* latency_us);
/* Batches launched, requests served, largest batch; returns a sticky device error if any. */
int buff_server_stats(buff_server* server, uint64_t* batches, uint64_t* queries, uint64_t* largest_batch);
int buff_server_destroy(buff_server* server);

/* ---- reference-shaped entry points (byte streams in host memory) ----------
 * Each takes the SplitBDDoubleCompress fields {chunksize, batchsize, scale}
 * (split_double.rs:29-39) and the stream as ptr+len, mirrors one method, and
 * runs it on the GPU of a process-wide default context (device 0). */
typedef struct {
  uint64_t chunksize;
  uint64_t batchsize;
  uint64_t scale;
} buff_split_double;

/* split_double.rs:298 byte_fixed_encode / buff_simd.rs:3116 buff_simd256_encode.
 * out == NULL: size query. */
int buff_sd_byte_fixed_encode(const buff_split_double* sd, const double* data, uint64_t n, uint8_t* out,
                              uint64_t cap, uint64_t* out_len);
int buff_sd_buff_simd256_encode(const buff_split_double* sd, const double* data, uint64_t n, uint8_t* out,
                                uint64_t cap, uint64_t* out_len);
/* split_double.rs:1243 byte_fixed_decode / buff_simd.rs:3252 buff_simd256_decode.
 * out: n doubles (n is in the stream header). */
int buff_sd_byte_fixed_decode(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* out,
                              uint64_t cap);
int buff_sd_buff_simd256_decode(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* out,
                                uint64_t cap);
/* buff_simd.rs:4507 buff_range_filter / 4925 buff_equal_filter, and the
 * buff-simd twins buff_simd.rs:3361 buff_simd_range_filter / 3549
 * buff_simd_equal_filter over a flipped stream. The reference prints the count;
 * here it is returned with the bitmap (words: ceil(n/32) u32, nullable). */
int buff_sd_buff_range_filter(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double pred,
                              uint32_t* words, uint64_t nwords, uint64_t* count);
int buff_sd_buff_equal_filter(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double pred,
                              uint32_t* words, uint64_t nwords, uint64_t* count);
int buff_sd_buff_simd_range_filter(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double pred,
                                   uint32_t* words, uint64_t nwords, uint64_t* count);
int buff_sd_buff_simd_equal_filter(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double pred,
                                   uint32_t* words, uint64_t nwords, uint64_t* count);
/* split_double.rs:2127 byte_fixed_sum */
int buff_sd_byte_fixed_sum(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* sum);
/* split_double.rs:3919 byte_fixed_max (prints the max; returned here) */
int buff_sd_byte_fixed_max(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* max_value);
/* RAPG residue streams (BUFF_RESIDUE; with ilen <= 15 the byte_fixed stream):
 * split_double.rs:624 byte_residue_encode (out == NULL: size query), :1353
 * byte_residue_decode, :1612 byte_residue_decode_with_precision (n_out, may be
 * NULL: values written, 0 when the outlier columns take every bit read), :2210
 * byte_residue_sum, :2409 byte_residue_sum_with_precision. A reduced-precision
 * decode of a stream with outlier columns whose read ends in fields of the
 * packed tail is BUFF_E_UNSUPPORTED. */
int buff_sd_byte_residue_encode(const buff_split_double* sd, const double* data, uint64_t n, uint8_t* out,
                                uint64_t cap, uint64_t* out_len);
int buff_sd_byte_residue_decode(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* out,
                                uint64_t cap, uint64_t* n_out);
int buff_sd_byte_residue_decode_with_precision(const buff_split_double* sd, const uint8_t* bytes, uint64_t len,
                                               uint32_t precision, double* out, uint64_t cap, uint64_t* n_out);
int buff_sd_byte_residue_sum(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* sum);
int buff_sd_byte_residue_sum_with_precision(const buff_split_double* sd, const uint8_t* bytes, uint64_t len,
                                            uint32_t precision, double* sum);
/* buff-major: buff_simd.rs:317 buff_encode_majority (out == NULL: size query),
 * :633 buff_decode_majority (*n_out = values written; 0 when the stream ends in
 * sparse columns), :863 buff_range_filter_majority, :1916
 * buff_equal_filter_majority, :2294 buff_sum_majority, :2444 buff_max_majority. */
int buff_sd_buff_encode_majority(const buff_split_double* sd, const double* data, uint64_t n, uint8_t* out,
                                 uint64_t cap, uint64_t* out_len);
int buff_sd_buff_decode_majority(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* out,
                                 uint64_t cap, uint64_t* n_out);
int buff_sd_buff_range_filter_majority(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double pred,
                                       uint32_t* words, uint64_t nwords, uint64_t* count);
int buff_sd_buff_equal_filter_majority(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double pred,
                                       uint32_t* words, uint64_t nwords, uint64_t* count);
int buff_sd_buff_sum_majority(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* sum);
int buff_sd_buff_max_majority(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* max_value);

/* ---- CSV ingest ---------------------------------------------------------------
 * construct_file_iterator_skip_newline::<f64>(file, skip_val, delim)
 * (database/src/client.rs:163-182), the reader every benchmark driver starts
 * with: lines split on '\n' (a trailing "\r" removed), items split on delim
 * (ASCII, not CR), the first skip_val items of each line skipped, and every item
 * f64::from_str accepts kept in file order, correctly rounded; other items are
 * dropped. *count = values found (BUFF_E_SMALL_BUFFER when above cap).
 * BUFF_E_UNSUPPORTED: a byte >= 0x80 (the reference drops lines that are not
 * UTF-8), or an item whose rounding the 128-bit power of 5 leaves open.
 * The _device form parses text already in device memory into device memory
 * (e.g. straight into buff_encode_device); buff_csv_parse takes host buffers;
 * buff_construct_file_iterator_skip_newline reads the file (BUFF_E_INVALID
 * where the reference returns Err(())). */
int buff_csv_parse_device(buff_ctx* ctx, const uint8_t* text_dev, uint64_t len, uint8_t delim, uint32_t skip_val,
                          double* out_dev, uint64_t cap, uint64_t* count);
int buff_csv_parse(buff_ctx* ctx, const uint8_t* text, uint64_t len, uint8_t delim, uint32_t skip_val, double* out,
                   uint64_t cap, uint64_t* count);
int buff_construct_file_iterator_skip_newline(const char* file, uint32_t skip_val, uint8_t delim, double* out,
                                              uint64_t cap, uint64_t* count);

/* byte-slice (BuffSliceCompress, buff_slice.rs:222-1970; its fields are
 * SplitBDDoubleCompress's {chunksize, batchsize, scale}, buff_slice.rs:222-232):
 * :298 buff_slice_encode (out == NULL: size query), :439 buff_slice_decode,
 * :1184 buff_slice_decode_bs, :1348 buff_slice_range_filter, :1665
 * buff_slice_equal_filter, :1753 buff_slice_range_filter_nosimd, :1873
 * buff_slice_equal_filter_nosimd, :702 buff_slice_buffsum, :1101
 * buff_slice_sum, :548 buff_slice_buffmax, :816 buff_slice_max; and over a
 * buff_simd256 stream buff_simd.rs:3736 buff_simd_range_filter_with_slice /
 * :3888 buff_simd_equal_filter_with_slice. Sums: the exact sum rounded once
 * (the reference's f64 running sum agrees within 1e-12 relative on the tested
 * columns). Maxima: the printed value and the printed argmax count. */
typedef buff_split_double buff_slice_compress;
int buff_sd_buff_slice_encode(const buff_slice_compress* sc, const double* data, uint64_t n, uint8_t* out,
                              uint64_t cap, uint64_t* out_len);
int buff_sd_buff_slice_decode(const buff_slice_compress* sc, const uint8_t* bytes, uint64_t len, double* out,
                              uint64_t cap);
int buff_sd_buff_slice_decode_bs(const buff_slice_compress* sc, const uint8_t* bytes, uint64_t len, double* out,
                                 uint64_t cap);
int buff_sd_buff_slice_range_filter(const buff_slice_compress* sc, const uint8_t* bytes, uint64_t len, double pred,
                                    uint32_t* words, uint64_t nwords, uint64_t* count);
int buff_sd_buff_slice_equal_filter(const buff_slice_compress* sc, const uint8_t* bytes, uint64_t len, double pred,
                                    uint32_t* words, uint64_t nwords, uint64_t* count);
int buff_sd_buff_slice_range_filter_nosimd(const buff_slice_compress* sc, const uint8_t* bytes, uint64_t len,
                                           double pred, uint32_t* words, uint64_t nwords, uint64_t* count);
int buff_sd_buff_slice_equal_filter_nosimd(const buff_slice_compress* sc, const uint8_t* bytes, uint64_t len,
                                           double pred, uint32_t* words, uint64_t nwords, uint64_t* count);
int buff_sd_buff_simd_range_filter_with_slice(const buff_split_double* sd, const uint8_t* bytes, uint64_t len,
                                              double pred, uint32_t* words, uint64_t nwords, uint64_t* count);
int buff_sd_buff_simd_equal_filter_with_slice(const buff_split_double* sd, const uint8_t* bytes, uint64_t len,
                                              double pred, uint32_t* words, uint64_t nwords, uint64_t* count);
int buff_sd_buff_slice_buffsum(const buff_slice_compress* sc, const uint8_t* bytes, uint64_t len, double* sum);
int buff_sd_buff_slice_sum(const buff_slice_compress* sc, const uint8_t* bytes, uint64_t len, double* sum);
int buff_sd_buff_slice_buffmax(const buff_slice_compress* sc, const uint8_t* bytes, uint64_t len, double* max_value,
                               uint64_t* argmax_count);
int buff_sd_buff_slice_max(const buff_slice_compress* sc, const uint8_t* bytes, uint64_t len, double* max_value,
                           uint64_t* argmax_count);

/* ---- the Segment container ---------------------------------------------------
 * Segment<T> (segment.rs:36-44) as bincode 1.x writes it (segment.rs:84-101):
 * little-endian, u64 lengths, u32 enum variants, u8 Option tags, fields in
 * declaration order. With method, time_lapse and prev_seg_offset None a
 * Segment<f64> is 31 + 8n bytes. Host code only (buff_segment_encode aside). */
typedef struct {
  int32_t method;             /* -1 None; Methods (methods/mod.rs:40-45): 0 Fourier, 1 Kernel(id),
                                 2 SparseLearning(id), 3 DeepLearning(model) */
  uint32_t dictionary_id;     /* Kernel / SparseLearning */
  const char* model;          /* DeepLearning: model_len UTF-8 bytes */
  uint64_t model_len;
  uint64_t timestamp_secs;    /* timestamp: SystemTime since UNIX_EPOCH */
  uint32_t timestamp_nanos;
  uint64_t signal;            /* SignalId */
  uint32_t elem_bytes;        /* T: 8 = f64, 4 = f32 */
  const void* data;           /* data: n values of T, little-endian */
  uint64_t n;
  int32_t has_time_lapse;     /* time_lapse: Option<Vec<Duration>> */
  const uint8_t* lapse;       /* n_lapse records of 12 bytes: u64 secs, u32 nanos */
  uint64_t n_lapse;
  int32_t has_prev_offset;    /* prev_seg_offset: Option<Duration> */
  uint64_t prev_secs;
  uint32_t prev_nanos;
  uint64_t consumed;          /* buff_segment_parse: bytes the Segment took (trailing bytes are allowed) */
} buff_segment;
/* Segment::get_byte_size (segment.rs:90-95). */
int buff_segment_byte_size(const buff_segment* seg, uint64_t* size);
/* Segment::convert_to_bytes (segment.rs:84-89). dst == NULL: only *len is set;
 * BUFF_E_SMALL_BUFFER (with *len) when cap is short. */
int buff_segment_serialize(const buff_segment* seg, uint8_t* dst, uint64_t cap, uint64_t* len);
/* Segment::convert_from_bytes (segment.rs:96-101): BUFF_E_INVALID where it
 * returns Err (early end, an Option tag above 1, a Methods variant above 3, a
 * model name that is not UTF-8, a time whose seconds overflow). The pointers
 * (model, data, lapse) point into `bytes`. */
int buff_segment_parse(const uint8_t* bytes, uint64_t len, uint32_t elem_bytes, buff_segment* seg);
/* byte_fixed_encode (split_double.rs:298) on a serialized Segment<T>: its data,
 * as float64, encoded into the resident form (flags as buff_encode). */
int buff_segment_encode(buff_ctx* ctx, const uint8_t* bytes, uint64_t len, uint32_t elem_bytes, uint64_t scale,
                        uint32_t flags, buff_chunk** out);

#ifdef __cplusplus
}
#endif

#endif /* BUFFGPU_H */
