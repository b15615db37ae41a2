//! The C ABI of libbuffgpu as include/buffgpu.h declares it: opaque handles,
//! the plain structs passed by pointer, the flag, operation and status values,
//! and every entry point. Layouts are #[repr(C)] copies of the header's.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

#[repr(C)]
pub struct BuffCtx {
    _private: [u8; 0],
}
#[repr(C)]
pub struct BuffChunk {
    _private: [u8; 0],
}
#[repr(C)]
pub struct BuffBatch {
    _private: [u8; 0],
}
#[repr(C)]
pub struct BuffServer {
    _private: [u8; 0],
}

/// buff_chunk_info: header fields of a chunk (split_double.rs:349-354) plus derived widths.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct BuffChunkInfo {
    pub base: i64,
    pub n: u32,
    pub ilen: u32,
    pub dlen: u32,
    pub width: u32,
    pub planes: u32,
    pub tail_bits: u32,
    pub flags: u32,
    pub scale: u64,
    pub stream_bytes: u64,
    pub device_bytes: u64,
    pub sparse_cols: u32,
    pub reserved: u32,
    pub decoded_rows: u64,
}

/// buff_tune_choice
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct BuffTuneChoice {
    pub op: u32,
    pub waves_per_wg: u32,
    pub max_wg: u32,
    pub tiles_per_wave: u32,
    pub prefetch: u32,
    pub best_us: f32,
    pub default_us: f32,
}

/// buff_query: one query of a batch or one served request.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct BuffQuery {
    pub op: u32,
    pub chunk: u32,
    pub pred: f64,
    pub precision: i32,
    pub flags: u32,
}

/// buff_query_result
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct BuffQueryResult {
    pub status: i32,
    pub reserved: u32,
    pub count: u64,
    pub value: f64,
    pub exact: f64,
    pub max_off: u64,
    pub out_offset: u64,
}

/// buff_server_cfg
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct BuffServerCfg {
    pub max_batch: u32,
    pub slots: u32,
    pub arena_bytes: u64,
}

/// buff_split_double (also buff_slice_compress): SplitBDDoubleCompress's fields
/// (split_double.rs:29-39; BuffSliceCompress, buff_slice.rs:222-232).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct BuffSplitDouble {
    pub chunksize: u64,
    pub batchsize: u64,
    pub scale: u64,
}

/// buff_segment: Segment<T> (segment.rs:36-44) in the fields its bincode form holds.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct BuffSegment {
    pub method: i32,
    pub dictionary_id: u32,
    pub model: *const c_char,
    pub model_len: u64,
    pub timestamp_secs: u64,
    pub timestamp_nanos: u32,
    pub signal: u64,
    pub elem_bytes: u32,
    pub data: *const c_void,
    pub n: u64,
    pub has_time_lapse: i32,
    pub lapse: *const u8,
    pub n_lapse: u64,
    pub has_prev_offset: i32,
    pub prev_secs: u64,
    pub prev_nanos: u32,
    pub consumed: u64,
}

pub const BUFF_OK: c_int = 0;
pub const BUFF_E_INVALID: c_int = 1;
pub const BUFF_E_OOM: c_int = 2;
pub const BUFF_E_DEVICE: c_int = 3;
pub const BUFF_E_REF_PANIC: c_int = 4;
pub const BUFF_E_UNSUPPORTED: c_int = 5;
pub const BUFF_E_SMALL_BUFFER: c_int = 6;
pub const BUFF_FLIP: u32 = 0x1;
pub const BUFF_SEMANTIC: u32 = 0x2;
pub const BUFF_MAJORITY: u32 = 0x4;
pub const BUFF_SLICE: u32 = 0x8;
pub const BUFF_SCALAR: u32 = 0x10;
pub const BUFF_WITH_SLICE: u32 = 0x20;
pub const BUFF_RESIDUE: u32 = 0x40;
pub const BUFF_OP_GT: c_int = 0;
pub const BUFF_OP_EQ: c_int = 1;
pub const BUFF_OP_LT: c_int = 2;
pub const BUFF_Q_GT: u32 = 0;
pub const BUFF_Q_EQ: u32 = 1;
pub const BUFF_Q_SUM: u32 = 2;
pub const BUFF_Q_MAX: u32 = 3;
pub const BUFF_Q_MAT: u32 = 4;

extern "C" {
    pub fn buff_ctx_create(device: c_int, hbm_budget_bytes: u64, out: *mut *mut BuffCtx) -> c_int;
    pub fn buff_ctx_destroy(ctx: *mut BuffCtx) -> c_int;
    pub fn buff_ctx_synchronize(ctx: *mut BuffCtx) -> c_int;
    pub fn buff_ctx_hbm_used(ctx: *mut BuffCtx, bytes: *mut u64) -> c_int;
    pub fn buff_last_error() -> *const c_char;
    pub fn buff_dev_alloc(ctx: *mut BuffCtx, bytes: u64, dptr: *mut *mut c_void) -> c_int;
    pub fn buff_dev_free(ctx: *mut BuffCtx, dptr: *mut c_void) -> c_int;
    pub fn buff_memcpy_h2d(ctx: *mut BuffCtx, dst_dev: *mut c_void, src_host: *const c_void, bytes: u64) -> c_int;
    pub fn buff_memcpy_d2h(ctx: *mut BuffCtx, dst_host: *mut c_void, src_dev: *const c_void, bytes: u64) -> c_int;
    pub fn buff_dev_fill(ctx: *mut BuffCtx, dptr: *mut c_void, value: u8, bytes: u64) -> c_int;
    pub fn buff_dev_touch(ctx: *mut BuffCtx, dptr: *const c_void, bytes: u64) -> c_int;
    pub fn buff_dev_copy(ctx: *mut BuffCtx, dst: *mut c_void, src: *const c_void, bytes: u64) -> c_int;
    pub fn buff_dev_read(ctx: *mut BuffCtx, src: *const c_void, bytes: u64) -> c_int;
    pub fn buff_prof_enable(ctx: *mut BuffCtx, on: c_int) -> c_int;
    pub fn buff_prof_reset(ctx: *mut BuffCtx) -> c_int;
    pub fn buff_prof_query(ctx: *mut BuffCtx, kernel: *const c_char, launches: *mut u64, total_ms: *mut f64) -> c_int;
    pub fn buff_encode_device(ctx: *mut BuffCtx, x_dev: *const f64, n: u64, scale: u64, flags: u32, out: *mut *mut BuffChunk) -> c_int;
    pub fn buff_encode(ctx: *mut BuffCtx, x_host: *const f64, n: u64, scale: u64, flags: u32, out: *mut *mut BuffChunk) -> c_int;
    pub fn buff_chunk_info_get(chunk: *const BuffChunk, info: *mut BuffChunkInfo) -> c_int;
    pub fn buff_chunk_serialize(ctx: *mut BuffCtx, chunk: *const BuffChunk, dst: *mut u8, len: *mut u64) -> c_int;
    pub fn buff_chunk_load(ctx: *mut BuffCtx, stream: *const u8, len: u64, scale: u64, flags: u32, out: *mut *mut BuffChunk) -> c_int;
    pub fn buff_chunk_free(chunk: *mut BuffChunk) -> c_int;
    pub fn buff_chunk_sparse_col(chunk: *const BuffChunk, j: u32, major: *mut u32, outliers: *mut u64, last_row: *mut u64) -> c_int;
    pub fn buff_filter(ctx: *mut BuffCtx, chunk: *const BuffChunk, op: c_int, pred: f64, flags: u32, words: *mut u32, nwords: u64, count: *mut u64) -> c_int;
    pub fn buff_filter_device(ctx: *mut BuffCtx, chunk: *const BuffChunk, op: c_int, pred: f64, flags: u32, words_dev: *mut u64, stats_dev: *mut u64) -> c_int;
    pub fn buff_filter_cond(ctx: *mut BuffCtx, chunk: *const BuffChunk, op: c_int, pred: f64, flags: u32, cond_words: *const u32, words: *mut u32, nwords: u64, count: *mut u64) -> c_int;
    pub fn buff_filter_cond_device(ctx: *mut BuffCtx, chunk: *const BuffChunk, op: c_int, pred: f64, flags: u32, cond_dev: *const u64, words_dev: *mut u64) -> c_int;
    pub fn buff_materialize_cond(ctx: *mut BuffCtx, chunk: *const BuffChunk, flags: u32, cond_words: *const u32, out: *mut f64, cap: u64, rows: *mut u64) -> c_int;
    pub fn buff_materialize_cond_device(ctx: *mut BuffCtx, chunk: *const BuffChunk, flags: u32, cond_dev: *const u64, out_dev: *mut f64, cap: u64, rows: *mut u64) -> c_int;
    pub fn buff_max_range(ctx: *mut BuffCtx, chunk: *const BuffChunk, s: u64, e: u64, window: u64, flags: u32, max_values: *mut f64, cap: u64, nwindows: *mut u64, argmax_words: *mut u32, nwords: u64) -> c_int;
    pub fn buff_max_range_device(ctx: *mut BuffCtx, chunk: *const BuffChunk, s: u64, e: u64, window: u64, flags: u32, max_values_dev: *mut f64, cap: u64, nwindows: *mut u64, argmax_dev: *mut u64) -> c_int;
    pub fn buff_sum(ctx: *mut BuffCtx, chunk: *const BuffChunk, ref_sum: *mut f64, exact_sum: *mut f64, plane_sums: *mut u64) -> c_int;
    pub fn buff_sum_ex(ctx: *mut BuffCtx, chunk: *const BuffChunk, flags: u32, ref_sum: *mut f64, exact_sum: *mut f64, plane_sums: *mut u64) -> c_int;
    pub fn buff_sum_noalloc(ctx: *mut BuffCtx, chunk: *const BuffChunk, ref_sum: *mut f64, exact_sum: *mut f64, plane_sums: *mut u64) -> c_int;
    pub fn buff_sum_precision(ctx: *mut BuffCtx, chunk: *const BuffChunk, precision: i32, flags: u32, ref_sum: *mut f64, exact_sum: *mut f64, sums: *mut u64) -> c_int;
    pub fn buff_max(ctx: *mut BuffCtx, chunk: *const BuffChunk, max_value: *mut f64, max_off: *mut u64) -> c_int;
    pub fn buff_max_ex(ctx: *mut BuffCtx, chunk: *const BuffChunk, flags: u32, max_value: *mut f64, max_off: *mut u64, argmax_count: *mut u64) -> c_int;
    pub fn buff_materialize(ctx: *mut BuffCtx, chunk: *const BuffChunk, precision: i32, flags: u32, out: *mut f64) -> c_int;
    pub fn buff_materialize_device(ctx: *mut BuffCtx, chunk: *const BuffChunk, precision: i32, flags: u32, out_dev: *mut f64) -> c_int;
    pub fn buff_tune_chunk(ctx: *mut BuffCtx, chunk: *const BuffChunk, pred: f64, choices: *mut BuffTuneChoice) -> c_int;
    pub fn buff_batch_encode_device(ctx: *mut BuffCtx, x_dev: *const f64, row_off: *const u64, scales: *const u64, nchunks: u32, flags: u32, out: *mut *mut BuffBatch) -> c_int;
    pub fn buff_batch_size(batch: *const BuffBatch, nchunks: *mut u32) -> c_int;
    pub fn buff_batch_chunk(batch: *mut BuffBatch, i: u32, chunk: *mut *mut BuffChunk) -> c_int;
    pub fn buff_batch_free(batch: *mut BuffBatch) -> c_int;
    pub fn buff_batch_query(ctx: *mut BuffCtx, batch: *mut BuffBatch, qs: *const BuffQuery, nq: u32, results: *mut BuffQueryResult) -> c_int;
    pub fn buff_batch_arena(batch: *mut BuffBatch, bitmaps_dev: *mut *const u64, values_dev: *mut *const f64) -> c_int;
    pub fn buff_batch_tune(ctx: *mut BuffCtx, batch: *mut BuffBatch, qs: *const BuffQuery, nq: u32, grid: *mut u32, tiles_per_item: *mut u32, first_planes: *mut u32) -> c_int;
    pub fn buff_batch_set_tuning(ctx: *mut BuffCtx, grid: u32, tiles_per_item: u32, first_planes: u32) -> c_int;
    pub fn buff_server_create(ctx: *mut BuffCtx, cfg: *const BuffServerCfg, out: *mut *mut BuffServer) -> c_int;
    pub fn buff_server_add_chunk(server: *mut BuffServer, chunk: *mut BuffChunk, id: *mut u32) -> c_int;
    pub fn buff_server_submit(server: *mut BuffServer, q: *const BuffQuery, out_dev: *mut c_void, ticket: *mut u64) -> c_int;
    pub fn buff_server_wait(server: *mut BuffServer, ticket: u64, result: *mut BuffQueryResult, latency_us: *mut f64) -> c_int;
    pub fn buff_server_poll(server: *mut BuffServer, ticket: u64, done: *mut c_int) -> c_int;
    pub fn buff_server_release(server: *mut BuffServer, out: *const c_void) -> c_int;
    pub fn buff_server_arena_stats(server: *mut BuffServer, in_use: *mut u64, peak: *mut u64, capacity: *mut u64) -> c_int;
    pub fn buff_server_run_trace(server: *mut BuffServer, qs: *const BuffQuery, arrival_s: *const f64, outs: *const *mut c_void, n: u32, results: *mut BuffQueryResult, latency_us: *mut f64) -> c_int;
    pub fn buff_trace_per_call(ctx: *mut BuffCtx, chunks: *const *mut BuffChunk, nchunks: u32, qs: *const BuffQuery, arrival_s: *const f64, outs: *const *mut c_void, n: u32, results: *mut BuffQueryResult, latency_us: *mut f64) -> c_int;
    pub fn buff_server_stats(server: *mut BuffServer, batches: *mut u64, queries: *mut u64, largest_batch: *mut u64) -> c_int;
    pub fn buff_server_destroy(server: *mut BuffServer) -> c_int;
    pub fn buff_sd_byte_fixed_encode(sd: *const BuffSplitDouble, data: *const f64, n: u64, out: *mut u8, cap: u64, out_len: *mut u64) -> c_int;
    pub fn buff_sd_buff_simd256_encode(sd: *const BuffSplitDouble, data: *const f64, n: u64, out: *mut u8, cap: u64, out_len: *mut u64) -> c_int;
    pub fn buff_sd_byte_fixed_decode(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, out: *mut f64, cap: u64) -> c_int;
    pub fn buff_sd_buff_simd256_decode(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, out: *mut f64, cap: u64) -> c_int;
    pub fn buff_sd_buff_range_filter(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, pred: f64, words: *mut u32, nwords: u64, count: *mut u64) -> c_int;
    pub fn buff_sd_buff_equal_filter(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, pred: f64, words: *mut u32, nwords: u64, count: *mut u64) -> c_int;
    pub fn buff_sd_buff_simd_range_filter(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, pred: f64, words: *mut u32, nwords: u64, count: *mut u64) -> c_int;
    pub fn buff_sd_buff_simd_equal_filter(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, pred: f64, words: *mut u32, nwords: u64, count: *mut u64) -> c_int;
    pub fn buff_sd_byte_fixed_sum(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, sum: *mut f64) -> c_int;
    pub fn buff_sd_byte_fixed_max(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, max_value: *mut f64) -> c_int;
    pub fn buff_sd_byte_residue_encode(sd: *const BuffSplitDouble, data: *const f64, n: u64, out: *mut u8, cap: u64, out_len: *mut u64) -> c_int;
    pub fn buff_sd_byte_residue_decode(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, out: *mut f64, cap: u64, n_out: *mut u64) -> c_int;
    pub fn buff_sd_byte_residue_decode_with_precision(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, precision: u32, out: *mut f64, cap: u64, n_out: *mut u64) -> c_int;
    pub fn buff_sd_byte_residue_sum(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, sum: *mut f64) -> c_int;
    pub fn buff_sd_byte_residue_sum_with_precision(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, precision: u32, sum: *mut f64) -> c_int;
    pub fn buff_sd_buff_encode_majority(sd: *const BuffSplitDouble, data: *const f64, n: u64, out: *mut u8, cap: u64, out_len: *mut u64) -> c_int;
    pub fn buff_sd_buff_decode_majority(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, out: *mut f64, cap: u64, n_out: *mut u64) -> c_int;
    pub fn buff_sd_buff_range_filter_majority(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, pred: f64, words: *mut u32, nwords: u64, count: *mut u64) -> c_int;
    pub fn buff_sd_buff_equal_filter_majority(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, pred: f64, words: *mut u32, nwords: u64, count: *mut u64) -> c_int;
    pub fn buff_sd_buff_sum_majority(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, sum: *mut f64) -> c_int;
    pub fn buff_sd_buff_max_majority(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, max_value: *mut f64) -> c_int;
    pub fn buff_csv_parse_device(ctx: *mut BuffCtx, text_dev: *const u8, len: u64, delim: u8, skip_val: u32, out_dev: *mut f64, cap: u64, count: *mut u64) -> c_int;
    pub fn buff_csv_parse(ctx: *mut BuffCtx, text: *const u8, len: u64, delim: u8, skip_val: u32, out: *mut f64, cap: u64, count: *mut u64) -> c_int;
    pub fn buff_construct_file_iterator_skip_newline(file: *const c_char, skip_val: u32, delim: u8, out: *mut f64, cap: u64, count: *mut u64) -> c_int;
    pub fn buff_sd_buff_slice_encode(sc: *const BuffSplitDouble, data: *const f64, n: u64, out: *mut u8, cap: u64, out_len: *mut u64) -> c_int;
    pub fn buff_sd_buff_slice_decode(sc: *const BuffSplitDouble, bytes: *const u8, len: u64, out: *mut f64, cap: u64) -> c_int;
    pub fn buff_sd_buff_slice_decode_bs(sc: *const BuffSplitDouble, bytes: *const u8, len: u64, out: *mut f64, cap: u64) -> c_int;
    pub fn buff_sd_buff_slice_range_filter(sc: *const BuffSplitDouble, bytes: *const u8, len: u64, pred: f64, words: *mut u32, nwords: u64, count: *mut u64) -> c_int;
    pub fn buff_sd_buff_slice_equal_filter(sc: *const BuffSplitDouble, bytes: *const u8, len: u64, pred: f64, words: *mut u32, nwords: u64, count: *mut u64) -> c_int;
    pub fn buff_sd_buff_slice_range_filter_nosimd(sc: *const BuffSplitDouble, bytes: *const u8, len: u64, pred: f64, words: *mut u32, nwords: u64, count: *mut u64) -> c_int;
    pub fn buff_sd_buff_slice_equal_filter_nosimd(sc: *const BuffSplitDouble, bytes: *const u8, len: u64, pred: f64, words: *mut u32, nwords: u64, count: *mut u64) -> c_int;
    pub fn buff_sd_buff_simd_range_filter_with_slice(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, pred: f64, words: *mut u32, nwords: u64, count: *mut u64) -> c_int;
    pub fn buff_sd_buff_simd_equal_filter_with_slice(sd: *const BuffSplitDouble, bytes: *const u8, len: u64, pred: f64, words: *mut u32, nwords: u64, count: *mut u64) -> c_int;
    pub fn buff_sd_buff_slice_buffsum(sc: *const BuffSplitDouble, bytes: *const u8, len: u64, sum: *mut f64) -> c_int;
    pub fn buff_sd_buff_slice_sum(sc: *const BuffSplitDouble, bytes: *const u8, len: u64, sum: *mut f64) -> c_int;
    pub fn buff_sd_buff_slice_buffmax(sc: *const BuffSplitDouble, bytes: *const u8, len: u64, max_value: *mut f64, argmax_count: *mut u64) -> c_int;
    pub fn buff_sd_buff_slice_max(sc: *const BuffSplitDouble, bytes: *const u8, len: u64, max_value: *mut f64, argmax_count: *mut u64) -> c_int;
    pub fn buff_segment_byte_size(seg: *const BuffSegment, size: *mut u64) -> c_int;
    pub fn buff_segment_serialize(seg: *const BuffSegment, dst: *mut u8, cap: u64, len: *mut u64) -> c_int;
    pub fn buff_segment_parse(bytes: *const u8, len: u64, elem_bytes: u32, seg: *mut BuffSegment) -> c_int;
    pub fn buff_segment_encode(ctx: *mut BuffCtx, bytes: *const u8, len: u64, elem_bytes: u32, scale: u64, flags: u32, out: *mut *mut BuffChunk) -> c_int;
}
