//! libbuffgpu from Rust. `SplitBDDoubleCompress` and `BuffSliceCompress` keep
//! the reference's struct fields and method names
//! (database/src/compress/split_double.rs:29-39, buff_simd.rs,
//! buff_slice.rs:222-232); every method runs its hand-written HIP kernels on
//! the GPU of a process-wide default context through the C ABI in [`ffi`].
//! Byte streams go in by value as in the reference (`Vec<u8>`). Where the
//! reference only prints (a filter's count, a max), the method returns it;
//! filters also return the result bitmap in `BitVec<u32>` storage order
//! (bit-vec's `BitVec::from_vec(words, n)` rebuilds the reference's type).
//! Inputs on which the reference panics panic here too, with the library's
//! message; `try_*` forms are not provided because the reference has none.
//!
//! The resident API (chunks kept in HBM, batched queries, the server) is the
//! plain C ABI in [`ffi`].

pub mod ffi;

use std::ffi::CStr;
use std::fmt;
use std::os::raw::c_int;
use std::ptr;

/// A failing call's status (`BUFF_E_*`) and the library's message for it.
#[derive(Clone, Debug)]
pub struct BuffError {
    pub code: c_int,
    pub message: String,
}

impl fmt::Display for BuffError {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        write!(f, "buffgpu status {}: {}", self.code, self.message)
    }
}

impl std::error::Error for BuffError {}

/// Ok for BUFF_OK, else the status with buff_last_error()'s message.
pub fn check(rc: c_int) -> Result<(), BuffError> {
    if rc == ffi::BUFF_OK {
        return Ok(());
    }
    let message = unsafe {
        let p = ffi::buff_last_error();
        if p.is_null() {
            String::new()
        } else {
            CStr::from_ptr(p).to_string_lossy().into_owned()
        }
    };
    Err(BuffError { code: rc, message })
}

fn ok(rc: c_int) {
    if let Err(e) = check(rc) {
        panic!("{}", e); // the reference panics on these inputs too
    }
}

/// Rows of a stream: the u32 after the i64 base in the header (split_double.rs:349-354).
fn stream_rows(bytes: &[u8]) -> usize {
    if bytes.len() < 12 {
        return 0;
    }
    u32::from_le_bytes([bytes[8], bytes[9], bytes[10], bytes[11]]) as usize
}

/// A filter's rows: `BitVec<u32>` storage words (row i is bit i % 32 of word i / 32)
/// and the count the reference prints.
#[derive(Clone, Debug, Default)]
pub struct FilterResult {
    pub words: Vec<u32>,
    pub count: u64,
}

type EncodeFn = unsafe extern "C" fn(*const ffi::BuffSplitDouble, *const f64, u64, *mut u8, u64, *mut u64) -> c_int;
type DecodeFn = unsafe extern "C" fn(*const ffi::BuffSplitDouble, *const u8, u64, *mut f64, u64) -> c_int;
type DecodeNFn =
    unsafe extern "C" fn(*const ffi::BuffSplitDouble, *const u8, u64, *mut f64, u64, *mut u64) -> c_int;
type FilterFn =
    unsafe extern "C" fn(*const ffi::BuffSplitDouble, *const u8, u64, f64, *mut u32, u64, *mut u64) -> c_int;
type ScalarFn = unsafe extern "C" fn(*const ffi::BuffSplitDouble, *const u8, u64, *mut f64) -> c_int;
type ArgmaxFn = unsafe extern "C" fn(*const ffi::BuffSplitDouble, *const u8, u64, *mut f64, *mut u64) -> c_int;

fn encode_with(sd: &ffi::BuffSplitDouble, data: &[f64], f: EncodeFn) -> Vec<u8> {
    let n = data.len() as u64;
    let mut len = 0u64;
    ok(unsafe { f(sd, data.as_ptr(), n, ptr::null_mut(), 0, &mut len) }); // size query
    let mut out = vec![0u8; len as usize];
    ok(unsafe { f(sd, data.as_ptr(), n, out.as_mut_ptr(), out.len() as u64, &mut len) });
    out.truncate(len as usize);
    out
}

fn decode_with(sd: &ffi::BuffSplitDouble, bytes: &[u8], f: DecodeFn) -> Vec<f64> {
    let n = stream_rows(bytes);
    let mut out = vec![0f64; n.max(1)];
    ok(unsafe { f(sd, bytes.as_ptr(), bytes.len() as u64, out.as_mut_ptr(), out.len() as u64) });
    out.truncate(n);
    out
}

fn decode_n_with(sd: &ffi::BuffSplitDouble, bytes: &[u8], f: DecodeNFn) -> Vec<f64> {
    let mut out = vec![0f64; stream_rows(bytes).max(1)];
    let mut got = 0u64;
    ok(unsafe { f(sd, bytes.as_ptr(), bytes.len() as u64, out.as_mut_ptr(), out.len() as u64, &mut got) });
    out.truncate(got as usize);
    out
}

fn filter_with(sd: &ffi::BuffSplitDouble, bytes: &[u8], pred: f64, f: FilterFn) -> FilterResult {
    let mut r = FilterResult { words: vec![0u32; (stream_rows(bytes) + 31) / 32], count: 0 };
    ok(unsafe {
        f(sd, bytes.as_ptr(), bytes.len() as u64, pred, r.words.as_mut_ptr(), r.words.len() as u64, &mut r.count)
    });
    r
}

fn scalar_with(sd: &ffi::BuffSplitDouble, bytes: &[u8], f: ScalarFn) -> f64 {
    let mut v = 0f64;
    ok(unsafe { f(sd, bytes.as_ptr(), bytes.len() as u64, &mut v) });
    v
}

fn argmax_with(sd: &ffi::BuffSplitDouble, bytes: &[u8], f: ArgmaxFn) -> (f64, u64) {
    let (mut v, mut count) = (0f64, 0u64);
    ok(unsafe { f(sd, bytes.as_ptr(), bytes.len() as u64, &mut v, &mut count) });
    (v, count)
}

/// split_double.rs:29-39: the compressor's configuration; `scale` is 10^decimals.
#[derive(Clone, Copy, Debug)]
pub struct SplitBDDoubleCompress {
    chunksize: usize,
    batchsize: usize,
    scale: usize,
}

impl SplitBDDoubleCompress {
    pub fn new(chunksize: usize, batchsize: usize, scale: usize) -> Self {
        SplitBDDoubleCompress { chunksize, batchsize, scale }
    }

    fn sd(&self) -> ffi::BuffSplitDouble {
        ffi::BuffSplitDouble { chunksize: self.chunksize as u64, batchsize: self.batchsize as u64, scale: self.scale as u64 }
    }

    /// split_double.rs:298
    pub fn byte_fixed_encode(&self, data: &[f64]) -> Vec<u8> {
        encode_with(&self.sd(), data, ffi::buff_sd_byte_fixed_encode)
    }
    /// buff_simd.rs:3116
    pub fn buff_simd256_encode(&self, data: &[f64]) -> Vec<u8> {
        encode_with(&self.sd(), data, ffi::buff_sd_buff_simd256_encode)
    }
    /// split_double.rs:1243
    pub fn byte_fixed_decode(&self, bytes: Vec<u8>) -> Vec<f64> {
        decode_with(&self.sd(), &bytes, ffi::buff_sd_byte_fixed_decode)
    }
    /// buff_simd.rs:3252
    pub fn buff_simd256_decode(&self, bytes: Vec<u8>) -> Vec<f64> {
        decode_with(&self.sd(), &bytes, ffi::buff_sd_buff_simd256_decode)
    }
    /// buff_simd.rs:4507 (the reference prints the count)
    pub fn buff_range_filter(&self, bytes: Vec<u8>, pred: f64) -> FilterResult {
        filter_with(&self.sd(), &bytes, pred, ffi::buff_sd_buff_range_filter)
    }
    /// buff_simd.rs:4925
    pub fn buff_equal_filter(&self, bytes: Vec<u8>, pred: f64) -> FilterResult {
        filter_with(&self.sd(), &bytes, pred, ffi::buff_sd_buff_equal_filter)
    }
    /// buff_simd.rs:3361
    pub fn buff_simd_range_filter(&self, bytes: Vec<u8>, pred: f64) -> FilterResult {
        filter_with(&self.sd(), &bytes, pred, ffi::buff_sd_buff_simd_range_filter)
    }
    /// buff_simd.rs:3549
    pub fn buff_simd_equal_filter(&self, bytes: Vec<u8>, pred: f64) -> FilterResult {
        filter_with(&self.sd(), &bytes, pred, ffi::buff_sd_buff_simd_equal_filter)
    }
    /// buff_simd.rs:3736, over a buff_simd256 stream
    pub fn buff_simd_range_filter_with_slice(&self, bytes: Vec<u8>, pred: f64) -> FilterResult {
        filter_with(&self.sd(), &bytes, pred, ffi::buff_sd_buff_simd_range_filter_with_slice)
    }
    /// buff_simd.rs:3888, over a buff_simd256 stream
    pub fn buff_simd_equal_filter_with_slice(&self, bytes: Vec<u8>, pred: f64) -> FilterResult {
        filter_with(&self.sd(), &bytes, pred, ffi::buff_sd_buff_simd_equal_filter_with_slice)
    }
    /// split_double.rs:2127
    pub fn byte_fixed_sum(&self, bytes: Vec<u8>) -> f64 {
        scalar_with(&self.sd(), &bytes, ffi::buff_sd_byte_fixed_sum)
    }
    /// split_double.rs:3919 (the reference prints the max)
    pub fn byte_fixed_max(&self, bytes: Vec<u8>) -> f64 {
        scalar_with(&self.sd(), &bytes, ffi::buff_sd_byte_fixed_max)
    }

    /// split_double.rs:624: the RAPG residue stream when the integer part
    /// needs more than 15 bits, else the byte_fixed stream.
    pub fn byte_residue_encode(&self, data: &[f64]) -> Vec<u8> {
        encode_with(&self.sd(), data, ffi::buff_sd_byte_residue_encode)
    }
    /// split_double.rs:1353
    pub fn byte_residue_decode(&self, bytes: Vec<u8>) -> Vec<f64> {
        decode_n_with(&self.sd(), &bytes, ffi::buff_sd_byte_residue_decode)
    }
    /// split_double.rs:1612 (empty when the outlier columns take every bit read)
    pub fn byte_residue_decode_with_precision(&self, bytes: Vec<u8>, precision: u32) -> Vec<f64> {
        let sd = self.sd();
        let mut out = vec![0f64; stream_rows(&bytes).max(1)];
        let mut got = 0u64;
        ok(unsafe {
            ffi::buff_sd_byte_residue_decode_with_precision(&sd, bytes.as_ptr(), bytes.len() as u64, precision,
                                                            out.as_mut_ptr(), out.len() as u64, &mut got)
        });
        out.truncate(got as usize);
        out
    }
    /// split_double.rs:2210
    pub fn byte_residue_sum(&self, bytes: Vec<u8>) -> f64 {
        scalar_with(&self.sd(), &bytes, ffi::buff_sd_byte_residue_sum)
    }
    /// split_double.rs:2409
    pub fn byte_residue_sum_with_precision(&self, bytes: Vec<u8>, precision: u32) -> f64 {
        let sd = self.sd();
        let mut v = 0f64;
        ok(unsafe {
            ffi::buff_sd_byte_residue_sum_with_precision(&sd, bytes.as_ptr(), bytes.len() as u64, precision, &mut v)
        });
        v
    }

    /// buff_simd.rs:317
    pub fn buff_encode_majority(&self, data: &[f64]) -> Vec<u8> {
        encode_with(&self.sd(), data, ffi::buff_sd_buff_encode_majority)
    }
    /// buff_simd.rs:633 (empty when the sparse columns end the stream)
    pub fn buff_decode_majority(&self, bytes: Vec<u8>) -> Vec<f64> {
        decode_n_with(&self.sd(), &bytes, ffi::buff_sd_buff_decode_majority)
    }
    /// buff_simd.rs:863
    pub fn buff_range_filter_majority(&self, bytes: Vec<u8>, pred: f64) -> FilterResult {
        filter_with(&self.sd(), &bytes, pred, ffi::buff_sd_buff_range_filter_majority)
    }
    /// buff_simd.rs:1916
    pub fn buff_equal_filter_majority(&self, bytes: Vec<u8>, pred: f64) -> FilterResult {
        filter_with(&self.sd(), &bytes, pred, ffi::buff_sd_buff_equal_filter_majority)
    }
    /// buff_simd.rs:2294
    pub fn buff_sum_majority(&self, bytes: Vec<u8>) -> f64 {
        scalar_with(&self.sd(), &bytes, ffi::buff_sd_buff_sum_majority)
    }
    /// buff_simd.rs:2444 (the reference prints the max)
    pub fn buff_max_majority(&self, bytes: Vec<u8>) -> f64 {
        scalar_with(&self.sd(), &bytes, ffi::buff_sd_buff_max_majority)
    }
}

/// buff_slice.rs:222-232: the byte-slice compressor (SplitBDDoubleCompress's fields).
#[derive(Clone, Copy, Debug)]
pub struct BuffSliceCompress {
    chunksize: usize,
    batchsize: usize,
    scale: usize,
}

impl BuffSliceCompress {
    pub fn new(chunksize: usize, batchsize: usize, scale: usize) -> Self {
        BuffSliceCompress { chunksize, batchsize, scale }
    }

    fn sc(&self) -> ffi::BuffSplitDouble {
        ffi::BuffSplitDouble { chunksize: self.chunksize as u64, batchsize: self.batchsize as u64, scale: self.scale as u64 }
    }

    /// buff_slice.rs:298
    pub fn buff_slice_encode(&self, data: &[f64]) -> Vec<u8> {
        encode_with(&self.sc(), data, ffi::buff_sd_buff_slice_encode)
    }
    /// buff_slice.rs:439
    pub fn buff_slice_decode(&self, bytes: Vec<u8>) -> Vec<f64> {
        decode_with(&self.sc(), &bytes, ffi::buff_sd_buff_slice_decode)
    }
    /// buff_slice.rs:1184
    pub fn buff_slice_decode_bs(&self, bytes: Vec<u8>) -> Vec<f64> {
        decode_with(&self.sc(), &bytes, ffi::buff_sd_buff_slice_decode_bs)
    }
    /// buff_slice.rs:1348
    pub fn buff_slice_range_filter(&self, bytes: Vec<u8>, pred: f64) -> FilterResult {
        filter_with(&self.sc(), &bytes, pred, ffi::buff_sd_buff_slice_range_filter)
    }
    /// buff_slice.rs:1665
    pub fn buff_slice_equal_filter(&self, bytes: Vec<u8>, pred: f64) -> FilterResult {
        filter_with(&self.sc(), &bytes, pred, ffi::buff_sd_buff_slice_equal_filter)
    }
    /// buff_slice.rs:1753
    pub fn buff_slice_range_filter_nosimd(&self, bytes: Vec<u8>, pred: f64) -> FilterResult {
        filter_with(&self.sc(), &bytes, pred, ffi::buff_sd_buff_slice_range_filter_nosimd)
    }
    /// buff_slice.rs:1873
    pub fn buff_slice_equal_filter_nosimd(&self, bytes: Vec<u8>, pred: f64) -> FilterResult {
        filter_with(&self.sc(), &bytes, pred, ffi::buff_sd_buff_slice_equal_filter_nosimd)
    }
    /// buff_slice.rs:702 (the exact sum, rounded once)
    pub fn buff_slice_buffsum(&self, bytes: Vec<u8>) -> f64 {
        scalar_with(&self.sc(), &bytes, ffi::buff_sd_buff_slice_buffsum)
    }
    /// buff_slice.rs:1101
    pub fn buff_slice_sum(&self, bytes: Vec<u8>) -> f64 {
        scalar_with(&self.sc(), &bytes, ffi::buff_sd_buff_slice_sum)
    }
    /// buff_slice.rs:548: the max and the number of rows holding it (both printed there)
    pub fn buff_slice_buffmax(&self, bytes: Vec<u8>) -> (f64, u64) {
        argmax_with(&self.sc(), &bytes, ffi::buff_sd_buff_slice_buffmax)
    }
    /// buff_slice.rs:816
    pub fn buff_slice_max(&self, bytes: Vec<u8>) -> (f64, u64) {
        argmax_with(&self.sc(), &bytes, ffi::buff_sd_buff_slice_max)
    }
}
