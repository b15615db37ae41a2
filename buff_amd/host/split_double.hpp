// C++ host mirror of the reference's `SplitBDDoubleCompress` operator API over
// the buffgpu C ABI (include/buffgpu.h).
//
// The reference is a Rust crate (database/src/compress/split_double.rs:29-39 for
// the struct, the methods in split_double.rs and buff_simd.rs). This header
// keeps its method names, argument meaning and by-value byte streams, so a C++
// caller (or the comp_profiler CLI next to this file) reads like the
// reference's drivers. Every method runs on the GPU; a status that the reference
// would have turned into a panic is thrown as buff::Panic, other failures as
// buff::Error.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "buffgpu.h"

namespace buff {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
struct Panic : Error {
  using Error::Error;
};

inline void check(int rc) {
  if (rc == BUFF_OK) return;
  const std::string msg = std::string("buffgpu status ") + std::to_string(rc) + ": " + buff_last_error();
  if (rc == BUFF_E_REF_PANIC) throw Panic(rc, msg);
  throw Error(rc, msg);
}

// BitVec<u32> words (bit-vec/src/lib.rs layout) plus the count the reference prints.
struct Bitmap {
  std::vector<uint32_t> words;
  uint64_t count = 0;
};

// database/src/segment.rs:35-43; only `data` matters to the codec. Its bincode
// size (segment.rs:90-95) is the reference's throughput denominator:
// Option tag (1) + SystemTime (12) + u64 signal (8) + Vec length (8) + 8n data
// + two Option tags (2).
struct Segment {
  std::vector<double> data;
  uint64_t byte_size() const { return 31 + 8 * (uint64_t)data.size(); }
};

class SplitBDDoubleCompress {
 public:
  // split_double.rs:37
  SplitBDDoubleCompress(uint64_t chunksize, uint64_t batchsize, uint64_t scale) : sd_{chunksize, batchsize, scale} {}

  uint64_t scale() const { return sd_.scale; }

  // split_double.rs:298 byte_fixed_encode
  std::vector<uint8_t> byte_fixed_encode(const Segment& seg) const { return encode(seg, false); }
  // buff_simd.rs:3116 buff_simd256_encode
  std::vector<uint8_t> buff_simd256_encode(const Segment& seg) const { return encode(seg, true); }

  // split_double.rs:1243 byte_fixed_decode
  std::vector<double> byte_fixed_decode(std::vector<uint8_t> bytes) const {
    std::vector<double> out(rows(bytes));
    check(buff_sd_byte_fixed_decode(&sd_, bytes.data(), bytes.size(), out.data(), out.size()));
    return out;
  }
  // buff_simd.rs:3252 buff_simd256_decode
  std::vector<double> buff_simd256_decode(std::vector<uint8_t> bytes) const {
    std::vector<double> out(rows(bytes));
    check(buff_sd_buff_simd256_decode(&sd_, bytes.data(), bytes.size(), out.data(), out.size()));
    return out;
  }
  // split_double.rs:624 byte_residue_encode (the RAPG residue stream)
  std::vector<uint8_t> byte_residue_encode(const Segment& seg) const {
    uint64_t len = 0;
    check(buff_sd_byte_residue_encode(&sd_, seg.data.data(), seg.data.size(), nullptr, 0, &len));
    std::vector<uint8_t> out(len);
    check(buff_sd_byte_residue_encode(&sd_, seg.data.data(), seg.data.size(), out.data(), out.size(), &len));
    return out;
  }
  // split_double.rs:1353 byte_residue_decode
  std::vector<double> byte_residue_decode(std::vector<uint8_t> bytes) const {
    std::vector<double> out(rows(bytes));
    uint64_t got = 0;
    check(buff_sd_byte_residue_decode(&sd_, bytes.data(), bytes.size(), out.data(), out.size(), &got));
    out.resize(got);
    return out;
  }
  // split_double.rs:1612 byte_residue_decode_with_precision
  std::vector<double> byte_residue_decode_with_precision(std::vector<uint8_t> bytes, uint32_t precision) const {
    std::vector<double> out(rows(bytes));
    uint64_t got = 0;
    check(buff_sd_byte_residue_decode_with_precision(&sd_, bytes.data(), bytes.size(), precision, out.data(),
                                                     out.size(), &got));
    out.resize(got);
    return out;
  }

  // buff_simd.rs:4507 buff_range_filter: prints the count like the reference
  void buff_range_filter(std::vector<uint8_t> bytes, double pred) const {
    std::printf("Number of qualified items:%llu\n", (unsigned long long)range_bitmap(std::move(bytes), pred).count);
  }
  // buff_simd.rs:4925 buff_equal_filter
  void buff_equal_filter(std::vector<uint8_t> bytes, double pred) const {
    std::printf("Number of qualified int items:%llu\n", (unsigned long long)equal_bitmap(std::move(bytes), pred).count);
  }
  // buff_simd.rs:3361 / 3549, flipped stream
  void buff_simd_range_filter(std::vector<uint8_t> bytes, double pred) const {
    std::printf("Number of qualified items:%llu\n",
                (unsigned long long)filter(bytes, pred, buff_sd_buff_simd_range_filter).count);
  }
  void buff_simd_equal_filter(std::vector<uint8_t> bytes, double pred) const {
    std::printf("Number of qualified int items:%llu\n",
                (unsigned long long)filter(bytes, pred, buff_sd_buff_simd_equal_filter).count);
  }
  // Bitmap-returning siblings of the printing filters.
  Bitmap range_bitmap(std::vector<uint8_t> bytes, double pred) const {
    return filter(bytes, pred, buff_sd_buff_range_filter);
  }
  Bitmap equal_bitmap(std::vector<uint8_t> bytes, double pred) const {
    return filter(bytes, pred, buff_sd_buff_equal_filter);
  }

  // split_double.rs:2127 byte_fixed_sum
  double byte_fixed_sum(std::vector<uint8_t> bytes) const {
    double s = 0;
    check(buff_sd_byte_fixed_sum(&sd_, bytes.data(), bytes.size(), &s));
    return s;
  }
  // split_double.rs:2210 byte_residue_sum
  double byte_residue_sum(std::vector<uint8_t> bytes) const {
    double s = 0;
    check(buff_sd_byte_residue_sum(&sd_, bytes.data(), bytes.size(), &s));
    return s;
  }
  // split_double.rs:2409 byte_residue_sum_with_precision
  double byte_residue_sum_with_precision(std::vector<uint8_t> bytes, uint32_t precision) const {
    double s = 0;
    check(buff_sd_byte_residue_sum_with_precision(&sd_, bytes.data(), bytes.size(), precision, &s));
    return s;
  }
  // split_double.rs:3919 byte_fixed_max: prints the max like the reference
  void byte_fixed_max(std::vector<uint8_t> bytes) const { std::printf("Max value:%.17g\n", max_value(std::move(bytes))); }
  double max_value(std::vector<uint8_t> bytes) const {
    double m = 0;
    check(buff_sd_byte_fixed_max(&sd_, bytes.data(), bytes.size(), &m));
    return m;
  }

  // buff_simd.rs:3736 / 3888: the "BS way" filters over a flipped stream
  void buff_simd_range_filter_with_slice(std::vector<uint8_t> bytes, double pred) const {
    std::printf("Number of qualified int items:%llu\n",
                (unsigned long long)filter(bytes, pred, buff_sd_buff_simd_range_filter_with_slice).count);
  }
  void buff_simd_equal_filter_with_slice(std::vector<uint8_t> bytes, double pred) const {
    std::printf("Number of qualified int items:%llu\n",
                (unsigned long long)filter(bytes, pred, buff_sd_buff_simd_equal_filter_with_slice).count);
  }

  // buff-major (buff_simd.rs:317-2861)
  std::vector<uint8_t> buff_encode_majority(const Segment& seg) const {
    uint64_t len = 0;
    check(buff_sd_buff_encode_majority(&sd_, seg.data.data(), seg.data.size(), nullptr, 0, &len));
    std::vector<uint8_t> out(len);
    check(buff_sd_buff_encode_majority(&sd_, seg.data.data(), seg.data.size(), out.data(), out.size(), &len));
    out.resize(len);
    return out;
  }
  std::vector<double> buff_decode_majority(std::vector<uint8_t> bytes) const {
    std::vector<double> out(rows(bytes));
    uint64_t n = 0;
    check(buff_sd_buff_decode_majority(&sd_, bytes.data(), bytes.size(), out.data(), out.size(), &n));
    out.resize(n);
    return out;
  }
  void buff_range_filter_majority(std::vector<uint8_t> bytes, double pred) const {
    std::printf("Number of qualified items:%llu\n",
                (unsigned long long)filter(bytes, pred, buff_sd_buff_range_filter_majority).count);
  }
  void buff_equal_filter_majority(std::vector<uint8_t> bytes, double pred) const {
    std::printf("Number of qualified items:%llu\n",
                (unsigned long long)filter(bytes, pred, buff_sd_buff_equal_filter_majority).count);
  }
  double buff_sum_majority(std::vector<uint8_t> bytes) const {
    double s = 0;
    check(buff_sd_buff_sum_majority(&sd_, bytes.data(), bytes.size(), &s));
    return s;
  }
  void buff_max_majority(std::vector<uint8_t> bytes) const {
    double m = 0;
    check(buff_sd_buff_max_majority(&sd_, bytes.data(), bytes.size(), &m));
    std::printf("Max value:%.17g\n", m);
  }

 protected:
  buff_split_double sd_;

  static uint64_t rows(const std::vector<uint8_t>& b) {  // the u32 row count of the header
    if (b.size() < 20) throw Error(BUFF_E_INVALID, "stream shorter than its 20-byte header");
    return (uint64_t)b[8] | ((uint64_t)b[9] << 8) | ((uint64_t)b[10] << 16) | ((uint64_t)b[11] << 24);
  }

  std::vector<uint8_t> encode(const Segment& seg, bool flip) const {
    auto fn = flip ? buff_sd_buff_simd256_encode : buff_sd_byte_fixed_encode;
    // a stream never exceeds its header plus 8 bytes per row (F <= 63)
    std::vector<uint8_t> out(20 + 8 * seg.data.size());
    uint64_t len = 0;
    check(fn(&sd_, seg.data.data(), seg.data.size(), out.data(), out.size(), &len));
    out.resize(len);
    return out;
  }

  template <typename Fn>
  Bitmap filter(const std::vector<uint8_t>& bytes, double pred, Fn fn) const {
    Bitmap bm;
    bm.words.assign((rows(bytes) + 31) / 32, 0u);
    check(fn(&sd_, bytes.data(), bytes.size(), pred, bm.words.data(), bm.words.size(), &bm.count));
    return bm;
  }
};

// database/src/compress/buff_slice.rs:222-1970 (BuffSliceCompress: the same
// three fields). Printing methods print what the reference prints.
class BuffSliceCompress : public SplitBDDoubleCompress {
 public:
  using SplitBDDoubleCompress::SplitBDDoubleCompress;

  std::vector<uint8_t> buff_slice_encode(const Segment& seg) const {  // :298
    uint64_t len = 0;
    check(buff_sd_buff_slice_encode(&sd_, seg.data.data(), seg.data.size(), nullptr, 0, &len));
    std::vector<uint8_t> out(len);
    check(buff_sd_buff_slice_encode(&sd_, seg.data.data(), seg.data.size(), out.data(), out.size(), &len));
    out.resize(len);
    return out;
  }
  std::vector<double> buff_slice_decode(std::vector<uint8_t> bytes) const {  // :439
    std::vector<double> out(rows(bytes));
    check(buff_sd_buff_slice_decode(&sd_, bytes.data(), bytes.size(), out.data(), out.size()));
    return out;
  }
  std::vector<double> buff_slice_decode_bs(std::vector<uint8_t> bytes) const {  // :1184
    std::vector<double> out(rows(bytes));
    check(buff_sd_buff_slice_decode_bs(&sd_, bytes.data(), bytes.size(), out.data(), out.size()));
    return out;
  }
  void buff_slice_range_filter(std::vector<uint8_t> bytes, double pred) const {  // :1348
    print_count(filter(bytes, pred, buff_sd_buff_slice_range_filter));
  }
  void buff_slice_equal_filter(std::vector<uint8_t> bytes, double pred) const {  // :1665
    print_count(filter(bytes, pred, buff_sd_buff_slice_equal_filter));
  }
  void buff_slice_range_filter_nosimd(std::vector<uint8_t> bytes, double pred) const {  // :1753
    print_count(filter(bytes, pred, buff_sd_buff_slice_range_filter_nosimd));
  }
  void buff_slice_equal_filter_nosimd(std::vector<uint8_t> bytes, double pred) const {  // :1873
    print_count(filter(bytes, pred, buff_sd_buff_slice_equal_filter_nosimd));
  }
  void buff_slice_buffsum(std::vector<uint8_t> bytes) const {  // :702
    double s = 0;
    check(buff_sd_buff_slice_buffsum(&sd_, bytes.data(), bytes.size(), &s));
    std::printf("sum is:%.17g\n", s);
  }
  void buff_slice_sum(std::vector<uint8_t> bytes) const {  // :1101
    double s = 0;
    check(buff_sd_buff_slice_sum(&sd_, bytes.data(), bytes.size(), &s));
    std::printf("sum is:%.17g\n", s);
  }
  void buff_slice_buffmax(std::vector<uint8_t> bytes) const {  // :548
    max_line(bytes, buff_sd_buff_slice_buffmax);
  }
  void buff_slice_max(std::vector<uint8_t> bytes) const {  // :816
    max_line(bytes, buff_sd_buff_slice_max);
  }

 private:
  static void print_count(const Bitmap& bm) {
    std::printf("Number of qualified items:%llu\n", (unsigned long long)bm.count);
  }
  template <typename Fn>
  void max_line(const std::vector<uint8_t>& bytes, Fn fn) const {
    double m = 0;
    uint64_t cnt = 0;
    check(fn(&sd_, bytes.data(), bytes.size(), &m, &cnt));
    std::printf("Number of qualified max items:%llu\nMax value:%.17g\n", (unsigned long long)cnt, m);
  }
};

// database/src/client.rs:163-182, construct_file_iterator_skip_newline::<f64>,
// parsed on the GPU; the reference's Err(()) (a file it cannot open) is thrown.
inline std::vector<double> construct_file_iterator_skip_newline(const char* file, uint32_t skip_val, char delim) {
  std::FILE* f = std::fopen(file, "rb");
  if (!f) throw Error(BUFF_E_INVALID, std::string("cannot open ") + file);
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fclose(f);
  std::vector<double> out((size > 0 ? (uint64_t)size / 2 : 0) + 1);
  uint64_t n = 0;
  check(buff_construct_file_iterator_skip_newline(file, skip_val, (uint8_t)delim, out.data(), out.size(), &n));
  out.resize(n);
  return out;
}

}  // namespace buff
