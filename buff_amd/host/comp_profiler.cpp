// comp_profiler: the reference's benchmark CLI (database/src/compress/main.rs:7-36)
// on the GPU engine.
//
//   comp_profiler <input_file> <method> <int_scale> <pred>
//
// The input is read by construct_file_iterator_skip_newline::<f64>(file, 0, ',')
// (client.rs:163-182), parsed on the GPU. Each method times the six operations
// of its reference driver and ends with the reference's `Performance:` line:
// compression ratio and MiB/s of the bincode Segment size per operation.
//   buff              run_buff_encoding_decoding_mybitvec (buff_simd.rs:167-243):
//                     encode, decode, range filter, equality filter, sum, max;
//   buff-simd         run_buff_simd_encoding_decoding (buff_simd.rs:28-95): the
//                     flipped layout's encode, decode and filters, then the
//                     "BS way" _with_slice filters;
//   buff-slice        run_buff_slice_encoding_decoding (buff_slice.rs:30-101):
//                     slice encode, decode, range and equality filters, buffsum,
//                     buffmax;
//   buff-slice-scalar run_buff_slice_scalar_encoding_decoding (buff_slice.rs:
//                     103-171): encode, decode_bs, the nosimd filters, sum, max;
//   buff-major        run_buff_majority_encoding_decoding (buff_simd.rs:246-311):
//                     majority encode, decode, filters, sum, max.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <vector>

#include "split_double.hpp"

namespace {

double timed_ns(const std::function<void()>& fn) {
  const auto t0 = std::chrono::steady_clock::now();
  fn();
  const auto t1 = std::chrono::steady_clock::now();
  return (double)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s <input_file> <buff|buff-simd|buff-slice|buff-slice-scalar|buff-major> "
                         "<int_scale> <pred>\n", argv[0]);
    return 2;
  }
  const char* file = argv[1];
  const std::string method = argv[2];
  const uint64_t scl = std::strtoull(argv[3], nullptr, 10);
  const double pred = std::strtod(argv[4], nullptr);
  std::printf("ARGS: %s, %s, %llu, %s, \n", file, method.c_str(), (unsigned long long)scl, argv[4]);
  try {
    buff::Segment seg{buff::construct_file_iterator_skip_newline(file, 0, ',')};
    const double org_size = (double)seg.byte_size();
    buff::BuffSliceCompress comp(10, 10, scl);
    std::vector<uint8_t> c;
    double d[6];
    if (method == "buff") {
      d[0] = timed_ns([&] { c = comp.byte_fixed_encode(seg); });
      d[1] = timed_ns([&] { comp.byte_fixed_decode(c); });
      d[2] = timed_ns([&] { comp.buff_range_filter(c, pred); });
      d[3] = timed_ns([&] { comp.buff_equal_filter(c, pred); });
      d[4] = timed_ns([&] { std::printf("sum is: %.17g\n", comp.byte_fixed_sum(c)); });
      d[5] = timed_ns([&] { comp.byte_fixed_max(c); });
    } else if (method == "buff-simd") {
      d[0] = timed_ns([&] { c = comp.buff_simd256_encode(seg); });
      d[1] = timed_ns([&] { comp.buff_simd256_decode(c); });
      d[2] = timed_ns([&] { comp.buff_simd_range_filter(c, pred); });
      d[3] = timed_ns([&] { comp.buff_simd_equal_filter(c, pred); });
      d[4] = timed_ns([&] { comp.buff_simd_range_filter_with_slice(c, pred); });
      d[5] = timed_ns([&] { comp.buff_simd_equal_filter_with_slice(c, pred); });
    } else if (method == "buff-slice") {
      d[0] = timed_ns([&] { c = comp.buff_slice_encode(seg); });
      d[1] = timed_ns([&] { comp.buff_slice_decode(c); });
      d[2] = timed_ns([&] { comp.buff_slice_range_filter(c, pred); });
      d[3] = timed_ns([&] { comp.buff_slice_equal_filter(c, pred); });
      d[4] = timed_ns([&] { comp.buff_slice_buffsum(c); });
      d[5] = timed_ns([&] { comp.buff_slice_buffmax(c); });
    } else if (method == "buff-slice-scalar") {
      d[0] = timed_ns([&] { c = comp.buff_slice_encode(seg); });
      d[1] = timed_ns([&] { comp.buff_slice_decode_bs(c); });
      d[2] = timed_ns([&] { comp.buff_slice_range_filter_nosimd(c, pred); });
      d[3] = timed_ns([&] { comp.buff_slice_equal_filter_nosimd(c, pred); });
      d[4] = timed_ns([&] { comp.buff_slice_sum(c); });
      d[5] = timed_ns([&] { comp.buff_slice_max(c); });
    } else if (method == "buff-major") {
      d[0] = timed_ns([&] { c = comp.buff_encode_majority(seg); });
      d[1] = timed_ns([&] { comp.buff_decode_majority(c); });
      d[2] = timed_ns([&] { comp.buff_range_filter_majority(c, pred); });
      d[3] = timed_ns([&] { comp.buff_equal_filter_majority(c, pred); });
      d[4] = timed_ns([&] { std::printf("sum is: %.17g\n", comp.buff_sum_majority(c)); });
      d[5] = timed_ns([&] { comp.buff_max_majority(c); });
    } else {
      throw buff::Panic(BUFF_E_REF_PANIC, "Compression not supported yet.");
    }
    std::printf("Performance:%s,%llu,%s,%.17g", file, (unsigned long long)scl, argv[4], (double)c.size() / org_size);
    for (double ns : d) std::printf(",%.17g", 1e9 * org_size / ns / 1024.0 / 1024.0);
    std::printf("\n");
  } catch (const buff::Panic& e) {
    std::fprintf(stderr, "panicked: %s\n", e.what());
    return 101;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
