// comp_profiler: the reference's benchmark CLI (database/src/compress/main.rs:7-36)
// on the GPU engine.
//
//   comp_profiler <input_file> <method> <int_scale> <pred>
//
// method "buff" runs the sequence of run_buff_encoding_decoding_mybitvec
// (buff_simd.rs:167-243): encode, decode, range filter, equality filter, sum,
// max. "buff-simd" runs run_buff_simd_encoding_decoding (buff_simd.rs:28-95):
// the flipped-layout encode, decode and filters; its fifth and sixth timings
// are the reference's fused "BS way" filters, which compute the same bitmaps
// as the plain ones and are timed here with the same GPU filters. The input is
// read like construct_file_iterator_skip_newline (client.rs:163-182): every
// comma-separated field of every line that parses as f64. The last line is the
// reference's `Performance:` line: compression ratio and MiB/s of the bincode
// Segment size per operator (buff_simd.rs:232-242).
#include <chrono>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "split_double.hpp"

namespace {

// Rust's str::parse::<f64> accepts an optional sign, digits with an optional
// dot and exponent, "inf"/"infinity"/"nan" (any case); no surrounding spaces.
bool parse_f64(const std::string& s, double* out) {
  if (s.empty()) return false;
  for (char ch : s)
    if (ch == ' ' || ch == '\t' || ch == '\r' || ch == '\n' || ch == '_') return false;
  if (s.size() > 1 && (s[0] == '0') && (s[1] == 'x' || s[1] == 'X')) return false;
  char* end = nullptr;
  errno = 0;
  const double v = std::strtod(s.c_str(), &end);
  if (end != s.c_str() + s.size()) return false;
  *out = v;
  return true;
}

std::vector<double> read_csv(const char* path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  std::vector<double> v;
  std::string line;
  while (std::getline(f, line)) {
    std::stringstream ss(line);
    std::string field;
    while (std::getline(ss, field, ',')) {
      double x;
      if (parse_f64(field, &x)) v.push_back(x);
    }
  }
  return v;
}

double timed_ns(const std::function<void()>& fn) {
  const auto t0 = std::chrono::steady_clock::now();
  fn();
  const auto t1 = std::chrono::steady_clock::now();
  return (double)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s <input_file> <buff|buff-simd> <int_scale> <pred>\n", argv[0]);
    return 2;
  }
  const char* file = argv[1];
  const std::string method = argv[2];
  const uint64_t scl = std::strtoull(argv[3], nullptr, 10);
  const double pred = std::strtod(argv[4], nullptr);
  std::printf("ARGS: %s, %s, %llu, %s, \n", file, method.c_str(), (unsigned long long)scl, argv[4]);
  try {
    buff::Segment seg{read_csv(file)};
    const double org_size = (double)seg.byte_size();
    buff::SplitBDDoubleCompress comp(10, 10, scl);
    const bool simd = method == "buff-simd";
    if (!simd && method != "buff") throw std::runtime_error("Compression not supported yet.");
    std::vector<uint8_t> compressed;
    double d[6];
    d[0] = timed_ns([&] { compressed = simd ? comp.buff_simd256_encode(seg) : comp.byte_fixed_encode(seg); });
    const double comp_size = (double)compressed.size();
    d[1] = timed_ns([&] {
      if (simd) comp.buff_simd256_decode(compressed);
      else comp.byte_fixed_decode(compressed);
    });
    d[2] = timed_ns([&] {
      if (simd) comp.buff_simd_range_filter(compressed, pred);
      else comp.buff_range_filter(compressed, pred);
    });
    d[3] = timed_ns([&] {
      if (simd) comp.buff_simd_equal_filter(compressed, pred);
      else comp.buff_equal_filter(compressed, pred);
    });
    if (simd) {
      d[4] = timed_ns([&] { comp.buff_simd_range_filter(compressed, pred); });
      d[5] = timed_ns([&] { comp.buff_simd_equal_filter(compressed, pred); });
    } else {
      d[4] = timed_ns([&] { std::printf("sum is: %.17g\n", comp.byte_fixed_sum(compressed)); });
      d[5] = timed_ns([&] { comp.byte_fixed_max(compressed); });
    }
    std::printf("Performance:%s,%llu,%s,%.17g", file, (unsigned long long)scl, argv[4], comp_size / org_size);
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
