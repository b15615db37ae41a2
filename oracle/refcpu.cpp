// refcpu — CPU restatement of the reference BUFF ("SplitBD") codec and query path.
//
// TEST INFRASTRUCTURE ONLY. This file is the oracle that the HIP path is checked
// against, and (its AVX2 "buff-simd" functions) the CPU baseline timed beside the
// GPU. It is loaded only by tests/, by __graft_entry__.smoke() as the checker and
// by bench.py's cpu_baseline leg. Nothing in buff_amd/ links or calls it.
//
// Every function follows the cited Rust function of the reference line by line,
// with Rust release-build arithmetic made explicit:
//   * shifts mask their amount to the operand width (`x >> (s & 63)`),
//   * signed add/sub wrap (done in uint64_t),
//   * `f64 as usize` / `f32 as i32` saturate and map NaN to 0,
//   * integer -> float conversions round to nearest even (C casts do).
// Where the reference panics (index out of bounds, unwrap of None/Err, assert!)
// the restatement throws RefPanic and the C entry point returns REF_PANIC.
//
// Citations are relative to the reference repository root.

#include <immintrin.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "snappy_ref.h"

namespace {

constexpr int REF_OK = 0;
constexpr int REF_PANIC = 4;
constexpr int REF_UNSUPPORTED = 5;
constexpr int REF_SMALL_BUFFER = 6;

struct RefPanic : std::runtime_error {
  explicit RefPanic(const char* m) : std::runtime_error(m) {}
};

// ---------------------------------------------------------------------------
// database/src/methods/prec_double.rs:14-16
constexpr uint64_t EXP_MASK = 0x7FF0000000000000ull;
constexpr uint64_t FIRST_ONE = 0x8000000000000000ull;

inline uint64_t f2u(double x) { uint64_t u; std::memcpy(&u, &x, 8); return u; }

// Rust `f64 as usize` (saturating, NaN -> 0).
inline uint64_t f64_as_usize(double v) {
  if (std::isnan(v) || v <= 0.0) return 0;
  if (v >= 18446744073709551616.0) return UINT64_MAX;
  return (uint64_t)v;
}
// Rust `f32 as i32` (saturating, NaN -> 0).
inline int32_t f32_as_i32(float v) {
  if (std::isnan(v)) return 0;
  if (v <= -2147483648.0f) return INT32_MIN;
  if (v >= 2147483648.0f) return INT32_MAX;
  return (int32_t)v;
}

// database/src/compress/mod.rs:29-46 (PRECISION_MAP); a missing key is the
// `.unwrap()` panic at split_double.rs:316.
int32_t precision_map(int32_t p) {
  static const int32_t tab[16] = {-1, 5, 8, 11, 15, 18, 21, 25, 28, 31, 35, 38, 50, 10, 10, 10};
  if (p < 1 || p > 15) throw RefPanic("PRECISION_MAP has no such precision");
  return tab[p];
}

// `(self.scale as f32).log10() as i32` (split_double.rs:304-310, 1244 and every
// operator's first line). usize -> f32 rounds to nearest.
int32_t prec_of_scale(uint64_t scale) {
  float s = (float)scale;
  return f32_as_i32(log10f(s));
}

// database/src/methods/prec_double.rs:273-281
double get_precision_bound(int32_t precision) {
  std::string s = "0.";
  for (int32_t pos = 0; pos < precision; ++pos) s.push_back('0');
  s += "49";
  return std::strtod(s.c_str(), nullptr);  // Rust's str::parse::<f64> is correctly rounded, as is glibc strtod
}

// database/src/methods/prec_double.rs:28-33 (PrecisionBound::new -> precision_exp)
int32_t precision_exp_of(double precision) {
  uint64_t xu = f2u(precision);
  return (int32_t)((xu & EXP_MASK) >> 52) - 1023;
}

// database/src/methods/prec_double.rs:190-205 (fetch_fixed_aligned)
int64_t fetch_fixed_aligned(double bd, int32_t precision_exp, uint64_t decimal_length) {
  uint64_t bdu = f2u(bd);
  int32_t exp = (int32_t)((bdu & EXP_MASK) >> 52) - 1023;
  uint64_t sign = bdu & FIRST_ONE;
  uint64_t fixed = 0;
  if (exp < precision_exp) {
    fixed = 0;
  } else {
    int32_t sh32 = 63 - exp - (int32_t)decimal_length;   // `as i32` of a u64 truncates
    uint64_t sh = (uint64_t)(int64_t)sh32;               // i32 as u64 sign-extends
    fixed = ((bdu << 11) | FIRST_ONE) >> (sh & 63);       // release-mode shift masking
    if (sign != 0) fixed = ~(fixed - 1);
  }
  return (int64_t)fixed;
}

// compiler-rt __powidf2, which `f64::powi` lowers to.
double powi(double a, int32_t b) {
  const bool recip = b < 0;
  double r = 1;
  int64_t bb = b;
  while (true) {
    if (bb & 1) r *= a;
    bb /= 2;
    if (bb == 0) break;
    a *= a;
  }
  return recip ? 1 / r : r;
}

// ---------------------------------------------------------------------------
// database/src/methods/bit_packing.rs:17-388 (BitPack)
constexpr size_t BYTE_BITS = 8;
constexpr size_t MAX_BITS = 32;

struct BitPackWriter {  // BitPack<Vec<u8>>
  std::vector<uint8_t> buff;
  size_t cursor = 0, bits = 0;

  size_t sum_bits() const { return cursor * BYTE_BITS + bits; }

  // slice write, bit_packing.rs:73-104 / 109-138 (`limit32` selects write vs write_bits)
  void slice_write(uint32_t value, size_t nbits, bool limit32) {
    if ((limit32 && nbits > MAX_BITS) || buff.size() * BYTE_BITS < sum_bits() + nbits)
      throw RefPanic("BitPack write: Err(bits) unwrapped");
    // `((1 << bits) - 1) as u32` is i32 arithmetic: the shift amount wraps at 32.
    if (!limit32 || nbits < MAX_BITS) value &= (uint32_t)((1u << (nbits & 31)) - 1u);
    while (true) {
      size_t bits_left = BYTE_BITS - bits;
      if (nbits <= bits_left) {
        if (cursor >= buff.size()) throw RefPanic("index out of bounds in BitPack write");
        buff[cursor] |= (uint8_t)((uint8_t)value << (uint8_t)bits);
        bits += nbits;
        if (bits >= BYTE_BITS) { cursor += 1; bits = 0; }
        break;
      }
      uint32_t bb = value & (uint32_t)((1ull << bits_left) - 1);
      if (cursor >= buff.size()) throw RefPanic("index out of bounds in BitPack write");
      buff[cursor] |= (uint8_t)((uint8_t)bb << (uint8_t)bits);
      cursor += 1;
      bits = 0;
      value >>= bits_left;
      nbits -= bits_left;
    }
  }
  // bit_packing.rs:317-337 / 339-358 (Vec variants: grow, then slice write)
  void grow_then_write(uint32_t value, size_t nbits, bool limit32) {
    size_t len = buff.size();
    size_t need = sum_bits() + nbits;
    if (need >= len * BYTE_BITS) {
      size_t over = need - len * BYTE_BITS;
      buff.resize(len + (over + BYTE_BITS - 1) / BYTE_BITS, 0);
    }
    slice_write(value, nbits, limit32);
  }
  void write(uint32_t v, size_t nb) { grow_then_write(v, nb, true); }
  void write_bits(uint32_t v, size_t nb) { grow_then_write(v, nb, false); }
  void write_byte(uint8_t v) { buff.push_back(v); }                    // 369-372
  void finish_write_byte() {                                            // 375-381
    size_t len = buff.size();
    buff.resize(len + 1, 0);
    bits = 0;
    cursor = len;
  }
};

struct BitPackReader {  // BitPack<&[u8]>
  const uint8_t* buff;
  size_t len;
  size_t cursor = 0, bits = 0;
  BitPackReader(const uint8_t* b, size_t l) : buff(b), len(l) {}

  size_t sum_bits() const { return cursor * BYTE_BITS + bits; }
  uint8_t at(size_t i) const {
    if (i >= len) throw RefPanic("index out of bounds in BitPack read");
    return buff[i];
  }
  // bit_packing.rs:154-187 (read) and 194-228 (read_bits)
  uint32_t read_impl(size_t nbits, bool limit32) {
    if ((limit32 && nbits > MAX_BITS) || len * BYTE_BITS < sum_bits() + nbits)
      throw RefPanic("BitPack read: Err(bits) unwrapped");
    uint32_t bits_left = 0, output = 0;
    while (true) {
      size_t byte_left = BYTE_BITS - bits;
      if (nbits <= byte_left) {
        uint32_t bb = at(cursor);
        bb >>= (uint32_t)bits;
        bb &= (uint32_t)((1ull << nbits) - 1);
        output |= bb << bits_left;
        bits += nbits;
        break;
      }
      uint32_t bb = at(cursor);
      bb >>= (uint32_t)bits;
      bb &= (uint32_t)((1ull << byte_left) - 1);
      output |= bb << bits_left;
      bits += byte_left;
      bits_left += (uint32_t)byte_left;
      nbits -= byte_left;
      if (bits >= BYTE_BITS) { cursor += 1; bits -= BYTE_BITS; }
    }
    return output;
  }
  uint32_t read(size_t nb) { return read_impl(nb, true); }
  uint8_t read_bits(size_t nb) { return (uint8_t)read_impl(nb, false); }
  void finish_read_byte() { cursor += 1; bits = 0; }                    // 231-235
  uint8_t read_byte() { cursor += 1; return at(cursor); }               // 238-242
  const uint8_t* read_n_byte(size_t n) {                                 // 245-251
    cursor += 1;
    size_t end = cursor + n;
    if (end > len || end < cursor) throw RefPanic("slice out of range in read_n_byte");
    const uint8_t* out = buff + cursor;
    cursor += n - 1;   // usize wrap when n == 0, as in release Rust
    return out;
  }
  void skip_n_byte(size_t n) { cursor += n; }                            // 262-266
  uint32_t peek8() const {   // what read(8) would return, reader unchanged
    BitPackReader c = *this;
    return c.read(8);
  }
  void skip(size_t nb) {                                                 // 268-283 (Err ignored by callers)
    if (len * BYTE_BITS < sum_bits() + nb) return;
    size_t bytes = nb / BYTE_BITS, left = nb % BYTE_BITS;
    size_t cur_bits = bits + left;
    cursor = cursor + bytes + cur_bits / BYTE_BITS;
    bits = cur_bits % BYTE_BITS;
  }
};

// ---------------------------------------------------------------------------
// bit-vec/src/lib.rs (BitVec<u32>): word i/32, bit i%32, unused high bits 0.
struct BitVec {
  std::vector<uint32_t> storage;
  size_t nbits = 0;
  explicit BitVec(size_t n = 0) : storage((n + 31) / 32, 0u), nbits(n) {}
  void set(size_t i) {                                                   // lib.rs:619-627
    if (i >= nbits) throw RefPanic("BitVec::set index out of bounds");
    storage[i / 32] |= 1u << (i % 32);
  }
  uint32_t cardinality() const {                                         // lib.rs:434-436
    uint32_t c = 0;
    for (uint32_t w : storage) c += (uint32_t)__builtin_popcount(w);
    return c;
  }
  void fix_last_block() {                                                // lib.rs:537-541
    size_t extra = nbits % 32;
    if (extra && !storage.empty()) storage.back() &= (1u << extra) - 1;
  }
  // lib.rs:466-480 `process`: assert_eq! on the bit lengths, then a zip over the blocks.
  size_t process(const BitVec& o) const {
    if (o.nbits != nbits) throw RefPanic("BitVec length mismatch");
    return storage.size() < o.storage.size() ? storage.size() : o.storage.size();
  }
  void and_with(const BitVec& o) { size_t m = process(o); for (size_t i = 0; i < m; ++i) storage[i] &= o.storage[i]; }
  void or_with(const BitVec& o) { size_t m = process(o); for (size_t i = 0; i < m; ++i) storage[i] |= o.storage[i]; }
  void negate() {                                                        // lib.rs:696-700
    for (uint32_t& w : storage) w = ~w;
    fix_last_block();
  }
  void clear() { for (uint32_t& w : storage) w = 0; }                    // lib.rs:1491-1494
  std::vector<uint8_t> to_binary() const {                               // lib.rs:399-407
    std::vector<uint8_t> b(8 + 4 * storage.size());
    uint64_t nb = nbits;
    std::memcpy(b.data(), &nb, 8);
    if (!storage.empty()) std::memcpy(b.data() + 8, storage.data(), 4 * storage.size());
    return b;
  }
  static BitVec from_binary(const std::vector<uint8_t>& data) {          // lib.rs:411-428
    if (data.size() < 8) throw RefPanic("from_binary: short input");
    uint64_t nb;
    std::memcpy(&nb, data.data(), 8);
    BitVec bv;
    bv.nbits = (size_t)nb;
    size_t cnt = 8, len = data.size() - 4;
    while (cnt <= len) {
      uint32_t w;
      std::memcpy(&w, data.data() + cnt, 4);
      bv.storage.push_back(w);
      cnt += 4;
    }
    bv.fix_last_block();
    return bv;
  }
};

// database/src/query/bit_vec_iter.rs:35-47 (BVIter: set bits in ascending order)
struct BVIter {
  const BitVec& bv;
  size_t pointer = 0;
  uint32_t cached;
  explicit BVIter(const BitVec& b) : bv(b) {
    if (b.storage.empty()) throw RefPanic("BVIter over empty storage");
    cached = b.storage[0];
  }
  bool next(size_t* out) {
    while (cached == 0) {
      if (pointer == bv.storage.size() - 1) return false;
      pointer += 1;
      cached = bv.storage[pointer];
    }
    uint32_t t = cached & ~(cached - 1);
    cached ^= t;
    *out = (pointer << 5) + (size_t)__builtin_popcount(t - 1);
    return true;
  }
};

// database/src/compress/buff_slice.rs:174-177
inline uint8_t flip(uint8_t x) { return x ^ (uint8_t)0x80; }

// ---------------------------------------------------------------------------
struct Header {
  int64_t base;
  uint32_t len, ilen, dlen;
};

// The five 32-bit header reads every operator starts with
// (e.g. split_double.rs:1249-1258).
Header read_header(BitPackReader& bp) {
  Header h;
  uint32_t lower = bp.read(32);
  uint32_t higher = bp.read(32);
  h.base = (int64_t)((uint64_t)lower | ((uint64_t)higher << 32));
  h.len = bp.read(32);
  h.ilen = bp.read(32);
  h.dlen = bp.read(32);
  return h;
}

// PrecisionBound as the query operators build it: precision_exp from the
// caller's scale, decimal_length from the stream header.
struct Bound {
  int32_t precision_exp;
  uint64_t dlen;
  int64_t fixed(double x) const { return fetch_fixed_aligned(x, precision_exp, dlen); }
};
Bound bound_for(uint64_t scale, uint32_t dlen) {
  int32_t prec = prec_of_scale(scale);
  return Bound{precision_exp_of(get_precision_bound(prec)), (uint64_t)dlen};
}

// ---------------------------------------------------------------------------
// split_double.rs:298-432 (byte_fixed_encode); flip=true is the
// buff_simd256_encode twin (buff_simd.rs:3116-3250): plane bytes XOR 0x80,
// tail bits not flipped.
std::vector<uint8_t> byte_fixed_encode(const double* data, size_t n, uint64_t scale, bool flp,
                                       int64_t* out_min, int64_t* out_max) {
  std::vector<int64_t> fixed_vec;
  fixed_vec.reserve(n);
  uint32_t t = (uint32_t)n;
  int32_t prec = scale == 0 ? 0 : prec_of_scale(scale);
  double prec_delta = get_precision_bound(prec);
  int32_t pexp = precision_exp_of(prec_delta);
  uint64_t dec_len = (uint64_t)precision_map(prec);
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  for (size_t i = 0; i < n; ++i) {
    int64_t f = fetch_fixed_aligned(data[i], pexp, dec_len);
    if (f < mn) mn = f;
    else if (f > mx) mx = f;   // `else if`: quirk kept
    fixed_vec.push_back(f);
  }
  if (out_min) *out_min = mn;
  if (out_max) *out_max = mx;
  int64_t delta = (int64_t)((uint64_t)mx - (uint64_t)mn);
  int64_t base_fixed = mn;
  uint64_t ubase = (uint64_t)base_fixed;
  double cal_int_length = 0.0;
  if (delta != 0) cal_int_length = std::ceil(std::log2((double)delta));
  size_t fixed_len = (size_t)f64_as_usize(cal_int_length);
  size_t ilen = fixed_len - (size_t)dec_len;
  size_t dlen = (size_t)dec_len;

  BitPackWriter bp;
  bp.write((uint32_t)ubase, 32);
  bp.write((uint32_t)(ubase >> 32), 32);
  bp.write(t, 32);
  bp.write((uint32_t)ilen, 32);
  bp.write((uint32_t)dlen, 32);

  size_t remain = fixed_len;
  if (remain < 8) {
    for (size_t i = 0; i < n; ++i)
      bp.write_bits((uint32_t)((uint64_t)fixed_vec[i] - (uint64_t)base_fixed), remain);
    remain = 0;
  } else {
    remain -= 8;
    std::vector<uint64_t> fixed_u64(n);
    for (size_t i = 0; i < n; ++i) {
      uint64_t cur = (uint64_t)fixed_vec[i] - (uint64_t)base_fixed;
      uint8_t b = remain > 0 ? (uint8_t)(cur >> (remain & 63)) : (uint8_t)cur;
      bp.write_byte(flp ? flip(b) : b);
      fixed_u64[i] = cur;
    }
    while (remain >= 8) {
      remain -= 8;
      for (size_t i = 0; i < n; ++i) {
        uint8_t b = remain > 0 ? (uint8_t)(fixed_u64[i] >> (remain & 63)) : (uint8_t)fixed_u64[i];
        bp.write_byte(flp ? flip(b) : b);
      }
    }
    if (remain > 0) {
      bp.finish_write_byte();
      for (size_t i = 0; i < n; ++i) bp.write_bits((uint32_t)fixed_u64[i], remain);
    }
  }
  return bp.buff;
}

// split_double.rs:1243-1350 (byte_fixed_decode); flip twin buff_simd.rs:3252-3357.
std::vector<double> byte_fixed_decode(const uint8_t* bytes, size_t blen, uint64_t scale, bool flp) {
  (void)scale;  // the reference computes prec/precision bound here but never uses them for decode
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  std::vector<double> out;
  std::vector<uint64_t> fixed_vec;
  double dec_scl = powi(2.0, (int32_t)h.dlen);
  uint32_t remain = h.dlen + h.ilen;
  auto F = [&](uint8_t x) { return flp ? flip(x) : x; };
  if (remain < 8) {
    for (uint32_t i = 0; i < h.len; ++i) {
      uint8_t cur = bp.read_bits(remain);
      out.push_back((double)(int64_t)((uint64_t)h.base + (uint64_t)cur) / dec_scl);
    }
  } else {
    remain -= 8;
    const uint8_t* chunk = bp.read_n_byte(h.len);
    if (remain == 0) {
      for (uint32_t i = 0; i < h.len; ++i)
        out.push_back((double)(int64_t)((uint64_t)h.base + (uint64_t)F(chunk[i])) / dec_scl);
    } else {
      for (uint32_t i = 0; i < h.len; ++i) fixed_vec.push_back((uint64_t)F(chunk[i]) << (remain & 63));
    }
    while (remain >= 8) {
      remain -= 8;
      chunk = bp.read_n_byte(h.len);
      if (remain == 0) {
        size_t m = std::min(fixed_vec.size(), (size_t)h.len);   // zip stops at the shorter
        for (size_t i = 0; i < m; ++i)
          out.push_back((double)(int64_t)((uint64_t)h.base + (fixed_vec[i] | (uint64_t)F(chunk[i]))) / dec_scl);
      } else {
        for (size_t i = 0; i < fixed_vec.size(); ++i) {
          if (i >= h.len) throw RefPanic("chunk iterator exhausted");
          fixed_vec[i] |= (uint64_t)F(chunk[i]) << (remain & 63);
        }
      }
    }
    if (remain > 0) {
      bp.finish_read_byte();
      for (size_t i = 0; i < fixed_vec.size(); ++i) {
        uint64_t v = fixed_vec[i] | (uint64_t)bp.read_bits(remain);
        out.push_back((double)(int64_t)((uint64_t)h.base + v) / dec_scl);
      }
    }
  }
  return out;
}

// The qualifying rows a filter reports: the printed count and its bitmap.
struct FilterOut {
  BitVec res;
  uint64_t count = 0;
  uint64_t cand[9] = {0};   // rows entering plane k (k>=1) / the tail (index P) as candidates
};

// buff_simd.rs:4507-4679 (buff_range_filter): x > pred.
FilterOut buff_range_filter(const uint8_t* bytes, size_t blen, uint64_t scale, double pred) {
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  size_t len = h.len;
  uint32_t remain = h.dlen + h.ilen;
  Bound bound = bound_for(scale, h.dlen);
  BitVec rb1(len), res(len);
  FilterOut fo;
  int64_t fixed_part = bound.fixed(pred);
  if (fixed_part < h.base) {            // 4531-4534: prints len, returns no bitmap
    fo.res = BitVec(len);
    for (size_t i = 0; i < len; ++i) fo.res.set(i);
    fo.count = len;
    return fo;
  }
  uint64_t fixed_target = (uint64_t)fixed_part - (uint64_t)h.base;
  uint8_t cur_tar = 0;
  int stage = 0;
  if (remain < 8) {
    for (size_t i = 0; i < len; ++i) {
      uint8_t cur = bp.read_bits(remain);
      if ((uint64_t)cur > fixed_target) res.set(i);
    }
    remain = 0;
  } else {
    remain -= 8;
    const uint8_t* chunk = bp.read_n_byte(len);
    cur_tar = (uint8_t)(fixed_target >> (remain & 63));
    for (size_t i = 0; i < len; ++i) {
      uint8_t c = chunk[i];
      if (c > cur_tar) res.set(i);
      else if (c == cur_tar) rb1.set(i);
    }
  }
  while (remain > 0) {
    ++stage;
    if (remain >= 8) {
      remain -= 8;
      BitVec cur_rb(len);
      if (rb1.cardinality() != 0) {
        fo.cand[stage] = rb1.cardinality();
        BVIter it(rb1);
        size_t dec_cur = 0, dec_pre = 0, idx;
        cur_tar = (uint8_t)(fixed_target >> (remain & 63));
        bool have = it.next(&idx);
        if (have) {
          dec_cur = idx;
          if (dec_cur != 0) bp.skip_n_byte(dec_cur);
          uint8_t dec = bp.read_byte();
          if (dec > cur_tar) res.set(dec_cur);
          else if (dec == cur_tar) cur_rb.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        while (have) {
          dec_cur = idx;
          size_t delta = dec_cur - dec_pre;
          if (delta != 1) bp.skip_n_byte(delta - 1);
          uint8_t dec = bp.read_byte();
          if (dec > cur_tar) res.set(dec_cur);
          else if (dec == cur_tar) cur_rb.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        if (len - dec_pre > 1) bp.skip_n_byte(len - dec_pre - 1);
      } else {
        bp.skip_n_byte(len);
        break;
      }
      rb1 = cur_rb;
    } else {
      uint8_t sh = (uint8_t)(BYTE_BITS - remain);
      cur_tar = (uint8_t)((uint8_t)((uint8_t)fixed_target << sh) >> sh);
      bp.finish_read_byte();
      if (rb1.cardinality() != 0) {
        fo.cand[stage] = rb1.cardinality();
        BVIter it(rb1);
        size_t dec_cur = 0, dec_pre = 0, idx;
        bool have = it.next(&idx);
        if (have) {
          dec_cur = idx;
          if (dec_cur != 0) bp.skip(dec_cur * remain);
          uint8_t dec = bp.read_bits(remain);
          if (dec > cur_tar) res.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        while (have) {
          dec_cur = idx;
          size_t delta = dec_cur - dec_pre;
          if (delta != 1) bp.skip((delta - 1) * remain);
          uint8_t dec = bp.read_bits(remain);
          if (dec > cur_tar) res.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        remain = 0;
      } else {
        break;
      }
    }
  }
  fo.count = res.cardinality();
  fo.res = std::move(res);
  return fo;
}

// buff_simd.rs:4925-5093 (buff_equal_filter): x == pred.
FilterOut buff_equal_filter(const uint8_t* bytes, size_t blen, uint64_t scale, double pred) {
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  size_t len = h.len;
  uint32_t remain = h.dlen + h.ilen;
  Bound bound = bound_for(scale, h.dlen);
  BitVec rb1(len), res(len);
  FilterOut fo;
  int64_t fixed_part = bound.fixed(pred);
  if (fixed_part < h.base) {            // 4949-4952: prints 0
    fo.res = BitVec(len);
    return fo;
  }
  uint64_t fixed_target = (uint64_t)fixed_part - (uint64_t)h.base;
  uint8_t dec_byte = (uint8_t)fixed_target;
  int stage = 0;
  if (remain < 8) {
    for (size_t i = 0; i < len; ++i) {
      uint8_t cur = bp.read_bits(remain);
      if ((uint64_t)cur == fixed_target) res.set(i);
    }
    remain = 0;
  } else {
    remain -= 8;
    const uint8_t* chunk = bp.read_n_byte(len);
    dec_byte = (uint8_t)(fixed_target >> (remain & 63));
    for (size_t i = 0; i < len; ++i)
      if (chunk[i] == dec_byte) rb1.set(i);
  }
  while (remain > 0) {
    ++stage;
    if (remain >= 8) {
      remain -= 8;
      BitVec cur_rb(len);
      if (rb1.cardinality() != 0) {
        fo.cand[stage] = rb1.cardinality();
        BVIter it(rb1);
        size_t dec_cur = 0, dec_pre = 0, idx;
        dec_byte = (uint8_t)(fixed_target >> (remain & 63));
        bool have = it.next(&idx);
        if (have) {
          dec_cur = idx;
          if (dec_cur != 0) bp.skip_n_byte(dec_cur);
          uint8_t dec = bp.read_byte();
          if (dec == dec_byte) cur_rb.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        while (have) {
          dec_cur = idx;
          size_t delta = dec_cur - dec_pre;
          if (delta != 1) bp.skip_n_byte(delta - 1);
          uint8_t dec = bp.read_byte();
          if (dec == dec_byte) cur_rb.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        if (len - dec_pre > 1) bp.skip_n_byte(len - dec_pre - 1);
      } else {
        bp.skip_n_byte(len);
        break;
      }
      rb1 = cur_rb;
    } else {
      uint8_t sh = (uint8_t)(BYTE_BITS - remain);
      dec_byte = (uint8_t)((uint8_t)((uint8_t)fixed_target << sh) >> sh);
      bp.finish_read_byte();
      if (rb1.cardinality() != 0) {
        fo.cand[stage] = rb1.cardinality();
        BVIter it(rb1);
        size_t dec_cur = 0, dec_pre = 0, idx;
        bool have = it.next(&idx);
        if (have) {
          dec_cur = idx;
          if (dec_cur != 0) bp.skip(dec_cur * remain);
          uint8_t dec = bp.read_bits(remain);
          if (dec == dec_byte) res.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        while (have) {
          dec_cur = idx;
          size_t delta = dec_cur - dec_pre;
          if (delta != 1) bp.skip((delta - 1) * remain);
          uint8_t dec = bp.read_bits(remain);
          if (dec == dec_byte) res.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        remain = 0;
      } else {
        break;
      }
    }
  }
  // F >= 8 and F % 8 == 0: the matches are left in rb1 and `res` (what the
  // reference reports) stays empty.
  fo.count = res.cardinality();
  fo.res = std::move(res);
  return fo;
}

// split_double.rs:2127-2208 (byte_fixed_sum). `plane_sums` receives the exact
// per-sub-column u64 sums (index P = tail).
double byte_fixed_sum(const uint8_t* bytes, size_t blen, uint64_t scale, uint64_t* plane_sums) {
  (void)scale;
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  uint32_t remain = h.ilen + h.dlen;
  int32_t processed = 0;
  double sum_base = (double)h.len * (double)h.base;
  uint64_t sum_fixed = 0;
  double sum = 0.0;
  double dec_scl = powi(2.0, (int32_t)h.dlen);
  sum_base = sum_base / dec_scl;
  int k = 0;
  if (remain >= 8) {
    remain -= 8;
    processed += 8;
    dec_scl = powi(2.0, (int32_t)((uint32_t)processed - h.ilen));
    const uint8_t* chunk = bp.read_n_byte(h.len);
    for (uint32_t i = 0; i < h.len; ++i) sum_fixed += chunk[i];
    if (plane_sums) plane_sums[k] = sum_fixed;
    ++k;
    sum = sum + (double)sum_fixed / dec_scl;
    sum_fixed = 0;
    if (remain == 0) return sum;   // returns without the base term
  }
  while (remain >= 8) {
    remain -= 8;
    processed += 8;
    dec_scl = powi(2.0, (int32_t)((uint32_t)processed - h.ilen));
    const uint8_t* chunk = bp.read_n_byte(h.len);
    for (uint32_t i = 0; i < h.len; ++i) sum_fixed += chunk[i];
    if (plane_sums) plane_sums[k] = sum_fixed;
    ++k;
    sum = sum + (double)sum_fixed / dec_scl;
    sum_fixed = 0;
    if (remain == 0) return sum;
  }
  if (remain > 0) {
    if (processed >= 8) bp.finish_read_byte();
    processed += (int32_t)remain;
    dec_scl = powi(2.0, (int32_t)((uint32_t)processed - h.ilen));
    for (uint32_t i = 0; i < h.len; ++i) sum_fixed += bp.read_bits(remain);
    if (plane_sums) plane_sums[k] = sum_fixed;
    sum = sum + (double)sum_fixed / dec_scl;
  }
  sum += sum_base;
  return sum;
}

// split_double.rs:3919-4105 (byte_fixed_max). Returns the printed max value.
double byte_fixed_max(const uint8_t* bytes, size_t blen, uint64_t scale, uint64_t* max_off,
                      uint64_t* argmax_count) {
  (void)scale;
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  uint32_t remain = h.dlen + h.ilen;
  uint32_t len = h.len;
  // The reference keeps candidates in croaring bitmaps; only their ascending
  // iteration order matters here, which a sorted index vector gives too.
  std::vector<uint32_t> rb1, res;
  uint64_t mx = 0;
  uint8_t byte_max = 0;
  if (remain < 8) {
    for (uint32_t i = 0; i < len; ++i) {
      uint64_t cur = bp.read_bits(remain);
      if (cur > mx) { mx = cur; res.clear(); res.push_back(i); }
      else if (cur == mx) res.push_back(i);
    }
    remain = 0;
  } else {
    remain -= 8;
    const uint8_t* chunk = bp.read_n_byte(len);
    for (uint32_t i = 0; i < len; ++i) {
      uint8_t c = chunk[i];
      if (c > byte_max) { byte_max = c; rb1.clear(); rb1.push_back(i); }
      else if (c == byte_max) rb1.push_back(i);
    }
    mx = (uint64_t)byte_max << (remain & 63);
  }
  while (remain > 0) {
    byte_max = 0;
    if (remain >= 8) {
      remain -= 8;
      std::vector<uint32_t> cur_rb;
      uint32_t dec_pre = 0;
      bool first = true;
      for (uint32_t dec_cur : rb1) {
        if (first) {
          if (dec_cur != 0) bp.skip_n_byte(dec_cur);
          uint8_t dec = bp.read_byte();
          if (dec > byte_max) { byte_max = dec; cur_rb.push_back(dec_cur); }
          else if (dec == byte_max) cur_rb.push_back(dec_cur);
          first = false;
        } else {
          uint32_t delta = dec_cur - dec_pre;
          if (delta != 1) bp.skip_n_byte(delta - 1);
          uint8_t dec = bp.read_byte();
          if (dec > byte_max) { byte_max = dec; cur_rb.clear(); cur_rb.push_back(dec_cur); }
          else if (dec == byte_max) cur_rb.push_back(dec_cur);
        }
        dec_pre = dec_cur;
      }
      if (len - dec_pre > 1) bp.skip_n_byte(len - dec_pre - 1);
      mx = mx | ((uint64_t)byte_max << (remain & 63));
      rb1 = cur_rb;
    } else {
      byte_max = 0;
      bp.finish_read_byte();
      uint32_t dec_pre = 0;
      bool first = true;
      for (uint32_t dec_cur : rb1) {
        if (first) {
          if (dec_cur != 0) bp.skip((size_t)(dec_cur * remain));
          uint8_t dec = bp.read_bits(remain);
          if (dec > byte_max) { byte_max = dec; res.push_back(dec_cur); }
          else if (dec == byte_max) res.push_back(dec_cur);
          first = false;
        } else {
          uint32_t delta = dec_cur - dec_pre;
          if (delta != 1) bp.skip((size_t)((delta - 1) * remain));
          uint8_t dec = bp.read_bits(remain);
          if (dec > byte_max) { byte_max = dec; res.clear(); res.push_back(dec_cur); }
          else if (dec == byte_max) res.push_back(dec_cur);
        }
        dec_pre = dec_cur;
      }
      remain = 0;
      mx = mx | (uint64_t)byte_max;
    }
  }
  if (max_off) *max_off = mx;
  if (argmax_count) *argmax_count = res.size();
  return (double)(int64_t)(mx + (uint64_t)h.base) / powi(2.0, (int32_t)h.dlen);
}

// split_double.rs:1612-1893 (byte_residue_decode_with_precision), the branch
// taken when the stored ilen <= 15 (1904-1893: the plain byte layout, identical
// to the byte_fixed stream). With ilen > 15 the reference parses an outlier
// section that a byte_fixed stream does not have: UNSUPPORTED.
std::vector<double> decode_with_precision(const uint8_t* bytes, size_t blen, uint64_t scale,
                                          uint32_t precision, bool* unsupported) {
  (void)scale;
  *unsupported = false;
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  std::vector<double> out;
  std::vector<uint64_t> fixed_vec;
  double dec_scl = powi(2.0, (int32_t)h.dlen);
  uint32_t remain = h.dlen + h.ilen;
  if (precision == 0) {
    remain = h.ilen;
    uint32_t byte_needed = h.ilen / 8;
    if (h.ilen % 8 > 0) { byte_needed += 1; remain = byte_needed * 8; }
  } else {
    uint32_t bits_needed = (uint32_t)precision_map((int32_t)precision);
    if (!(h.dlen >= bits_needed)) throw RefPanic("assert!(dlen >= bits_needed)");
    remain = h.ilen + bits_needed;
    uint32_t fixed_byte = (h.dlen + h.ilen) / 8;
    uint32_t byte_needed = (h.ilen + bits_needed) / 8;
    uint32_t extra_bits = (h.ilen + bits_needed) % 8;
    if (byte_needed < fixed_byte && extra_bits > 0) { byte_needed += 1; remain = byte_needed * 8; }
  }
  if (h.ilen > 15) { *unsupported = true; return out; }
  if (remain < 8) {
    for (uint32_t i = 0; i < h.len; ++i) {
      uint8_t cur = bp.read_bits(remain);
      out.push_back((double)(int64_t)((uint64_t)h.base + (uint64_t)cur) / dec_scl);
    }
  } else {
    remain -= 8;
    const uint8_t* chunk = bp.read_n_byte(h.len);
    if (remain == 0) {
      for (uint32_t i = 0; i < h.len; ++i)
        out.push_back((double)(int64_t)((uint64_t)h.base + (uint64_t)chunk[i]) / dec_scl);
    } else {
      for (uint32_t i = 0; i < h.len; ++i) fixed_vec.push_back((uint64_t)chunk[i] << (remain & 63));
    }
    while (remain >= 8) {
      remain -= 8;
      chunk = bp.read_n_byte(h.len);
      if (remain == 0) {
        size_t m = std::min(fixed_vec.size(), (size_t)h.len);
        for (size_t i = 0; i < m; ++i)
          out.push_back((double)(int64_t)((uint64_t)h.base + (fixed_vec[i] | (uint64_t)chunk[i])) / dec_scl);
      } else {
        for (size_t i = 0; i < fixed_vec.size(); ++i) fixed_vec[i] |= (uint64_t)chunk[i] << (remain & 63);
      }
    }
    if (remain > 0) {
      bp.finish_read_byte();
      for (size_t i = 0; i < fixed_vec.size(); ++i) {
        uint64_t v = fixed_vec[i] | (uint64_t)bp.read_bits(remain);
        out.push_back((double)(int64_t)((uint64_t)h.base + v) / dec_scl);
      }
    }
  }
  return out;
}

// split_double.rs:2409-2629 (byte_residue_sum_with_precision), the branch taken
// when the stored ilen <= 15 (2576-2624: the plain byte layout, identical to the
// byte_fixed stream). The read width R follows the same rule as the
// reduced-precision decode (2439-2462). `plane_sums` receives the u64 sum of
// every sub-column read (index k) and, when R % 8 != 0, the sum of the R % 8-bit
// fields read_bits returns (index R / 8). With ilen > 15 the reference parses an
// outlier section that a byte_fixed stream does not have: UNSUPPORTED.
double sum_with_precision(const uint8_t* bytes, size_t blen, uint64_t scale, uint32_t precision,
                          bool* unsupported, uint64_t* plane_sums) {
  (void)scale;
  *unsupported = false;
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  uint32_t remain = h.ilen + h.dlen;
  int32_t processed = 0;
  double sum_base = (double)h.len * (double)h.base;
  uint64_t sum_fixed = 0;
  double sum = 0.0;
  double dec_scl = powi(2.0, (int32_t)h.dlen);
  sum_base = sum_base / dec_scl;
  if (precision == 0) {
    remain = h.ilen;
    uint32_t byte_needed = h.ilen / 8;
    if (h.ilen % 8 > 0) { byte_needed += 1; remain = byte_needed * 8; }
  } else {
    const uint32_t bits_needed = (uint32_t)precision_map((int32_t)precision);
    if (!(h.dlen >= bits_needed)) throw RefPanic("assert!(dlen >= bits_needed)");
    remain = h.ilen + bits_needed;
    const uint32_t fixed_byte = (h.dlen + h.ilen) / 8;
    uint32_t byte_needed = (h.ilen + bits_needed) / 8;
    const uint32_t extra_bits = (h.ilen + bits_needed) % 8;
    if (byte_needed < fixed_byte && extra_bits > 0) { byte_needed += 1; remain = byte_needed * 8; }
  }
  if (h.ilen > 15) { *unsupported = true; return 0.0; }
  int k = 0;
  // 2577-2593 and 2594-2610: one whole sub-column per step; the loop returns
  // without the base term as soon as the width is used up on a byte boundary
  while (remain >= 8) {
    remain -= 8;
    processed += 8;
    dec_scl = powi(2.0, (int32_t)((uint32_t)processed - h.ilen));
    const uint8_t* chunk = bp.read_n_byte(h.len);
    for (uint32_t i = 0; i < h.len; ++i) sum_fixed += chunk[i];
    if (plane_sums && k < 9) plane_sums[k] = sum_fixed;
    ++k;
    sum = sum + (double)sum_fixed / dec_scl;
    sum_fixed = 0;
    if (remain == 0) return sum;
  }
  if (remain > 0) {   // 2611-2622
    if (processed >= 8) bp.finish_read_byte();
    processed += (int32_t)remain;
    dec_scl = powi(2.0, (int32_t)((uint32_t)processed - h.ilen));
    for (uint32_t i = 0; i < h.len; ++i) sum_fixed += bp.read_bits(remain);
    if (plane_sums && k < 9) plane_sums[k] = sum_fixed;
    sum = sum + (double)sum_fixed / dec_scl;
  }
  sum += sum_base;
  return sum;
}

// ---------------------------------------------------------------------------
// AVX2 "buff-simd" query path — the reference's fast CPU path, timed as the
// CPU baseline. database/src/simd/vectorize_query.rs:168-222 and
// database/src/compress/buff_simd.rs:3361-3733.
constexpr double SIMD_THRESHOLD = 0.06;   // buff_simd.rs:26

// vectorize_query.rs:168-195: fills res with the `>` words of the first
// floor(len/32) blocks, returns the `==` bitmap.
__attribute__((target("avx2"))) BitVec range_simd_mybitvec(const uint8_t* x, size_t len, BitVec& res, uint8_t pred) {
  std::vector<uint32_t> resv, midv;
  resv.reserve(len / 32);
  midv.reserve(len / 32 + 1);
  const __m256i pw = _mm256_set1_epi8((char)flip(pred));
  for (size_t off = 0; off + 32 <= len; off += 32) {
    __m256i w = _mm256_lddqu_si256((const __m256i*)(x + off));
    resv.push_back((uint32_t)_mm256_movemask_epi8(_mm256_cmpgt_epi8(w, pw)));
    midv.push_back((uint32_t)_mm256_movemask_epi8(_mm256_cmpeq_epi8(w, pw)));
  }
  std::memcpy(res.storage.data(), resv.data(), resv.size() * 4);   // set_storage
  BitVec eq(len);                                                    // from_vec pads with zeros
  std::memcpy(eq.storage.data(), midv.data(), midv.size() * 4);
  eq.fix_last_block();
  return eq;
}

// vectorize_query.rs:198-222
__attribute__((target("avx2"))) BitVec equal_simd_mybitvec(const uint8_t* x, size_t len, uint8_t pred) {
  BitVec eq(len);
  const __m256i pw = _mm256_set1_epi8((char)flip(pred));
  size_t k = 0;
  for (size_t off = 0; off + 32 <= len; off += 32, ++k) {
    __m256i w = _mm256_lddqu_si256((const __m256i*)(x + off));
    eq.storage[k] = (uint32_t)_mm256_movemask_epi8(_mm256_cmpeq_epi8(w, pw));
  }
  eq.fix_last_block();
  return eq;
}

// buff_simd.rs:3361-3547 (buff_simd_range_filter) over a flipped stream.
FilterOut buff_simd_range_filter(const uint8_t* bytes, size_t blen, uint64_t scale, double pred) {
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  size_t len = h.len;
  uint32_t remain = h.dlen + h.ilen;
  Bound bound = bound_for(scale, h.dlen);
  BitVec rb1(len), res(len);
  FilterOut fo;
  int64_t fixed_part = bound.fixed(pred);
  if (fixed_part < h.base) {
    fo.res = BitVec(len);
    for (size_t i = 0; i < len; ++i) fo.res.set(i);
    fo.count = len;
    return fo;
  }
  uint64_t fixed_target = (uint64_t)fixed_part - (uint64_t)h.base;
  uint8_t cur_tar = 0;
  double check_ratio = 0.0;
  if (remain < 8) {
    for (size_t i = 0; i < len; ++i) {
      uint8_t cur = bp.read_bits(remain);
      if ((uint64_t)cur > fixed_target) res.set(i);
    }
    remain = 0;
  } else {
    remain -= 8;
    const uint8_t* chunk = bp.read_n_byte(len);
    cur_tar = (uint8_t)(fixed_target >> (remain & 63));
    rb1 = range_simd_mybitvec(chunk, len, res, cur_tar);
    check_ratio = (double)rb1.cardinality() / (double)len;
  }
  while (remain > 0) {
    if (remain >= 8) {
      remain -= 8;
      if (check_ratio > SIMD_THRESHOLD) {
        BitVec temp_res(len);
        const uint8_t* chunk = bp.read_n_byte(len);
        cur_tar = (uint8_t)(fixed_target >> (remain & 63));
        BitVec eq_bm = range_simd_mybitvec(chunk, len, temp_res, cur_tar);
        temp_res.and_with(rb1);
        res.or_with(temp_res);
        rb1.and_with(eq_bm);
        check_ratio = (double)rb1.cardinality() / (double)len;
      } else {
        BitVec cur_rb(len);
        if (rb1.cardinality() != 0) {
          BVIter it(rb1);
          size_t dec_cur = 0, dec_pre = 0, idx;
          cur_tar = (uint8_t)(fixed_target >> (remain & 63));
          bool have = it.next(&idx);
          if (have) {
            dec_cur = idx;
            if (dec_cur != 0) bp.skip_n_byte(dec_cur);
            uint8_t dec = flip(bp.read_byte());
            if (dec > cur_tar) res.set(dec_cur);
            else if (dec == cur_tar) cur_rb.set(dec_cur);
            have = it.next(&idx);
            dec_pre = dec_cur;
          }
          while (have) {
            dec_cur = idx;
            size_t delta = dec_cur - dec_pre;
            if (delta != 1) bp.skip_n_byte(delta - 1);
            uint8_t dec = flip(bp.read_byte());
            if (dec > cur_tar) res.set(dec_cur);
            else if (dec == cur_tar) cur_rb.set(dec_cur);
            have = it.next(&idx);
            dec_pre = dec_cur;
          }
          if (len - dec_pre > 1) bp.skip_n_byte(len - dec_pre - 1);
        } else {
          bp.skip_n_byte(len);
          break;
        }
        rb1 = cur_rb;
      }
    } else {
      uint8_t sh = (uint8_t)(BYTE_BITS - remain);
      cur_tar = (uint8_t)((uint8_t)((uint8_t)fixed_target << sh) >> sh);
      bp.finish_read_byte();
      if (rb1.cardinality() != 0) {
        BVIter it(rb1);
        size_t dec_cur = 0, dec_pre = 0, idx;
        bool have = it.next(&idx);
        if (have) {
          dec_cur = idx;
          if (dec_cur != 0) bp.skip(dec_cur * remain);
          uint8_t dec = bp.read_bits(remain);
          if (dec > cur_tar) res.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        while (have) {
          dec_cur = idx;
          size_t delta = dec_cur - dec_pre;
          if (delta != 1) bp.skip((delta - 1) * remain);
          uint8_t dec = bp.read_bits(remain);
          if (dec > cur_tar) res.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        remain = 0;
      } else {
        break;
      }
    }
  }
  fo.count = res.cardinality();
  fo.res = std::move(res);
  return fo;
}

// buff_simd.rs:3549-3733 (buff_simd_equal_filter) over a flipped stream.
FilterOut buff_simd_equal_filter(const uint8_t* bytes, size_t blen, uint64_t scale, double pred) {
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  size_t len = h.len;
  uint32_t remain = h.dlen + h.ilen;
  Bound bound = bound_for(scale, h.dlen);
  BitVec rb1(len), res(len);
  FilterOut fo;
  int64_t fixed_part = bound.fixed(pred);
  if (fixed_part < h.base) {
    fo.res = BitVec(len);
    return fo;
  }
  uint64_t fixed_target = (uint64_t)fixed_part - (uint64_t)h.base;
  uint8_t dec_byte = (uint8_t)fixed_target;
  double check_ratio = 0.0;
  if (remain < 8) {
    for (size_t i = 0; i < len; ++i) {
      uint8_t cur = bp.read_bits(remain);
      if ((uint64_t)cur == fixed_target) res.set(i);
    }
    remain = 0;
  } else {
    remain -= 8;
    const uint8_t* chunk = bp.read_n_byte(len);
    dec_byte = (uint8_t)(fixed_target >> (remain & 63));
    rb1 = equal_simd_mybitvec(chunk, len, dec_byte);
    check_ratio = (double)rb1.cardinality() / (double)len;
  }
  while (remain > 0) {
    if (remain >= 8) {
      remain -= 8;
      if (check_ratio > SIMD_THRESHOLD) {
        const uint8_t* chunk = bp.read_n_byte(len);
        dec_byte = (uint8_t)(fixed_target >> (remain & 63));
        rb1.and_with(equal_simd_mybitvec(chunk, len, dec_byte));
        check_ratio = (double)rb1.cardinality() / (double)len;
      } else {
        BitVec cur_rb(len);
        if (rb1.cardinality() != 0) {
          BVIter it(rb1);
          size_t dec_cur = 0, dec_pre = 0, idx;
          dec_byte = (uint8_t)(fixed_target >> (remain & 63));
          bool have = it.next(&idx);
          if (have) {
            dec_cur = idx;
            if (dec_cur != 0) bp.skip_n_byte(dec_cur);
            uint8_t dec = flip(bp.read_byte());
            if (dec == dec_byte) cur_rb.set(dec_cur);
            have = it.next(&idx);
            dec_pre = dec_cur;
          }
          while (have) {
            dec_cur = idx;
            size_t delta = dec_cur - dec_pre;
            if (delta != 1) bp.skip_n_byte(delta - 1);
            uint8_t dec = flip(bp.read_byte());
            if (dec == dec_byte) cur_rb.set(dec_cur);
            have = it.next(&idx);
            dec_pre = dec_cur;
          }
          if (len - dec_pre > 1) bp.skip_n_byte(len - dec_pre - 1);
        } else {
          bp.skip_n_byte(len);
          break;
        }
        rb1 = cur_rb;
      }
    } else {
      uint8_t sh = (uint8_t)(BYTE_BITS - remain);
      dec_byte = (uint8_t)((uint8_t)((uint8_t)fixed_target << sh) >> sh);
      bp.finish_read_byte();
      if (rb1.cardinality() != 0) {
        BVIter it(rb1);
        size_t dec_cur = 0, dec_pre = 0, idx;
        bool have = it.next(&idx);
        if (have) {
          dec_cur = idx;
          if (dec_cur != 0) bp.skip(dec_cur * remain);
          uint8_t dec = bp.read_bits(remain);
          if (dec == dec_byte) res.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        while (have) {
          dec_cur = idx;
          size_t delta = dec_cur - dec_pre;
          if (delta != 1) bp.skip((delta - 1) * remain);
          uint8_t dec = bp.read_bits(remain);
          if (dec == dec_byte) res.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        remain = 0;
      } else {
        break;
      }
    }
  }
  fo.count = res.cardinality();
  fo.res = std::move(res);
  return fo;
}

// buff_simd.rs:4683-4921 (buff_range_filter_condition): x > pred over the
// rows the iterator yields (`rows`, ascending). The result keeps no other row.
// major: buff_range_filter_majority_condition (buff_simd.rs:1279-1526), the
// same after the column-count byte, and no rows when there are sparse columns
// (1313-1315 log "not implemented" and skip everything).
FilterOut buff_range_filter_condition(const uint8_t* bytes, size_t blen, uint64_t scale, double pred,
                                      const uint64_t* rows, size_t nrows, bool major = false) {
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  size_t len = h.len;
  size_t remain = (size_t)(uint32_t)(h.dlen + h.ilen);   // `(dlen+ilen) as usize`: the u32 sum wraps first
  Bound bound = bound_for(scale, h.dlen);
  BitVec rb1(len), res(len);
  FilterOut fo;
  int64_t fixed_part = bound.fixed(pred);
  if (fixed_part < h.base) {            // 4706-4709: prints len, returns the empty bitmap
    fo.res = std::move(res);
    return fo;
  }
  uint64_t fixed_target = (uint64_t)fixed_part - (uint64_t)h.base;
  uint8_t cur_tar = 0;
  if (major && bp.read(8) > 0) {        // 1308-1315
    fo.res = std::move(res);
    return fo;
  }
  size_t ri = 0;
  auto next_row = [&](size_t* out) {
    if (ri >= nrows) return false;
    *out = (size_t)rows[ri++];
    return true;
  };
  if (remain < 8) {                     // 4714-4752
    size_t dec_cur = 0, dec_pre = 0, idx;
    cur_tar = (uint8_t)fixed_target;    // `fixed_target as u8`
    bool have = next_row(&idx);
    if (have) {
      dec_cur = idx;
      if (dec_cur != 0) bp.skip(dec_cur * remain);
      uint8_t dec = bp.read_bits(remain);
      if (dec > cur_tar) res.set(dec_cur);
      have = next_row(&idx);
      dec_pre = dec_cur;
    }
    while (have) {
      dec_cur = idx;
      size_t delta = dec_cur - dec_pre;
      if (delta != 1) bp.skip((delta - 1) * remain);
      uint8_t dec = bp.read_bits(remain);
      if (dec > cur_tar) res.set(dec_cur);
      have = next_row(&idx);
      dec_pre = dec_cur;
    }
    remain = 0;
  } else {                              // 4753-4800
    remain -= 8;
    size_t dec_cur = 0, dec_pre = 0, idx;
    cur_tar = (uint8_t)(fixed_target >> (remain & 63));
    bool have = next_row(&idx);
    if (have) {
      dec_cur = idx;
      if (dec_cur != 0) bp.skip_n_byte(dec_cur);
      uint8_t dec = bp.read_byte();
      if (dec > cur_tar) res.set(dec_cur);
      else if (dec == cur_tar) rb1.set(dec_cur);
      have = next_row(&idx);
      dec_pre = dec_cur;
    }
    while (have) {
      dec_cur = idx;
      size_t delta = dec_cur - dec_pre;
      if (delta != 1) bp.skip_n_byte(delta - 1);
      uint8_t dec = bp.read_byte();
      if (dec > cur_tar) res.set(dec_cur);
      else if (dec == cur_tar) rb1.set(dec_cur);
      have = next_row(&idx);
      dec_pre = dec_cur;
    }
    if (rb1.cardinality() == 0) {
      fo.count = res.cardinality();
      fo.res = std::move(res);
      return fo;
    }
    if (len - dec_pre > 1) bp.skip_n_byte(len - dec_pre - 1);
  }
  while (remain > 0) {                  // 4805-4918
    if (remain >= 8) {
      remain -= 8;
      if (rb1.cardinality() != 0) {
        BitVec cur_rb(len);
        BVIter it(rb1);
        size_t dec_cur = 0, dec_pre = 0, idx;
        cur_tar = (uint8_t)(fixed_target >> (remain & 63));
        bool have = it.next(&idx);
        if (have) {
          dec_cur = idx;
          if (dec_cur != 0) bp.skip_n_byte(dec_cur);
          uint8_t dec = bp.read_byte();
          if (dec > cur_tar) res.set(dec_cur);
          else if (dec == cur_tar) cur_rb.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        while (have) {
          dec_cur = idx;
          size_t delta = dec_cur - dec_pre;
          if (delta != 1) bp.skip_n_byte(delta - 1);
          uint8_t dec = bp.read_byte();
          if (dec > cur_tar) res.set(dec_cur);
          else if (dec == cur_tar) cur_rb.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        if (len - dec_pre > 1) bp.skip_n_byte(len - dec_pre - 1);
        rb1 = cur_rb;
      } else {
        bp.skip_n_byte(len);
        break;
      }
    } else {
      uint8_t sh = (uint8_t)(BYTE_BITS - remain);
      cur_tar = (uint8_t)((uint8_t)((uint8_t)fixed_target << sh) >> sh);
      bp.finish_read_byte();
      if (rb1.cardinality() != 0) {
        BVIter it(rb1);
        size_t dec_cur = 0, dec_pre = 0, idx;
        bool have = it.next(&idx);
        if (have) {
          dec_cur = idx;
          if (dec_cur != 0) bp.skip(dec_cur * remain);
          uint8_t dec = bp.read_bits(remain);
          if (dec > cur_tar) res.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        while (have) {
          dec_cur = idx;
          size_t delta = dec_cur - dec_pre;
          if (delta != 1) bp.skip((delta - 1) * remain);
          uint8_t dec = bp.read_bits(remain);
          if (dec > cur_tar) res.set(dec_cur);
          have = it.next(&idx);
          dec_pre = dec_cur;
        }
        remain = 0;
      } else {
        break;
      }
    }
  }
  fo.count = res.cardinality();
  fo.res = std::move(res);
  return fo;
}

// buff_simd.rs:1528-1712 (buff_decode_condition): the values of the rows the
// iterator yields (`rows`, ascending), in that order. An empty iterator is the
// reference's unwrap of None. When F % 8 != 0 the tail bits are OR-ed in
// shifted left by the tail width (1693, 1701), not in place.
// major: buff_major_decode_condition (buff_simd.rs:1714-1914): after the
// column-count byte, nothing when there are sparse columns (1738-1740); the
// tail bits in place; an empty condition gives no values instead of a panic.
std::vector<double> buff_decode_condition(const uint8_t* bytes, size_t blen, uint64_t scale, const uint64_t* rows,
                                          size_t nrows, bool major = false) {
  (void)scale;   // 1529-1533 build a PrecisionBound from it that the decode never reads
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  size_t len = h.len;
  std::vector<double> out;
  std::vector<uint64_t> fixed_vec;
  const double dec_scl = powi(2.0, (int32_t)h.dlen);
  size_t remain = (size_t)(uint32_t)(h.dlen + h.ilen);   // `(dlen+ilen) as usize`: the u32 sum wraps first
  if (major) {
    if (bp.read(8) > 0 || nrows == 0) return out;
  } else if (nrows == 0) {
    throw RefPanic("iter.next().unwrap() on an empty condition");
  }
  auto value = [&](uint64_t off) { return (double)(int64_t)((uint64_t)h.base + off) / dec_scl; };
  size_t pre, delta;
  if (remain < 8) {                     // 1559-1577
    pre = rows[0];
    bp.skip(pre * remain);
    uint8_t cur = bp.read_bits(remain);
    out.push_back(value(cur));
    for (size_t i = 1; i < nrows; ++i) {
      size_t entry = rows[i];
      delta = entry - pre;
      if (delta != 1) bp.skip((delta - 1) * remain);
      cur = bp.read_bits(remain);
      out.push_back(value(cur));
      pre = entry;
    }
    return out;
  }
  remain -= 8;
  if (remain == 0) {                    // 1583-1602
    pre = rows[0];
    bp.skip_n_byte(pre);
    uint8_t cur = bp.read_byte();
    out.push_back(value(cur));
    for (size_t i = 1; i < nrows; ++i) {
      size_t entry = rows[i];
      delta = entry - pre;
      if (delta != 1) bp.skip_n_byte(delta - 1);
      cur = bp.read_byte();
      out.push_back(value(cur));
      pre = entry;
    }
    if (len - pre > 1) bp.skip_n_byte(len - pre - 1);
    return out;
  }
  pre = rows[0];                        // 1604-1623
  bp.skip_n_byte(pre);
  uint8_t cur = bp.read_byte();
  fixed_vec.push_back((uint64_t)cur << (remain & 63));
  for (size_t i = 1; i < nrows; ++i) {
    size_t entry = rows[i];
    delta = entry - pre;
    if (delta != 1) bp.skip_n_byte(delta - 1);
    cur = bp.read_byte();
    fixed_vec.push_back((uint64_t)cur << (remain & 63));
    pre = entry;
  }
  if (len - pre > 1) bp.skip_n_byte(len - pre - 1);
  while (remain >= 8) {                 // 1627-1678
    remain -= 8;
    if (remain == 0) {
      pre = rows[0];
      bp.skip_n_byte(pre);
      cur = bp.read_byte();
      out.push_back(value((uint64_t)cur | fixed_vec[0]));
      for (size_t i = 1; i < nrows; ++i) {
        size_t entry = rows[i];
        delta = entry - pre;
        if (delta != 1) bp.skip_n_byte(delta - 1);
        cur = bp.read_byte();
        out.push_back(value((uint64_t)cur | fixed_vec[i]));
        pre = entry;
      }
    } else {
      pre = rows[0];
      bp.skip_n_byte(pre);
      cur = bp.read_byte();
      fixed_vec[0] |= (uint64_t)cur << (remain & 63);
      for (size_t i = 1; i < nrows; ++i) {
        size_t entry = rows[i];
        delta = entry - pre;
        if (delta != 1) bp.skip_n_byte(delta - 1);
        cur = bp.read_byte();
        fixed_vec[i] |= (uint64_t)cur << (remain & 63);
        pre = entry;
      }
      if (len - pre > 1) bp.skip_n_byte(len - pre - 1);
    }
  }
  if (remain > 0) {                     // 1685-1706
    bp.finish_read_byte();
    pre = rows[0];
    bp.skip(pre * remain);
    const uint32_t tsh = major ? 0u : (uint32_t)(remain & 63);   // 1693/1701 shift; 1895/1909 do not
    uint8_t c = bp.read_bits(remain);
    out.push_back(value(fixed_vec[0] | ((uint64_t)c << tsh)));
    for (size_t i = 1; i < nrows; ++i) {
      size_t entry = rows[i];
      delta = entry - pre;
      if (delta != 1) bp.skip((delta - 1) * remain);
      c = bp.read_bits(remain);
      out.push_back(value(fixed_vec[i] | ((uint64_t)c << tsh)));
      pre = entry;
    }
  }
  return out;
}

// croaring Bitmap::remove_range(a..b) on the sorted row set.
void remove_range(std::set<uint32_t>& b, uint32_t lo, uint32_t hi) {
  if (lo >= hi) return;
  b.erase(b.lower_bound(lo), b.lower_bound(hi));
}

// buff_simd.rs:5096-5334 (buff_max_range): the max of each window of `window`
// rows of [s, e) (max_vec, stored offsets) and the croaring bitmap `res` of the
// rows holding it, which only the tail stage (or an F < 8 column) fills. u32
// arithmetic wraps as in release Rust; Vec::get_mut().unwrap(), assert_eq! and
// division by zero are panics.
struct MaxRangeOut {
  std::vector<uint64_t> maxv;
  std::set<uint32_t> res;
};
// major: buff_max_majority_range (buff_simd.rs:2863-3114), the same after the
// column-count byte, and nothing when there are sparse columns (2894-2896).
MaxRangeOut buff_max_range(const uint8_t* bytes, size_t blen, uint32_t s, uint32_t e, uint32_t window,
                           bool major = false) {
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  if (major && bp.read(8) > 0) return MaxRangeOut{};
  const uint32_t len = h.len;
  uint32_t remain = h.dlen + h.ilen;
  std::set<uint32_t> rb1;
  MaxRangeOut o;
  uint64_t max = 0;
  uint8_t byte_max = 0;
  uint32_t cur_s = s;
  auto at = [&](uint32_t cs) -> uint64_t& {
    if (window == 0) throw RefPanic("attempt to divide by zero");
    const size_t k = (size_t)((uint32_t)(cs - s) / window);
    if (k >= o.maxv.size()) throw RefPanic("max_vec.get_mut(..).unwrap() on None");
    return o.maxv[k];
  };
  auto check = [&](uint32_t cs) {
    if (window == 0) throw RefPanic("attempt to divide by zero");
    if ((uint32_t)(cs - s) / window + 1 != (uint32_t)(e - s) / window) throw RefPanic("assert_eq! in buff_max_range");
  };
  if (remain < 8) {                                                   // 5118-5143
    bp.skip((size_t)(uint32_t)(s * remain));
    for (uint32_t i = s; i < e; ++i) {
      if (i == (uint32_t)(cur_s + window)) {
        cur_s = i;
        o.maxv.push_back(max);
        max = 0;
      }
      const uint64_t cur = bp.read_bits(remain);
      if (cur > max) {
        max = cur;
        remove_range(o.res, cur_s, i);
        o.res.insert(i);
      } else if (cur == max) {
        o.res.insert(i);
      }
    }
    if (e == (uint32_t)(cur_s + window)) {
      cur_s = e;
      o.maxv.push_back(max);
      max = 0;
    }
    remain = 0;
  } else {                                                            // 5144-5175
    remain -= 8;
    bp.skip_n_byte(s);
    const uint32_t cnt = e - s;
    const uint8_t* chunk = bp.read_n_byte(cnt);
    uint32_t i = s;
    for (uint32_t k = 0; k < cnt; ++k) {
      const uint8_t c = chunk[k];
      if (i == (uint32_t)(cur_s + window)) {
        cur_s = i;
        o.maxv.push_back((uint64_t)byte_max << (remain & 63));
        byte_max = 0;
      }
      if (c > byte_max) {
        byte_max = c;
        remove_range(rb1, cur_s, i);
        rb1.insert(i);
      } else if (c == byte_max) {
        rb1.insert(i);
      }
      i += 1;
    }
    if (i == (uint32_t)(cur_s + window)) {
      cur_s = i;
      o.maxv.push_back((uint64_t)byte_max << (remain & 63));
      byte_max = 0;
    }
    bp.skip_n_byte((uint32_t)(len - e));
  }
  while (remain > 0) {                                                // 5177-5331
    byte_max = 0;
    uint32_t cs = s;
    if (remain >= 8) {
      remain -= 8;
      uint32_t dec_cur = 0, dec_pre = 0;
      std::set<uint32_t> cur_rb;
      auto it = rb1.begin();
      if (it != rb1.end()) {
        dec_cur = *it;
        if (dec_cur != 0) bp.skip_n_byte(dec_cur);
        const uint8_t dec = bp.read_byte();
        if (dec > byte_max) {
          byte_max = dec;
          cur_rb.insert(dec_cur);
        } else if (dec == byte_max) {
          cur_rb.insert(dec_cur);
        }
        ++it;
        dec_pre = dec_cur;
      }
      for (; it != rb1.end(); ++it) {
        dec_cur = *it;
        if (dec_cur >= (uint32_t)(cs + window)) {
          at(cs) |= (uint64_t)byte_max << (remain & 63);
          byte_max = 0;
          cs = cs + window;
        }
        const uint32_t delta = dec_cur - dec_pre;
        if (delta != 1) bp.skip_n_byte((uint32_t)(delta - 1));
        const uint8_t dec = bp.read_byte();
        if (dec > byte_max) {
          byte_max = dec;
          remove_range(cur_rb, cs, dec_cur);
          cur_rb.insert(dec_cur);
        } else if (dec == byte_max) {
          cur_rb.insert(dec_cur);
        }
        dec_pre = dec_cur;
      }
      check(cs);
      at(cs) |= (uint64_t)byte_max << (remain & 63);
      byte_max = 0;
      cs = cs + window;
      if ((uint32_t)(len - dec_pre) > 1) bp.skip_n_byte((uint32_t)(len - dec_pre - 1));
      rb1 = cur_rb;
    } else {
      byte_max = 0;
      bp.finish_read_byte();
      uint32_t dec_cur = 0, dec_pre = 0;
      auto it = rb1.begin();
      if (it != rb1.end()) {
        dec_cur = *it;
        if (dec_cur != 0) bp.skip((size_t)(uint32_t)(dec_cur * remain));
        const uint8_t dec = bp.read_bits(remain);
        if (dec > byte_max) {
          byte_max = dec;
          o.res.insert(dec_cur);
        } else if (dec == byte_max) {
          o.res.insert(dec_cur);
        }
        ++it;
        dec_pre = dec_cur;
      }
      for (; it != rb1.end(); ++it) {
        dec_cur = *it;
        if (dec_cur >= (uint32_t)(cs + window)) {
          at(cs) |= (uint64_t)byte_max;
          byte_max = 0;
          cs = cs + window;
        }
        const uint32_t delta = dec_cur - dec_pre;
        if (delta != 1) bp.skip((size_t)(uint32_t)((delta - 1) * remain));
        const uint8_t dec = bp.read_bits(remain);
        if (dec > byte_max) {
          byte_max = dec;
          remove_range(o.res, cs, dec_cur);
          o.res.insert(dec_cur);
        } else if (dec == byte_max) {
          o.res.insert(dec_cur);
        }
        dec_pre = dec_cur;
      }
      check(cs);
      at(cs) |= (uint64_t)byte_max;
      byte_max = 0;
      cs = cs + window;
      remain = 0;
    }
  }
  return o;
}

int copy_filter(const FilterOut& fo, uint32_t* words, uint64_t nwords_cap, uint64_t* count, uint64_t* cand) {
  if (words) {
    if (fo.res.storage.size() > nwords_cap) return REF_SMALL_BUFFER;
    std::memcpy(words, fo.res.storage.data(), fo.res.storage.size() * 4);
  }
  if (count) *count = fo.count;
  if (cand) std::memcpy(cand, fo.cand, sizeof(fo.cand));
  return REF_OK;
}


// ===========================================================================
// buff-major: majority / outlier sparse columns (buff_simd.rs:317-2861).

constexpr size_t SAMPLE = 2000;          // split_double.rs:25
constexpr float MAJOR_R = 0.6f;          // split_double.rs:27

std::vector<uint8_t> snappy_compress(const std::vector<uint8_t>& in) {   // bit_vec_iter.rs:61-64
  return snappy_ref::compress(in.data(), in.size());
}
std::vector<uint8_t> snappy_decompress(const uint8_t* p, size_t n) {     // bit_vec_iter.rs:69-71 (unwrap)
  try {
    return snappy_ref::decompress(p, n);
  } catch (const snappy_ref::Error& e) {
    throw RefPanic(e.what());
  }
}

// The Boyer-Moore vote of buff_simd.rs:402-412 / 424-436 over `(sample >> rsf) as u8`.
void majority_vote(const std::vector<uint64_t>& head, size_t rsf, uint8_t* major, int32_t* frequency) {
  int32_t f = 0;
  uint8_t m = 0;
  for (uint64_t smp : head) {
    uint8_t u = (uint8_t)(smp >> (rsf & 63));
    if (f == 0) { m = u; f = 1; }
    else if (m == u) f += 1;
    else f -= 1;
  }
  *major = m;
  *frequency = f;
}

// buff_simd.rs:317-630 (buff_encode_majority).
std::vector<uint8_t> buff_encode_majority(const double* data, size_t n, uint64_t scale, int64_t* out_min,
                                          int64_t* out_max) {
  std::vector<int64_t> fixed_vec;
  fixed_vec.reserve(n);
  uint32_t t = (uint32_t)n;
  int32_t prec = scale == 0 ? 0 : prec_of_scale(scale);
  double prec_delta = get_precision_bound(prec);
  int32_t pexp = precision_exp_of(prec_delta);
  uint64_t dec_len = (uint64_t)precision_map(prec);
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  for (size_t i = 0; i < n; ++i) {          // 340-349: `if` / `if`
    int64_t f = fetch_fixed_aligned(data[i], pexp, dec_len);
    if (f < mn) mn = f;
    if (f > mx) mx = f;
    fixed_vec.push_back(f);
  }
  if (out_min) *out_min = mn;
  if (out_max) *out_max = mx;
  int64_t delta = (int64_t)((uint64_t)mx - (uint64_t)mn);
  int64_t base_fixed = mn;
  uint64_t ubase = (uint64_t)base_fixed;
  double cal_int_length = 0.0;
  if (delta != 0) cal_int_length = std::ceil(std::log2((double)delta));
  size_t fixed_len = (size_t)f64_as_usize(cal_int_length);
  size_t ilen = fixed_len - (size_t)dec_len;
  size_t dlen = (size_t)dec_len;

  BitPackWriter bp;
  bp.write((uint32_t)ubase, 32);
  bp.write((uint32_t)(ubase >> 32), 32);
  bp.write(t, 32);
  bp.write((uint32_t)ilen, 32);
  bp.write((uint32_t)dlen, 32);

  size_t remain = fixed_len;
  std::vector<uint8_t> freq_vec;
  size_t rsf = fixed_len >= 8 ? fixed_len - 8 : 0;
  int32_t m_limit = (int32_t)((float)SAMPLE * MAJOR_R);
  if (n < SAMPLE) throw RefPanic("fixed_vec[0..SAMPLE] out of range");     // 397
  std::vector<uint64_t> head(SAMPLE);
  for (size_t i = 0; i < SAMPLE; ++i) head[i] = (uint64_t)fixed_vec[i] - (uint64_t)base_fixed;
  uint8_t major;
  int32_t frequency;
  majority_vote(head, rsf, &major, &frequency);
  if (frequency > m_limit) {
    freq_vec.push_back(major);
    while (rsf > 7) {
      rsf -= 8;
      majority_vote(head, rsf, &major, &frequency);
      if (frequency > m_limit) freq_vec.push_back(major);
      else break;
    }
  }
  uint32_t cols_outlier = (uint32_t)freq_vec.size();
  if (cols_outlier > 0) {
    // With fixed_len < 8 `remain -= 8` wraps: a debug build panics, a release
    // build's `while remain >= 8` then writes ~2^61 planes.
    if (remain < 8) throw RefPanic("remain -= 8 underflows (F < 8 with a majority byte)");
    bp.write(cols_outlier, 8);
    std::vector<uint64_t> fixed_u64(n);
    for (size_t i = 0; i < n; ++i) fixed_u64[i] = (uint64_t)fixed_vec[i] - (uint64_t)base_fixed;
    for (uint32_t col = 0; col < cols_outlier; ++col) {
      uint8_t f_item = freq_vec[col];
      remain -= 8;
      BitVec bm_outlier(n);
      std::vector<uint8_t> outlier;
      for (size_t i = 0; i < n; ++i) {
        uint8_t cur = (uint8_t)(fixed_u64[i] >> (remain & 63));
        if (cur != f_item) {
          outlier.push_back(cur);
          bm_outlier.set(i);
        }
      }
      uint32_t n_outlier = bm_outlier.cardinality();
      std::vector<uint8_t> bm_vec = snappy_compress(bm_outlier.to_binary());
      bp.write(f_item, 8);
      bp.write((uint32_t)bm_vec.size(), 32);
      bp.buff.insert(bp.buff.end(), bm_vec.begin(), bm_vec.end());   // write_bytes: append, cursor kept
      bp.finish_write_byte();
      bp.write(n_outlier, 32);
      bp.buff.insert(bp.buff.end(), outlier.begin(), outlier.end());
      bp.finish_write_byte();
    }
    while (remain >= 8) {
      remain -= 8;
      for (size_t i = 0; i < n; ++i)
        bp.write_byte(remain > 0 ? (uint8_t)(fixed_u64[i] >> (remain & 63)) : (uint8_t)fixed_u64[i]);
    }
    if (remain > 0) {
      bp.finish_write_byte();
      for (size_t i = 0; i < n; ++i) bp.write_bits((uint32_t)fixed_u64[i], remain);
    }
  } else {
    bp.write_byte(0);
    if (remain < 8) {
      // write_bits starts at the cursor, which still points at the byte just pushed.
      for (size_t i = 0; i < n; ++i) bp.write_bits((uint32_t)((uint64_t)fixed_vec[i] - (uint64_t)base_fixed), remain);
    } else {
      remain -= 8;
      std::vector<uint64_t> fixed_u64(n);
      for (size_t i = 0; i < n; ++i) {
        uint64_t cur = (uint64_t)fixed_vec[i] - (uint64_t)base_fixed;
        bp.write_byte(remain > 0 ? (uint8_t)(cur >> (remain & 63)) : (uint8_t)cur);
        fixed_u64[i] = cur;
      }
      while (remain >= 8) {
        remain -= 8;
        for (size_t i = 0; i < n; ++i)
          bp.write_byte(remain > 0 ? (uint8_t)(fixed_u64[i] >> (remain & 63)) : (uint8_t)fixed_u64[i]);
      }
      if (remain > 0) {
        bp.finish_write_byte();
        for (size_t i = 0; i < n; ++i) bp.write_bits((uint32_t)fixed_u64[i], remain);
      }
    }
  }
  return bp.buff;
}

// One sparse column as the readers walk it (e.g. buff_simd.rs:672-681).
struct SparseCol {
  uint8_t major;
  BitVec bitmap;
  uint32_t n_outlier;
  const uint8_t* outliers;
};
SparseCol read_sparse_col(BitPackReader& bp, bool skip_outliers) {
  SparseCol c;
  c.major = (uint8_t)bp.read(8);
  uint32_t bm_len = bp.read(32);
  const uint8_t* bm_mat = bp.read_n_byte(bm_len);
  c.bitmap = BitVec::from_binary(snappy_decompress(bm_mat, bm_len));
  bp.finish_read_byte();
  c.n_outlier = bp.read(32);
  if (skip_outliers) {
    bp.skip_n_byte(c.n_outlier);
    c.outliers = nullptr;
  } else {
    c.outliers = bp.read_n_byte(c.n_outlier);
  }
  bp.finish_read_byte();
  return c;
}

// A majority stream whose bit width is below 8: only the no-majority encoder
// branch writes one, and it ORs the tail bits into the outlier-column count
// byte (buff_simd.rs:550-555), so the readers misparse it.
void reject_narrow(uint32_t remain) {
  if (remain < 8) throw RefPanic("buff-major stream with F < 8 cannot be read back");
}

// buff_simd.rs:633-861 (buff_decode_majority).
std::vector<double> buff_decode_majority(const uint8_t* bytes, size_t blen, uint64_t scale) {
  (void)scale;
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  std::vector<double> out;
  std::vector<uint64_t> fixed_vec;
  double dec_scl = powi(2.0, (int32_t)h.dlen);
  uint32_t remain = h.dlen + h.ilen;
  reject_narrow(remain);
  auto value = [&](uint64_t f) { return (double)(int64_t)((uint64_t)h.base + f) / dec_scl; };
  uint32_t outlier_cols = bp.read(8);
  if (outlier_cols > 0) {
    for (uint32_t col = 0; col < outlier_cols; ++col) {
      SparseCol c = read_sparse_col(bp, false);
      remain -= 8;
      if (col == 0) fixed_vec.assign(h.len, (uint64_t)c.major << (remain & 63));
      uint64_t major_u64 = (uint64_t)c.major << (remain & 63);
      if (c.n_outlier == 0) throw RefPanic("outlier_iter.next().unwrap() on no outliers");
      size_t oi = 0;
      uint8_t outlier_value = c.outliers[oi++];
      BVIter bm(c.bitmap);
      size_t ind;
      if (!bm.next(&ind)) throw RefPanic("bm_iterator.next().unwrap()");
      size_t i = 0, counter = 0;
      for (uint64_t& x : fixed_vec) {
        if (i == ind) {
          if (col == 0) x = (uint64_t)outlier_value << (remain & 63);
          else x |= (uint64_t)outlier_value << (remain & 63);
          counter += 1;
          if (counter == c.n_outlier) break;
          if (!bm.next(&ind)) throw RefPanic("bm_iterator.next().unwrap()");
          outlier_value = c.outliers[oi++];
        } else if (col > 0) {
          x |= major_u64 << (remain & 63);   // 733: the major byte is shifted twice
        }
        i += 1;
      }
    }
  } else if (remain >= 8) {
    remain -= 8;
    const uint8_t* chunk = bp.read_n_byte(h.len);
    if (remain == 0) {
      for (uint32_t i = 0; i < h.len; ++i) out.push_back(value(chunk[i]));
    } else {
      for (uint32_t i = 0; i < h.len; ++i) fixed_vec.push_back((uint64_t)chunk[i] << (remain & 63));
    }
  }
  while (remain >= 8) {
    remain -= 8;
    const uint8_t* chunk = bp.read_n_byte(h.len);
    if (remain == 0) {
      size_t m = std::min(fixed_vec.size(), (size_t)h.len);
      for (size_t i = 0; i < m; ++i) out.push_back(value(fixed_vec[i] | (uint64_t)chunk[i]));
    } else {
      for (size_t i = 0; i < fixed_vec.size(); ++i) {
        if (i >= h.len) throw RefPanic("chunk iterator exhausted");
        fixed_vec[i] |= (uint64_t)chunk[i] << (remain & 63);
      }
    }
  }
  if (remain > 0) {
    bp.finish_read_byte();
    for (size_t i = 0; i < fixed_vec.size(); ++i) out.push_back(value(fixed_vec[i] | (uint64_t)bp.read_bits(remain)));
  }
  return out;
}

// The candidate walk over one dense plane shared by the majority filters
// (buff_simd.rs:990-1052): reads the plane bytes of the rb1 rows through the
// reader, returns false (the caller breaks) when rb1 is empty.
template <typename Visit>
bool walk_plane(BitPackReader& bp, const BitVec& rb1, size_t len, Visit visit) {
  if (rb1.cardinality() == 0) {
    bp.skip_n_byte(len);
    return false;
  }
  BVIter it(rb1);
  size_t idx, dec_pre = 0;
  bool first = true;
  while (it.next(&idx)) {
    if (first) {
      if (idx != 0) bp.skip_n_byte(idx);
      first = false;
    } else {
      size_t delta = idx - dec_pre;
      if (delta != 1) bp.skip_n_byte(delta - 1);
    }
    visit(idx, bp.read_byte());
    dec_pre = idx;
  }
  if (len - dec_pre > 1) bp.skip_n_byte(len - dec_pre - 1);
  return true;
}
// The tail walk (buff_simd.rs:1092-1130): `remain`-bit values of the rb1 rows.
template <typename Visit>
bool walk_tail(BitPackReader& bp, const BitVec& rb1, uint32_t remain, Visit visit) {
  bp.finish_read_byte();
  if (rb1.cardinality() == 0) return false;
  BVIter it(rb1);
  size_t idx, dec_pre = 0;
  bool first = true;
  while (it.next(&idx)) {
    if (first) {
      if (idx != 0) bp.skip(idx * remain);
      first = false;
    } else {
      size_t delta = idx - dec_pre;
      if (delta != 1) bp.skip((delta - 1) * remain);
    }
    visit(idx, bp.read_bits(remain));
    dec_pre = idx;
  }
  return true;
}

// buff_simd.rs:863-1277 (buff_range_filter_majority): x > pred.
FilterOut buff_range_filter_majority(const uint8_t* bytes, size_t blen, uint64_t scale, double pred) {
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  size_t len = h.len;
  uint32_t remain = h.dlen + h.ilen;
  Bound bound = bound_for(scale, h.dlen);
  BitVec rb1(len), res(len);
  FilterOut fo;
  int64_t fixed_part = bound.fixed(pred);
  if (fixed_part < h.base) {              // 886-889: prints len, no bitmap
    fo.res = BitVec(len);
    for (size_t i = 0; i < len; ++i) fo.res.set(i);
    fo.count = len;
    return fo;
  }
  reject_narrow(remain);
  uint64_t fixed_target = (uint64_t)fixed_part - (uint64_t)h.base;
  uint32_t outlier_cols = bp.read(8);
  if (outlier_cols > 0) {
    for (uint32_t col = 0; col < outlier_cols; ++col) {
      remain -= 8;
      uint8_t cur_tar = (uint8_t)(fixed_target >> (remain & 63));
      uint8_t major = (uint8_t)bp.peek8();
      if (cur_tar == 0 && major == 0) {
        SparseCol c = read_sparse_col(bp, true);
        if (col == 0) {                    // 903-915
          res.or_with(c.bitmap);
          c.bitmap.negate();
          rb1.or_with(c.bitmap);
        } else {                           // 971-984: rb1 ends empty
          rb1.and_with(c.bitmap);
          res.or_with(rb1);
          c.bitmap.negate();
          rb1.and_with(c.bitmap);
        }
      } else {
        SparseCol c = read_sparse_col(bp, false);
        BVIter bm(c.bitmap);
        BitVec cur_rb(len);
        BitVec& eq_target = col == 0 ? rb1 : cur_rb;
        for (uint32_t k = 0; k < c.n_outlier; ++k) {
          size_t ind;
          if (!bm.next(&ind)) throw RefPanic("bm_iterator.next().unwrap()");
          uint8_t val = c.outliers[k];
          if (val == cur_tar) eq_target.set(ind);
          else if (val > cur_tar) res.set(ind);   // rows after column 0: not masked by rb1
        }
        c.bitmap.negate();
        if (col == 0) {                    // 941-946
          if (cur_tar < major) res.or_with(c.bitmap);
          else if (cur_tar == major) rb1.or_with(c.bitmap);
        } else {                           // 1011-1019
          if (cur_tar < major) {
            c.bitmap.and_with(rb1);
            res.or_with(c.bitmap);
          } else if (cur_tar == major) {
            cur_rb.or_with(c.bitmap);
          }
          rb1.and_with(cur_rb);
        }
      }
    }
  } else if (remain < 8) {
    for (size_t i = 0; i < len; ++i) {
      uint8_t cur = bp.read_bits(remain);
      if ((uint64_t)cur > fixed_target) res.set(i);
    }
    remain = 0;
  } else {
    remain -= 8;
    const uint8_t* chunk = bp.read_n_byte(len);
    uint8_t cur_tar = (uint8_t)(fixed_target >> (remain & 63));
    for (size_t i = 0; i < len; ++i) {
      if (chunk[i] > cur_tar) res.set(i);
      else if (chunk[i] == cur_tar) rb1.set(i);
    }
  }
  while (remain > 0) {
    if (remain >= 8) {
      remain -= 8;
      uint8_t cur_tar = (uint8_t)(fixed_target >> (remain & 63));
      BitVec cur_rb(len);
      if (!walk_plane(bp, rb1, len, [&](size_t i, uint8_t dec) {
            if (dec > cur_tar) res.set(i);
            else if (dec == cur_tar) cur_rb.set(i);
          }))
        break;
      rb1 = std::move(cur_rb);
    } else {
      uint8_t sh = (uint8_t)(8 - remain);
      uint8_t cur_tar = (uint8_t)((uint8_t)((uint8_t)fixed_target << sh) >> sh);
      if (!walk_tail(bp, rb1, remain, [&](size_t i, uint8_t dec) {
            if (dec > cur_tar) res.set(i);
          }))
        break;
      remain = 0;
    }
  }
  fo.count = res.cardinality();
  fo.res = std::move(res);
  return fo;
}

// buff_simd.rs:1916-2292 (buff_equal_filter_majority).
FilterOut buff_equal_filter_majority(const uint8_t* bytes, size_t blen, uint64_t scale, double pred) {
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  size_t len = h.len;
  uint32_t remain = h.dlen + h.ilen;
  Bound bound = bound_for(scale, h.dlen);
  BitVec rb1(len), res(len);
  FilterOut fo;
  int64_t fixed_part = bound.fixed(pred);
  if (fixed_part < h.base) {              // 1939-1942
    fo.res = BitVec(len);
    fo.count = 0;
    return fo;
  }
  reject_narrow(remain);
  uint64_t fixed_target = (uint64_t)fixed_part - (uint64_t)h.base;
  uint32_t outlier_cols = bp.read(8);
  if (outlier_cols > 0) {
    for (uint32_t col = 0; col < outlier_cols; ++col) {
      remain -= 8;
      uint8_t dec_byte = (uint8_t)(fixed_target >> (remain & 63));
      uint8_t major = (uint8_t)bp.peek8();
      if (dec_byte == major) {
        SparseCol c = read_sparse_col(bp, true);
        c.bitmap.negate();
        if (col == 0) rb1.or_with(c.bitmap);
        else rb1.and_with(c.bitmap);
      } else {
        SparseCol c = read_sparse_col(bp, false);
        BVIter bm(c.bitmap);
        BitVec cur_rb(len);
        BitVec& hit = col == 0 ? rb1 : cur_rb;
        for (uint32_t k = 0; k < c.n_outlier; ++k) {
          size_t ind;
          if (!bm.next(&ind)) throw RefPanic("bm_iterator.next().unwrap()");
          if (c.outliers[k] == dec_byte) hit.set(ind);
        }
        if (col > 0) rb1.and_with(cur_rb);
      }
    }
  } else if (remain < 8) {
    for (size_t i = 0; i < len; ++i) {
      uint8_t cur = bp.read_bits(remain);
      if ((uint64_t)cur == fixed_target) res.set(i);
    }
    remain = 0;
  } else {
    remain -= 8;
    const uint8_t* chunk = bp.read_n_byte(len);
    uint8_t dec_byte = (uint8_t)(fixed_target >> (remain & 63));
    for (size_t i = 0; i < len; ++i)
      if (chunk[i] == dec_byte) rb1.set(i);
  }
  while (remain > 0) {
    if (remain >= 8) {
      remain -= 8;
      uint8_t dec_byte = (uint8_t)(fixed_target >> (remain & 63));
      BitVec cur_rb(len);
      if (!walk_plane(bp, rb1, len, [&](size_t i, uint8_t dec) {
            if (dec == dec_byte) cur_rb.set(i);
          }))
        break;
      rb1 = std::move(cur_rb);
    } else {
      uint8_t sh = (uint8_t)(8 - remain);
      uint8_t dec_byte = (uint8_t)((uint8_t)((uint8_t)fixed_target << sh) >> sh);
      if (!walk_tail(bp, rb1, remain, [&](size_t i, uint8_t dec) {
            if (dec == dec_byte) res.set(i);
          }))
        break;
      remain = 0;
    }
  }
  fo.count = res.cardinality();
  fo.res = std::move(res);
  return fo;
}

// buff_simd.rs:2294-2442 (buff_sum_majority). `col_sums` receives the exact
// per-byte-column sums (sparse and dense columns in stream order, then the tail).
double buff_sum_majority(const uint8_t* bytes, size_t blen, uint64_t scale, uint64_t* col_sums) {
  (void)scale;
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  uint32_t remain = h.ilen + h.dlen;
  reject_narrow(remain);
  int32_t processed = 0;
  double sum_base = (double)h.len * (double)h.base;
  double sum = 0.0;
  double dec_scl = powi(2.0, (int32_t)h.dlen);
  sum_base = sum_base / dec_scl;
  int k = 0;
  uint32_t outlier_cols = bp.read(8);
  for (uint32_t col = 0; col < outlier_cols; ++col) {
    uint8_t major = (uint8_t)bp.read(8);
    remain -= 8;
    processed += 8;
    uint32_t bm_len = bp.read(32);
    bp.read_n_byte(bm_len);
    bp.finish_read_byte();
    uint32_t n_outlier = bp.read(32);
    const uint8_t* outliers = bp.read_n_byte(n_outlier);
    dec_scl = powi(2.0, processed - (int32_t)h.ilen);
    uint64_t sum_fixed = 0;
    for (uint32_t j = 0; j < n_outlier; ++j) sum_fixed += outliers[j];
    sum_fixed += (uint64_t)major * ((uint64_t)h.len - (uint64_t)n_outlier);
    if (col_sums) col_sums[k++] = sum_fixed;
    sum = sum + (double)sum_fixed / dec_scl;
    bp.finish_read_byte();
  }
  while (remain >= 8) {
    remain -= 8;
    processed += 8;
    dec_scl = powi(2.0, processed - (int32_t)h.ilen);
    const uint8_t* chunk = bp.read_n_byte(h.len);
    uint64_t sum_fixed = 0;
    for (uint32_t i = 0; i < h.len; ++i) sum_fixed += chunk[i];
    if (col_sums) col_sums[k++] = sum_fixed;
    sum = sum + (double)sum_fixed / dec_scl;
    if (remain == 0) return sum;           // 2370-2372 / 2412-2414: no base term
  }
  if (remain > 0) {
    if (processed >= 8) bp.finish_read_byte();
    processed += (int32_t)remain;
    dec_scl = powi(2.0, processed - (int32_t)h.ilen);
    uint64_t sum_fixed = 0;
    for (uint32_t i = 0; i < h.len; ++i) sum_fixed += bp.read_bits(remain);
    if (col_sums) col_sums[k++] = sum_fixed;
    sum = sum + (double)sum_fixed / dec_scl;
  }
  sum += sum_base;
  return sum;
}

// buff_simd.rs:2444-2861 (buff_max_majority). Returns the printed max value.
double buff_max_majority(const uint8_t* bytes, size_t blen, uint64_t scale, uint64_t* max_off,
                         uint64_t* argmax_count) {
  (void)scale;
  BitPackReader bp(bytes, blen);
  Header h = read_header(bp);
  size_t len = h.len;
  uint32_t remain = h.dlen + h.ilen;
  reject_narrow(remain);
  BitVec rb1(len), res(len);
  uint64_t max = 0;
  uint32_t outlier_cols = bp.read(8);
  if (outlier_cols > 0) {
    for (uint32_t col = 0; col < outlier_cols; ++col) {
      SparseCol c = read_sparse_col(bp, false);
      remain -= 8;
      uint8_t byte_max = c.major;
      BVIter bm(c.bitmap);
      if (col == 0) {                      // 2481-2511
        rb1 = c.bitmap;
        rb1.negate();
        for (uint32_t k = 0; k < c.n_outlier; ++k) {
          size_t ind;
          if (!bm.next(&ind)) throw RefPanic("bm_iterator.next().unwrap()");
          uint8_t val = c.outliers[k];
          if (val > byte_max) {
            byte_max = val;
            rb1.clear();
            rb1.set(ind);
          } else if (val == byte_max) {
            rb1.set(ind);
          }
        }
        max = max | ((uint64_t)byte_max << (remain & 63));
        continue;
      }
      BitVec cur_rb = c.bitmap;              // 2522-2532
      cur_rb.negate();
      cur_rb.and_with(rb1);
      rb1.and_with(c.bitmap);
      if (rb1.cardinality() == 0) {          // 2534-2538: no candidate is an outlier
        rb1 = std::move(cur_rb);
        continue;
      }
      BVIter mi(rb1);
      size_t max_cur = 0;
      mi.next(&max_cur);
      for (uint32_t k = 0; k < c.n_outlier; ++k) {
        size_t ind;
        if (!bm.next(&ind)) throw RefPanic("bm_iterator.next().unwrap()");
        if (ind == max_cur) {
          uint8_t val = c.outliers[k];
          if (val > byte_max) {
            byte_max = val;
            cur_rb.clear();
            cur_rb.set(ind);
          } else if (val == byte_max) {
            cur_rb.set(ind);
          }
          if (!mi.next(&max_cur)) break;
        }
      }
      rb1 = std::move(cur_rb);
      max = max | ((uint64_t)byte_max << (remain & 63));
    }
  } else if (remain < 8) {
    for (uint32_t i = 0; i < len; ++i) {
      uint64_t cur = bp.read_bits(remain);
      if (cur > max) {
        max = cur;
        res.clear();
        res.set(i);
      } else if (cur == max) {
        res.set(i);
      }
    }
    remain = 0;
  } else {
    remain -= 8;
    const uint8_t* chunk = bp.read_n_byte(len);
    uint8_t byte_max = 0;
    for (size_t i = 0; i < len; ++i) {
      if (chunk[i] > byte_max) {
        byte_max = chunk[i];
        rb1.clear();
        rb1.set(i);
      } else if (chunk[i] == byte_max) {
        rb1.set(i);
      }
    }
    max = (uint64_t)byte_max << (remain & 63);
  }
  while (remain > 0) {
    uint8_t byte_max = 0;
    if (remain >= 8) {
      remain -= 8;
      BitVec cur_rb(len);
      // 2597-2650: an empty rb1 still skips len - 1 bytes
      BVIter it(rb1);
      size_t idx, dec_pre = 0;
      bool first = true;
      while (it.next(&idx)) {
        if (first) {
          if (idx != 0) bp.skip_n_byte(idx);
        } else {
          size_t delta = idx - dec_pre;
          if (delta != 1) bp.skip_n_byte(delta - 1);
        }
        uint8_t dec = bp.read_byte();
        if (dec > byte_max) {
          byte_max = dec;
          if (!first) cur_rb.clear();
          cur_rb.set(idx);
        } else if (dec == byte_max) {
          cur_rb.set(idx);
        }
        first = false;
        dec_pre = idx;
      }
      if (len - dec_pre > 1) bp.skip_n_byte(len - dec_pre - 1);
      max = max | ((uint64_t)byte_max << (remain & 63));
      rb1 = std::move(cur_rb);
    } else {
      bp.finish_read_byte();
      BVIter it(rb1);
      size_t idx, dec_pre = 0;
      bool first = true;
      while (it.next(&idx)) {
        if (first) {
          if (idx != 0) bp.skip(idx * remain);
        } else {
          size_t delta = idx - dec_pre;
          if (delta != 1) bp.skip((delta - 1) * remain);
        }
        uint8_t dec = bp.read_bits(remain);
        if (dec > byte_max) {
          byte_max = dec;
          if (!first) res.clear();
          res.set(idx);
        } else if (dec == byte_max) {
          res.set(idx);
        }
        first = false;
        dec_pre = idx;
      }
      remain = 0;
      max = max | (uint64_t)byte_max;
    }
  }
  if (max_off) *max_off = max;
  if (argmax_count) *argmax_count = res.cardinality();
  return (double)(int64_t)(max + (uint64_t)h.base) / powi(2.0, (int32_t)h.dlen);
}

#include "refslice.inc"
#include "refrapg.inc"

}  // namespace

// ===========================================================================
// C entry points (loaded with ctypes by the tests and the bench's CPU leg).
extern "C" {

int ref_prec_of_scale(uint64_t scale) { return prec_of_scale(scale); }
double ref_get_precision_bound(int32_t p) { return get_precision_bound(p); }
int32_t ref_precision_exp(double precision) { return precision_exp_of(precision); }
int ref_precision_map(int32_t p, int32_t* dlen) {
  try { *dlen = precision_map(p); return REF_OK; } catch (const RefPanic&) { return REF_PANIC; }
}
int64_t ref_fetch_fixed_aligned(double x, int32_t precision_exp, uint64_t dlen) {
  return fetch_fixed_aligned(x, precision_exp, dlen);
}
double ref_powi(double a, int32_t b) { return powi(a, b); }

// Sequence of BitPack::<Vec<u8>>::write(v, bits) calls (bits<=32; kind[i]==1
// selects write_bits). Returns the buffer in out (cap bytes).
int ref_bitpack_write(const uint32_t* values, const uint32_t* bits, const uint8_t* kind, uint64_t k,
                      uint8_t* out, uint64_t cap, uint64_t* out_len) {
  try {
    BitPackWriter w;
    for (uint64_t i = 0; i < k; ++i) {
      if (kind && kind[i] == 1) w.write_bits(values[i], bits[i]);
      else if (kind && kind[i] == 2) w.write_byte((uint8_t)values[i]);
      else if (kind && kind[i] == 3) w.finish_write_byte();
      else w.write(values[i], bits[i]);
    }
    *out_len = w.buff.size();
    if (w.buff.size() > cap) return REF_SMALL_BUFFER;
    std::memcpy(out, w.buff.data(), w.buff.size());
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

int ref_bitpack_read(const uint8_t* buf, uint64_t len, const uint32_t* bits, uint64_t k, uint32_t* values) {
  try {
    BitPackReader r(buf, len);
    for (uint64_t i = 0; i < k; ++i) values[i] = r.read(bits[i]);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

// byte_fixed_encode / buff_simd256_encode. Call with out==NULL for the size.
int ref_encode(const double* x, uint64_t n, uint64_t scale, int flip_layout, uint8_t* out, uint64_t cap,
               uint64_t* out_len, int64_t* minmax) {
  try {
    int64_t mn, mx;
    std::vector<uint8_t> s = byte_fixed_encode(x, n, scale, flip_layout != 0, &mn, &mx);
    if (minmax) { minmax[0] = mn; minmax[1] = mx; }
    *out_len = s.size();
    if (!out) return REF_OK;
    if (s.size() > cap) return REF_SMALL_BUFFER;
    std::memcpy(out, s.data(), s.size());
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

int ref_decode(const uint8_t* s, uint64_t len, uint64_t scale, int flip_layout, double* out, uint64_t cap,
               uint64_t* n_out) {
  try {
    std::vector<double> v = byte_fixed_decode(s, len, scale, flip_layout != 0);
    *n_out = v.size();
    if (v.size() > cap) return REF_SMALL_BUFFER;
    std::memcpy(out, v.data(), v.size() * 8);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

// op: 0 = range (x > pred), 1 = equal. words: ceil(n/32) u32; cand: 9 u64.
int ref_filter(const uint8_t* s, uint64_t len, uint64_t scale, int op, double pred, uint32_t* words,
               uint64_t nwords_cap, uint64_t* count, uint64_t* cand) {
  try {
    FilterOut fo = op == 0 ? buff_range_filter(s, len, scale, pred) : buff_equal_filter(s, len, scale, pred);
    return copy_filter(fo, words, nwords_cap, count, cand);
  } catch (const RefPanic&) { return REF_PANIC; }
}

// buff_range_filter_condition (major: _majority_condition) over the ascending row ids `rows`.
int ref_filter_cond(const uint8_t* s, uint64_t len, uint64_t scale, double pred, const uint64_t* rows, uint64_t nrows,
                    int major, uint32_t* words, uint64_t nwords_cap, uint64_t* count) {
  try {
    FilterOut fo = buff_range_filter_condition(s, len, scale, pred, rows, nrows, major != 0);
    return copy_filter(fo, words, nwords_cap, count, nullptr);
  } catch (const RefPanic&) { return REF_PANIC; }
}

// buff_decode_condition (major: buff_major_decode_condition) over the ascending row ids `rows`.
int ref_decode_cond(const uint8_t* s, uint64_t len, uint64_t scale, const uint64_t* rows, uint64_t nrows, int major,
                    double* out, uint64_t cap, uint64_t* n_out) {
  try {
    std::vector<double> v = buff_decode_condition(s, len, scale, rows, nrows, major != 0);
    *n_out = v.size();
    if (v.size() > cap) return REF_SMALL_BUFFER;
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * 8);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

// AVX2 buff-simd filters over a flipped stream.
int ref_simd_filter(const uint8_t* s, uint64_t len, uint64_t scale, int op, double pred, uint32_t* words,
                    uint64_t nwords_cap, uint64_t* count) {
  try {
    FilterOut fo = op == 0 ? buff_simd_range_filter(s, len, scale, pred)
                           : buff_simd_equal_filter(s, len, scale, pred);
    return copy_filter(fo, words, nwords_cap, count, nullptr);
  } catch (const RefPanic&) { return REF_PANIC; }
}

int ref_sum(const uint8_t* s, uint64_t len, uint64_t scale, double* sum, uint64_t* plane_sums) {
  try {
    if (plane_sums) std::memset(plane_sums, 0, 9 * 8);
    *sum = byte_fixed_sum(s, len, scale, plane_sums);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

int ref_max(const uint8_t* s, uint64_t len, uint64_t scale, double* mx, uint64_t* max_off, uint64_t* argmax_count) {
  try {
    *mx = byte_fixed_max(s, len, scale, max_off, argmax_count);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

// buff_max_range (major: buff_max_majority_range): the window maxima (stored offsets, and as the reference's
// f64 `(x as i64 + base) / 2^dlen`) and the argmax rows as BitVec<u32> words.
int ref_max_range(const uint8_t* s, uint64_t len, uint32_t rs, uint32_t re, uint32_t window, int major, uint64_t* offs,
                  double* vals, uint64_t cap, uint64_t* nwin, uint32_t* words, uint64_t nwords_cap) {
  try {
    MaxRangeOut o = buff_max_range(s, len, rs, re, window, major != 0);
    BitPackReader bp(s, len);
    Header h = read_header(bp);
    *nwin = o.maxv.size();
    if (o.maxv.size() > cap || (size_t)(h.len + 31) / 32 > nwords_cap) return REF_SMALL_BUFFER;
    const double dec_scl = powi(2.0, (int32_t)h.dlen);
    for (size_t k = 0; k < o.maxv.size(); ++k) {
      offs[k] = o.maxv[k];
      vals[k] = (double)(int64_t)(o.maxv[k] + (uint64_t)h.base) / dec_scl;
    }
    std::memset(words, 0, 4 * ((size_t)(h.len + 31) / 32));
    for (uint32_t r : o.res) {
      if (r >= h.len) throw RefPanic("argmax row past the column");
      words[r / 32] |= 1u << (r % 32);
    }
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

int ref_decode_with_precision(const uint8_t* s, uint64_t len, uint64_t scale, uint32_t precision, double* out,
                              uint64_t cap, uint64_t* n_out) {
  try {
    bool unsup = false;
    std::vector<double> v = decode_with_precision(s, len, scale, precision, &unsup);
    if (unsup) return REF_UNSUPPORTED;
    *n_out = v.size();
    if (v.size() > cap) return REF_SMALL_BUFFER;
    std::memcpy(out, v.data(), v.size() * 8);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

int ref_sum_with_precision(const uint8_t* s, uint64_t len, uint64_t scale, uint32_t precision, double* sum,
                           uint64_t* plane_sums) {
  try {
    if (plane_sums) std::memset(plane_sums, 0, 9 * 8);
    bool unsup = false;
    *sum = sum_with_precision(s, len, scale, precision, &unsup, plane_sums);
    return unsup ? REF_UNSUPPORTED : REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

// buff-major (buff_simd.rs:317-2861). Call ref_major_encode with out==NULL for the size.
int ref_major_encode(const double* x, uint64_t n, uint64_t scale, uint8_t* out, uint64_t cap, uint64_t* out_len,
                     int64_t* minmax) {
  try {
    int64_t mn, mx;
    std::vector<uint8_t> s = buff_encode_majority(x, n, scale, &mn, &mx);
    if (minmax) { minmax[0] = mn; minmax[1] = mx; }
    *out_len = s.size();
    if (!out) return REF_OK;
    if (s.size() > cap) return REF_SMALL_BUFFER;
    std::memcpy(out, s.data(), s.size());
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

int ref_major_decode(const uint8_t* s, uint64_t len, uint64_t scale, double* out, uint64_t cap, uint64_t* n_out) {
  try {
    std::vector<double> v = buff_decode_majority(s, len, scale);
    *n_out = v.size();
    if (v.size() > cap) return REF_SMALL_BUFFER;
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * 8);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

int ref_major_filter(const uint8_t* s, uint64_t len, uint64_t scale, int op, double pred, uint32_t* words,
                     uint64_t nwords_cap, uint64_t* count) {
  try {
    FilterOut fo = op == 0 ? buff_range_filter_majority(s, len, scale, pred)
                           : buff_equal_filter_majority(s, len, scale, pred);
    return copy_filter(fo, words, nwords_cap, count, nullptr);
  } catch (const RefPanic&) { return REF_PANIC; }
}

// col_sums: up to 9 u64 (byte columns in stream order, then the tail).
int ref_major_sum(const uint8_t* s, uint64_t len, uint64_t scale, double* sum, uint64_t* col_sums) {
  try {
    if (col_sums) std::memset(col_sums, 0, 9 * 8);
    *sum = buff_sum_majority(s, len, scale, col_sums);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

int ref_major_max(const uint8_t* s, uint64_t len, uint64_t scale, double* mx, uint64_t* max_off,
                  uint64_t* argmax_count) {
  try {
    *mx = buff_max_majority(s, len, scale, max_off, argmax_count);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

// Raw snappy of the restatement above (cap bytes of output; out_len always set).
int ref_snappy_compress(const uint8_t* in, uint64_t n, uint8_t* out, uint64_t cap, uint64_t* out_len) {
  std::vector<uint8_t> v = snappy_ref::compress(in, n);
  *out_len = v.size();
  if (v.size() > cap) return REF_SMALL_BUFFER;
  std::memcpy(out, v.data(), v.size());
  return REF_OK;
}
int ref_snappy_decompress(const uint8_t* in, uint64_t n, uint8_t* out, uint64_t cap, uint64_t* out_len) {
  try {
    std::vector<uint8_t> v = snappy_ref::decompress(in, n);
    *out_len = v.size();
    if (v.size() > cap) return REF_SMALL_BUFFER;
    if (!v.empty()) std::memcpy(out, v.data(), v.size());
    return REF_OK;
  } catch (const snappy_ref::Error&) { return REF_PANIC; }
}

// ---- buff-slice (buff_slice.rs:222-1970, buff_simd.rs:3736-4025) ----------
int ref_slice_encode(const double* data, uint64_t n, uint64_t scale, uint8_t* out, uint64_t cap, uint64_t* out_len) {
  try {
    std::vector<uint8_t> v = buff_slice_encode(data, n, scale);
    *out_len = v.size();
    if (v.size() > cap) return REF_SMALL_BUFFER;
    if (!v.empty()) std::memcpy(out, v.data(), v.size());
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

// kind 0: buff_slice_decode, 1: buff_slice_decode_bs, 2: buff_slice_decode_condition (rows).
int ref_slice_decode(const uint8_t* s, uint64_t len, int kind, const uint64_t* rows, uint64_t nrows, double* out,
                     uint64_t cap, uint64_t* n_out) {
  try {
    std::vector<double> v = kind == 0 ? buff_slice_decode(s, len)
                            : kind == 1 ? buff_slice_decode_bs(s, len)
                                        : buff_slice_decode_condition(s, len, rows, nrows);
    *n_out = v.size();
    if (v.size() > cap) return REF_SMALL_BUFFER;
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * 8);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

// scalar 0: buff_slice_buffsum, 1: buff_slice_sum.
int ref_slice_sum(const uint8_t* s, uint64_t len, int scalar, double* sum) {
  try {
    *sum = scalar ? buff_slice_sum(s, len) : buff_slice_buffsum(s, len);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

// scalar 0: buff_slice_buffmax, 1: buff_slice_max. words: ceil(n/32) u32.
int ref_slice_max(const uint8_t* s, uint64_t len, int scalar, double* mx, uint64_t* count, uint32_t* words,
                  uint64_t nwords_cap) {
  try {
    SliceMaxOut o = scalar ? buff_slice_max(s, len) : buff_slice_buffmax(s, len);
    *mx = o.max;
    *count = o.res.cardinality();
    if (o.res.storage.size() > nwords_cap) return REF_SMALL_BUFFER;
    if (!o.res.storage.empty()) std::memcpy(words, o.res.storage.data(), o.res.storage.size() * 4);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

int ref_slice_max_range(const uint8_t* s, uint64_t len, uint32_t rs, uint32_t re, uint32_t window, double* vals,
                        uint64_t cap, uint64_t* nwin, uint32_t* words, uint64_t nwords_cap) {
  try {
    SliceMaxRangeOut o = buff_slice_max_range(s, len, rs, re, window);
    *nwin = o.max_vec.size();
    if (o.max_vec.size() > cap || o.res.storage.size() > nwords_cap) return REF_SMALL_BUFFER;
    if (!o.max_vec.empty()) std::memcpy(vals, o.max_vec.data(), o.max_vec.size() * 8);
    if (!o.res.storage.empty()) std::memcpy(words, o.res.storage.data(), o.res.storage.size() * 4);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

// kind: SliceFilterKind; rows/nrows: the condition of kinds 4 and 5.
int ref_slice_filter(const uint8_t* s, uint64_t len, uint64_t scale, double pred, int kind, const uint64_t* rows,
                     uint64_t nrows, uint32_t* words, uint64_t nwords_cap, uint64_t* count) {
  try {
    return copy_filter(buff_slice_filter(s, len, scale, pred, kind, rows, nrows), words, nwords_cap, count,
                       nullptr);
  } catch (const RefPanic&) { return REF_PANIC; }
}


// RAPG residue format (split_double.rs:624-954, 1353-1893, 2210-2629). precision
// < 0: the full-width reads. Call ref_residue_encode with out==NULL for the size.
int ref_residue_encode(const double* x, uint64_t n, uint64_t scale, uint8_t* out, uint64_t cap, uint64_t* out_len,
                       uint32_t* cols) {
  try {
    std::vector<uint8_t> s = byte_residue_encode(x, n, scale, cols);
    *out_len = s.size();
    if (!out) return REF_OK;
    if (s.size() > cap) return REF_SMALL_BUFFER;
    std::memcpy(out, s.data(), s.size());
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

int ref_residue_decode(const uint8_t* s, uint64_t len, int32_t precision, double* out, uint64_t cap, uint64_t* n_out) {
  try {
    std::vector<double> v = byte_residue_decode(s, len, precision);
    *n_out = v.size();
    if (v.size() > cap) return REF_SMALL_BUFFER;
    std::memcpy(out, v.data(), v.size() * 8);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

int ref_residue_sum(const uint8_t* s, uint64_t len, int32_t precision, double* sum, uint64_t* plane_sums,
                    uint32_t* cols) {
  try {
    if (plane_sums) std::memset(plane_sums, 0, 9 * 8);
    *sum = byte_residue_sum(s, len, precision, plane_sums, cols);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

int ref_roaring_serialize(const uint32_t* vals, uint64_t n, uint8_t* out, uint64_t cap, uint64_t* out_len) {
  std::vector<uint8_t> s = roaring_ref::serialize(std::vector<uint32_t>(vals, vals + n));
  *out_len = s.size();
  if (!out) return REF_OK;
  if (s.size() > cap) return REF_SMALL_BUFFER;
  std::memcpy(out, s.data(), s.size());
  return REF_OK;
}

int ref_roaring_deserialize(const uint8_t* p, uint64_t n, uint32_t* vals, uint64_t cap, uint64_t* n_out) {
  try {
    std::vector<uint32_t> v = roaring_ref::deserialize(p, n);
    *n_out = v.size();
    if (v.size() > cap) return REF_SMALL_BUFFER;
    if (!v.empty()) std::memcpy(vals, v.data(), v.size() * 4);
    return REF_OK;
  } catch (const RefPanic&) { return REF_PANIC; }
}

}  // extern "C"
