// CSV -> f64 ingest on the device: construct_file_iterator_skip_newline::<f64>
// (database/src/client.rs:163-182) over a file's bytes resident in HBM.
//
// The reference reads lines (split on '\n', a trailing "\r" removed), splits
// each line on `delim`, skips the first skip_val items of the line, and keeps
// every item `f64::from_str` accepts, in file order. Here:
//   k_csv_blocks  one workgroup per 4 KiB block: items ending in the block,
//                 items after its last newline, its last delimiter;
//   k_csv_scan    one workgroup: prefixes over the blocks (item base, index of
//                 the block's first item within its line, previous delimiter),
//                 and after the parse the prefix of parsed items;
//   k_csv_parse   one workgroup per block: every item ending in the block is
//                 parsed by the lane owning its last byte; values are written
//                 compacted within the block at its item base;
//   k_csv_compact moves each block's values to their final place when some
//                 item was rejected (otherwise the parse wrote them in place).
// Parsing is exact: Rust's grammar (optional sign; digits with an optional
// point and at least one digit; optional exponent; or inf / infinity / nan in
// any case), the value correctly rounded (ties to even) by the exact fast path
// (at most 2^53 for the digits and |exponent| <= 22: one IEEE multiply or
// divide by an exact power of ten) or else by the 192-bit product of the
// normalized digits with 5^q truncated to 128 bits, whose rounding is decided
// unless a carry from the truncated part could reach the rounding bit (then
// the item is reported undecided). More than 19 significant digits: both
// 19-digit bounds must round alike.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "buff_csv.h"
#include "buff_device.h"
#include "buff_launch.h"

namespace buffgpu {

// ---- host: the table of 5^q ------------------------------------------------------

namespace {

struct Big {  // little-endian base-2^32 magnitude
  std::vector<uint32_t> d;
  uint32_t bits() const {
    for (size_t i = d.size(); i-- > 0;)
      if (d[i]) return (uint32_t)(32 * i + 32 - __builtin_clz(d[i]));
    return 0;
  }
  void mul_small(uint32_t m) {
    uint64_t c = 0;
    for (uint32_t& x : d) {
      const uint64_t v = (uint64_t)x * m + c;
      x = (uint32_t)v;
      c = v >> 32;
    }
    if (c) d.push_back((uint32_t)c);
  }
  void div_small(uint32_t m) {  // floor division
    uint64_t r = 0;
    for (size_t i = d.size(); i-- > 0;) {
      const uint64_t v = (r << 32) | d[i];
      d[i] = (uint32_t)(v / m);
      r = v % m;
    }
  }
  bool bit(uint32_t k) const { return k / 32 < d.size() && ((d[k / 32] >> (k % 32)) & 1u); }
  // bits [k, k+64) as an integer
  uint64_t window64(int64_t k) const {
    uint64_t v = 0;
    for (int j = 63; j >= 0; --j) v = (v << 1) | (k + j >= 0 && bit((uint32_t)(k + j)) ? 1u : 0u);
    return v;
  }
};

}  // namespace

void csv_pow5_table(Pow5* out) {
  for (int q = kPow5Min; q <= kPow5Max; ++q) {
    Pow5& t = out[q - kPow5Min];
    const uint32_t k = (uint32_t)(q < 0 ? -q : q);
    Big p5{{1u}};
    for (uint32_t i = 0; i < k; ++i) p5.mul_small(5);
    const uint32_t L = p5.bits();
    if (q >= 0) {
      // 5^q = (5^q << (128 - L)) * 2^(L - 128); exact while it fits 128 bits
      const int64_t sh = (int64_t)L - 128;
      t.hi = p5.window64(sh + 64);
      t.lo = p5.window64(sh);
      t.e = (int32_t)sh;
      t.exact = L <= 128 ? 1u : 0u;
    } else {
      // 5^q = 2^(127+L) / 5^k * 2^-(127+L); floor of the quotient by repeated /5
      Big a;
      a.d.assign((127 + L) / 32 + 1, 0u);
      a.d[(127 + L) / 32] = 1u << ((127 + L) % 32);
      for (uint32_t i = 0; i < k; ++i) a.div_small(5);
      t.hi = a.window64(64);
      t.lo = a.window64(0);
      t.e = -(int32_t)(127 + L);
      t.exact = 0;
    }
  }
}

// ---- device: one item -> f64 -------------------------------------------------------

namespace {

__host__ __device__ __forceinline__ bool csv_digit(uint32_t c) { return c - '0' < 10u; }

// The correctly rounded f64 bits of w * 10^q (w >= 1), or false when the
// truncated power leaves the rounding open.
__device__ bool round_w10q(uint64_t w, int64_t q, const Pow5* __restrict__ tab, uint64_t* bits) {
  if (q < kPow5Min) {  // below half the smallest subnormal
    *bits = 0;
    return true;
  }
  if (q > kPow5Max) {
    *bits = 0x7FF0000000000000ull;
    return true;
  }
  const Pow5 t = tab[q - kPow5Min];
  const int lz = __clzll((long long)w);
  const uint64_t wn = w << lz;
  const uint64_t a_lo = wn * t.hi, a_hi = __umul64hi(wn, t.hi);
  const uint64_t b_lo = wn * t.lo, b_hi = __umul64hi(wn, t.lo);
  const uint64_t lo = b_lo;
  const uint64_t mid = a_lo + b_hi;
  const uint64_t hi = a_hi + (mid < a_lo ? 1u : 0u);
  const int m = (hi >> 63) ? 191 : 190;  // msb of the 192-bit product
  const int64_t ex = (int64_t)m + t.e + q - lz;
  if (ex > 1023) {
    *bits = 0x7FF0000000000000ull;
    return true;
  }
  const int64_t K = ex >= -1022 ? 53 : ex + 1075;  // mantissa bits kept
  if (K < 0) {
    *bits = 0;
    return true;
  }
  const int hs = (int)(m - K + 1 - 128);  // lsb of the kept bits within hi
  const int hr = hs - 1;                   // rounding bit within hi (>= 9)
  uint64_t mant = hs >= 64 ? 0 : (hi >> hs);
  const uint64_t round = (hi >> hr) & 1u;
  const uint64_t below_mask = (1ull << hr) - 1;
  const uint64_t below = hi & below_mask;
  bool sticky;
  if (t.exact) {
    sticky = below || mid || lo;
  } else {
    if (below == below_mask && mid == ~0ull) return false;
    sticky = true;  // the true product exceeds the truncated one
  }
  if (round && (sticky || (mant & 1u))) mant += 1;
  uint64_t b = ex >= -1022 ? ((uint64_t)(ex + 1022) << 52) + mant : mant;
  if (b >= 0x7FF0000000000000ull) b = 0x7FF0000000000000ull;
  *bits = b;
  return true;
}

__constant__ double kPow10Exact[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                       1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

// The correctly rounded bits of w * 10^q (w >= 1), or false when undecided:
// the exact fast path; then, for q < 0 with 5^-q dividing w (the only way the
// value can be a dyadic rational, so sit exactly on a double or a tie),
// (w / 5^-q) * 2^q from one correctly rounded u64 -> f64 conversion and an
// exact scaling; otherwise the 192-bit product, where the value cannot sit on
// a boundary and only a carry out of the truncated part is left open.
__device__ bool w10q_bits(uint64_t w, int64_t q, const Pow5* __restrict__ tab, uint64_t* bits) {
  if (q >= -22 && q <= 22 && w <= (1ull << 53)) {
    const double v = (double)w;
    *bits = (uint64_t)__double_as_longlong(q >= 0 ? v * kPow10Exact[q] : v / kPow10Exact[-q]);
    return true;
  }
  if (q < 0 && q >= -27) {
    uint64_t p5 = 1;
    for (int64_t j = 0; j < -q; ++j) p5 *= 5;
    if (w % p5 == 0) {
      *bits = (uint64_t)__double_as_longlong(ldexp((double)(w / p5), (int)q));
      return true;
    }
  }
  return round_w10q(w, q, tab, bits);
}

__device__ __forceinline__ bool ci_eq(const uint8_t* p, const char* word, uint64_t n) {
  for (uint64_t j = 0; j < n; ++j)
    if ((uint32_t)(p[j] | 0x20u) != (uint32_t)word[j]) return false;
  return true;
}

// 0: not an f64 in Rust's grammar; 1: *out holds the value; 2: undecided.
__device__ int parse_item(const uint8_t* __restrict__ p, uint64_t n, const Pow5* __restrict__ tab, double* out) {
  if (n == 0) return 0;
  uint64_t i = 0;
  bool neg = false;
  if (p[0] == '+' || p[0] == '-') {
    neg = p[0] == '-';
    i = 1;
  }
  const uint64_t sign = neg ? 0x8000000000000000ull : 0ull;
  const uint64_t rest = n - i;
  if ((rest == 3 && ci_eq(p + i, "inf", 3)) || (rest == 8 && ci_eq(p + i, "infinity", 8))) {
    *out = __longlong_as_double((long long)(sign | 0x7FF0000000000000ull));
    return 1;
  }
  if (rest == 3 && ci_eq(p + i, "nan", 3)) {
    *out = __longlong_as_double((long long)(sign | 0x7FF8000000000000ull));
    return 1;
  }
  uint64_t w = 0;
  int nsig = 0;
  int64_t q = 0;
  bool trunc = false;
  uint64_t ndig = 0;
  while (i < n && csv_digit(p[i])) {
    const uint32_t d = p[i] - '0';
    ++ndig;
    if (w == 0 && d == 0) {
    } else if (nsig < 19) {
      w = w * 10 + d;
      ++nsig;
    } else {
      ++q;
      trunc |= d != 0;
    }
    ++i;
  }
  if (i < n && p[i] == '.') {
    ++i;
    while (i < n && csv_digit(p[i])) {
      const uint32_t d = p[i] - '0';
      ++ndig;
      if (w == 0 && d == 0) {
        --q;
      } else if (nsig < 19) {
        w = w * 10 + d;
        ++nsig;
        --q;
      } else {
        trunc |= d != 0;
      }
      ++i;
    }
  }
  if (ndig == 0) return 0;
  if (i < n && (p[i] | 0x20u) == 'e') {
    ++i;
    bool eneg = false;
    if (i < n && (p[i] == '+' || p[i] == '-')) {
      eneg = p[i] == '-';
      ++i;
    }
    if (i >= n || !csv_digit(p[i])) return 0;
    int64_t ev = 0;
    while (i < n && csv_digit(p[i])) {
      if (ev < 100000000) ev = ev * 10 + (p[i] - '0');
      ++i;
    }
    q += eneg ? -ev : ev;
  }
  if (i != n) return 0;
  uint64_t bits = 0;
  if (w != 0) {
    if (!w10q_bits(w, q, tab, &bits)) return 2;
    if (trunc) {  // the digits past the 19th put the value strictly between w and w + 1 units
      uint64_t b2 = 0;
      if (!w10q_bits(w + 1, q, tab, &b2) || b2 != bits) return 2;
    }
  }
  *out = __longlong_as_double((long long)(sign | bits));
  return 1;
}

__device__ __forceinline__ bool csv_delim(uint32_t c, uint32_t delim) { return c == delim || c == '\n'; }

// The 16 bytes of `lane` of block b (zero past the end).
__device__ __forceinline__ void load16(const uint8_t* __restrict__ text, uint64_t len, uint64_t at, uint8_t* by) {
  if (at + 16 <= len) {
    const uint4 v = *reinterpret_cast<const uint4*>(text + at);  // cached: the parse reads the items' bytes again
    const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 16; ++j) by[j] = (uint8_t)(w4[j / 4] >> (8 * (j % 4)));
  } else {
#pragma unroll
    for (int j = 0; j < 16; ++j) by[j] = at + j < len ? text[at + j] : 0;
  }
}

// Lane shuffles of the scanned types.
__device__ __forceinline__ uint64_t shfl_up_t(uint64_t v, unsigned d) {
  return (uint64_t)__shfl_up((unsigned long long)v, d, 64);
}
__device__ __forceinline__ int64_t shfl_up_t(int64_t v, unsigned d) { return (int64_t)__shfl_up((long long)v, d, 64); }

struct Seg {  // items of a run: reset = it contains a newline, v = items since the last one
  uint32_t reset, v;
};
struct SegOp {
  __device__ Seg operator()(Seg a, Seg b) const { return b.reset ? b : Seg{a.reset, a.v + b.v}; }
};
struct SumOp {
  __device__ uint64_t operator()(uint64_t a, uint64_t b) const { return a + b; }
};
struct MaxOp {
  __device__ int64_t operator()(int64_t a, int64_t b) const { return a > b ? a : b; }
};
__device__ __forceinline__ Seg shfl_up_t(Seg v, unsigned d) {
  return Seg{(uint32_t)__shfl_up((unsigned)v.reset, d, 64), (uint32_t)__shfl_up((unsigned)v.v, d, 64)};
}

// Exclusive scan over the 256 lanes of a workgroup (in lane order; `op` need
// not commute): shuffles within each wave, then the four wave totals from LDS.
// *total gets the workgroup's combined value.
template <class T, class Op>
__device__ T wg_exclusive(T x, T ident, Op op, T* lds4, T* total) {
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  T incl = x;
#pragma unroll
  for (unsigned d = 1; d < 64; d <<= 1) {
    const T y = shfl_up_t(incl, d);
    if (lane >= d) incl = op(y, incl);
  }
  T ex_in_wave = shfl_up_t(incl, 1);
  if (lane == 0) ex_in_wave = ident;
  if (lane == 63) lds4[wave] = incl;
  __syncthreads();
  T before = ident, all = ident;
#pragma unroll
  for (unsigned w = 0; w < 4; ++w) {
    if (w < wave) before = op(before, lds4[w]);
    all = op(all, lds4[w]);
  }
  __syncthreads();
  if (total) *total = all;
  return op(before, ex_in_wave);
}

// Per-lane facts of 16 bytes at absolute position `at`.
struct LaneFacts {
  uint32_t ntok, since, has_nl, nonascii;
  int64_t last_delim;
};

__device__ LaneFacts lane_facts(const uint8_t* by, uint64_t at, uint64_t len, uint32_t delim) {
  LaneFacts f{0, 0, 0, 0, -1};
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const uint64_t pos = at + j;
    if (pos >= len) break;
    const uint32_t c = by[j];
    f.nonascii |= c >= 0x80u;
    const bool dl = csv_delim(c, delim);
    if (dl) f.last_delim = (int64_t)pos;
    // a delimiter ends an item; so does the end of the text after any other byte
    const uint32_t ends = (dl || pos + 1 == len) ? 1u : 0u;
    f.ntok += ends;
    f.since += ends;
    if (c == '\n') {
      f.has_nl = 1;
      f.since = 0;  // the newline's own item belongs to the line it ends
    }
  }
  return f;
}

}  // namespace

__global__ __launch_bounds__(256) void k_csv_blocks(const uint8_t* __restrict__ text, uint64_t len, uint32_t delim,
                                                    CsvBlock* __restrict__ blk) {
  __shared__ uint64_t lds_u[4];
  __shared__ int64_t lds_i[4];
  __shared__ Seg lds_s[4];
  const uint64_t b = blockIdx.x;
  const uint64_t at = b * kCsvBlock + 16ull * threadIdx.x;
  uint8_t by[16];
  load16(text, len, at, by);
  const LaneFacts f = lane_facts(by, at, len, delim);
  uint64_t ntot = 0;
  int64_t dmax = -1;
  Seg stot{0, 0};
  (void)wg_exclusive<uint64_t>(f.ntok, 0ull, SumOp{}, lds_u, &ntot);
  (void)wg_exclusive<int64_t>(f.last_delim, -1ll, MaxOp{}, lds_i, &dmax);
  (void)wg_exclusive<Seg>(Seg{f.has_nl, f.since}, Seg{0, 0}, SegOp{}, lds_s, &stot);
  __shared__ uint32_t any_nonascii;
  if (threadIdx.x == 0) any_nonascii = 0;
  __syncthreads();
  if (f.nonascii) atomicOr(&any_nonascii, 1u);
  __syncthreads();
  if (threadIdx.x == 0) {
    CsvBlock o;
    o.ntok = (uint32_t)ntot;
    o.since = stot.reset ? stot.v : (uint32_t)ntot;
    o.has_nl = stot.reset;
    o.nonascii = any_nonascii;
    o.last_delim = dmax;
    blk[b] = o;
  }
}

// mode 0: item base, line index, previous delimiter, totals[0..1];
// mode 1: parsed-item base, totals[2]. Each lane takes 16 consecutive blocks
// (one pass over its blocks, a workgroup scan of the lane totals, a second pass
// writing the prefixes), 4096 blocks per step.
__global__ __launch_bounds__(256) void k_csv_scan(const CsvBlock* __restrict__ blk, CsvPre* __restrict__ pre,
                                                  uint64_t nblocks, int mode, unsigned long long* __restrict__ totals) {
  constexpr uint32_t kPer = 16;
  __shared__ uint64_t lds_u[4];
  __shared__ int64_t lds_i[4];
  __shared__ Seg lds_s[4];
  __shared__ uint32_t any_nonascii;
  if (threadIdx.x == 0) any_nonascii = 0;
  __syncthreads();
  uint64_t carry_tok = 0;
  Seg carry_seg{0, 0};
  int64_t carry_delim = -1;
  uint32_t nonascii = 0;
  for (uint64_t base = 0; base < nblocks; base += 256ull * kPer) {
    const uint64_t i0 = base + (uint64_t)threadIdx.x * kPer;
    const uint64_t i1 = i0 + kPer < nblocks ? i0 + kPer : nblocks;
    if (mode == 0) {
      // the lane's blocks in registers: all loads issued before any is used
      CsvBlock cb[kPer];
#pragma unroll
      for (uint32_t q = 0; q < kPer; ++q) cb[q] = i0 + q < i1 ? blk[i0 + q] : CsvBlock{0, 0, 0, 0, -1};
      uint64_t tok = 0;
      Seg seg{0, 0};
      int64_t dl = -1;
#pragma unroll
      for (uint32_t q = 0; q < kPer; ++q) {
        const CsvBlock& c = cb[q];
        tok += c.ntok;
        seg = SegOp{}(seg, Seg{c.has_nl, c.has_nl ? c.since : c.ntok});
        dl = c.last_delim > dl ? c.last_delim : dl;
        nonascii |= c.nonascii;
      }
      uint64_t tsum = 0;
      int64_t dtot = -1;
      Seg stot{0, 0};
      const uint64_t ex_tok = wg_exclusive<uint64_t>(tok, 0ull, SumOp{}, lds_u, &tsum);
      const int64_t ex_delim = wg_exclusive<int64_t>(dl, -1ll, MaxOp{}, lds_i, &dtot);
      const Seg ex_seg = wg_exclusive<Seg>(seg, Seg{0, 0}, SegOp{}, lds_s, &stot);
      uint64_t t = carry_tok + ex_tok;
      Seg sg = SegOp{}(carry_seg, ex_seg);
      int64_t pd = ex_delim > carry_delim ? ex_delim : carry_delim;
#pragma unroll
      for (uint32_t q = 0; q < kPer; ++q) {
        const CsvBlock& c = cb[q];
        if (i0 + q < i1) {
          CsvPre p;
          p.tok_base = t;
          p.ok_base = 0;
          p.line_idx = sg.v;
          p.ok = 0;
          p.prev_delim = pd;
          pre[i0 + q] = p;
        }
        t += c.ntok;
        sg = SegOp{}(sg, Seg{c.has_nl, c.has_nl ? c.since : c.ntok});
        pd = c.last_delim > pd ? c.last_delim : pd;
      }
      carry_tok += tsum;
      carry_seg = SegOp{}(carry_seg, stot);
      carry_delim = dtot > carry_delim ? dtot : carry_delim;
    } else {
      uint32_t oks[kPer];
#pragma unroll
      for (uint32_t q = 0; q < kPer; ++q) oks[q] = i0 + q < i1 ? pre[i0 + q].ok : 0u;
      uint64_t ok = 0;
#pragma unroll
      for (uint32_t q = 0; q < kPer; ++q) ok += oks[q];
      uint64_t tsum = 0;
      uint64_t acc = carry_tok + wg_exclusive<uint64_t>(ok, 0ull, SumOp{}, lds_u, &tsum);
#pragma unroll
      for (uint32_t q = 0; q < kPer; ++q) {
        if (i0 + q < i1) pre[i0 + q].ok_base = acc;
        acc += oks[q];
      }
      carry_tok += tsum;
    }
  }
  if (nonascii) atomicOr(&any_nonascii, 1u);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (mode == 0) {
      totals[0] = carry_tok;
      totals[1] = any_nonascii;
      totals[3] = 0;
      totals[4] = 0;
      totals[5] = ~0ull;
    } else {
      totals[2] = carry_tok;
    }
  }
}

__global__ __launch_bounds__(256) void k_csv_parse(const uint8_t* __restrict__ text, uint64_t len, uint32_t delim,
                                                   uint32_t skip, const Pow5* __restrict__ tab,
                                                   CsvPre* __restrict__ pre, double* __restrict__ dst,
                                                   uint8_t* __restrict__ okb, CsvSlow* __restrict__ slow,
                                                   uint64_t slow_cap, unsigned long long* __restrict__ totals) {
  __shared__ uint64_t lds_u[4];
  __shared__ int64_t lds_i[4];
  __shared__ Seg lds_s[4];
  const uint64_t b = blockIdx.x;
  const CsvPre bp = pre[b];
  const uint64_t at = b * kCsvBlock + 16ull * threadIdx.x;
  uint8_t by[16];
  load16(text, len, at, by);
  // the block and the 512 bytes before it in LDS: items are parsed from there
  constexpr uint32_t kHalo = 512;
  __shared__ __align__(16) uint8_t txt[kHalo + kCsvBlock];
  const uint64_t blk0 = b * kCsvBlock;
  {
    uint32_t w4[4] = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 16; ++j) w4[j / 4] |= (uint32_t)by[j] << (8 * (j % 4));
    *reinterpret_cast<uint4*>(txt + kHalo + 16 * threadIdx.x) = make_uint4(w4[0], w4[1], w4[2], w4[3]);
    if (b > 0 && threadIdx.x < kHalo / 16)
      *reinterpret_cast<uint4*>(txt + 16 * threadIdx.x) =
          *reinterpret_cast<const uint4*>(text + blk0 - kHalo + 16 * threadIdx.x);
  }
  __syncthreads();
  const uint64_t staged = b > 0 ? blk0 - kHalo : 0;  // first byte held in LDS
  const uint8_t* lds_at = txt + kHalo;                 // LDS address of byte blk0
  const LaneFacts f = lane_facts(by, at, len, delim);
  uint64_t dummy_u = 0;
  int64_t dummy_i = 0;
  Seg dummy_s{0, 0};
  const uint64_t ex_tok = wg_exclusive<uint64_t>(f.ntok, 0ull, SumOp{}, lds_u, &dummy_u);
  const int64_t ex_delim = wg_exclusive<int64_t>(f.last_delim, -1ll, MaxOp{}, lds_i, &dummy_i);
  const Seg ex_seg = wg_exclusive<Seg>(Seg{f.has_nl, f.since}, Seg{0, 0}, SegOp{}, lds_s, &dummy_s);
  uint32_t li = ex_seg.reset ? ex_seg.v : bp.line_idx + ex_seg.v;
  int64_t prev = ex_delim > bp.prev_delim ? ex_delim : bp.prev_delim;
  uint64_t k = bp.tok_base + ex_tok;  // this lane's first item
  uint32_t nok = 0;
  // the lane's item ends first, then one item per step: a wave takes as many
  // parse steps as its busiest lane has items, not one per byte position.
  // Each value goes to its item's slot; okb marks the items that parsed.
  uint32_t ends = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const uint64_t pos = at + j;
    if (pos < len && (csv_delim(by[j], delim) || pos + 1 == len)) ends |= 1u << j;
  }
  const uint8_t* lane_txt = txt + kHalo + 16 * threadIdx.x;
  while (ends) {
    const int j = __builtin_ctz(ends);
    ends &= ends - 1;
    const uint64_t pos = at + j;
    const uint32_t c = lane_txt[j];
    const bool eof = pos + 1 == len && !csv_delim(c, delim);
    const uint64_t start = (uint64_t)(prev + 1);
    uint64_t end = eof ? len : pos;
    if (c == '\n' && end > start && lds_at[(int64_t)(end - 1) - (int64_t)blk0] == '\r') end -= 1;  // "\r\n"
    uint8_t ok = 0;
    if (li >= skip) {
      double v = 0.0;
      const uint8_t* src = start >= staged ? lds_at + ((int64_t)start - (int64_t)blk0) : text + start;
      const int r = parse_item(src, end - start, tab, &v);
      if (r == 1) {
        dst[k] = v;
        ok = 1;
      } else if (r == 2) {  // a valid item: its exact value is filled in by k_csv_slow
        ok = 1;
        const unsigned long long e = atomicAdd(&totals[3], 1ull);
        if (e < slow_cap) {
          slow[e] = CsvSlow{start, end, k};
        } else {
          atomicAdd(&totals[4], 1ull);
          atomicMin(&totals[5], (unsigned long long)start);
        }
      }
    }
    okb[k] = ok;
    nok += ok;
    ++k;
    prev = (int64_t)pos;
    li = c == '\n' ? 0 : li + 1;
  }
  uint64_t ok_tot = 0;
  (void)wg_exclusive<uint64_t>(nok, 0ull, SumOp{}, lds_u, &ok_tot);
  if (threadIdx.x == 0) pre[b].ok = (uint32_t)ok_tot;
}

// ---- the exact slow path: a per-thread big integer ---------------------------------

namespace {

constexpr int kBigLimbs = 72;  // 2304 bits: 300 significant digits at any exponent
constexpr int kMaxSigDigits = 300;

struct DevBig {
  uint32_t d[kBigLimbs];
  int n;  // limbs in use (top limb nonzero)
};

__host__ __device__ void big_mul_add(DevBig& b, uint32_t m, uint32_t a) {
  uint64_t c = a;
  for (int i = 0; i < b.n; ++i) {
    const uint64_t v = (uint64_t)b.d[i] * m + c;
    b.d[i] = (uint32_t)v;
    c = v >> 32;
  }
  if (c && b.n < kBigLimbs) b.d[b.n++] = (uint32_t)c;
}

__host__ __device__ uint32_t big_div_small(DevBig& b, uint32_t m) {
  uint64_t r = 0;
  for (int i = b.n - 1; i >= 0; --i) {
    const uint64_t v = (r << 32) | b.d[i];
    b.d[i] = (uint32_t)(v / m);
    r = v % m;
  }
  while (b.n > 0 && b.d[b.n - 1] == 0) --b.n;
  return (uint32_t)r;
}

__host__ __device__ int big_bits(const DevBig& b) { return b.n ? 32 * (b.n - 1) + 32 - __builtin_clz(b.d[b.n - 1]) : 0; }

__host__ __device__ void big_shl(DevBig& b, int s) {
  const int ls = s / 32, bs = s % 32;
  if (b.n == 0 || s == 0) return;
  int nn = b.n + ls + 1;
  if (nn > kBigLimbs) nn = kBigLimbs;
  for (int i = nn - 1; i >= 0; --i) {
    const int src = i - ls;
    const uint32_t hi = src >= 0 && src < b.n ? b.d[src] : 0u;
    const uint32_t lo = src - 1 >= 0 && src - 1 < b.n ? b.d[src - 1] : 0u;
    b.d[i] = bs ? (hi << bs) | (lo >> (32 - bs)) : hi;
  }
  b.n = nn;
  while (b.n > 0 && b.d[b.n - 1] == 0) --b.n;
}

__host__ __device__ uint32_t big_bit(const DevBig& b, int k) { return k >= 0 && k / 32 < b.n ? (b.d[k / 32] >> (k % 32)) & 1u : 0u; }

__host__ __device__ bool big_any_below(const DevBig& b, int k) {  // any bit < k
  for (int i = 0; i < b.n && 32 * i < k; ++i) {
    const int top = k - 32 * i;
    const uint32_t mask = top >= 32 ? ~0u : ((1u << top) - 1u);
    if (b.d[i] & mask) return true;
  }
  return false;
}

// The f64 bits nearest to (A + f) * 2^E2, where 0 <= f < 1 is nonzero iff
// `extra` (A has at least 55 bits whenever extra is set).
__host__ __device__ uint64_t big_round(const DevBig& A, int64_t E2, bool extra) {
  const int L = big_bits(A);
  const int64_t ex = (int64_t)L - 1 + E2;
  if (ex > 1023) return 0x7FF0000000000000ull;
  const int64_t K = ex >= -1022 ? 53 : ex + 1075;
  if (K < 0) return 0;
  const int64_t lsb = L - K;
  uint64_t mant = 0;
  for (int64_t j = K - 1; j >= 0; --j) mant = (mant << 1) | big_bit(A, (int)(lsb + j));
  const uint32_t round = lsb >= 1 ? big_bit(A, (int)(lsb - 1)) : 0u;
  const bool sticky = (lsb >= 2 && big_any_below(A, (int)(lsb - 1))) || extra;
  if (round && (sticky || (mant & 1u))) mant += 1;
  uint64_t b = ex >= -1022 ? ((uint64_t)(ex + 1022) << 52) + mant : mant;
  if (b >= 0x7FF0000000000000ull) b = 0x7FF0000000000000ull;
  return b;
}

// The exact value of a valid item (the grammar was checked by parse_item).
__host__ __device__ bool exact_item(const uint8_t* __restrict__ p, uint64_t n, uint64_t* out) {
  uint64_t i = 0;
  uint64_t sign = 0;
  if (p[0] == '+' || p[0] == '-') {
    sign = p[0] == '-' ? 0x8000000000000000ull : 0ull;
    i = 1;
  }
  DevBig D;
  D.n = 0;
  int nsd = 0;
  int64_t Q = 0;
  while (i < n && csv_digit(p[i])) {
    const uint32_t d = p[i] - '0';
    if (D.n || d) {
      if (++nsd > kMaxSigDigits) return false;
      if (D.n == 0) {
        D.d[0] = d;
        D.n = 1;
      } else {
        big_mul_add(D, 10, d);
      }
    }
    ++i;
  }
  if (i < n && p[i] == '.') {
    ++i;
    while (i < n && csv_digit(p[i])) {
      const uint32_t d = p[i] - '0';
      if (D.n || d) {
        if (++nsd > kMaxSigDigits) return false;
        if (D.n == 0) {
          D.d[0] = d;
          D.n = 1;
        } else {
          big_mul_add(D, 10, d);
        }
      }
      --Q;
      ++i;
    }
  }
  if (i < n && (p[i] | 0x20u) == 'e') {
    ++i;
    bool eneg = false;
    if (i < n && (p[i] == '+' || p[i] == '-')) {
      eneg = p[i] == '-';
      ++i;
    }
    int64_t ev = 0;
    while (i < n && csv_digit(p[i])) {
      if (ev < 100000000) ev = ev * 10 + (p[i] - '0');
      ++i;
    }
    Q += eneg ? -ev : ev;
  }
  uint64_t bits;
  if (D.n == 0) {
    bits = 0;
  } else if (Q >= 0) {
    if (Q + nsd > 310) {
      bits = 0x7FF0000000000000ull;
    } else {
      for (int64_t j = 0; j < Q; ++j) big_mul_add(D, 10, 0);
      bits = big_round(D, 0, false);
    }
  } else {
    const int64_t k = -Q;
    if (k - nsd > 345) {
      bits = 0;
    } else {
      const int64_t need = 56 + (k * 3322 + 999) / 1000 - big_bits(D);  // >= 55 quotient bits
      big_shl(D, need > 0 ? (int)need : 0);
      bool rem = false;
      for (int64_t j = 0; j < k; ++j) rem |= big_div_small(D, 10) != 0;
      bits = big_round(D, -(need > 0 ? need : 0), rem);
    }
  }
  *out = sign | bits;
  return true;
}

}  // namespace

__global__ __launch_bounds__(64) void k_csv_slow(const uint8_t* __restrict__ text, const CsvSlow* __restrict__ slow,
                                                 uint64_t count, double* __restrict__ dst,
                                                 unsigned long long* __restrict__ totals) {
  const uint64_t e = (uint64_t)blockIdx.x * 64 + threadIdx.x;
  if (e >= count) return;
  const CsvSlow it = slow[e];
  uint64_t bits = 0;
  if (exact_item(text + it.start, it.end - it.start, &bits)) {
    dst[it.slot] = __longlong_as_double((long long)bits);
  } else {
    atomicAdd(&totals[4], 1ull);
    atomicMin(&totals[5], (unsigned long long)it.start);
  }
}

// Moves block b's parsed values (items [tok_base, tok_base + ntok), marked in
// okb) to out[ok_base ...] in item order: each lane takes 17 consecutive items.
__global__ __launch_bounds__(256) void k_csv_compact(const CsvBlock* __restrict__ blk, const CsvPre* __restrict__ pre,
                                                     const uint8_t* __restrict__ okb, const double* __restrict__ src,
                                                     double* __restrict__ out) {
  __shared__ uint64_t lds_u[4];
  constexpr uint32_t kPer = (kCsvBlock + 1 + 255) / 256;  // at most kCsvBlock + 1 items end in a block
  const CsvPre bp = pre[blockIdx.x];
  const uint32_t ntok = blk[blockIdx.x].ntok;
  const uint32_t i0 = threadIdx.x * kPer, i1 = i0 + kPer < ntok ? i0 + kPer : ntok;
  uint64_t cnt = 0;
  for (uint32_t i = i0; i < i1; ++i) cnt += okb[bp.tok_base + i];
  uint64_t tot = 0;
  uint64_t r = bp.ok_base + wg_exclusive<uint64_t>(cnt, 0ull, SumOp{}, lds_u, &tot);
  for (uint32_t i = i0; i < i1; ++i)
    if (okb[bp.tok_base + i]) out[r++] = src[bp.tok_base + i];
}

// ---- launchers ------------------------------------------------------------------------

static inline uint64_t rup256(uint64_t x) { return (x + 255) / 256 * 256; }

uint64_t csv_totals_offset(uint64_t nblocks) {
  return rup256(nblocks * sizeof(CsvBlock)) + rup256(nblocks * sizeof(CsvPre));
}
uint64_t csv_scratch_bytes(uint64_t nblocks) { return csv_totals_offset(nblocks) + 256; }

static inline CsvBlock* blk_of(uint8_t* s) { return reinterpret_cast<CsvBlock*>(s); }
static inline CsvPre* pre_of(uint8_t* s, uint64_t nb) {
  return reinterpret_cast<CsvPre*>(s + rup256(nb * sizeof(CsvBlock)));
}

hipError_t launch_csv_count(const uint8_t* text, uint64_t len, uint8_t delim, uint8_t* scratch, uint64_t nblocks,
                            hipStream_t s) {
  unsigned long long* totals = reinterpret_cast<unsigned long long*>(scratch + csv_totals_offset(nblocks));
  launch_k(k_csv_blocks, dim3((uint32_t)nblocks), dim3(256), 0, s, text, len, (uint32_t)delim, blk_of(scratch));
  launch_k(k_csv_scan, dim3(1), dim3(256), 0, s, (const CsvBlock*)blk_of(scratch), pre_of(scratch, nblocks), nblocks, 0,
           totals);
  return hipGetLastError();
}

hipError_t launch_csv_parse(const uint8_t* text, uint64_t len, uint8_t delim, uint32_t skip, const Pow5* tab,
                            uint8_t* scratch, uint64_t nblocks, double* dst, uint8_t* okb, CsvSlow* slow,
                            uint64_t slow_cap, hipStream_t s) {
  unsigned long long* totals = reinterpret_cast<unsigned long long*>(scratch + csv_totals_offset(nblocks));
  launch_k(k_csv_parse, dim3((uint32_t)nblocks), dim3(256), 0, s, text, len, (uint32_t)delim, skip, tab,
           pre_of(scratch, nblocks), dst, okb, slow, slow_cap, totals);
  launch_k(k_csv_scan, dim3(1), dim3(256), 0, s, (const CsvBlock*)blk_of(scratch), pre_of(scratch, nblocks), nblocks, 1,
           totals);
  return hipGetLastError();
}

hipError_t launch_csv_slow(const uint8_t* text, const CsvSlow* slow, uint64_t count, uint8_t* scratch,
                           uint64_t nblocks, double* dst, hipStream_t s) {
  unsigned long long* totals = reinterpret_cast<unsigned long long*>(scratch + csv_totals_offset(nblocks));
  if (count) launch_k(k_csv_slow, dim3((uint32_t)((count + 63) / 64)), dim3(64), 0, s, text, slow, count, dst, totals);
  return hipGetLastError();
}

hipError_t launch_csv_compact(const uint8_t* scratch, uint64_t nblocks, const uint8_t* okb, const double* src,
                              double* out, hipStream_t s) {
  uint8_t* sc = const_cast<uint8_t*>(scratch);
  launch_k(k_csv_compact, dim3((uint32_t)nblocks), dim3(256), 0, s, (const CsvBlock*)blk_of(sc),
           (const CsvPre*)pre_of(sc, nblocks), okb, src, out);
  return hipGetLastError();
}

}  // namespace buffgpu

// ---- host entry points ------------------------------------------------------------------

#include <cstdio>

#include "../../include/buffgpu.h"
#include "buff_runtime.h"

namespace buffgpu {

namespace {

std::once_flag g_pow5_once;
std::vector<Pow5> g_pow5;

const std::vector<Pow5>& pow5_table() {
  std::call_once(g_pow5_once, [] {
    g_pow5.resize(kPow5Max - kPow5Min + 1);
    csv_pow5_table(g_pow5.data());
  });
  return g_pow5;
}

}  // namespace

int csv_parse_device(Ctx* ctx, const uint8_t* text, uint64_t len, uint8_t delim, uint32_t skip, double* out,
                     uint64_t cap, uint64_t* count) {
  if (!ctx || !count || (len && !text)) return fail(BUFF_E_INVALID, "null argument");
  if (delim == '\r' || delim >= 0x80) return fail(BUFF_E_INVALID, "delimiter 0x%02x: an ASCII byte other than CR", delim);
  *count = 0;
  if (len == 0) return BUFF_OK;
  const uint64_t nb = (len + kCsvBlock - 1) / kCsvBlock;
  if (nb > 0x7FFFFFFFull) return fail(BUFF_E_INVALID, "text of %llu bytes", (unsigned long long)len);
  if (!ctx->csv_tab_ready) {
    const std::vector<Pow5>& tab = pow5_table();
    BUFF_TRY(ctx->scratch(ctx->s_csvtab, tab.size() * sizeof(Pow5)));
    HIP_TRY(hipMemcpyAsync(ctx->s_csvtab.ptr, tab.data(), tab.size() * sizeof(Pow5), hipMemcpyHostToDevice,
                           ctx->stream));
    ctx->csv_tab_ready = true;
  }
  BUFF_TRY(ctx->scratch(ctx->s_csv, csv_scratch_bytes(nb)));
  uint8_t* scr = (uint8_t*)ctx->s_csv.ptr;
  unsigned long long tot[kCsvTotals] = {0, 0, 0, 0, 0, 0};
  {
    ProfScope ps(ctx, "k_csv_count");
    HIP_TRY(launch_csv_count(text, len, delim, scr, nb, ctx->stream));
  }
  BUFF_TRY(d2h_small(ctx, tot, scr + csv_totals_offset(nb), sizeof(tot)));
  if (tot[1])
    return fail(BUFF_E_UNSUPPORTED, "bytes >= 0x80: the reference drops every line that is not UTF-8");
  const uint64_t items = tot[0];
  if (items == 0) return BUFF_OK;
  const bool direct = out && cap >= items;
  double* dst = out;
  if (!direct) {
    BUFF_TRY(ctx->scratch(ctx->s_csvtmp, items * 8));
    dst = (double*)ctx->s_csvtmp.ptr;
  }
  // room for the items the 19-digit rounding leaves open (more than 19 digits
  // close to a rounding boundary: rare); they are decided exactly by k_csv_slow
  const uint64_t slow_cap = std::max<uint64_t>(4096, items / 64);
  // [queue of undecided items | one ok byte per item]
  const uint64_t slow_bytes = (slow_cap * sizeof(CsvSlow) + 255) / 256 * 256;
  BUFF_TRY(ctx->scratch(ctx->s_csvslow, slow_bytes + items));
  CsvSlow* slow = (CsvSlow*)ctx->s_csvslow.ptr;
  uint8_t* okb = (uint8_t*)ctx->s_csvslow.ptr + slow_bytes;
  {
    ProfScope ps(ctx, "k_csv_parse");
    HIP_TRY(launch_csv_parse(text, len, delim, skip, (const Pow5*)ctx->s_csvtab.ptr, scr, nb, dst, okb, slow,
                             slow_cap, ctx->stream));
  }
  BUFF_TRY(d2h_small(ctx, tot, scr + csv_totals_offset(nb), sizeof(tot)));
  if (tot[3] && !tot[4]) {
    ProfScope ps(ctx, "k_csv_slow");
    HIP_TRY(launch_csv_slow(text, slow, std::min<uint64_t>(tot[3], slow_cap), scr, nb, dst, ctx->stream));
    BUFF_TRY(d2h_small(ctx, tot, scr + csv_totals_offset(nb), sizeof(tot)));
  }
  if (tot[4])
    return fail(BUFF_E_UNSUPPORTED, "%llu items (the first at byte %llu) have more than %d significant digits or "
                "exceed the exact path's queue", (unsigned long long)tot[4], (unsigned long long)tot[5], 300);
  const uint64_t ok = tot[2];
  *count = ok;
  if (ok > cap) return fail(BUFF_E_SMALL_BUFFER, "%llu values, room for %llu", (unsigned long long)ok,
                            (unsigned long long)cap);
  if (ok && !out) return fail(BUFF_E_INVALID, "null output");
  if (ok < items || !direct) {
    const double* src = dst;
    if (direct) {  // the values sit in out at their items' slots: move them through scratch
      BUFF_TRY(ctx->scratch(ctx->s_csvtmp, items * 8));
      HIP_TRY(hipMemcpyAsync(ctx->s_csvtmp.ptr, out, items * 8, hipMemcpyDeviceToDevice, ctx->stream));
      src = (const double*)ctx->s_csvtmp.ptr;
    }
    ProfScope ps(ctx, "k_csv_compact");
    HIP_TRY(launch_csv_compact(scr, nb, okb, src, out, ctx->stream));
  }
  return BUFF_OK;
}

int csv_parse_host(Ctx* ctx, const uint8_t* text, uint64_t len, uint8_t delim, uint32_t skip, double* out,
                   uint64_t cap, uint64_t* count) {
  if (!ctx || !count || (len && !text)) return fail(BUFF_E_INVALID, "null argument");
  *count = 0;
  if (len == 0) return BUFF_OK;
  BUFF_TRY(ctx->scratch(ctx->s_stream, len));
  HIP_TRY(hipMemcpyAsync(ctx->s_stream.ptr, text, len, hipMemcpyHostToDevice, ctx->stream));
  // every parsed item takes at least one byte and a delimiter
  const uint64_t most = len / 2 + 1;
  BUFF_TRY(ctx->scratch(ctx->s_values, most * 8));
  uint64_t n = 0;
  BUFF_TRY(csv_parse_device(ctx, (const uint8_t*)ctx->s_stream.ptr, len, delim, skip, (double*)ctx->s_values.ptr,
                            most, &n));
  *count = n;
  if (n > cap) return fail(BUFF_E_SMALL_BUFFER, "%llu values, room for %llu", (unsigned long long)n,
                           (unsigned long long)cap);
  if (n) {
    if (!out) return fail(BUFF_E_INVALID, "null output");
    HIP_TRY(hipMemcpyAsync(out, ctx->s_values.ptr, n * 8, hipMemcpyDeviceToHost, ctx->stream));
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return BUFF_OK;
}

}  // namespace buffgpu

using namespace buffgpu;

extern "C" {

int buff_csv_parse_device(buff_ctx* ctx, const uint8_t* text_dev, uint64_t len, uint8_t delim, uint32_t skip_val,
                          double* out_dev, uint64_t cap, uint64_t* count) {
  return csv_parse_device(reinterpret_cast<Ctx*>(ctx), text_dev, len, delim, skip_val, out_dev, cap, count);
}

int buff_csv_parse(buff_ctx* ctx, const uint8_t* text, uint64_t len, uint8_t delim, uint32_t skip_val, double* out,
                   uint64_t cap, uint64_t* count) {
  return csv_parse_host(reinterpret_cast<Ctx*>(ctx), text, len, delim, skip_val, out, cap, count);
}

int buff_construct_file_iterator_skip_newline(const char* file, uint32_t skip_val, uint8_t delim, double* out,
                                              uint64_t cap, uint64_t* count) {
  if (!file || !count) return fail(BUFF_E_INVALID, "null argument");
  FILE* f = std::fopen(file, "rb");
  if (!f) return fail(BUFF_E_INVALID, "cannot open %s", file);  // the reference's Err(())
  std::vector<uint8_t> text;
  uint8_t buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.insert(text.end(), buf, buf + got);
  const bool err = std::ferror(f) != 0;
  std::fclose(f);
  if (err) return fail(BUFF_E_INVALID, "read error on %s", file);
  Ctx* c = nullptr;
  BUFF_TRY(default_ctx(&c));
  return csv_parse_host(c, text.data(), text.size(), delim, skip_val, out, cap, count);
}

}  // extern "C"
