// Byte-slice chunks: BuffSliceCompress (database/src/compress/buff_slice.rs:
// 222-1970) and the buff-simd "_with_slice" filters (buff_simd.rs:3736-4025).
//
// A byte-slice stream carries byte_fixed's header and offsets; only its body
// differs (every sub-column a full byte XOR 0x80, the last one right-padded;
// one plane of 0x80 bytes when F == 0). The resident form is therefore the
// byte_fixed one, and every query runs the byte_fixed kernels: what differs is
// the target the host derives from each reference function's comparison rule,
// and the rows it covers:
//   * the filters read at most four planes: with more, rows that tie the
//     target on those planes are neither greater nor equal, which is a
//     comparison against the target with its low F - 32 bits set (greater) or
//     a range of offsets (equal);
//   * the AVX2 forms cover the first 32*floor(n/32) rows and crash below 32
//     rows (the u32 bound `len - 32` wraps);
//   * the "num" readers (decode_bs, decode_condition, max, sum, max_range)
//     panic unless 1 <= F <= 32;
//   * values are compared as f64, which equals comparing offsets while
//     |base + offset| <= 2^53 (checked; otherwise BUFF_E_UNSUPPORTED).
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstring>
#include <new>
#include <vector>
#include "../../include/buffgpu.h"
#include "buff_common.h"
#include "buff_runtime.h"

namespace buffgpu {

namespace {

enum Kind {
  kRange = 0,           // buff_slice.rs:1348-1438
  kEqual = 1,           // buff_slice.rs:1665-1750
  kRangeNosimd = 2,     // buff_slice.rs:1753-1870
  kEqualNosimd = 3,     // buff_slice.rs:1873-1969
  kRangeCond = 4,       // buff_slice.rs:1441-1550
  kSmallerCond = 5,     // buff_slice.rs:1553-1662
  kRangeWithSlice = 6,  // buff_simd.rs:3736-3886
  kEqualWithSlice = 7,  // buff_simd.rs:3888-4025
};

bool slice_chunk(const Chunk* ch) { return (ch->flags & BUFF_SLICE) != 0; }

// Every stored value is an exact f64 (|base + offset| <= 2^53), so the
// reference's f64 comparisons order the rows as their offsets do.
bool values_exact(const Chunk* ch) {
  const __int128 lo = ch->view.base;
  const __int128 hi = lo + (((__int128)1 << ch->F) - 1);
  const __int128 lim = (__int128)1 << 53;
  return lo >= -lim && hi <= lim;
}

// The per-plane digits of offset target T (T < 2^F).
FilterTarget target_of(const ChunkView& v, uint64_t T) {
  FilterTarget tg;
  memset(&tg, 0, sizeof(tg));
  if (v.F >= 8) {
    for (uint32_t p = 0; p < v.P; ++p) tg.digit[p] = (uint32_t)((T >> (v.F - 8 * (p + 1))) & 0xFF);
    tg.tail = v.r ? (T & ((1ull << v.r) - 1)) : 0;
  } else {
    tg.tail = T;
  }
  return tg;
}

// offset > T into out (T < 2^F; F >= 1).
int run_gt(Ctx* ctx, const Chunk* ch, uint64_t T, uint64_t* out, uint64_t nwords) {
  ProfScope ps(ctx, "k_filter_gt");
  HIP_TRY(launch_filter(ch->view, 0, target_of(ch->view, T), out, nwords, nullptr, ctx->stream,
                        ctx->shape(Op::FilterGt, ch->view)));
  return BUFF_OK;
}

// offset == T into out (T < 2^F; F >= 1).
int run_eq(Ctx* ctx, const Chunk* ch, uint64_t T, uint64_t* out, uint64_t nwords) {
  ProfScope ps(ctx, "k_filter_eq");
  HIP_TRY(launch_filter(ch->view, 1, target_of(ch->view, T), out, nwords, nullptr, ctx->stream,
                        ctx->shape(Op::FilterEq, ch->view)));
  return BUFF_OK;
}

// The comparison a kind makes, reduced to offsets: rows with lo <= offset <= hi
// (`range`), or offset > gt (`greater`), or none / all.
struct Plan {
  enum { kNone, kAll, kGreater, kEqualTo, kBetween, kBelow } what = kNone;
  uint64_t a = 0, b = 0;  // kGreater: a; kEqualTo: a; kBetween: [a, b]; kBelow: offset < a
};

int slice_filter_device(Ctx* ctx, const Chunk* ch, int kind, double pred, uint32_t flags, const uint64_t* cond,
                        uint64_t* words) {
  const ChunkView& v = ch->view;
  const uint64_t n = ch->n, nwords = (n + 63) / 64;
  const bool semantic = (flags & BUFF_SEMANTIC) != 0;
  const bool is_cond = kind == kRangeCond || kind == kSmallerCond;
  const bool with_slice = kind == kRangeWithSlice || kind == kEqualWithSlice;
  const bool nosimd = kind == kRangeNosimd || kind == kEqualNosimd;
  const int rel = (kind == kRange || kind == kRangeNosimd || kind == kRangeCond || kind == kRangeWithSlice) ? 0
                  : kind == kSmallerCond                                                             ? 2
                                                                                                     : 1;
  const uint32_t F = ch->F;
  if (nwords && !words) return fail(BUFF_E_INVALID, "null bitmap");
  if (is_cond && nwords && !cond) return fail(BUFF_E_INVALID, "null condition bitmap");
  const int64_t fixed = fixed_of_pred(ch, pred);
  Plan pl;
  uint64_t cover = n;
  if (fixed < v.base) {
    // every filter prints len and keeps no bitmap (buff_slice.rs:1376-1379 and
    // twins); the condition forms return their empty bitmap (1473-1476)
    if (semantic) pl.what = rel == 0 ? Plan::kAll : Plan::kNone;
    else pl.what = is_cond ? Plan::kNone : Plan::kAll;
  } else {
    const uint64_t t = (uint64_t)fixed - (uint64_t)v.base;
    if (semantic) {
      if (F < 64 && t >= (1ull << F)) {
        pl.what = rel == 2 ? Plan::kAll : Plan::kNone;
      } else if (rel == 0) {
        pl.what = Plan::kGreater, pl.a = t;
      } else if (rel == 1) {
        pl.what = Plan::kEqualTo, pl.a = t;
      } else if (t == 0) {
        pl.what = Plan::kNone;
      } else {
        pl.what = Plan::kBelow, pl.a = t;
      }
    } else {
      // F == 0: ceil(0, 8) wraps to 2^29 slices, the second read is past the
      // stream; with_slice needs a full plane (slice_ptr.get(0).unwrap())
      if (F == 0 || (with_slice && F < 8)) {
        if (nosimd && n == 0) return BUFF_OK;
        return fail(BUFF_E_REF_PANIC, "the reference reads a plane the stream does not have (F = %u)", F);
      }
      if (!nosimd && n < 32)
        return fail(BUFF_E_REF_PANIC, "%llu rows: the reference's u32 bound len - 32 wraps and its AVX2 loop "
                                      "reads past the planes", (unsigned long long)n);
      const uint64_t tm = F < 64 ? t & ((1ull << F) - 1) : t;
      uint32_t d = 0;  // low offset bits below the planes the filter reads
      if (with_slice) {
        const uint32_t P = F / 8, r = F % 8;
        if (P > 4 && r > 0)
          return fail(BUFF_E_UNSUPPORTED, "F = %u: the reference compares the tail after four of %u planes", F, P);
        if (P > 4) d = F - 32;
        if (rel == 1 && r == 0) {
          pl.what = Plan::kNone;  // matches stay in rb1; res is never set (3977-4022)
        } else if (rel == 1) {
          pl.what = Plan::kEqualTo, pl.a = tm;
        } else {
          pl.what = Plan::kGreater, pl.a = tm | ((1ull << d) - 1);
        }
      } else {
        const uint32_t ns = (F + 7) / 8;
        if (ns > 4) d = F - 32;
        const uint64_t lo = d ? (tm >> d) << d : tm, hi = lo | ((1ull << d) - 1);
        if (rel == 0) {
          pl.what = Plan::kGreater, pl.a = hi;
        } else if (rel == 1) {
          if (d == 0) pl.what = Plan::kEqualTo, pl.a = tm;
          else pl.what = Plan::kBetween, pl.a = lo, pl.b = hi;
        } else {
          if (lo == 0) pl.what = Plan::kNone;
          else pl.what = Plan::kBelow, pl.a = lo;
        }
      }
      if (!nosimd) cover = n / 32 * 32;
    }
  }
  if (nwords == 0) return BUFF_OK;
  const uint64_t* cm = is_cond ? cond : nullptr;
  // F == 0: every offset is 0
  if (F == 0 && pl.what != Plan::kNone && pl.what != Plan::kAll) {
    bool all = pl.what == Plan::kGreater ? false
               : pl.what == Plan::kEqualTo ? pl.a == 0
               : pl.what == Plan::kBetween ? pl.a == 0
                                           : pl.a > 0;
    pl.what = all ? Plan::kAll : Plan::kNone;
  }
  switch (pl.what) {
    case Plan::kNone:
      HIP_TRY(launch_fill_bitmap(words, nwords, n, 0, ctx->stream));
      return BUFF_OK;
    case Plan::kAll:
      HIP_TRY(launch_bitmap_logic(words, nullptr, nullptr, cm, nwords, cover, 0, 0, ctx->stream));
      return BUFF_OK;
    case Plan::kGreater:
      BUFF_TRY(run_gt(ctx, ch, pl.a, words, nwords));
      break;
    case Plan::kEqualTo:
      BUFF_TRY(run_eq(ctx, ch, pl.a, words, nwords));
      break;
    case Plan::kBelow:  // offset < a  <=>  not (offset > a - 1), a >= 1
      BUFF_TRY(run_gt(ctx, ch, pl.a - 1, words, nwords));
      HIP_TRY(launch_bitmap_logic(words, words, nullptr, cm, nwords, cover, 1, 0, ctx->stream));
      return BUFF_OK;
    case Plan::kBetween: {  // offset > a - 1 and not offset > b
      BUFF_TRY(ctx->scratch(ctx->s_slice, nwords * 8));
      uint64_t* hiw = (uint64_t*)ctx->s_slice.ptr;
      BUFF_TRY(run_gt(ctx, ch, pl.b, hiw, nwords));
      if (pl.a == 0) {
        HIP_TRY(launch_bitmap_logic(words, hiw, nullptr, cm, nwords, cover, 1, 0, ctx->stream));
      } else {
        BUFF_TRY(run_gt(ctx, ch, pl.a - 1, words, nwords));
        HIP_TRY(launch_bitmap_logic(words, words, hiw, cm, nwords, cover, 0, 1, ctx->stream));
      }
      return BUFF_OK;
    }
  }
  if (cm || cover < n) HIP_TRY(launch_bitmap_logic(words, words, nullptr, cm, nwords, cover, 0, 0, ctx->stream));
  return BUFF_OK;
}

}  // namespace

uint64_t slice_stream_bytes(uint64_t n, uint32_t F) { return kStreamHeader + n * (F == 0 ? 1 : (F + 7) / 8); }

// buff_slice.rs:298-437 layout of a resident byte-slice chunk.
int slice_serialize(Ctx* ctx, const Chunk* ch, uint8_t* dst, uint64_t* len) {
  const uint64_t total = slice_stream_bytes(ch->n, ch->F);
  const uint64_t need = *len;
  *len = total;
  if (!dst) return BUFF_OK;
  if (need < total) return fail(BUFF_E_SMALL_BUFFER, "stream needs %llu bytes", (unsigned long long)total);
  const uint64_t ubase = (uint64_t)ch->view.base;
  const uint32_t hdr[5] = {(uint32_t)ubase, (uint32_t)(ubase >> 32), (uint32_t)ch->n, ch->ilen, ch->dlen};
  memcpy(dst, hdr, 20);
  const uint64_t body = total - kStreamHeader;
  if (body == 0) return BUFF_OK;
  BUFF_TRY(ctx->scratch(ctx->s_stream, body));
  uint8_t* dev = (uint8_t*)ctx->s_stream.ptr;
  {
    ProfScope ps(ctx, "k_serialize");
    if (ch->F == 0) {
      HIP_TRY(launch_fill_bytes(dev, body, 0x80, ctx->stream));  // flip(0 << 8) per row
    } else {
      ChunkView planes = ch->view;
      planes.r = 0;
      if (planes.P) HIP_TRY(launch_serialize(planes, 0x80, dev, 0, ctx->stream));
      if (ch->view.r) HIP_TRY(launch_slice_tail_serialize(ch->view, dev + (uint64_t)ch->view.P * ch->n, ctx->stream));
    }
  }
  HIP_TRY(hipMemcpyAsync(dst + kStreamHeader, dev, body, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return BUFF_OK;
}

int slice_load(Ctx* ctx, const uint8_t* s, uint64_t len, uint64_t scale, uint32_t flags, Chunk** out) {
  (void)flags;
  if (len < kStreamHeader)
    return fail(BUFF_E_INVALID, "stream of %llu bytes has no 20-byte header", (unsigned long long)len);
  uint32_t hdr[5];
  memcpy(hdr, s, 20);
  const uint64_t n = hdr[2];
  const uint32_t F = hdr[3] + hdr[4];  // u32 wrap, as the readers compute `dlen + ilen`
  if (F > 63) return fail(BUFF_E_INVALID, "stored width dlen+ilen = %u is not a BUFF stream width", F);
  const uint64_t expect = slice_stream_bytes(n, F);
  if (len < expect)
    return fail(BUFF_E_INVALID, "stream is %llu bytes, its header needs %llu", (unsigned long long)len,
                (unsigned long long)expect);
  Chunk* ch = new (std::nothrow) Chunk();
  if (!ch) return fail(BUFF_E_OOM, "host allocation");
  ch->ctx = ctx;
  ch->n = n;
  ch->scale = scale;
  ch->flags = BUFF_SLICE;
  ch->ilen = hdr[3];
  ch->dlen = hdr[4];
  ch->F = F;
  ch->view.base = (int64_t)((uint64_t)hdr[0] | ((uint64_t)hdr[1] << 32));
  ch->view.dlen = ch->dlen;
  chunk_set_width(ch, F);
  int rc = chunk_alloc(ctx, ch);
  if (rc != BUFF_OK) {
    delete ch;
    return rc;
  }
  if (n > 0 && F > 0) {
    const uint64_t body = expect - kStreamHeader;
    rc = ctx->scratch(ctx->s_stream, body);
    if (rc == BUFF_OK) {
      uint8_t* dev = (uint8_t*)ctx->s_stream.ptr;
      hipError_t e = hipMemcpyAsync(dev, s + kStreamHeader, body, hipMemcpyHostToDevice, ctx->stream);
      if (e == hipSuccess) {
        ProfScope ps(ctx, "k_load");
        ChunkView planes = ch->view;
        planes.r = 0;
        if (planes.P) e = launch_load(planes, 0x80, dev, 0, ctx->stream);
        if (e == hipSuccess && ch->view.r) e = launch_slice_tail_load(ch->view, dev + (uint64_t)ch->view.P * n, ctx->stream);
      }
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // the host stream may go away
      if (e != hipSuccess) rc = fail(BUFF_E_DEVICE, "stream upload: %s", hipGetErrorString(e));
    }
    if (rc != BUFF_OK) {
      (void)ctx->release(ch->mem);
      delete ch;
      return rc;
    }
  }
  *out = ch;
  return BUFF_OK;
}

// Filters of byte-slice chunks (op GT / EQ; BUFF_SCALAR: the nosimd forms; with
// a condition: GT / BUFF_OP_LT) and the _with_slice filters of byte_fixed
// chunks (BUFF_WITH_SLICE). Returns -1 when the call is not one of these.
int slice_filter_dispatch(Ctx* ctx, const Chunk* ch, int op, double pred, uint32_t flags, bool conditional,
                          const uint64_t* cond, uint64_t* words) {
  const bool scalar = (flags & BUFF_SCALAR) != 0;
  if (slice_chunk(ch)) {
    if (flags & BUFF_WITH_SLICE) return fail(BUFF_E_INVALID, "BUFF_WITH_SLICE reads a byte_fixed chunk");
    int kind;
    if (conditional) {
      if (op == BUFF_OP_GT) kind = kRangeCond;
      else if (op == BUFF_OP_LT) kind = kSmallerCond;
      else return fail(BUFF_E_UNSUPPORTED, "the reference has no conditional equality filter of a byte-slice stream");
    } else {
      if (op == BUFF_OP_GT) kind = scalar ? kRangeNosimd : kRange;
      else if (op == BUFF_OP_EQ) kind = scalar ? kEqualNosimd : kEqual;
      else return fail(BUFF_E_INVALID, "filter op %d takes a condition", op);
    }
    return slice_filter_device(ctx, ch, kind, pred, flags, cond, words);
  }
  if (flags & BUFF_WITH_SLICE) {
    if (ch->flags & BUFF_MAJORITY) return fail(BUFF_E_UNSUPPORTED, "BUFF_WITH_SLICE reads a byte_fixed chunk");
    if (conditional) return fail(BUFF_E_UNSUPPORTED, "the _with_slice filters take no condition");
    if (op != BUFF_OP_GT && op != BUFF_OP_EQ) return fail(BUFF_E_INVALID, "unknown filter op %d", op);
    return slice_filter_device(ctx, ch, op == BUFF_OP_GT ? kRangeWithSlice : kEqualWithSlice, pred, flags, nullptr,
                               words);
  }
  if (op == BUFF_OP_LT) return fail(BUFF_E_UNSUPPORTED, "BUFF_OP_LT is the byte-slice smaller filter");
  return -1;
}

// buff_slice_decode (buff_slice.rs:439-545) is the exact value of every row;
// decode_bs (1184-1265, BUFF_SCALAR) and decode_condition (1266-1347) panic
// unless 1 <= F <= 32. No reduced-precision form.
int slice_materialize_check(const Chunk* ch, int32_t precision, uint32_t flags) {
  if (precision >= 0)
    return fail(BUFF_E_UNSUPPORTED, "the reference has no reduced-precision decode of a byte-slice stream");
  if ((flags & BUFF_SCALAR) && !(flags & BUFF_SEMANTIC) && (ch->F == 0 || ch->F > 32))
    return fail(BUFF_E_REF_PANIC, "bit length greater than 32 is not supported yet (F = %u)", ch->F);
  return BUFF_OK;
}

// buff_slice_buffsum (702-814) / buff_slice_sum (1101-1182, BUFF_SCALAR: panics
// unless 1 <= F <= 32): the f64 running sum of the row values. ref_sum is the
// exact sum rounded once; the running sum differs from it by the rounding of
// its n additions (the oracle's agrees within 1e-12 relative on the tested columns).
int slice_sum(Ctx* ctx, const Chunk* ch, uint32_t flags, double* ref_sum, double* exact_sum, uint64_t* plane_sums) {
  if ((flags & BUFF_SCALAR) && !(flags & BUFF_SEMANTIC) && (ch->F == 0 || ch->F > 32))
    return fail(BUFF_E_REF_PANIC, "bit length greater than 32 is not supported yet (F = %u)", ch->F);
  double ex = 0.0;
  BUFF_TRY(sum_op(ctx, ch, nullptr, &ex, plane_sums));
  if (ref_sum) *ref_sum = ex;
  if (exact_sum) *exact_sum = ex;
  return BUFF_OK;
}

// buff_slice_buffmax (548-700) / buff_slice_max (816-929, BUFF_SCALAR): the max
// f64 value (f64::MIN for an empty column) and the rows holding it.
int slice_max(Ctx* ctx, const Chunk* ch, uint32_t flags, double* value, uint64_t* max_off, uint64_t* argmax) {
  const bool semantic = (flags & BUFF_SEMANTIC) != 0;
  if ((flags & BUFF_SCALAR) && !semantic && (ch->F == 0 || ch->F > 32))
    return fail(BUFF_E_REF_PANIC, "bit length greater than 32 is not supported yet (F = %u)", ch->F);
  if (!semantic && !values_exact(ch))
    return fail(BUFF_E_UNSUPPORTED, "values beyond 2^53: the reference compares rounded f64 values");
  if (ch->n == 0) {
    if (value) *value = -DBL_MAX;
    if (max_off) *max_off = 0;
    if (argmax) *argmax = 0;
    return BUFF_OK;
  }
  uint64_t best = 0;
  BUFF_TRY(max_op(ctx, ch, 0, value, &best, nullptr));
  if (max_off) *max_off = best;
  if (argmax) {
    const uint64_t nwords = (ch->n + 63) / 64;
    uint64_t cnt = ch->n;  // F == 0: every row holds the max
    if (ch->F > 0) {
      BUFF_TRY(ctx->scratch(ctx->s_slice, nwords * 8));
      BUFF_TRY(run_eq(ctx, ch, best, (uint64_t*)ctx->s_slice.ptr, nwords));
      std::vector<uint64_t> host(nwords);
      HIP_TRY(hipMemcpyAsync(host.data(), ctx->s_slice.ptr, nwords * 8, hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      cnt = 0;
      for (uint64_t w : host) cnt += (uint64_t)__builtin_popcountll(w);
    }
    *argmax = cnt;
  }
  return BUFF_OK;
}

// buff_slice_max_range (930-1098): a panic unless 1 <= F <= 32 and the windows
// tile a non-empty [s, e); then every window's max and argmax rows, which is
// the BUFF_SEMANTIC windowed max.
int slice_max_range_check(const Chunk* ch, uint64_t s, uint64_t e, uint64_t W, uint32_t flags) {
  if (flags & BUFF_SEMANTIC) return BUFF_OK;
  if (ch->F == 0 || ch->F > 32)
    return fail(BUFF_E_REF_PANIC, "bit length greater than 32 is not supported yet (F = %u)", ch->F);
  if (e == s || (e - s) % W)
    return fail(BUFF_E_REF_PANIC, "the reference asserts that windows of %llu rows tile %llu rows",
                (unsigned long long)W, (unsigned long long)(e - s));
  if (!values_exact(ch))
    return fail(BUFF_E_UNSUPPORTED, "values beyond 2^53: the reference compares rounded f64 values");
  return BUFF_OK;
}

}  // namespace buffgpu

// ---- reference-shaped entry points (BuffSliceCompress methods on host streams) ----

using namespace buffgpu;

namespace {

int sl_load(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, uint32_t flags, Ctx** c,
            buff_chunk** ch) {
  if (!sd || !bytes) return fail(BUFF_E_INVALID, "null argument");
  BUFF_TRY(default_ctx(c));
  return buff_chunk_load(reinterpret_cast<buff_ctx*>(*c), bytes, len, sd->scale, flags, ch);
}

int sl_decode(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, uint32_t qflags, double* out,
              uint64_t cap) {
  Ctx* c = nullptr;
  buff_chunk* ch = nullptr;
  BUFF_TRY(sl_load(sd, bytes, len, BUFF_SLICE, &c, &ch));
  const uint64_t n = reinterpret_cast<Chunk*>(ch)->n;
  int rc = cap < n ? fail(BUFF_E_SMALL_BUFFER, "output needs %llu doubles", (unsigned long long)n) : BUFF_OK;
  if (rc == BUFF_OK) rc = buff_materialize(reinterpret_cast<buff_ctx*>(c), ch, -1, qflags, out);
  (void)buff_chunk_free(ch);
  return rc;
}

int sl_filter(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, uint32_t lflags, int op,
              uint32_t qflags, double pred, uint32_t* words, uint64_t nwords, uint64_t* count) {
  Ctx* c = nullptr;
  buff_chunk* ch = nullptr;
  BUFF_TRY(sl_load(sd, bytes, len, lflags, &c, &ch));
  int rc = buff_filter(reinterpret_cast<buff_ctx*>(c), ch, op, pred, qflags, words, nwords, count);
  (void)buff_chunk_free(ch);
  return rc;
}

int sl_sum(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, uint32_t qflags, double* sum) {
  Ctx* c = nullptr;
  buff_chunk* ch = nullptr;
  BUFF_TRY(sl_load(sd, bytes, len, BUFF_SLICE, &c, &ch));
  int rc = buff_sum_ex(reinterpret_cast<buff_ctx*>(c), ch, qflags, sum, nullptr, nullptr);
  (void)buff_chunk_free(ch);
  return rc;
}

int sl_max(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, uint32_t qflags, double* max_value,
           uint64_t* argmax_count) {
  Ctx* c = nullptr;
  buff_chunk* ch = nullptr;
  BUFF_TRY(sl_load(sd, bytes, len, BUFF_SLICE, &c, &ch));
  int rc = buff_max_ex(reinterpret_cast<buff_ctx*>(c), ch, qflags, max_value, nullptr, argmax_count);
  (void)buff_chunk_free(ch);
  return rc;
}

}  // namespace

extern "C" {

int buff_sd_buff_slice_encode(const buff_split_double* sd, const double* data, uint64_t n, uint8_t* out, uint64_t cap,
                              uint64_t* out_len) {
  if (!sd || !out_len) return fail(BUFF_E_INVALID, "null argument");
  Ctx* c = nullptr;
  BUFF_TRY(default_ctx(&c));
  buff_chunk* ch = nullptr;
  BUFF_TRY(buff_encode(reinterpret_cast<buff_ctx*>(c), data, n, sd->scale, BUFF_SLICE, &ch));
  uint64_t len = cap;
  int rc = buff_chunk_serialize(reinterpret_cast<buff_ctx*>(c), ch, out, &len);
  *out_len = len;
  (void)buff_chunk_free(ch);
  return rc;
}

int buff_sd_buff_slice_decode(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* out,
                              uint64_t cap) {
  return sl_decode(sd, bytes, len, 0, out, cap);
}

int buff_sd_buff_slice_decode_bs(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* out,
                                 uint64_t cap) {
  return sl_decode(sd, bytes, len, BUFF_SCALAR, out, cap);
}

int buff_sd_buff_slice_range_filter(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double pred,
                                    uint32_t* words, uint64_t nwords, uint64_t* count) {
  return sl_filter(sd, bytes, len, BUFF_SLICE, BUFF_OP_GT, 0, pred, words, nwords, count);
}

int buff_sd_buff_slice_equal_filter(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double pred,
                                    uint32_t* words, uint64_t nwords, uint64_t* count) {
  return sl_filter(sd, bytes, len, BUFF_SLICE, BUFF_OP_EQ, 0, pred, words, nwords, count);
}

int buff_sd_buff_slice_range_filter_nosimd(const buff_split_double* sd, const uint8_t* bytes, uint64_t len,
                                           double pred, uint32_t* words, uint64_t nwords, uint64_t* count) {
  return sl_filter(sd, bytes, len, BUFF_SLICE, BUFF_OP_GT, BUFF_SCALAR, pred, words, nwords, count);
}

int buff_sd_buff_slice_equal_filter_nosimd(const buff_split_double* sd, const uint8_t* bytes, uint64_t len,
                                           double pred, uint32_t* words, uint64_t nwords, uint64_t* count) {
  return sl_filter(sd, bytes, len, BUFF_SLICE, BUFF_OP_EQ, BUFF_SCALAR, pred, words, nwords, count);
}

int buff_sd_buff_simd_range_filter_with_slice(const buff_split_double* sd, const uint8_t* bytes, uint64_t len,
                                              double pred, uint32_t* words, uint64_t nwords, uint64_t* count) {
  return sl_filter(sd, bytes, len, BUFF_FLIP, BUFF_OP_GT, BUFF_WITH_SLICE, pred, words, nwords, count);
}

int buff_sd_buff_simd_equal_filter_with_slice(const buff_split_double* sd, const uint8_t* bytes, uint64_t len,
                                              double pred, uint32_t* words, uint64_t nwords, uint64_t* count) {
  return sl_filter(sd, bytes, len, BUFF_FLIP, BUFF_OP_EQ, BUFF_WITH_SLICE, pred, words, nwords, count);
}

int buff_sd_buff_slice_buffsum(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* sum) {
  return sl_sum(sd, bytes, len, 0, sum);
}

int buff_sd_buff_slice_sum(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* sum) {
  return sl_sum(sd, bytes, len, BUFF_SCALAR, sum);
}

int buff_sd_buff_slice_buffmax(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* max_value,
                               uint64_t* argmax_count) {
  return sl_max(sd, bytes, len, 0, max_value, argmax_count);
}

int buff_sd_buff_slice_max(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* max_value,
                           uint64_t* argmax_count) {
  return sl_max(sd, bytes, len, BUFF_SCALAR, max_value, argmax_count);
}

}  // extern "C"
