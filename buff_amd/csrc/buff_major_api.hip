// Host side of the buff-major (majority / outlier) layout: encode, stream
// conversion and the queries of database/src/compress/buff_simd.rs:317-2861 on
// the resident form of buff_major.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/buffgpu.h"
#include "buff_common.h"
#include "buff_major.h"
#include "buff_runtime.h"
#include "roaring_portable.h"
#include "snappy_raw.h"

namespace buffgpu {

namespace {

constexpr uint64_t kSample = 2000;   // split_double.rs:25 SAMPLE
// split_double.rs:27 MAJOR_R = 0.6f32; (2000 as f32 * 0.6f32) as i32 == 1200 (buff_simd.rs:393)
constexpr int32_t kMajorLimit = 1200;
// readers fetch a lane's outlier bytes as one 20-byte window past its position
constexpr uint64_t kOutlierPad = 64;

// Boyer-Moore vote over `(offset >> rsf) as u8` (buff_simd.rs:402-412, 424-436).
void vote(const std::vector<uint64_t>& head, uint32_t rsf, uint8_t* major, int32_t* freq) {
  int32_t f = 0;
  uint8_t m = 0;
  for (uint64_t v : head) {
    const uint8_t u = (uint8_t)(v >> (rsf & 63));
    if (f == 0) {
      m = u;
      f = 1;
    } else if (m == u) {
      f += 1;
    } else {
      f -= 1;
    }
  }
  *major = m;
  *freq = f;
}

uint64_t mask_bytes(uint64_t ntiles) { return ntiles * 64 * 2; }
uint64_t prefix_bytes(uint64_t ntiles) { return round_up((ntiles + 1) * 4, 256); }

// Places the bitmaps and prefixes of S sparse columns in one allocation.
int sparse_alloc(Ctx* ctx, Chunk* ch) {
  const uint64_t ntiles = ch->view.ntiles;
  const uint64_t per = mask_bytes(ntiles) + prefix_bytes(ntiles);
  void* p = nullptr;
  BUFF_TRY(ctx->alloc(std::max<uint64_t>(per * ch->S, 256), &p));
  ch->mem_sparse = p;
  ch->device_bytes += per * ch->S;
  for (uint32_t j = 0; j < ch->S; ++j) {
    uint8_t* b = (uint8_t*)p + per * j;
    ch->sp[j].mask16 = (uint16_t*)b;
    ch->sp[j].prefix = (uint32_t*)(b + mask_bytes(ntiles));
  }
  return BUFF_OK;
}

// Outlier bytes and row positions of every column, each at a 256-byte aligned offset.
int outlier_alloc(Ctx* ctx, Chunk* ch) {
  uint64_t tot = 0;
  for (uint32_t j = 0; j < ch->S; ++j)
    tot += round_up(ch->sp[j].n_out + kOutlierPad, 256) + round_up(2 * ch->sp[j].n_out + kOutlierPad, 256);
  void* p = nullptr;
  BUFF_TRY(ctx->alloc(std::max<uint64_t>(tot, 256), &p));
  ch->mem_outl = p;
  ch->device_bytes += tot;
  uint64_t off = 0;
  for (uint32_t j = 0; j < ch->S; ++j) {
    ch->sp[j].outl = (uint8_t*)p + off;
    off += round_up(ch->sp[j].n_out + kOutlierPad, 256);
    ch->sp[j].pos = (uint16_t*)((uint8_t*)p + off);
    off += round_up(2 * ch->sp[j].n_out + kOutlierPad, 256);
  }
  return BUFF_OK;
}

// The row positions of every sparse column (after its bitmap and prefixes).
hipError_t build_positions(Ctx* ctx, Chunk* ch) {
  for (uint32_t j = 0; j < ch->S; ++j) {
    const hipError_t e =
        launch_major_positions(ch->sp[j].mask16, ch->sp[j].prefix, ch->view.ntiles, ch->sp[j].pos, ctx->stream);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Encode scratch in ctx->s_major: staging slots and tile counts of every column,
// workgroup sums, then the per-column totals.
struct EncScratch {
  uint64_t counts, stage, wgsum, totals, bytes;
};

EncScratch enc_layout(uint32_t S, uint64_t ntiles, uint32_t nwg) {
  EncScratch l;
  l.stage = 0;
  l.counts = (uint64_t)S * ntiles * kTileRows;
  l.wgsum = l.counts + round_up((uint64_t)S * ntiles * 4, 256);
  l.totals = l.wgsum + round_up((uint64_t)S * nwg * 4, 256);
  l.bytes = l.totals + 16ull * kMaxSparse;
  return l;
}

MajorEnc enc_of(const Chunk* ch, uint8_t* scratch) {
  MajorEnc me;
  memset(&me, 0, sizeof(me));
  me.S = ch->S;
  major_enc_geometry(ch->view.ntiles, &me);
  const EncScratch l = enc_layout(ch->S, ch->view.ntiles, me.nwg);
  for (uint32_t j = 0; j < ch->S; ++j) {
    me.major[j] = ch->sp[j].major;
    me.shift[j] = ch->sp[j].shift;
    me.mask16[j] = ch->sp[j].mask16;
    me.prefix[j] = ch->sp[j].prefix;
    me.outl[j] = ch->sp[j].outl;
    if (scratch) {
      me.stage[j] = scratch + l.stage + (uint64_t)j * ch->view.ntiles * kTileRows;
      me.counts[j] = (uint32_t*)(scratch + l.counts) + (uint64_t)j * ch->view.ntiles;
      me.wgsum[j] = (uint32_t*)(scratch + l.wgsum) + (uint64_t)j * me.nwg;
    }
  }
  return me;
}

Chunk* new_major_chunk(Ctx* ctx, uint64_t n, uint64_t scale, uint32_t flags, uint32_t dlen, uint32_t ilen, int64_t base,
                       uint32_t F, uint32_t S) {
  Chunk* ch = new (std::nothrow) Chunk();
  if (!ch) return nullptr;
  ch->ctx = ctx;
  ch->n = n;
  ch->scale = scale;
  ch->flags = (flags & (BUFF_FLIP | BUFF_RESIDUE)) | BUFF_MAJORITY;
  ch->dlen = dlen;
  ch->ilen = ilen;
  ch->F = F;
  ch->S = S;
  ch->view.base = base;
  ch->view.dlen = dlen;
  chunk_set_width(ch, S ? F - 8 * S : F);
  return ch;
}

void free_partial(Ctx* ctx, Chunk* ch) {
  if (ch->mem) (void)ctx->release(ch->mem);
  if (ch->mem_sparse) (void)ctx->release(ch->mem_sparse);
  if (ch->mem_outl) (void)ctx->release(ch->mem_outl);
  delete ch;
}

}  // namespace

MajorView major_view(const Chunk* ch) {
  MajorView mv;
  memset(&mv, 0, sizeof(mv));
  mv.dense = ch->view;
  mv.S = ch->S;
  mv.F = ch->F;
  for (uint32_t j = 0; j < ch->S; ++j) {
    const SparseCol& c = ch->sp[j];
    mv.col[j].mask16 = c.mask16;
    mv.col[j].prefix = c.prefix;
    mv.col[j].outl = c.outl;
    mv.col[j].pos = c.pos;
    mv.col[j].last = c.last;
    mv.col[j].major = c.major;
    mv.col[j].shift = c.shift;
  }
  return mv;
}


void major_release(Chunk* ch) {
  if (ch->mem_sparse) (void)ch->ctx->release(ch->mem_sparse);
  if (ch->mem_outl) (void)ch->ctx->release(ch->mem_outl);
  ch->mem_sparse = ch->mem_outl = nullptr;
}

// split_double.rs:680-716: the leading-zero histogram of the first 2000
// offsets (clz - (64 - F), usize, counted below 32), summed byte by byte over
// the ilen / 8 leading bytes; a byte becomes an outlier column while fewer than
// (2000 as f32 * 0.1f32) as u32 = 200 samples have a nonzero bit up to it. The
// histogram has 32 slots: reading slot 32 panics.
int residue_columns(const std::vector<uint64_t>& head, uint32_t F, uint64_t ilen64, uint32_t* S) {
  const uint64_t head_len = 64 - (uint64_t)F;
  uint32_t counter[32] = {0};
  for (uint64_t v : head) {
    const uint64_t lz = (uint64_t)(v ? __builtin_clzll(v) : 64) - head_len;
    if (lz < 32) counter[lz] += 1;
  }
  const uint32_t limit = (uint32_t)((float)kSample * 0.1f);
  const uint64_t adjusted = (ilen64 / 8) * 8;
  uint32_t sum = 0, cols = 0;
  for (uint64_t i = 0; i < adjusted; ++i) {
    if (i >= 32) return fail(BUFF_E_REF_PANIC, "byte_residue_encode reads bits_counter[%llu] of 32", (unsigned long long)i);
    sum += counter[i];
    if (i % 8 == 7) {
      if (sum < limit) cols += 1;
      else break;
    }
  }
  if (cols > kMaxSparse) return fail(BUFF_E_UNSUPPORTED, "%u outlier columns", cols);
  *S = cols;
  return BUFF_OK;
}

// ---------------------------------------------------------------------------
// buff_simd.rs:317-630 (buff_encode_majority)

int major_build(Ctx* ctx, const double* x, uint64_t n, uint64_t scale, uint32_t flags, int32_t pexp, uint32_t dlen,
                int64_t mn, uint32_t F, uint32_t S, const uint8_t* majors, Chunk** out);

int major_encode_device(Ctx* ctx, const double* x, uint64_t n, uint64_t scale, uint32_t flags, Chunk** out) {
  if (!ctx || !out) return fail(BUFF_E_INVALID, "null argument");
  if (n > 0xFFFFFFFFull) return fail(BUFF_E_INVALID, "a chunk holds at most 2^32-1 rows (n is stored as u32)");
  if (n && !x) return fail(BUFF_E_INVALID, "null input");
  const int prec = scale == 0 ? 0 : prec_of_scale(scale);
  const int dl = precision_map(prec);
  if (dl < 0)
    return fail(BUFF_E_REF_PANIC, "precision %d (scale %llu) is not in PRECISION_MAP", prec, (unsigned long long)scale);
  const bool residue = (flags & BUFF_RESIDUE) != 0;
  if (n < kSample && !residue)
    return fail(BUFF_E_REF_PANIC, "buff-major samples the first %llu rows (fixed_vec[0..SAMPLE]); the column has %llu",
                (unsigned long long)kSample, (unsigned long long)n);
  if (residue && n == 0)   // min/max stay at their initial values, ilen wraps above 15, the sample slice panics
    return fail(BUFF_E_REF_PANIC, "byte_residue_encode of an empty column samples fixed_vec[0..SAMPLE]");
  const uint32_t dlen = (uint32_t)dl;
  const int32_t pexp = precision_exp_of_prec(prec);

  // plain min / max (buff_simd.rs:340-349: `if` / `if`)
  MinMaxOut mm;
  BUFF_TRY(ctx->minmax_pass(x, n, pexp, dlen, &mm));
  const int64_t mn = mm.min_all;
  const int64_t mx = std::max(mm.first, mm.max_rest);
  const uint32_t F = width_of_delta((int64_t)((uint64_t)mx - (uint64_t)mn));
  if (F > 63) return fail(BUFF_E_UNSUPPORTED, "bit width %u exceeds 63", F);

  // byte_residue_encode (split_double.rs:624-954): ilen = fixed_len - dec_len in
  // usize (wrapping for a constant column); with ilen > 15 the leading byte
  // columns whose top byte is nonzero in fewer than 10% of the first 2000
  // offsets become outlier columns (majority byte 0), else the byte_fixed layout
  const uint64_t ilen64 = (uint64_t)F - dlen;
  if (residue && ilen64 <= 15)
    return major_build(ctx, x, n, scale, flags, pexp, dlen, mn, F, 0, nullptr, out);
  if (residue && n < kSample)
    return fail(BUFF_E_REF_PANIC, "byte_residue_encode samples fixed_vec[0..SAMPLE] (ilen %llu > 15); the column "
                "has %llu rows", (unsigned long long)ilen64, (unsigned long long)n);
  // the vote reads the offsets of the first 2000 rows
  std::vector<double> head_x(kSample);
  HIP_TRY(hipMemcpyAsync(head_x.data(), x, kSample * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  std::vector<uint64_t> head(kSample);
  for (uint64_t i = 0; i < kSample; ++i) {
    uint64_t u;
    memcpy(&u, &head_x[i], 8);
    head[i] = (uint64_t)fetch_fixed_aligned_bits(u, pexp, dlen) - (uint64_t)mn;
  }
  uint8_t majors[kMaxSparse] = {0};
  uint32_t S = 0;
  if (residue) {
    BUFF_TRY(residue_columns(head, F, ilen64, &S));
  } else {
    uint32_t rsf = F >= 8 ? F - 8 : 0;
    uint8_t m;
    int32_t fq;
    vote(head, rsf, &m, &fq);
    if (fq > kMajorLimit) {
      majors[S++] = m;
      while (rsf > 7) {
        rsf -= 8;
        vote(head, rsf, &m, &fq);
        if (fq > kMajorLimit) majors[S++] = m;
        else break;
      }
    }
    if (S > 0 && F < 8)
      return fail(BUFF_E_REF_PANIC, "F = %u < 8 with a majority byte: `remain -= 8` underflows (buff_simd.rs:465)", F);
  }
  return major_build(ctx, x, n, scale, flags, pexp, dlen, mn, F, S, majors, out);
}

// The resident chunk of S sparse columns with the given majority bytes (the
// dense planes and tail below them) from the float64 column x.
int major_build(Ctx* ctx, const double* x, uint64_t n, uint64_t scale, uint32_t flags, int32_t pexp, uint32_t dlen,
                int64_t mn, uint32_t F, uint32_t S, const uint8_t* majors, Chunk** out) {
  Chunk* ch = new_major_chunk(ctx, n, scale, flags, dlen, (uint32_t)((uint64_t)F - dlen), mn, F, S);
  if (!ch) return fail(BUFF_E_OOM, "host allocation");
  int rc = chunk_alloc(ctx, ch);
  if (rc != BUFF_OK) {
    delete ch;
    return rc;
  }
  if (S == 0) {
    if (ch->view.P || ch->view.r) {
      ProfScope ps(ctx, "k_encode");
      hipError_t e = launch_encode(x, ch->view, pexp, ctx->stream, ctx->shape(Op::Encode, ch->view));
      if (e != hipSuccess) {
        free_partial(ctx, ch);
        return fail(BUFF_E_DEVICE, "k_encode launch: %s", hipGetErrorString(e));
      }
    }
    *out = ch;
    return BUFF_OK;
  }
  for (uint32_t j = 0; j < S; ++j) {
    ch->sp[j].major = majors[j];
    ch->sp[j].shift = F - 8 * (j + 1);
  }
  rc = sparse_alloc(ctx, ch);
  MajorEnc me;
  major_enc_geometry(ch->view.ntiles, &me);
  const EncScratch lay = enc_layout(S, ch->view.ntiles, me.nwg);
  if (rc == BUFF_OK) rc = ctx->scratch(ctx->s_major, lay.bytes);
  if (rc != BUFF_OK) {
    free_partial(ctx, ch);
    return rc;
  }
  uint8_t* scr = (uint8_t*)ctx->s_major.ptr;
  me = enc_of(ch, scr);
  hipError_t e;
  {
    ProfScope ps(ctx, "k_major_encode");
    e = launch_major_encode(x, n, pexp, dlen, mn, ch->view, me, ctx->stream);
    if (e == hipSuccess) e = launch_major_scan(me, ch->view.ntiles, (unsigned long long*)(scr + lay.totals), ctx->stream);
  }
  uint64_t t[2 * kMaxSparse];
  if (e == hipSuccess) rc = d2h_small(ctx, t, scr + lay.totals, 16ull * S);
  if (e != hipSuccess || rc != BUFF_OK) {
    free_partial(ctx, ch);
    return e != hipSuccess ? fail(BUFF_E_DEVICE, "k_major_encode: %s", hipGetErrorString(e)) : rc;
  }
  for (uint32_t j = 0; j < S; ++j) {
    ch->sp[j].n_out = t[2 * j];
    ch->sp[j].last = t[2 * j + 1];
  }
  rc = outlier_alloc(ctx, ch);
  if (rc != BUFF_OK) {
    free_partial(ctx, ch);
    return rc;
  }
  me = enc_of(ch, scr);
  {
    ProfScope ps(ctx, "k_major_place");
    e = launch_major_place(me, ch->view.ntiles, ctx->stream);
    if (e == hipSuccess) e = build_positions(ctx, ch);
  }
  if (e != hipSuccess) {
    free_partial(ctx, ch);
    return fail(BUFF_E_DEVICE, "k_major_place: %s", hipGetErrorString(e));
  }
  *out = ch;
  return BUFF_OK;
}

// ---------------------------------------------------------------------------
// Stream form (buff_simd.rs:451-555 writer, 633-700 reader):
//   header | u8 S | S x {u8 major, u32 bm_len, snappy(u64 n | u32 words), u32 n_out, n_out bytes}
//   | u8 0 (only when S > 0) | dense planes | LSB-first tail.
// With S == 0 and F < 8 the tail starts at byte 20 (write_bits ORs into the count byte).

namespace {

uint64_t dense_body_bytes(const Chunk* ch) {
  return (uint64_t)ch->view.P * ch->n + (ch->n * ch->view.r + 7) / 8;
}

int bitmap_words(Ctx* ctx, const Chunk* ch, uint32_t j, std::vector<uint8_t>* binary) {
  const uint64_t nw = (ch->n + 31) / 32;
  binary->assign(8 + 4 * nw, 0);
  const uint64_t n = ch->n;
  memcpy(binary->data(), &n, 8);
  if (nw) {
    HIP_TRY(hipMemcpyAsync(binary->data() + 8, ch->sp[j].mask16, 4 * nw, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  return BUFF_OK;
}

}  // namespace

int major_serialize(Ctx* ctx, const Chunk* ch, uint8_t* dst, uint64_t* len) {
  if (!ctx || !ch || !len) return fail(BUFF_E_INVALID, "null argument");
  const uint64_t need = *len;
  const uint64_t body = dense_body_bytes(ch);
  std::vector<std::vector<uint8_t>> bms(ch->S);
  uint64_t total = kStreamHeader;
  if (ch->S == 0) {
    total += ch->F >= 8 ? 1 + body : std::max<uint64_t>(body, 1);
  } else {
    total += 1;
    for (uint32_t j = 0; j < ch->S; ++j) {
      std::vector<uint8_t> binary;
      BUFF_TRY(bitmap_words(ctx, ch, j, &binary));
      bms[j] = snappy::compress(binary.data(), binary.size());
      total += 1 + 4 + bms[j].size() + 4 + ch->sp[j].n_out;
    }
    total += 1 + body;
  }
  *len = total;
  if (!dst) return BUFF_OK;
  if (need < total) return fail(BUFF_E_SMALL_BUFFER, "stream needs %llu bytes", (unsigned long long)total);
  const uint64_t ubase = (uint64_t)ch->view.base;
  const uint32_t hdr[5] = {(uint32_t)ubase, (uint32_t)(ubase >> 32), (uint32_t)ch->n, ch->ilen, ch->dlen};
  memcpy(dst, hdr, 20);
  uint64_t pos = kStreamHeader;
  if (ch->S == 0 && ch->F < 8) {
    dst[pos] = 0;   // the tail bits are ORed into the count byte below
  } else {
    dst[pos++] = (uint8_t)ch->S;
  }
  for (uint32_t j = 0; j < ch->S; ++j) {
    const SparseCol& c = ch->sp[j];
    dst[pos++] = (uint8_t)c.major;
    const uint32_t bl = (uint32_t)bms[j].size();
    memcpy(dst + pos, &bl, 4);
    pos += 4;
    memcpy(dst + pos, bms[j].data(), bl);
    pos += bl;
    const uint32_t no = (uint32_t)c.n_out;
    memcpy(dst + pos, &no, 4);
    pos += 4;
    if (c.n_out) HIP_TRY(hipMemcpyAsync(dst + pos, c.outl, c.n_out, hipMemcpyDeviceToHost, ctx->stream));
    pos += c.n_out;
  }
  if (ch->S) dst[pos++] = 0;
  if (body) {
    BUFF_TRY(ctx->scratch(ctx->s_stream, body));
    {
      ProfScope ps(ctx, "k_serialize");
      HIP_TRY(launch_serialize(ch->view, (ch->flags & BUFF_FLIP) ? 0x80 : 0, (uint8_t*)ctx->s_stream.ptr,
                               (ch->n * ch->view.r + 7) / 8, ctx->stream));
    }
    if (ch->S == 0 && ch->F < 8) {
      std::vector<uint8_t> tmp(body);
      HIP_TRY(hipMemcpyAsync(tmp.data(), ctx->s_stream.ptr, body, hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      for (uint64_t i = 0; i < body; ++i) dst[pos + i] |= tmp[i];
      pos += body;
    } else {
      HIP_TRY(hipMemcpyAsync(dst + pos, ctx->s_stream.ptr, body, hipMemcpyDeviceToHost, ctx->stream));
      pos += body;
    }
  } else if (ch->S == 0 && ch->F < 8) {
    pos += 1;
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (pos != total) return fail(BUFF_E_DEVICE, "internal: wrote %llu of %llu stream bytes", (unsigned long long)pos,
                                (unsigned long long)total);
  return BUFF_OK;
}

int major_load(Ctx* ctx, const uint8_t* s, uint64_t len, uint64_t scale, uint32_t flags, Chunk** out) {
  if (!ctx || !s || !out) return fail(BUFF_E_INVALID, "null argument");
  if (len < kStreamHeader + 1)
    return fail(BUFF_E_INVALID, "buff-major stream of %llu bytes is shorter than its header", (unsigned long long)len);
  uint32_t hdr[5];
  memcpy(hdr, s, 20);
  const int64_t base = (int64_t)((uint64_t)hdr[0] | ((uint64_t)hdr[1] << 32));
  const uint64_t n = hdr[2];
  const uint32_t ilen = hdr[3], dlen = hdr[4];
  const uint32_t F = dlen + ilen;
  if (F > 63) return fail(BUFF_E_INVALID, "stored width dlen+ilen = %u is not a BUFF stream width", F);
  if (F < 8)
    return fail(BUFF_E_REF_PANIC, "buff-major stream with F = %u < 8: its tail overlaps the column count byte and "
                "the reference's readers run past the end", F);
  const uint32_t S = s[kStreamHeader];
  if (S > F / 8) return fail(BUFF_E_INVALID, "%u sparse columns in a %u-bit stream", S, F);
  const uint64_t ntiles = (n + kTileRows - 1) / kTileRows;
  const uint64_t nw = (n + 31) / 32;
  struct Parsed {
    uint8_t major;
    std::vector<uint8_t> binary;
    const uint8_t* outl;
    uint64_t n_out;
  };
  std::vector<Parsed> cols(S);
  uint64_t pos = kStreamHeader + 1;
  for (uint32_t j = 0; j < S; ++j) {
    if (pos + 5 > len) return fail(BUFF_E_INVALID, "truncated sparse column %u", j);
    cols[j].major = s[pos];
    uint32_t bl;
    memcpy(&bl, s + pos + 1, 4);
    pos += 5;
    if (bl > len - pos) return fail(BUFF_E_INVALID, "sparse column %u bitmap runs past the stream", j);
    std::string why;
    if (!snappy::decompress(s + pos, bl, &cols[j].binary, &why))
      return fail(BUFF_E_REF_PANIC, "sparse column %u bitmap: snappy %s (the reference unwraps)", j, why.c_str());
    pos += bl;
    if (cols[j].binary.size() != 8 + 4 * nw)
      return fail(BUFF_E_INVALID, "sparse column %u bitmap has %llu bytes, %llu rows need %llu", j,
                  (unsigned long long)cols[j].binary.size(), (unsigned long long)n, (unsigned long long)(8 + 4 * nw));
    uint64_t nbits;
    memcpy(&nbits, cols[j].binary.data(), 8);
    if (nbits != n) return fail(BUFF_E_INVALID, "sparse column %u bitmap is %llu bits long", j, (unsigned long long)nbits);
    if (pos + 4 > len) return fail(BUFF_E_INVALID, "truncated sparse column %u", j);
    uint32_t no;
    memcpy(&no, s + pos, 4);
    pos += 4;
    if (no > len - pos) return fail(BUFF_E_INVALID, "sparse column %u outliers run past the stream", j);
    uint64_t pop = 0;
    for (uint64_t w = 0; w < nw; ++w) {
      uint32_t v;
      memcpy(&v, cols[j].binary.data() + 8 + 4 * w, 4);
      if (w == nw - 1 && (n % 32)) v &= (1u << (n % 32)) - 1u;   // from_binary fixes the last block
      memcpy(cols[j].binary.data() + 8 + 4 * w, &v, 4);
      pop += (uint64_t)__builtin_popcount(v);
    }
    if (pop != no)
      return fail(BUFF_E_INVALID, "sparse column %u: %llu outlier bits, %u outlier bytes", j, (unsigned long long)pop, no);
    cols[j].outl = s + pos;
    cols[j].n_out = no;
    pos += no;
  }
  if (S) pos += 1;   // the zero byte left by finish_write_byte
  Chunk* ch = new_major_chunk(ctx, n, scale, flags, dlen, ilen, base, F, S);
  if (!ch) return fail(BUFF_E_OOM, "host allocation");
  const uint64_t body = dense_body_bytes(ch);
  if (pos > len || body > len - pos) {
    delete ch;
    return fail(BUFF_E_INVALID, "stream is %llu bytes, its layout needs %llu", (unsigned long long)len,
                (unsigned long long)(pos + body));
  }
  int rc = chunk_alloc(ctx, ch);
  for (uint32_t j = 0; j < S; ++j) {
    ch->sp[j].major = cols[j].major;
    ch->sp[j].shift = F - 8 * (j + 1);
    ch->sp[j].n_out = cols[j].n_out;
  }
  if (rc == BUFF_OK && S) rc = sparse_alloc(ctx, ch);
  if (rc == BUFF_OK && S) rc = outlier_alloc(ctx, ch);
  if (rc != BUFF_OK) {
    free_partial(ctx, ch);
    return rc;
  }
  // tile prefixes and last outlier rows from the bitmap words
  std::vector<std::vector<uint32_t>> pre(S);
  for (uint32_t j = 0; j < S; ++j) {
    pre[j].assign(ntiles + 1, 0);
    const uint8_t* wb = cols[j].binary.data() + 8;
    uint64_t run = 0, last = ~0ull;
    for (uint64_t c = 0; c < ntiles; ++c) {
      pre[j][c] = (uint32_t)run;
      for (uint64_t w = 32 * c; w < std::min<uint64_t>(32 * c + 32, nw); ++w) {
        uint32_t v;
        memcpy(&v, wb + 4 * w, 4);
        run += (uint64_t)__builtin_popcount(v);
        if (v) last = 32 * w + 31 - (uint64_t)__builtin_clz(v);
      }
    }
    pre[j][ntiles] = (uint32_t)run;
    ch->sp[j].last = last;
  }
  hipError_t e = hipSuccess;
  for (uint32_t j = 0; j < S && e == hipSuccess; ++j) {
    e = hipMemsetAsync(ch->sp[j].mask16, 0, mask_bytes(ntiles), ctx->stream);
    if (e == hipSuccess && nw)
      e = hipMemcpyAsync(ch->sp[j].mask16, cols[j].binary.data() + 8, 4 * nw, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(ch->sp[j].prefix, pre[j].data(), 4 * (ntiles + 1), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && cols[j].n_out)
      e = hipMemcpyAsync(ch->sp[j].outl, cols[j].outl, cols[j].n_out, hipMemcpyHostToDevice, ctx->stream);
  }
  if (e == hipSuccess) e = build_positions(ctx, ch);
  if (e == hipSuccess && body) {
    rc = ctx->scratch(ctx->s_stream, body);
    if (rc == BUFF_OK) {
      e = hipMemcpyAsync(ctx->s_stream.ptr, s + pos, body, hipMemcpyHostToDevice, ctx->stream);
      if (e == hipSuccess) {
        ProfScope ps(ctx, "k_load");
        e = launch_load(ch->view, (ch->flags & BUFF_FLIP) ? 0x80 : 0, (const uint8_t*)ctx->s_stream.ptr,
                        (n * ch->view.r + 7) / 8, ctx->stream);
      }
    }
  }
  if (e == hipSuccess && rc == BUFF_OK) e = hipStreamSynchronize(ctx->stream);   // host buffers may go away
  if (e != hipSuccess) rc = fail(BUFF_E_DEVICE, "buff-major upload: %s", hipGetErrorString(e));
  if (rc != BUFF_OK) {
    (void)hipStreamSynchronize(ctx->stream);   // copies queued before the failure still read the host vectors
    free_partial(ctx, ch);
    return rc;
  }
  *out = ch;
  return BUFF_OK;
}

// ---------------------------------------------------------------------------
// queries (S > 0; a buff-major chunk without sparse columns uses the plain path)

int major_filter_plan(const Chunk* ch, int op, double pred, uint32_t flags, MajorTarget* tg) {
  const int prec = prec_of_scale(ch->scale);
  const int32_t pexp = precision_exp_of_prec(prec);
  uint64_t pu;
  memcpy(&pu, &pred, 8);
  const int64_t fixed = fetch_fixed_aligned_bits(pu, pexp, ch->dlen);
  const bool semantic = (flags & BUFF_SEMANTIC) != 0;
  if (fixed < ch->view.base) return op == BUFF_OP_GT ? 1 : 0;   // buff_simd.rs:886-889 / 1939-1942
  const uint64_t t = (uint64_t)fixed - (uint64_t)ch->view.base;
  if (semantic && ch->F < 64 && t >= (1ull << ch->F)) return 0;
  // equality answers come from the tail stage only (buff_simd.rs:2110-2150)
  if (op == BUFF_OP_EQ && !semantic && ch->view.r == 0) return 0;
  memset(tg, 0, sizeof(*tg));
  for (uint32_t k = 0; k < ch->F / 8; ++k) tg->digit[k] = (uint32_t)((t >> (ch->F - 8 * (k + 1))) & 0xFF);
  tg->tail = ch->view.r ? (t & ((1ull << ch->view.r) - 1)) : 0;
  return -1;
}

int major_filter_device(Ctx* ctx, const Chunk* ch, int op, double pred, uint32_t flags, uint64_t* words) {
  const uint64_t nwords = (ch->n + 63) / 64;
  if (nwords == 0) return BUFF_OK;
  const bool semantic = (flags & BUFF_SEMANTIC) != 0;
  MajorTarget tg;
  const int fill = major_filter_plan(ch, op, pred, flags, &tg);
  if (fill >= 0) {
    ProfScope ps(ctx, "k_fill_bitmap");
    HIP_TRY(launch_fill_bitmap(words, nwords, ch->n, fill, ctx->stream));
    return BUFF_OK;
  }
  const MajorView mv = major_view(ch);
  ProfScope ps(ctx, op == BUFF_OP_GT ? "k_major_filter_gt" : "k_major_filter_eq");
  HIP_TRY(launch_major_filter(mv, op == BUFF_OP_GT ? 0 : 1, semantic, tg, words, nwords, ctx->stream,
                              ctx->shape(op == BUFF_OP_GT ? Op::FilterGt : Op::FilterEq, ch->view)));
  return BUFF_OK;
}

// buff_sum_majority's fold (buff_simd.rs:2311-2439) of the outlier byte sums of
// the sparse columns (outsum[j]) and the dense planes' and tail's sums
// (dense_ps[p], tail at index P): each sparse column sums to its outliers plus
// major * (n - outliers).
void major_sum_fold(const Chunk* ch, const uint64_t* outsum, const uint64_t* dense_ps, double* ref_sum,
                    double* exact_sum, uint64_t* col_sums) {
  const ChunkView& v = ch->view;
  const uint32_t S = ch->S;
  // byte columns in stream order, then the tail
  uint64_t cs[kMaxPlanes + 2] = {0};
  uint32_t k = 0;
  for (uint32_t j = 0; j < S; ++j) cs[k++] = outsum[j] + (uint64_t)ch->sp[j].major * (ch->n - ch->sp[j].n_out);
  for (uint32_t p = 0; p < v.P; ++p) cs[k++] = dense_ps[p];
  const uint64_t tail_sum = dense_ps[v.P];
  if (col_sums) {
    memset(col_sums, 0, 9 * 8);
    for (uint32_t i = 0; i < k && i < 9; ++i) col_sums[i] = cs[i];
    if (v.r && k < 9) col_sums[k] = tail_sum;
  }
  if (ref_sum) {   // buff_simd.rs:2311-2439
    const uint32_t ilen = ch->ilen;
    int32_t processed = 0;
    double sum_base = (double)(uint32_t)ch->n * (double)v.base;
    sum_base = sum_base / powi2((int32_t)ch->dlen);
    double sum = 0.0;
    bool early = false;
    for (uint32_t i = 0; i < k; ++i) {
      processed += 8;
      sum = sum + (double)cs[i] / powi2((int32_t)((uint32_t)processed - ilen));
      if (i >= S && i == k - 1 && v.r == 0) {   // a dense plane ends the stream: no base term
        early = true;
        break;
      }
    }
    if (!early) {
      if (v.r) {
        processed += (int32_t)v.r;
        sum = sum + (double)tail_sum / powi2((int32_t)((uint32_t)processed - ilen));
      }
      sum += sum_base;
    }
    *ref_sum = sum;
  }
  if (exact_sum) {
    __int128 tot = (__int128)ch->n * (__int128)v.base;
    for (uint32_t i = 0; i < k; ++i) tot += (__int128)cs[i] << (ch->F - 8 * (i + 1));
    tot += (__int128)tail_sum;
    *exact_sum = (double)tot / powi2((int32_t)ch->dlen);
  }
}

int major_sum(Ctx* ctx, const Chunk* ch, uint32_t flags, double* ref_sum, double* exact_sum, uint64_t* col_sums) {
  (void)flags;
  const ChunkView& v = ch->view;
  const uint32_t S = ch->S;
  uint64_t dense_ps[kMaxPlanes + 1] = {0};
  uint64_t outsum[kMaxSparse] = {0};
  const LaunchShape ls = ctx->shape(Op::Sum, v);
  // [outlier sums | pad | per-block outlier rows (4 KiB) | dense sums + k_sum rows]
  const uint64_t rows_off = 256, dense_off = rows_off + 4096;
  BUFF_TRY(ctx->scratch(ctx->s_sums, dense_off + sizeof(dense_ps) + sum_scratch_bytes(ls)));
  uint8_t* sb = (uint8_t*)ctx->s_sums.ptr;
  const MajorView mv = major_view(ch);
  uint64_t n_out[kMaxSparse];
  for (uint32_t j = 0; j < S; ++j) n_out[j] = ch->sp[j].n_out;
  const bool dense = ch->n > 0 && (v.P || v.r);
  {
    ProfScope sc(ctx, "k_major_sum");
    HIP_TRY(launch_major_outlier_sums(mv, n_out, (unsigned long long*)sb, (unsigned long long*)(sb + rows_off),
                                      ctx->stream));
    if (dense) HIP_TRY(launch_sum(v, (unsigned long long*)(sb + dense_off), ctx->stream, ls));
  }
  BUFF_TRY(d2h_small(ctx, outsum, sb, 8ull * S));
  if (dense) BUFF_TRY(d2h_small(ctx, dense_ps, sb + dense_off, sizeof(dense_ps)));
  major_sum_fold(ch, outsum, dense_ps, ref_sum, exact_sum, col_sums);
  return BUFF_OK;
}

int major_max(Ctx* ctx, const Chunk* ch, uint32_t flags, double* value, uint64_t* max_off, uint64_t* argmax) {
  const bool semantic = (flags & BUFF_SEMANTIC) != 0;
  const ChunkView& v = ch->view;
  BUFF_TRY(ctx->scratch(ctx->s_best,
                        256 + std::max<uint64_t>(major_max_scratch_bytes(v.ntiles), kMaxListScratch + v.ntiles)));
  uint8_t* sb = (uint8_t*)ctx->s_best.ptr;
  MajorMaxState* st = (MajorMaxState*)sb;
  const MajorView mv = major_view(ch);
  MajorMaxState hs;
  hs.fallback = 1;
  if (ch->S >= 1 && ch->n > 0) {   // the rows holding column 0's chosen byte, when they are few
    {
      ProfScope sc(ctx, "k_major_max");
      HIP_TRY(launch_major_max_list(mv, ch->sp[0].n_out, st, sb + 256, ctx->stream, semantic));
    }
    BUFF_TRY(d2h_small(ctx, &hs, st, sizeof(hs)));
  }
  if (hs.fallback) {
    {
      ProfScope sc(ctx, "k_major_max");
      HIP_TRY(launch_major_max(mv, st, sb + 256, ctx->stream, ctx->shape(Op::Max, v), semantic));
    }
    BUFF_TRY(d2h_small(ctx, &hs, st, sizeof(hs)));
  }
  if (max_off) *max_off = hs.acc;
  if (argmax) *argmax = hs.argmax;
  // buff_simd.rs:2855: (max as i64 + base_int) as f64 / 2^dlen
  if (value) *value = (double)(int64_t)((uint64_t)hs.acc + (uint64_t)v.base) / powi2((int32_t)ch->dlen);
  return BUFF_OK;
}

uint64_t major_decoded_rows(const Chunk* ch, uint32_t flags) {
  if (ch->S == 0 || (flags & BUFF_SEMANTIC)) return ch->n;
  return ch->F == 8 * ch->S ? 0 : ch->n;   // buff_simd.rs:788-845 push values only below the sparse columns
}

// The bits a BUFF_SEMANTIC reduced-precision read drops from a buff-major
// chunk's offsets: F - R, with R from the read rule of split_double.rs:1642-1664
// (every bit when more decimals are asked than stored).
int major_precision_drop(const Chunk* ch, int32_t precision, uint32_t* drop) {
  *drop = 0;
  if (precision < 0) return BUFF_OK;
  uint32_t R = 0;
  const int rc = precision_bits(ch, precision, &R);
  if (rc == BUFF_E_REF_PANIC) R = ch->F;
  else if (rc != BUFF_OK) return rc;
  *drop = ch->F - std::min(R, ch->F);
  return BUFF_OK;
}

int major_materialize_device(Ctx* ctx, const Chunk* ch, int32_t precision, uint32_t flags, double* out,
                             uint64_t* rows) {
  const bool semantic = (flags & BUFF_SEMANTIC) != 0;
  if (precision >= 0 && !semantic)
    return fail(BUFF_E_UNSUPPORTED, "the reference has no reduced-precision decode of a buff-major stream");
  uint32_t drop = 0;
  BUFF_TRY(major_precision_drop(ch, precision, &drop));
  if (!semantic) {
    for (uint32_t j = 0; j < ch->S; ++j)
      if (ch->sp[j].n_out == 0)
        return fail(BUFF_E_REF_PANIC, "sparse column %u has no outliers: outlier_iter.next().unwrap() panics "
                    "(buff_simd.rs:683)", j);
  }
  const uint64_t r = major_decoded_rows(ch, flags);
  if (rows) *rows = r;
  if (r == 0) return BUFF_OK;
  if (!out) return fail(BUFF_E_INVALID, "null output");
  ProfScope sc(ctx, "k_major_decode");
  HIP_TRY(launch_major_decode(major_view(ch), semantic ? 1 : 0, drop, std::ldexp(1.0, -(int)ch->dlen), out, ctx->stream,
                              ctx->shape(Op::Materialize, ch->view)));
  return BUFF_OK;
}

// ---------------------------------------------------------------------------
// RAPG residue format (split_double.rs:624-954 byte_residue_encode, 1353-1893
// byte_residue_decode[_with_precision], 2210-2629 byte_residue_sum
// [_with_precision]) on the same resident form, every majority byte 0.
//
// Stream: header; with ilen > 15 (header ilen, u32) a count byte S, per
// outlier column {u32 size, portable bitmap, u32 count, bytes}, and when S > 0
// one zero byte; then the remaining planes and the LSB-first tail, as
// byte_fixed_encode writes them (ilen <= 15: the byte_fixed stream itself).

namespace {

bool residue_section(const Chunk* ch) { return ch->ilen > 15; }

}  // namespace

int residue_serialize(Ctx* ctx, const Chunk* ch, uint8_t* dst, uint64_t* len) {
  if (!ctx || !ch || !len) return fail(BUFF_E_INVALID, "null argument");
  const uint64_t need = *len;
  const uint64_t body = dense_body_bytes(ch);
  std::vector<std::vector<uint8_t>> bms(ch->S);
  uint64_t total = kStreamHeader + body;
  if (residue_section(ch)) total += 1 + (ch->S ? 1 : 0);
  for (uint32_t j = 0; j < ch->S; ++j) {
    std::vector<uint8_t> binary;
    BUFF_TRY(bitmap_words(ctx, ch, j, &binary));
    roaring::serialize(reinterpret_cast<const uint32_t*>(binary.data() + 8), ch->n, &bms[j]);
    total += 4 + bms[j].size() + 4 + ch->sp[j].n_out;
  }
  *len = total;
  if (!dst) return BUFF_OK;
  if (need < total) return fail(BUFF_E_SMALL_BUFFER, "stream needs %llu bytes", (unsigned long long)total);
  const uint64_t ubase = (uint64_t)ch->view.base;
  const uint32_t hdr[5] = {(uint32_t)ubase, (uint32_t)(ubase >> 32), (uint32_t)ch->n, ch->ilen, ch->dlen};
  memcpy(dst, hdr, 20);
  uint64_t pos = kStreamHeader;
  if (residue_section(ch)) dst[pos++] = (uint8_t)ch->S;
  for (uint32_t j = 0; j < ch->S; ++j) {
    const uint32_t bl = (uint32_t)bms[j].size(), no = (uint32_t)ch->sp[j].n_out;
    memcpy(dst + pos, &bl, 4);
    memcpy(dst + pos + 4, bms[j].data(), bl);
    pos += 4 + bl;
    memcpy(dst + pos, &no, 4);
    pos += 4;
    if (no) HIP_TRY(hipMemcpyAsync(dst + pos, ch->sp[j].outl, no, hipMemcpyDeviceToHost, ctx->stream));
    pos += no;
  }
  if (ch->S) dst[pos++] = 0;
  if (body) {
    BUFF_TRY(ctx->scratch(ctx->s_stream, body));
    {
      ProfScope ps(ctx, "k_serialize");
      HIP_TRY(launch_serialize(ch->view, 0, (uint8_t*)ctx->s_stream.ptr, (ch->n * ch->view.r + 7) / 8, ctx->stream));
    }
    HIP_TRY(hipMemcpyAsync(dst + pos, ctx->s_stream.ptr, body, hipMemcpyDeviceToHost, ctx->stream));
    pos += body;
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (pos != total) return fail(BUFF_E_DEVICE, "internal: wrote %llu of %llu stream bytes", (unsigned long long)pos,
                                (unsigned long long)total);
  return BUFF_OK;
}

int residue_load(Ctx* ctx, const uint8_t* s, uint64_t len, uint64_t scale, uint32_t flags, Chunk** out) {
  if (!ctx || !s || !out) return fail(BUFF_E_INVALID, "null argument");
  if (len < kStreamHeader) return fail(BUFF_E_INVALID, "residue stream of %llu bytes", (unsigned long long)len);
  uint32_t hdr[5];
  memcpy(hdr, s, 20);
  const int64_t base = (int64_t)((uint64_t)hdr[0] | ((uint64_t)hdr[1] << 32));
  const uint64_t n = hdr[2];
  const uint32_t ilen = hdr[3], dlen = hdr[4];
  const uint32_t F = dlen + ilen;   // u32, as the readers add them
  if (F > 63) return fail(BUFF_E_INVALID, "stored width dlen+ilen = %u is not a BUFF stream width", F);
  struct Col {
    std::vector<uint32_t> words;
    const uint8_t* outl;
    uint64_t n_out;
  };
  std::vector<Col> cols;
  uint64_t pos = kStreamHeader;
  uint32_t S = 0;
  if (ilen > 15) {
    if (len < kStreamHeader + 1) return fail(BUFF_E_INVALID, "residue stream without its column count");
    S = s[pos++];
    if (S > F / 8 || S > kMaxSparse) return fail(BUFF_E_INVALID, "%u outlier columns in a %u-bit stream", S, F);
    cols.resize(S);
    for (uint32_t j = 0; j < S; ++j) {
      uint32_t bl, no;
      if (pos + 4 > len) return fail(BUFF_E_INVALID, "truncated outlier column %u", j);
      memcpy(&bl, s + pos, 4);
      pos += 4;
      if (bl > len - pos) return fail(BUFF_E_INVALID, "outlier column %u bitmap runs past the stream", j);
      uint64_t card = 0;
      const char* why = "";
      if (!roaring::deserialize(s + pos, bl, n, &cols[j].words, &card, &why))
        return fail(BUFF_E_INVALID, "outlier column %u bitmap: %s", j, why);
      pos += bl;
      if (pos + 4 > len) return fail(BUFF_E_INVALID, "truncated outlier column %u", j);
      memcpy(&no, s + pos, 4);
      pos += 4;
      if (no > len - pos) return fail(BUFF_E_INVALID, "outlier column %u bytes run past the stream", j);
      if (no != card)
        return fail(BUFF_E_UNSUPPORTED, "outlier column %u: %llu bitmap rows, %u bytes (the readers pair them in order)",
                    j, (unsigned long long)card, no);
      cols[j].outl = s + pos;
      cols[j].n_out = no;
      pos += no;
    }
    if (S) pos += 1;   // the zero byte finish_write_byte leaves
  }
  Chunk* ch = new_major_chunk(ctx, n, scale, flags | BUFF_RESIDUE, dlen, ilen, base, F, S);
  if (!ch) return fail(BUFF_E_OOM, "host allocation");
  const uint64_t body = dense_body_bytes(ch);
  if (pos > len || body > len - pos) {
    delete ch;
    return fail(BUFF_E_INVALID, "stream is %llu bytes, its layout needs %llu", (unsigned long long)len,
                (unsigned long long)(pos + body));
  }
  int rc = chunk_alloc(ctx, ch);
  for (uint32_t j = 0; j < S; ++j) {
    ch->sp[j].major = 0;
    ch->sp[j].shift = F - 8 * (j + 1);
    ch->sp[j].n_out = cols[j].n_out;
  }
  if (rc == BUFF_OK && S) rc = sparse_alloc(ctx, ch);
  if (rc == BUFF_OK && S) rc = outlier_alloc(ctx, ch);
  if (rc != BUFF_OK) {
    free_partial(ctx, ch);
    return rc;
  }
  const uint64_t ntiles = ch->view.ntiles, nw = (n + 31) / 32;
  std::vector<std::vector<uint32_t>> pre(S);
  for (uint32_t j = 0; j < S; ++j) {   // tile prefixes and last outlier rows
    pre[j].assign(ntiles + 1, 0);
    uint64_t run = 0, last = ~0ull;
    for (uint64_t c = 0; c < ntiles; ++c) {
      pre[j][c] = (uint32_t)run;
      for (uint64_t w = 32 * c; w < std::min<uint64_t>(32 * c + 32, nw); ++w) {
        const uint32_t v = cols[j].words[w];
        run += (uint64_t)__builtin_popcount(v);
        if (v) last = 32 * w + 31 - (uint64_t)__builtin_clz(v);
      }
    }
    pre[j][ntiles] = (uint32_t)run;
    ch->sp[j].last = last;
  }
  hipError_t e = hipSuccess;
  for (uint32_t j = 0; j < S && e == hipSuccess; ++j) {
    e = hipMemsetAsync(ch->sp[j].mask16, 0, mask_bytes(ntiles), ctx->stream);
    if (e == hipSuccess && nw)
      e = hipMemcpyAsync(ch->sp[j].mask16, cols[j].words.data(), 4 * nw, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(ch->sp[j].prefix, pre[j].data(), 4 * (ntiles + 1), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && cols[j].n_out)
      e = hipMemcpyAsync(ch->sp[j].outl, cols[j].outl, cols[j].n_out, hipMemcpyHostToDevice, ctx->stream);
  }
  if (e == hipSuccess && S) e = build_positions(ctx, ch);
  if (e == hipSuccess && body) {
    rc = ctx->scratch(ctx->s_stream, body);
    if (rc == BUFF_OK) {
      e = hipMemcpyAsync(ctx->s_stream.ptr, s + pos, body, hipMemcpyHostToDevice, ctx->stream);
      if (e == hipSuccess) {
        ProfScope ps(ctx, "k_load");
        e = launch_load(ch->view, 0, (const uint8_t*)ctx->s_stream.ptr, (n * ch->view.r + 7) / 8, ctx->stream);
      }
    }
  }
  if (e == hipSuccess && rc == BUFF_OK) e = hipStreamSynchronize(ctx->stream);   // host buffers may go away
  if (e != hipSuccess) rc = fail(BUFF_E_DEVICE, "residue upload: %s", hipGetErrorString(e));
  if (rc != BUFF_OK) {
    (void)hipStreamSynchronize(ctx->stream);   // copies queued before the failure still read the host vectors
    free_partial(ctx, ch);
    return rc;
  }
  *out = ch;
  return BUFF_OK;
}

namespace {

// What a reference read of R bits takes from a residue chunk with outlier
// columns: the S columns, k whole planes and q-bit fields after them. Reads
// past the stored planes (R > F) read whatever bytes follow: a panic where they
// run past the stream, else not reproduced. q-bit fields that are not the
// rows' tail values (0 < q < r) are fields of the packed tail.
struct ResidueRead {
  uint32_t R, k, q;
  bool fields;   // 0 < q < r: the misaligned tail fields
};

int residue_read(const Chunk* ch, int32_t precision, ResidueRead* rr) {
  uint32_t R = ch->F;
  if (precision >= 0) BUFF_TRY(precision_bits(ch, precision, &R));
  const ChunkView& v = ch->view;
  const uint32_t rem = R - 8 * ch->S;   // R >= 8 S: R >= ilen >= 8 S
  rr->R = R;
  rr->k = rem / 8;
  rr->q = rem % 8;
  rr->fields = false;
  if (rr->k > v.P || (rr->k == v.P && rr->q > v.r)) {
    const uint64_t avail = dense_body_bytes(ch);
    const uint64_t reads = (uint64_t)rr->k * ch->n + (ch->n * rr->q + 7) / 8;
    if (reads > avail) return fail(BUFF_E_REF_PANIC, "the reference reads %u bits of a %u-bit residue stream", R, ch->F);
    return fail(BUFF_E_UNSUPPORTED, "the reference would read %u bits of a %u-bit residue stream", R, ch->F);
  }
  if (rr->q > 0 && rr->q != v.r) {
    if (rr->k != v.P) return fail(BUFF_E_UNSUPPORTED, "q-bit fields of plane bytes");
    rr->fields = true;
  }
  return BUFF_OK;
}

}  // namespace

int residue_sum_fold(const Chunk* ch, const ResidueRead& rr, bool semantic, const uint64_t* osum, const uint64_t* ps,
                     const uint64_t* fields, double* ref_sum, double* exact_sum, uint64_t* sums_out);

uint64_t residue_decoded_rows(const Chunk* ch, int32_t precision) {
  if (ch->S == 0) return ch->n;
  ResidueRead rr;
  if (residue_read(ch, precision, &rr) != BUFF_OK) return 0;
  return (rr.k || rr.q) ? ch->n : 0;   // nothing is pushed unless a plane or the tail is read
}

// byte_residue_decode (precision < 0) / byte_residue_decode_with_precision on a
// residue chunk with outlier columns: (base + (off >> (F - R))) / 2^dlen for
// every row (mis-scaled below F bits, as the reference), nothing when the
// columns take every bit read; a column without outliers panics
// (outlier_iter.next().unwrap(), split_double.rs:1405).
int residue_materialize_device(Ctx* ctx, const Chunk* ch, int32_t precision, double* out, uint64_t* rows) {
  ResidueRead rr;
  BUFF_TRY(residue_read(ch, precision, &rr));
  for (uint32_t j = 0; j < ch->S; ++j)
    if (ch->sp[j].n_out == 0)
      return fail(BUFF_E_REF_PANIC, "outlier column %u has no outliers: outlier_iter.next().unwrap() panics "
                  "(split_double.rs:1405)", j);
  if (rr.fields)
    return fail(BUFF_E_UNSUPPORTED, "a %u-bit read ends in %u-bit fields of the %u-bit packed tail", rr.R, rr.q,
                ch->view.r);
  const uint64_t r = (rr.k || rr.q) ? ch->n : 0;
  if (rows) *rows = r;
  if (r == 0) return BUFF_OK;
  if (!out) return fail(BUFF_E_INVALID, "null output");
  ProfScope sc(ctx, "k_major_decode");
  HIP_TRY(launch_major_decode(major_view(ch), 2, ch->F - rr.R, std::ldexp(1.0, -(int)ch->dlen), out, ctx->stream,
                              ctx->shape(Op::Materialize, ch->view)));
  return BUFF_OK;
}

// byte_residue_sum (precision < 0) / byte_residue_sum_with_precision on a
// residue chunk: the outlier byte sums of the columns, the sums of the k planes
// read, the q-bit fields (the tail, or the packed tail's q-bit fields), folded in
// the reference's order (each term / 2^(bits read - ilen); no base term when the
// plane loop ends on a byte, split_double.rs:2271-2273). sums_out: the u64 sums in
// read order. exact_sum: the exact sum of the values truncated to their top R
// bits, base included, rounded once.
int residue_sum_op(Ctx* ctx, const Chunk* ch, int32_t precision, uint32_t flags, double* ref_sum, double* exact_sum,
                   uint64_t* sums_out) {
  const bool semantic = (flags & BUFF_SEMANTIC) != 0;
  ResidueRead rr;
  const int rc = residue_read(ch, precision, &rr);
  if (rc != BUFF_OK) {
    if (!semantic) return rc;
    rr = ResidueRead{ch->F, ch->view.P, ch->view.r, false};   // what cannot be read: every bit
  }
  const ChunkView& v = ch->view;
  uint64_t ps[kMaxPlanes + 1] = {0};
  uint64_t osum[kMaxSparse] = {0};
  uint64_t fields[2] = {0, 0};
  if (ch->n > 0) {
    const LaunchShape ls = ctx->shape(Op::Sum, v);
    const uint64_t rows_off = 256, dense_off = rows_off + 4096;
    BUFF_TRY(ctx->scratch(ctx->s_sums, dense_off + sizeof(ps) + sum_scratch_bytes(ls)));
    uint8_t* sb = (uint8_t*)ctx->s_sums.ptr;
    ChunkView pv = v;
    pv.P = rr.k;
    pv.r = (rr.q > 0 && !rr.fields) ? v.r : 0;
    {
      ProfScope sc(ctx, "k_residue_sum");
      if (ch->S) {
        uint64_t n_out[kMaxSparse];
        for (uint32_t j = 0; j < ch->S; ++j) n_out[j] = ch->sp[j].n_out;
        HIP_TRY(launch_major_outlier_sums(major_view(ch), n_out, (unsigned long long*)sb,
                                          (unsigned long long*)(sb + rows_off), ctx->stream));
      }
      if (pv.P || pv.r) HIP_TRY(launch_sum(pv, (unsigned long long*)(sb + dense_off), ctx->stream, ls));
    }
    if (rr.fields) {
      BUFF_TRY(ctx->scratch(ctx->s_best, 256));
      HIP_TRY(launch_fill_bytes(ctx->s_best.ptr, 16, 0, ctx->stream));
      ProfScope sc(ctx, "k_tail_fields");
      HIP_TRY(launch_tail_fields(v, rr.q, (unsigned long long*)ctx->s_best.ptr, ctx->stream));
    }
    if (ch->S) BUFF_TRY(d2h_small(ctx, osum, sb, 8ull * ch->S));
    if (pv.P || pv.r) BUFF_TRY(d2h_small(ctx, ps, sb + dense_off, sizeof(ps)));
    if (rr.fields) BUFF_TRY(d2h_small(ctx, fields, ctx->s_best.ptr, sizeof(fields)));
  }
  return residue_sum_fold(ch, rr, semantic, osum, ps, fields, ref_sum, exact_sum, sums_out);
}

// The fold of residue_sum_op from its u64 sums: osum, the columns' outlier byte
// sums; ps, the sums of the rr.k planes read and, at ps[rr.k], the tail's;
// fields, the q-bit fields' sums when rr.fields.
int residue_sum_fold(const Chunk* ch, const ResidueRead& rr, bool semantic, const uint64_t* osum, const uint64_t* ps,
                     const uint64_t* fields, double* ref_sum, double* exact_sum, uint64_t* sums_out) {
  const ChunkView& v = ch->view;
  // sums in read order: columns, whole planes, the q-bit part
  uint64_t sums[2 * kMaxPlanes + 1] = {0};
  uint32_t m = 0;
  for (uint32_t j = 0; j < ch->S; ++j) sums[m++] = osum[j];
  for (uint32_t k = 0; k < rr.k; ++k) sums[m++] = ps[k];
  uint64_t top_q = 0;
  if (rr.q > 0) {
    sums[m++] = rr.fields ? fields[0] : ps[rr.k];
    top_q = rr.fields ? fields[1] : ps[rr.k];
  }
  if (sums_out) memcpy(sums_out, sums, sizeof(uint64_t) * (kMaxPlanes + 1));
  // exact: the top R bits of every offset (columns and planes whole, q top tail bits)
  __int128 tot = (__int128)ch->n * (__int128)v.base;
  for (uint32_t j = 0; j < ch->S; ++j) tot += (__int128)osum[j] << ch->sp[j].shift;
  for (uint32_t k = 0; k < rr.k; ++k) tot += (__int128)ps[k] << (v.F - 8 * (k + 1));
  if (rr.q) tot += (__int128)top_q << (v.r - rr.q);
  const double exact = (double)tot / powi2((int32_t)ch->dlen);
  if (exact_sum) *exact_sum = exact;
  if (!ref_sum) return BUFF_OK;
  if (semantic) {
    *ref_sum = exact;
    return BUFF_OK;
  }
  // split_double.rs:2245-2304 / 2467-2533
  const uint32_t ilen = ch->ilen;
  uint32_t remain = rr.R;
  int32_t processed = 0;
  double sum = 0.0;
  const double sum_base = (double)(uint32_t)ch->n * (double)v.base / powi2((int32_t)ch->dlen);
  uint32_t i = 0;
  for (uint32_t j = 0; j < ch->S; ++j) {
    remain -= 8;
    processed += 8;
    sum = sum + (double)sums[i++] / powi2((int32_t)((uint32_t)processed - ilen));
  }
  while (remain >= 8) {
    remain -= 8;
    processed += 8;
    sum = sum + (double)sums[i++] / powi2((int32_t)((uint32_t)processed - ilen));
    if (remain == 0) {
      *ref_sum = sum;
      return BUFF_OK;
    }
  }
  if (remain > 0) {
    processed += (int32_t)remain;
    sum = sum + (double)sums[i] / powi2((int32_t)((uint32_t)processed - ilen));
  }
  *ref_sum = sum + sum_base;
  return BUFF_OK;
}

// byte_residue_sum (the full read) from the columns' outlier byte sums and the
// planes' and tail's sums (the batched sums of a residue chunk).
int residue_full_sum_fold(const Chunk* ch, const uint64_t* osum, const uint64_t* ps, double* ref_sum,
                          double* exact_sum) {
  ResidueRead rr;
  BUFF_TRY(residue_read(ch, -1, &rr));
  return residue_sum_fold(ch, rr, false, osum, ps, nullptr, ref_sum, exact_sum, nullptr);
}

}  // namespace buffgpu
