// The Segment container (database/src/segment.rs:36-44) in its serialized
// form: bincode 1.x with the default options of bincode::serialize /
// bincode::deserialize (segment.rs:84-101; `bincode = "1.1.3"`,
// database/Cargo.toml:34): little-endian fixed-width integers, u64 sequence
// and string lengths, u32 enum variant indices, u8 Option tags, fields in
// declaration order, trailing bytes allowed. serde writes SystemTime as
// {secs_since_epoch u64, nanos_since_epoch u32} and Duration as {secs u64,
// nanos u32}, and rejects values whose seconds plus nanos / 10^9 overflow
// (u64 for Duration; the platform's i64 seconds for SystemTime on Linux).
//
// Host code only, plus buff_segment_encode: the reference's encoders take a
// &mut Segment<T> (byte_fixed_encode, split_double.rs:298-318, reads
// seg.get_data()); this entry point takes the serialized Segment, uploads its
// data as float64 and runs the resident encode.
#include <hip/hip_runtime.h>

#include <cstring>
#include <limits>
#include <vector>

#include "../../include/buffgpu.h"
#include "buff_common.h"
#include "buff_runtime.h"

namespace buffgpu {
namespace {

constexpr uint64_t kNanosPerSec = 1000000000ull;
constexpr uint64_t kRecord = 12;   // a Duration: u64 secs, u32 nanos

bool valid_utf8(const uint8_t* s, uint64_t n) {
  uint64_t i = 0;
  while (i < n) {
    const uint8_t c = s[i];
    uint32_t len, cp;
    if (c < 0x80) {
      ++i;
      continue;
    } else if ((c & 0xE0) == 0xC0) {
      len = 2;
      cp = c & 0x1F;
    } else if ((c & 0xF0) == 0xE0) {
      len = 3;
      cp = c & 0x0F;
    } else if ((c & 0xF8) == 0xF0) {
      len = 4;
      cp = c & 0x07;
    } else {
      return false;
    }
    if (n - i < len) return false;
    for (uint32_t k = 1; k < len; ++k) {
      if ((s[i + k] & 0xC0) != 0x80) return false;
      cp = (cp << 6) | (s[i + k] & 0x3F);
    }
    // shortest form, no surrogates, at most U+10FFFF
    if ((len == 2 && cp < 0x80) || (len == 3 && cp < 0x800) || (len == 4 && cp < 0x10000) || cp > 0x10FFFF ||
        (cp >= 0xD800 && cp <= 0xDFFF))
      return false;
    i += len;
  }
  return true;
}

struct Writer {
  uint8_t* dst;
  uint64_t cap, len = 0;
  void put(const void* p, uint64_t k) {
    if (dst && len + k <= cap) memcpy(dst + len, p, k);
    len += k;
  }
  void u8(uint8_t v) { put(&v, 1); }
  void u32(uint32_t v) { put(&v, 4); }   // little-endian hosts only (x86-64, the reference's target)
  void u64(uint64_t v) { put(&v, 8); }
};

int check_segment(const buff_segment* s) {
  if (!s) return fail(BUFF_E_INVALID, "null argument");
  if (s->method < -1 || s->method > 3) return fail(BUFF_E_INVALID, "no Methods variant %d", s->method);
  if (s->method == 3 && s->model_len && !s->model) return fail(BUFF_E_INVALID, "null model name");
  if (s->elem_bytes != 8 && s->elem_bytes != 4) return fail(BUFF_E_INVALID, "T is f64 (8) or f32 (4), not %u bytes",
                                                            s->elem_bytes);
  if (s->n && !s->data) return fail(BUFF_E_INVALID, "null data");
  if (s->has_time_lapse && s->n_lapse && !s->lapse) return fail(BUFF_E_INVALID, "null time_lapse");
  return BUFF_OK;
}

void write_segment(const buff_segment* s, Writer& w) {
  if (s->method < 0) {
    w.u8(0);
  } else {
    w.u8(1);
    w.u32((uint32_t)s->method);
    if (s->method == 1 || s->method == 2) w.u32(s->dictionary_id);
    if (s->method == 3) {
      w.u64(s->model_len);
      w.put(s->model, s->model_len);
    }
  }
  w.u64(s->timestamp_secs);
  w.u32(s->timestamp_nanos);
  w.u64(s->signal);
  w.u64(s->n);
  w.put(s->data, s->n * s->elem_bytes);
  if (!s->has_time_lapse) {
    w.u8(0);
  } else {
    w.u8(1);
    w.u64(s->n_lapse);
    w.put(s->lapse, s->n_lapse * kRecord);
  }
  if (!s->has_prev_offset) {
    w.u8(0);
  } else {
    w.u8(1);
    w.u64(s->prev_secs);
    w.u32(s->prev_nanos);
  }
}

struct Reader {
  const uint8_t* b;
  uint64_t len, i = 0;
  bool take(uint64_t k, const uint8_t** p) {
    if (k > len - i) return false;
    *p = b + i;
    i += k;
    return true;
  }
  bool u8(uint8_t* v) {
    const uint8_t* p;
    if (!take(1, &p)) return false;
    *v = *p;
    return true;
  }
  bool u32(uint32_t* v) {
    const uint8_t* p;
    if (!take(4, &p)) return false;
    memcpy(v, p, 4);
    return true;
  }
  bool u64(uint64_t* v) {
    const uint8_t* p;
    if (!take(8, &p)) return false;
    memcpy(v, p, 8);
    return true;
  }
};

int eof() { return fail(BUFF_E_INVALID, "Segment bytes end early (unexpected end of input)"); }

int read_option(Reader& r, int32_t* some) {
  uint8_t t;
  if (!r.u8(&t)) return eof();
  if (t > 1) return fail(BUFF_E_INVALID, "invalid Option tag %u", t);
  *some = t;
  return BUFF_OK;
}

// serde's overflow check: secs + nanos / 10^9 must stay within `limit`
int read_time(Reader& r, uint64_t limit, uint64_t* secs, uint32_t* nanos, const char* what) {
  if (!r.u64(secs) || !r.u32(nanos)) return eof();
  const uint64_t carry = *nanos / kNanosPerSec;
  if (*secs > limit - carry) return fail(BUFF_E_INVALID, "overflow deserializing %s", what);
  return BUFF_OK;
}

int parse_segment(const uint8_t* bytes, uint64_t len, uint32_t elem_bytes, buff_segment* s) {
  if (!s || (len && !bytes)) return fail(BUFF_E_INVALID, "null argument");
  if (elem_bytes != 8 && elem_bytes != 4) return fail(BUFF_E_INVALID, "T is f64 (8) or f32 (4)");
  buff_segment z;
  memset(&z, 0, sizeof(z));
  z.method = -1;
  z.elem_bytes = elem_bytes;
  Reader r{bytes, len};
  int32_t some = 0;
  BUFF_TRY(read_option(r, &some));
  if (some) {
    uint32_t v;
    if (!r.u32(&v)) return eof();
    if (v > 3) return fail(BUFF_E_INVALID, "invalid Methods variant %u", v);
    z.method = (int32_t)v;
    if (v == 1 || v == 2) {
      if (!r.u32(&z.dictionary_id)) return eof();
    } else if (v == 3) {
      const uint8_t* p;
      if (!r.u64(&z.model_len) || !r.take(z.model_len, &p)) return eof();
      if (!valid_utf8(p, z.model_len)) return fail(BUFF_E_INVALID, "model name is not UTF-8");
      z.model = (const char*)p;
    }
  }
  BUFF_TRY(read_time(r, (uint64_t)std::numeric_limits<int64_t>::max(), &z.timestamp_secs, &z.timestamp_nanos,
                     "SystemTime"));
  if (!r.u64(&z.signal)) return eof();
  const uint8_t* p;
  if (!r.u64(&z.n) || z.n > (len - r.i) / elem_bytes || !r.take(z.n * elem_bytes, &p)) return eof();
  z.data = p;
  BUFF_TRY(read_option(r, &z.has_time_lapse));
  if (z.has_time_lapse) {
    if (!r.u64(&z.n_lapse) || z.n_lapse > (len - r.i) / kRecord) return eof();
    z.lapse = bytes + r.i;
    for (uint64_t k = 0; k < z.n_lapse; ++k) {
      uint64_t secs;
      uint32_t nanos;
      BUFF_TRY(read_time(r, std::numeric_limits<uint64_t>::max(), &secs, &nanos, "Duration"));
    }
  }
  BUFF_TRY(read_option(r, &z.has_prev_offset));
  if (z.has_prev_offset)
    BUFF_TRY(read_time(r, std::numeric_limits<uint64_t>::max(), &z.prev_secs, &z.prev_nanos, "Duration"));
  z.consumed = r.i;
  *s = z;
  return BUFF_OK;
}

}  // namespace
}  // namespace buffgpu

extern "C" {

int buff_segment_byte_size(const buff_segment* seg, uint64_t* size) {
  BUFF_TRY(buffgpu::check_segment(seg));
  if (!size) return buffgpu::fail(BUFF_E_INVALID, "null argument");
  buffgpu::Writer w{nullptr, 0};
  buffgpu::write_segment(seg, w);
  *size = w.len;
  return BUFF_OK;
}

int buff_segment_serialize(const buff_segment* seg, uint8_t* dst, uint64_t cap, uint64_t* len) {
  BUFF_TRY(buffgpu::check_segment(seg));
  if (!len) return buffgpu::fail(BUFF_E_INVALID, "null argument");
  buffgpu::Writer sz{nullptr, 0};
  buffgpu::write_segment(seg, sz);
  *len = sz.len;
  if (!dst) return BUFF_OK;
  if (cap < sz.len)
    return buffgpu::fail(BUFF_E_SMALL_BUFFER, "Segment needs %llu bytes, the buffer holds %llu",
                         (unsigned long long)sz.len, (unsigned long long)cap);
  buffgpu::Writer w{dst, cap};
  buffgpu::write_segment(seg, w);
  return BUFF_OK;
}

int buff_segment_parse(const uint8_t* bytes, uint64_t len, uint32_t elem_bytes, buff_segment* seg) {
  return buffgpu::parse_segment(bytes, len, elem_bytes, seg);
}

int buff_segment_encode(buff_ctx* ctx, const uint8_t* bytes, uint64_t len, uint32_t elem_bytes, uint64_t scale,
                        uint32_t flags, buff_chunk** out) {
  if (!ctx || !out) return buffgpu::fail(BUFF_E_INVALID, "null argument");
  buff_segment s;
  BUFF_TRY(buffgpu::parse_segment(bytes, len, elem_bytes, &s));
  if (elem_bytes == 8) {
    std::vector<double> x(s.n);   // the data section is not 8-byte aligned in general
    if (s.n) memcpy(x.data(), s.data, s.n * 8);
    return buff_encode(ctx, x.data(), s.n, scale, flags, out);
  }
  std::vector<double> x(s.n);   // T = f32: Into<f64> is exact
  for (uint64_t i = 0; i < s.n; ++i) {
    float f;
    memcpy(&f, (const uint8_t*)s.data + 4 * i, 4);
    x[i] = (double)f;
  }
  return buff_encode(ctx, x.data(), s.n, scale, flags, out);
}

}  // extern "C"
