// Raw-format snappy for the buff-major outlier bitmaps (host code).
//
// The reference stores each sparse column's outlier bitmap as
// `parity_snappy::compress(BitVec::to_binary())` (database/src/query/bit_vec_iter.rs:61-71,
// database/src/compress/buff_simd.rs:481,515) and reads it back with
// `decompress(..).unwrap()` (buff_simd.rs:677,714,909,924). parity-snappy wraps
// Google's C++ snappy. This is Google snappy's published level-1 algorithm:
// 64 KiB fragments compressed independently, a 16-bit position table per fragment
// of the smallest power of two in [256, 16384] not below the fragment size,
// hashed by the top log2(size) bits of load32 * 0x1e35a7bd, a match search that
// widens its stride by one byte every 32 misses, and the literal / 1-byte-offset /
// 2-byte-offset copy tags of the format. Decompression is format-defined.
#pragma once

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace buffgpu {
namespace snappy {

class Writer {
 public:
  explicit Writer(std::vector<uint8_t>& out) : out_(out) {}

  void varint32(uint32_t v) {
    while (v >= 0x80) {
      out_.push_back((uint8_t)(v | 0x80));
      v >>= 7;
    }
    out_.push_back((uint8_t)v);
  }
  void literal(const uint8_t* p, size_t len) {
    const size_t n = len - 1;
    if (n < 60) {
      out_.push_back((uint8_t)(n << 2));
    } else {
      const int count = n < (1u << 8) ? 1 : n < (1u << 16) ? 2 : n < (1u << 24) ? 3 : 4;
      out_.push_back((uint8_t)((59 + count) << 2));
      for (int k = 0; k < count; ++k) out_.push_back((uint8_t)(n >> (8 * k)));
    }
    out_.insert(out_.end(), p, p + len);
  }
  void copy(size_t offset, size_t len) {
    if (len < 12) {
      copy64(offset, len, true);
      return;
    }
    while (len >= 68) {
      copy64(offset, 64, false);
      len -= 64;
    }
    if (len > 64) {
      copy64(offset, 60, false);
      len -= 60;
    }
    copy64(offset, len, len < 12);
  }

 private:
  void copy64(size_t offset, size_t len, bool short_len) {
    if (short_len && offset < 2048) {
      out_.push_back((uint8_t)(0x01 | ((len - 4) << 2) | ((offset >> 8) << 5)));
      out_.push_back((uint8_t)offset);
    } else {
      out_.push_back((uint8_t)(0x02 | ((len - 1) << 2)));
      out_.push_back((uint8_t)offset);
      out_.push_back((uint8_t)(offset >> 8));
    }
  }
  std::vector<uint8_t>& out_;
};

inline uint32_t read32(const uint8_t* p) {
  uint32_t v;
  std::memcpy(&v, p, 4);
  return v;
}

inline void compress_fragment(const uint8_t* in, size_t size, Writer& w, uint16_t* table, int bits) {
  constexpr size_t kMargin = 15;
  auto slot = [bits](uint32_t v) { return (v * 0x1e35a7bdu) >> (32 - bits); };
  const uint8_t* const end = in + size;
  const uint8_t* emit = in;
  if (size >= kMargin) {
    const uint8_t* const limit = end - kMargin;
    const uint8_t* ip = in + 1;
    uint32_t next_slot = slot(read32(ip));
    for (;;) {
      uint32_t skip = 32;
      const uint8_t* next = ip;
      const uint8_t* cand;
      do {
        ip = next;
        const uint32_t h = next_slot;
        const uint32_t stride = skip >> 5;
        skip += stride;
        next = ip + stride;
        if (next > limit) goto tail;
        next_slot = slot(read32(next));
        cand = in + table[h];
        table[h] = (uint16_t)(ip - in);
      } while (read32(ip) != read32(cand));
      w.literal(emit, (size_t)(ip - emit));
      do {
        const uint8_t* m = ip + 4;
        const uint8_t* c = cand + 4;
        while (m < end && *m == *c) ++m, ++c;
        w.copy((size_t)(ip - cand), (size_t)(m - ip));
        ip = m;
        emit = ip;
        if (ip >= limit) goto tail;
        table[slot(read32(ip - 1))] = (uint16_t)(ip - in - 1);
        const uint32_t h = slot(read32(ip));
        cand = in + table[h];
        table[h] = (uint16_t)(ip - in);
      } while (read32(ip) == read32(cand));
      ++ip;
      next_slot = slot(read32(ip));
    }
  }
tail:
  if (emit < end) w.literal(emit, (size_t)(end - emit));
}

inline std::vector<uint8_t> compress(const uint8_t* in, size_t n) {
  std::vector<uint8_t> out;
  out.reserve(n + n / 6 + 32);
  Writer w(out);
  w.varint32((uint32_t)n);
  std::vector<uint16_t> table(1u << 14);
  for (size_t pos = 0; pos < n; pos += 1u << 16) {
    const size_t frag = n - pos < (1u << 16) ? n - pos : (1u << 16);
    int bits = 8;
    while (bits < 14 && ((size_t)1 << bits) < frag) ++bits;
    std::memset(table.data(), 0, sizeof(uint16_t) << bits);
    compress_fragment(in + pos, frag, w, table.data(), bits);
  }
  return out;
}

// Returns false (with a reason) on malformed input.
inline bool decompress(const uint8_t* in, size_t n, std::vector<uint8_t>* out, std::string* why) {
  size_t pos = 0;
  uint64_t ulen = 0;
  for (int shift = 0;; shift += 7) {
    if (pos >= n || shift > 28) {
      *why = "bad length varint";
      return false;
    }
    const uint8_t b = in[pos++];
    ulen |= (uint64_t)(b & 0x7f) << shift;
    if (!(b & 0x80)) break;
  }
  if (ulen > 0xffffffffull) {
    *why = "length overflow";
    return false;
  }
  out->clear();
  out->reserve((size_t)ulen);
  while (pos < n) {
    const uint8_t tag = in[pos++];
    size_t len = 0, off = 0;
    const uint32_t kind = tag & 3;
    if (kind == 0) {
      len = (tag >> 2) + 1;
      if (len > 60) {
        const size_t nb = len - 60;
        if (pos + nb > n) {
          *why = "truncated literal length";
          return false;
        }
        len = 0;
        for (size_t k = 0; k < nb; ++k) len |= (size_t)in[pos + k] << (8 * k);
        len += 1;
        pos += nb;
      }
      if (len > n - pos || out->size() + len > ulen) {
        *why = "truncated or oversized literal";
        return false;
      }
      out->insert(out->end(), in + pos, in + pos + len);
      pos += len;
      continue;
    }
    const size_t extra = kind == 1 ? 1 : kind == 2 ? 2 : 4;
    if (pos + extra > n) {
      *why = "truncated copy";
      return false;
    }
    if (kind == 1) {
      len = ((tag >> 2) & 7) + 4;
      off = ((size_t)(tag >> 5) << 8) | in[pos];
    } else {
      len = (tag >> 2) + 1;
      off = kind == 2 ? ((size_t)in[pos] | ((size_t)in[pos + 1] << 8)) : (size_t)read32(in + pos);
    }
    pos += extra;
    if (off == 0 || off > out->size() || out->size() + len > ulen) {
      *why = "bad copy";
      return false;
    }
    const size_t from = out->size() - off;
    for (size_t k = 0; k < len; ++k) out->push_back((*out)[from + k]);
  }
  if (out->size() != ulen) {
    *why = "length mismatch";
    return false;
  }
  return true;
}

}  // namespace snappy
}  // namespace buffgpu
