// Raw-format snappy, restated for the oracle (test infrastructure only).
//
// The reference compresses the buff-major outlier bitmaps with
// `parity_snappy::compress` / `decompress` (database/src/query/bit_vec_iter.rs:61-71,
// called at database/src/compress/buff_simd.rs:481,515 and :677,714,909,924).
// parity-snappy 0.1 wraps Google's C++ snappy library; the reference pins no
// version of it (no Cargo.lock). This file restates Google snappy's published
// level-1 compressor (snappy.cc: Compress -> CompressFragment over 64 KiB
// fragments, a fresh 16-bit hash table per fragment sized to the next power of two
// in [256, 16384], multiplicative hash 0x1e35a7bd keeping the top log2(size) bits
// as in the 1.1.x releases of 2017-2018, the skip-accelerating match search,
// EmitLiteral / EmitCopy tag rules) and the format-defined decompressor.
// The decompressed content is fixed by the format. The compressed bytes of
// compressible input depend on the library release (later releases changed the
// match search), so they are pinned only where every release agrees: literal-only
// and incompressible input, and runs of one byte.
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace snappy_ref {

constexpr size_t kBlockSize = 1u << 16;
constexpr int kMaxHashTableBits = 14;
constexpr size_t kMaxHashTableSize = 1u << kMaxHashTableBits;
constexpr size_t kMinHashTableSize = 1u << 8;
constexpr size_t kInputMarginBytes = 15;

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline uint32_t load32(const uint8_t* p) {
  uint32_t v;
  std::memcpy(&v, p, 4);
  return v;
}

// Index into a table of 2^bits entries: the top `bits` bits of the product.
inline uint32_t hash_index(uint32_t bytes, int bits) { return (bytes * 0x1e35a7bdu) >> (32 - bits); }

inline void emit_literal(std::vector<uint8_t>& out, const uint8_t* lit, size_t len) {
  size_t n = len - 1;
  if (n < 60) {
    out.push_back((uint8_t)(n << 2));
  } else {
    uint8_t extra[4];
    int count = 0;
    while (n > 0) {
      extra[count++] = (uint8_t)(n & 0xff);
      n >>= 8;
    }
    out.push_back((uint8_t)((59 + count) << 2));
    out.insert(out.end(), extra, extra + count);
  }
  out.insert(out.end(), lit, lit + len);
}

inline void emit_copy_at_most_64(std::vector<uint8_t>& out, size_t offset, size_t len, bool len_less_than_12) {
  if (len_less_than_12 && offset < 2048) {
    out.push_back((uint8_t)(1 + ((len - 4) << 2) + ((offset >> 8) << 5)));
    out.push_back((uint8_t)(offset & 0xff));
  } else {
    out.push_back((uint8_t)(2 + ((len - 1) << 2)));
    out.push_back((uint8_t)(offset & 0xff));
    out.push_back((uint8_t)((offset >> 8) & 0xff));
  }
}

inline void emit_copy(std::vector<uint8_t>& out, size_t offset, size_t len) {
  if (len < 12) {
    emit_copy_at_most_64(out, offset, len, true);
    return;
  }
  while (len >= 68) {
    emit_copy_at_most_64(out, offset, 64, false);
    len -= 64;
  }
  if (len > 64) {
    emit_copy_at_most_64(out, offset, 60, false);
    len -= 60;
  }
  emit_copy_at_most_64(out, offset, len, len < 12);
}

inline size_t match_length(const uint8_t* s1, const uint8_t* s2, const uint8_t* s2_limit) {
  size_t m = 0;
  while (s2 + m < s2_limit && s1[m] == s2[m]) ++m;
  return m;
}

inline void compress_fragment(const uint8_t* input, size_t input_size, std::vector<uint8_t>& out,
                              uint16_t* table, uint32_t table_size) {
  int bits = 0;
  while ((1u << bits) < table_size) ++bits;
  const uint8_t* ip = input;
  const uint8_t* ip_end = input + input_size;
  const uint8_t* base_ip = ip;
  const uint8_t* next_emit = ip;
  if (input_size >= kInputMarginBytes) {
    const uint8_t* ip_limit = input + input_size - kInputMarginBytes;
    ++ip;
    uint32_t next_hash = hash_index(load32(ip), bits);
    for (;;) {
      // Step 1: scan for a 4-byte match, accelerating over incompressible input.
      uint32_t skip = 32;
      const uint8_t* next_ip = ip;
      const uint8_t* candidate;
      do {
        ip = next_ip;
        uint32_t h = next_hash;
        uint32_t step = skip >> 5;
        skip += step;
        next_ip = ip + step;
        if (next_ip > ip_limit) goto emit_remainder;
        next_hash = hash_index(load32(next_ip), bits);
        candidate = base_ip + table[h];
        table[h] = (uint16_t)(ip - base_ip);
      } while (load32(ip) != load32(candidate));

      // Step 2: the bytes since the last emit go out as one literal.
      emit_literal(out, next_emit, (size_t)(ip - next_emit));

      // Step 3: copies while the next position matches too.
      do {
        const uint8_t* base = ip;
        size_t matched = 4 + match_length(candidate + 4, ip + 4, ip_end);
        ip += matched;
        emit_copy(out, (size_t)(base - candidate), matched);
        next_emit = ip;
        if (ip >= ip_limit) goto emit_remainder;
        table[hash_index(load32(ip - 1), bits)] = (uint16_t)(ip - base_ip - 1);
        uint32_t cur = hash_index(load32(ip), bits);
        candidate = base_ip + table[cur];
        table[cur] = (uint16_t)(ip - base_ip);
      } while (load32(ip) == load32(candidate));
      ++ip;
      next_hash = hash_index(load32(ip), bits);
    }
  }
emit_remainder:
  if (next_emit < ip_end) emit_literal(out, next_emit, (size_t)(ip_end - next_emit));
}

inline std::vector<uint8_t> compress(const uint8_t* in, size_t n) {
  std::vector<uint8_t> out;
  uint32_t v = (uint32_t)n;   // varint32 of the uncompressed length
  while (v >= 0x80) {
    out.push_back((uint8_t)(v | 0x80));
    v >>= 7;
  }
  out.push_back((uint8_t)v);
  std::vector<uint16_t> table(kMaxHashTableSize);
  size_t pos = 0;
  while (pos < n) {
    size_t frag = n - pos < kBlockSize ? n - pos : kBlockSize;
    uint32_t ts = (uint32_t)kMinHashTableSize;
    while (ts < kMaxHashTableSize && ts < frag) ts <<= 1;
    std::fill(table.begin(), table.begin() + ts, (uint16_t)0);
    compress_fragment(in + pos, frag, out, table.data(), ts);
    pos += frag;
  }
  return out;
}

// Format-defined decompression; malformed input throws Error (the reference unwraps).
inline std::vector<uint8_t> decompress(const uint8_t* in, size_t n) {
  size_t pos = 0;
  uint64_t ulen = 0;
  int shift = 0;
  for (;;) {
    if (pos >= n || shift > 28) throw Error("snappy: bad length varint");
    uint8_t b = in[pos++];
    ulen |= (uint64_t)(b & 0x7f) << shift;
    if (!(b & 0x80)) break;
    shift += 7;
  }
  if (ulen > 0xffffffffull) throw Error("snappy: length overflow");
  std::vector<uint8_t> out;
  out.reserve((size_t)ulen);
  while (pos < n) {
    uint8_t tag = in[pos++];
    size_t len, off = 0;
    switch (tag & 3) {
      case 0: {
        len = (tag >> 2) + 1;
        if (len > 60) {
          size_t nb = len - 60;
          if (pos + nb > n) throw Error("snappy: truncated literal length");
          len = 0;
          for (size_t k = 0; k < nb; ++k) len |= (size_t)in[pos + k] << (8 * k);
          len += 1;
          pos += nb;
        }
        if (pos + len > n) throw Error("snappy: truncated literal");
        out.insert(out.end(), in + pos, in + pos + len);
        pos += len;
        continue;
      }
      case 1:
        if (pos + 1 > n) throw Error("snappy: truncated copy");
        len = ((tag >> 2) & 7) + 4;
        off = ((size_t)(tag >> 5) << 8) | in[pos];
        pos += 1;
        break;
      case 2:
        if (pos + 2 > n) throw Error("snappy: truncated copy");
        len = (tag >> 2) + 1;
        off = (size_t)in[pos] | ((size_t)in[pos + 1] << 8);
        pos += 2;
        break;
      default:
        if (pos + 4 > n) throw Error("snappy: truncated copy");
        len = (tag >> 2) + 1;
        off = (size_t)load32(in + pos);
        pos += 4;
        break;
    }
    if (off == 0 || off > out.size()) throw Error("snappy: bad copy offset");
    size_t from = out.size() - off;
    for (size_t k = 0; k < len; ++k) out.push_back(out[from + k]);
  }
  if (out.size() != ulen) throw Error("snappy: length mismatch");
  return out;
}

}  // namespace snappy_ref
