// Roaring bitmaps in the portable serialized form, as croaring 0.5 (the
// binding of CRoaring that database/src/compress/split_double.rs uses for the
// residue format's outlier rows, 744-760 and 1398-1401) writes them after
// run_optimize() and reads them in Bitmap::deserialize:
//   * rows are grouped by their high 16 bits into containers (keys ascending);
//     a container built by add() is a sorted u16 array up to 4096 rows, else a
//     65536-bit bitset; run_optimize turns it into runs when that is smaller
//     (array: 2 + 4 runs < 2 card + 2; bitset: 2 + 4 runs < 8192);
//   * with a run container: u32 12347 | (containers - 1) << 16, then one flag
//     bit per container (run or not); without: u32 12346, u32 containers;
//   * per container u16 key, u16 card - 1; then, without runs or with four or
//     more containers, a u32 byte offset of each container's data from the
//     start of the serialization; then the data: array u16s, bitset 8192 bytes
//     (LE u64 words), run u16 count then (u16 start, u16 length - 1) pairs.
// Here a bitmap is BitVec<u32> words (row i = bit i % 32 of word i / 32) over
// `nbits` rows, the outlier masks' own layout.
#pragma once
#include <stdint.h>

#include <cstring>
#include <vector>

namespace buffgpu {
namespace roaring {

constexpr uint32_t kCookieNoRun = 12346, kCookieRun = 12347;
constexpr uint32_t kNoOffsetThreshold = 4, kArrayMax = 4096;
constexpr uint32_t kWordsPerKey = 65536 / 32;

struct Shape {
  uint32_t key, card, runs;
  int type;   // 0 array, 1 bitset, 2 run
};

// the words of container `key` (zero past nbits)
inline void key_words(const uint32_t* words, uint64_t nwords, uint32_t key, uint32_t* w) {
  const uint64_t w0 = (uint64_t)key * kWordsPerKey;
  for (uint32_t i = 0; i < kWordsPerKey; ++i) w[i] = w0 + i < nwords ? words[w0 + i] : 0u;
}

inline void serialize(const uint32_t* words, uint64_t nbits, std::vector<uint8_t>* out) {
  const uint64_t nwords = (nbits + 31) / 32;
  const uint32_t nkeys = (uint32_t)((nbits + 65535) / 65536);
  std::vector<Shape> cs;
  std::vector<uint32_t> w(kWordsPerKey);
  for (uint32_t k = 0; k < nkeys; ++k) {
    key_words(words, nwords, k, w.data());
    uint32_t card = 0, runs = 0, carry = 0;
    for (uint32_t i = 0; i < kWordsPerKey; ++i) {
      card += (uint32_t)__builtin_popcount(w[i]);
      runs += (uint32_t)__builtin_popcount(w[i] & ~((w[i] << 1) | carry));   // set bits after a clear one
      carry = w[i] >> 31;
    }
    if (!card) continue;
    int type = card <= kArrayMax ? 0 : 1;
    const uint32_t as_run = 2 + 4 * runs;
    if ((type == 0 && as_run < 2 * card + 2) || (type == 1 && as_run < 8192)) type = 2;
    cs.push_back(Shape{k, card, runs, type});
  }
  bool hasrun = false;
  for (const Shape& c : cs) hasrun |= c.type == 2;
  const uint32_t size = (uint32_t)cs.size();
  std::vector<uint8_t>& o = *out;
  o.clear();
  auto u16 = [&](uint32_t v) { o.push_back((uint8_t)v); o.push_back((uint8_t)(v >> 8)); };
  auto u32 = [&](uint32_t v) { u16(v & 0xFFFF); u16(v >> 16); };
  uint32_t at;   // where the first container's data starts
  if (hasrun) {
    u32(kCookieRun | ((size - 1) << 16));
    const uint32_t flag_bytes = (size + 7) / 8;
    const size_t f0 = o.size();
    o.resize(f0 + flag_bytes, 0);
    for (uint32_t i = 0; i < size; ++i)
      if (cs[i].type == 2) o[f0 + i / 8] |= (uint8_t)(1u << (i % 8));
    at = 4 + flag_bytes + 4 * size + (size >= kNoOffsetThreshold ? 4 * size : 0);
  } else {
    u32(kCookieNoRun);
    u32(size);
    at = 8 + 8 * size;
  }
  for (const Shape& c : cs) {
    u16(c.key);
    u16(c.card - 1);
  }
  if (!hasrun || size >= kNoOffsetThreshold) {
    for (const Shape& c : cs) {
      u32(at);
      at += c.type == 0 ? 2 * c.card : c.type == 1 ? 8192 : 2 + 4 * c.runs;
    }
  }
  for (const Shape& c : cs) {
    key_words(words, nwords, c.key, w.data());
    if (c.type == 1) {
      const size_t o0 = o.size();
      o.resize(o0 + 8192);
      memcpy(o.data() + o0, w.data(), 8192);   // LE u32 pairs = LE u64 words
      continue;
    }
    if (c.type == 2) u16(c.runs);
    uint32_t run_start = 0;
    bool in_run = false;
    for (uint32_t i = 0; i < 65536; ++i) {
      const bool b = (w[i >> 5] >> (i & 31)) & 1u;
      if (c.type == 0) {
        if (b) u16(i);
        continue;
      }
      if (b && !in_run) {
        run_start = i;
        in_run = true;
      } else if (!b && in_run) {
        u16(run_start);
        u16(i - 1 - run_start);
        in_run = false;
      }
    }
    if (c.type == 2 && in_run) {
      u16(run_start);
      u16(65535 - run_start);
    }
  }
}

// Parses a portable bitmap into BitVec words of nbits rows. Returns false
// (with `why`) on a malformed bitmap (truncated, an unknown cookie, runs or
// array values out of order, a run past its container) or a row at or past
// nbits.
inline bool deserialize(const uint8_t* p, uint64_t n, uint64_t nbits, std::vector<uint32_t>* words, uint64_t* card,
                        const char** why) {
  uint64_t at = 0;
  auto need = [&](uint64_t k) { return k <= n - at; };
  auto u16 = [&](uint32_t* v) {
    if (!need(2)) return false;
    *v = (uint32_t)p[at] | ((uint32_t)p[at + 1] << 8);
    at += 2;
    return true;
  };
  auto u32 = [&](uint32_t* v) {
    uint32_t lo, hi;
    if (!u16(&lo) || !u16(&hi)) return false;
    *v = lo | (hi << 16);
    return true;
  };
  words->assign((nbits + 31) / 32, 0u);
  *card = 0;
  uint32_t cookie, size;
  const uint8_t* flags = nullptr;
  *why = "truncated";
  if (!u32(&cookie)) return false;
  if ((cookie & 0xFFFF) == kCookieRun) {
    size = (cookie >> 16) + 1;
    if (!need((size + 7) / 8)) return false;
    flags = p + at;
    at += (size + 7) / 8;
  } else if (cookie == kCookieNoRun) {
    if (!u32(&size)) return false;
    if (size > 65536) {
      *why = "more than 65536 containers";
      return false;
    }
  } else {
    *why = "unknown cookie";
    return false;
  }
  std::vector<uint32_t> keys(size), cards(size);
  for (uint32_t i = 0; i < size; ++i) {
    if (!u16(&keys[i]) || !u16(&cards[i])) return false;
    cards[i] += 1;
  }
  if (!flags || size >= kNoOffsetThreshold) {
    if (!need(4ull * size)) return false;
    at += 4ull * size;
  }
  auto set = [&](uint64_t row) {
    if (row >= nbits) return false;
    uint32_t& w = (*words)[row >> 5];
    const uint32_t bit = 1u << (row & 31);
    *card += (w & bit) ? 0 : 1;
    w |= bit;
    return true;
  };
  for (uint32_t i = 0; i < size; ++i) {
    const uint64_t hi = (uint64_t)keys[i] << 16;
    bool ok = true;
    if (flags && ((flags[i / 8] >> (i % 8)) & 1)) {
      uint32_t nr;
      if (!u16(&nr)) return false;
      uint32_t next = 0;   // runs are sorted, disjoint and inside the container
      for (uint32_t r = 0; r < nr && ok; ++r) {
        uint32_t s, l;
        if (!u16(&s) || !u16(&l)) return false;
        if (s < next || s + l > 65535) {
          *why = "runs out of order or past their container";
          return false;
        }
        for (uint32_t v = s; v <= s + l && ok; ++v) ok = set(hi | v);
        next = s + l + 1;
      }
    } else if (cards[i] > kArrayMax) {
      if (!need(8192)) return false;
      for (uint32_t b = 0; b < 65536 && ok; ++b)
        if ((p[at + (b >> 3)] >> (b & 7)) & 1) ok = set(hi | b);
      at += 8192;
    } else {
      uint32_t next = 0;   // array values strictly increase
      for (uint32_t k = 0; k < cards[i] && ok; ++k) {
        uint32_t v;
        if (!u16(&v)) return false;
        if (v < next) {
          *why = "array values out of order";
          return false;
        }
        ok = set(hi | v);
        next = v + 1;
      }
    }
    if (!ok) {
      *why = "a row past the column";
      return false;
    }
  }
  return true;
}

}  // namespace roaring
}  // namespace buffgpu
