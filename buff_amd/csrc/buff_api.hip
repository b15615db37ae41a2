// Host runtime and C ABI of the BUFF engine (include/buffgpu.h).
//
// A buff_ctx owns one HIP stream on one device, a caching device allocator
// that enforces the context's HBM budget, small pinned result buffers, per-op
// launch shapes (set by the autotuner) and optional per-kernel event timing.
// A buff_chunk is one column in the resident layout of buff_common.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/buffgpu.h"
#include "buff_common.h"
#include "buff_cond.h"
#include "buff_kernels.h"
#include "buff_runtime.h"
#include "buff_launch.h"
#include "buff_wmax.h"

namespace buffgpu {

// ---------------------------------------------------------------------------
// errors

static thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

// ---------------------------------------------------------------------------
// precision helpers (the reference's per-call setup, split_double.rs:304-316)

int prec_of_scale(uint64_t scale) {
  const float s = (float)scale;          // usize as f32
  const float l = log10f(s);
  if (std::isnan(l)) return 0;           // f32 as i32 saturates, NaN -> 0
  if (l <= -2147483648.0f) return INT32_MIN;
  if (l >= 2147483648.0f) return INT32_MAX;
  return (int)l;
}

int precision_map(int p) {  // database/src/compress/mod.rs:29-46
  static const int tab[16] = {-1, 5, 8, 11, 15, 18, 21, 25, 28, 31, 35, 38, 50, 10, 10, 10};
  return (p >= 1 && p <= 15) ? tab[p] : -1;
}

int32_t precision_exp_of_prec(int prec) {  // prec_double.rs:28-33 of get_precision_bound(p), 273-281
  std::string s = "0.";
  for (int i = 0; i < prec; ++i) s.push_back('0');
  s += "49";
  const double b = strtod(s.c_str(), nullptr);
  uint64_t u;
  memcpy(&u, &b, 8);
  return (int32_t)((u & 0x7FF0000000000000ull) >> 52) - 1023;
}

double powi2(int32_t e) {  // 2.0f64.powi(e) as lowered by compiler-rt __powidf2
  double r = 1, a = 2.0;
  int64_t b = e;
  const bool recip = b < 0;
  while (true) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1 / r : r;
}

// `(delta as f64).log2().ceil() as usize` (split_double.rs:338-344)
uint32_t width_of_delta(int64_t delta) {
  if (delta == 0) return 0;
  const double c = std::ceil(std::log2((double)delta));
  if (std::isnan(c) || c <= 0) return 0;
  if (c >= 64) return 64;
  return (uint32_t)c;
}

// ---------------------------------------------------------------------------
// context

Ctx::~Ctx() {
  if (stream) {
    (void)hipStreamSynchronize(stream);
  }
  for (auto& kv : free_blocks) (void)hipFree(kv.second);
  free_blocks.clear();
  for (auto& p : prof_pool) (void)hipEventDestroy(p);
  if (pinned) (void)hipHostFree(pinned);
  if (mbox) (void)hipHostFree(mbox);
  if (stream) (void)hipStreamDestroy(stream);
}

int Ctx::alloc(uint64_t bytes, void** out) {
  const uint64_t need = round_up(std::max<uint64_t>(bytes, 256), 256);
  std::lock_guard<std::mutex> lk(mu);
  // reuse a cached block of close size
  auto it = free_blocks.lower_bound(need);
  if (it != free_blocks.end() && it->first <= need + need / 4 + (1u << 20)) {
    *out = it->second;
    live[it->second] = it->first;
    free_blocks.erase(it);
    return BUFF_OK;
  }
  if (budget && held + need > budget) {
    // give cached blocks back to the device first
    for (auto& kv : free_blocks) {
      (void)hipFree(kv.second);
      held -= kv.first;
    }
    free_blocks.clear();
    if (held + need > budget)
      return fail(BUFF_E_OOM, "HBM budget exceeded: %llu held + %llu requested > %llu",
                  (unsigned long long)held, (unsigned long long)need, (unsigned long long)budget);
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, need);
  if (e != hipSuccess) {
    for (auto& kv : free_blocks) {
      (void)hipFree(kv.second);
      held -= kv.first;
    }
    free_blocks.clear();
    e = hipMalloc(&p, need);
    if (e != hipSuccess) return fail(BUFF_E_OOM, "hipMalloc(%llu): %s", (unsigned long long)need, hipGetErrorString(e));
  }
  held += need;
  live[p] = need;
  *out = p;
  return BUFF_OK;
}

int Ctx::release(void* p) {
  if (!p) return BUFF_OK;
  std::lock_guard<std::mutex> lk(mu);
  auto it = live.find(p);
  if (it == live.end()) return fail(BUFF_E_INVALID, "pointer %p was not allocated by this context", p);
  free_blocks.emplace(it->second, p);
  live.erase(it);
  return BUFF_OK;
}

int Ctx::max_slots(unsigned int** slots, uint32_t* tag) {
  if (!s_maxslot.ptr || max_tag + 1 >= kMaxTagLimit) {
    BUFF_TRY(scratch(s_maxslot, kMaxSlotBytes));
    HIP_TRY(launch_fill_bytes(s_maxslot.ptr, kMaxSlotBytes, 0, stream));
    max_tag = 0;
  }
  *tag = ++max_tag;
  *slots = static_cast<unsigned int*>(s_maxslot.ptr);
  return BUFF_OK;
}

int Ctx::sum_slots(unsigned long long** set, unsigned long long** next) {
  if (!s_sumslot.ptr) {
    BUFF_TRY(scratch(s_sumslot, 2ull * kSumSetWords * 8));
    HIP_TRY(launch_fill_bytes(s_sumslot.ptr, 2ull * kSumSetWords * 8, 0, stream));
    sum_set = 0;
  }
  unsigned long long* base = static_cast<unsigned long long*>(s_sumslot.ptr);
  *set = base + (uint64_t)sum_set * kSumSetWords;
  *next = base + (uint64_t)(sum_set ^ 1u) * kSumSetWords;
  return BUFF_OK;
}

int Ctx::minmax_slots(int64_t** set, int64_t** next) {
  if (!s_mmslot.ptr) {
    BUFF_TRY(scratch(s_mmslot, 2ull * kMinMaxSetWords * 8));
    HIP_TRY(launch_minmax_slots_init(static_cast<int64_t*>(s_mmslot.ptr), 2, stream));
    mm_set = 0;
  }
  int64_t* base = static_cast<int64_t*>(s_mmslot.ptr);
  *set = base + (uint64_t)mm_set * kMinMaxSetWords;
  *next = base + (uint64_t)(mm_set ^ 1u) * kMinMaxSetWords;
  return BUFF_OK;
}

int Ctx::publish(const void* src, uint32_t mode, uint32_t nk, Publish* pub) {
  pub->mbox = mbox_dev;
  pub->seq = ++mbox_seq;
  HIP_TRY(launch_publish(src, mode, nk, *pub, stream));
  return BUFF_OK;
}

int Ctx::published(const Publish& pub, uint64_t* words, uint32_t nw) {
  // The launch is the last work on the stream: spin on the sequence word, and
  // now and then ask the stream, so a failed launch (or one that ended
  // without publishing) ends the wait with its error.
  for (uint32_t spin = 1;; ++spin) {
    if (__atomic_load_n(&mbox[0], __ATOMIC_ACQUIRE) == pub.seq) break;
    if (spin % 512 == 0) {
      const hipError_t e = hipStreamQuery(stream);
      if (e == hipSuccess) {
        if (__atomic_load_n(&mbox[0], __ATOMIC_ACQUIRE) == pub.seq) break;
        return fail(BUFF_E_DEVICE, "a reduction ended without publishing its result");
      }
      if (e != hipErrorNotReady) return fail(BUFF_E_DEVICE, "reduction: %s", hipGetErrorString(e));
    }
    __builtin_ia32_pause();
  }
  for (uint32_t i = 0; i < nw; ++i) words[i] = __atomic_load_n(&mbox[1 + i], __ATOMIC_RELAXED);
  return BUFF_OK;
}

int Ctx::minmax_pass(const double* x, uint64_t n, int32_t pexp, uint32_t dlen, MinMaxOut* mm) {
  const uint64_t ntiles = (n + kTileRows - 1) / kTileRows;
  const uint32_t nwg = minmax_grid(ntiles);
  BUFF_TRY(scratch(s_partial, 3ull * nwg * sizeof(int64_t)));
  int64_t *set = nullptr, *next = nullptr;
  BUFF_TRY(minmax_slots(&set, &next));
  {
    ProfScope ps(this, "k_minmax");
    HIP_TRY(launch_minmax_slots(x, n, pexp, dlen, (int64_t*)s_partial.ptr, set, next, nwg, stream));
    minmax_slots_used();
  }
  Publish pub;
  BUFF_TRY(publish(set, kPublishMinMax, 0, &pub));
  uint64_t w[3];
  BUFF_TRY(published(pub, w, 3));
  *mm = MinMaxOut{};
  mm->min_all = (int64_t)w[0];
  mm->max_rest = (int64_t)w[1];
  mm->first = (int64_t)w[2];
  return BUFF_OK;
}

int Ctx::scratch(Scratch& s, uint64_t bytes) {
  if (s.bytes >= bytes) return BUFF_OK;
  if (s.ptr) {
    // the previous buffer may still be read by queued kernels
    HIP_TRY(hipStreamSynchronize(stream));
    BUFF_TRY(release(s.ptr));
    s.ptr = nullptr;
    s.bytes = 0;
  }
  BUFF_TRY(alloc(bytes, &s.ptr));
  s.bytes = bytes;
  return BUFF_OK;
}

void Ctx::prof_begin(const char* name, hipEvent_t* ev, bool kernels) {
  ev[0] = ev[1] = nullptr;
  if (!prof_on) return;
  for (int i = 0; i < 2; ++i) {
    if (prof_pool.empty()) {
      (void)hipEventCreate(&ev[i]);
    } else {
      ev[i] = prof_pool.back();
      prof_pool.pop_back();
    }
  }
  if (kernels) {
    LaunchEvents& le = launch_events();
    le.start = ev[0];
    le.stop = ev[1];
    le.started = false;
  } else {
    (void)hipEventRecord(ev[0], stream);
  }
  (void)name;
}

void Ctx::prof_end(const char* name, hipEvent_t* ev, bool kernels) {
  if (!ev[0]) return;
  if (kernels) {
    LaunchEvents& le = launch_events();
    const bool started = le.started;
    le = LaunchEvents{};
    if (!started) {   // nothing launched: an empty interval
      (void)hipEventRecord(ev[0], stream);
      (void)hipEventRecord(ev[1], stream);
    }
  } else {
    (void)hipEventRecord(ev[1], stream);
  }
  prof_pending.push_back(ProfRec{name, ev[0], ev[1]});
}


int Ctx::prof_collect() {
  if (prof_pending.empty()) return BUFF_OK;
  HIP_TRY(hipStreamSynchronize(stream));
  for (auto& r : prof_pending) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    auto& s = prof_stats[r.name];
    s.first += 1;
    s.second += ms;
    prof_pool.push_back(r.a);
    prof_pool.push_back(r.b);
  }
  prof_pending.clear();
  return BUFF_OK;
}

LaunchShape Ctx::shape(Op op, const ChunkView& v) const {
  auto it = tuned.find(tune_key(op, v.ntiles, v.P));
  if (it != tuned.end()) return it->second;
  return LaunchShape{4, 0, 0};
}

uint32_t tune_bucket(uint64_t ntiles) {
  uint32_t b = 0;
  while ((1ull << (b + 1)) <= ntiles) ++b;
  return b;
}

// ---------------------------------------------------------------------------
// chunk helpers

uint64_t stream_bytes_of(uint64_t n, uint32_t P, uint32_t r) {
  return kStreamHeader + (uint64_t)P * n + (n * r + 7) / 8;
}

int chunk_alloc(Ctx* ctx, Chunk* ch) {
  const uint64_t ntiles = (ch->n + kTileRows - 1) / kTileRows;
  ch->view.n = ch->n;
  ch->view.ntiles = ntiles;
  ch->view.pitch = ntiles * kTileRows;
  const uint64_t plane_bytes = (uint64_t)ch->view.P * ch->view.pitch;
  const uint64_t tail_bytes = ntiles * 64ull * ch->view.r * 2ull;
  ch->device_bytes = std::max<uint64_t>(plane_bytes + tail_bytes, 256);
  void* p = nullptr;
  BUFF_TRY(ctx->alloc(ch->device_bytes, &p));
  ch->mem = p;
  ch->view.planes = (uint8_t*)p;
  ch->view.tail = (uint16_t*)((uint8_t*)p + plane_bytes);
  return BUFF_OK;
}

void chunk_set_width(Chunk* ch, uint32_t F) {
  ch->view.F = F;
  ch->view.P = F >= 8 ? F / 8 : 0;
  ch->view.r = F >= 8 ? F % 8 : F;
}

// Small results come back through the mailbox (whole u64 words, up to
// kPublishWords - 1), else with a copy and a stream wait.
int d2h_small(Ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (bytes > 0 && bytes % 8 == 0 && bytes / 8 < kPublishWords && (uintptr_t)src % 8 == 0) {
    uint64_t w[kPublishWords];
    Publish pub;
    BUFF_TRY(ctx->publish(src, kPublishCopy, (uint32_t)(bytes / 8), &pub));
    BUFF_TRY(ctx->published(pub, w, (uint32_t)(bytes / 8)));
    memcpy(dst, w, bytes);
    return BUFF_OK;
  }
  if (bytes > Ctx::kPinnedBytes) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return BUFF_OK;
  }
  HIP_TRY(hipMemcpyAsync(ctx->pinned, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  memcpy(dst, ctx->pinned, bytes);
  return BUFF_OK;
}

// ---------------------------------------------------------------------------
// operations

// The running min and `else if` max of split_double.rs:321-330 over n rows of
// device memory (n == 0: INT64_MAX / INT64_MIN, as the reference's initial values).
int quirk_minmax(Ctx* ctx, const double* x, uint64_t n, int32_t pexp, uint32_t dlen, int64_t* mn_out,
                 int64_t* mx_out) {
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  if (n > 0) {
    MinMaxOut mm;
    BUFF_TRY(ctx->minmax_pass(x, n, pexp, dlen, &mm));
    const uint64_t ntiles = (n + kTileRows - 1) / kTileRows;
    const uint32_t nwg = minmax_grid(ntiles);   // the grid the pass ran with: its partial rows
    BUFF_TRY(ctx->scratch(ctx->s_minmax, sizeof(MinMaxOut)));
    MinMaxOut* dmm = (MinMaxOut*)ctx->s_minmax.ptr;
    int64_t* part = (int64_t*)ctx->s_partial.ptr;
    mn = mm.min_all;
    if (n == 1) {
      mx = mm.first == INT64_MAX ? INT64_MAX : INT64_MIN;
    } else if (mm.max_rest >= mm.first) {
      mx = mm.max_rest;
    } else {
      // row 0 holds the unique maximum: replay the `else if` exactly
      {
        ProfScope ps(ctx, "k_minmax_prefix");
        HIP_TRY(launch_minmax(x, n, pexp, dlen, part, part + 2 * nwg, dmm, nwg, true, ctx->stream));
      }
      BUFF_TRY(d2h_small(ctx, &mm, dmm, sizeof(mm)));
      mx = mm.qmax;
      if (mm.first == INT64_MAX) mx = INT64_MAX;
    }
  }
  *mn_out = mn;
  *mx_out = mx;
  return BUFF_OK;
}

int encode_device(Ctx* ctx, const double* x, uint64_t n, uint64_t scale, uint32_t flags, Chunk** out) {
  if (!ctx || !out) return fail(BUFF_E_INVALID, "null argument");
  const bool slice = (flags & BUFF_SLICE) != 0;
  if (slice && (flags & (BUFF_FLIP | BUFF_MAJORITY)))
    return fail(BUFF_E_INVALID, "BUFF_SLICE excludes BUFF_FLIP and BUFF_MAJORITY");
  if ((flags & BUFF_RESIDUE) && (flags & (BUFF_FLIP | BUFF_MAJORITY | BUFF_SLICE)))
    return fail(BUFF_E_INVALID, "BUFF_RESIDUE excludes BUFF_FLIP, BUFF_MAJORITY and BUFF_SLICE");
  if (flags & (BUFF_MAJORITY | BUFF_RESIDUE)) return major_encode_device(ctx, x, n, scale, flags, out);
  if (n > 0xFFFFFFFFull) return fail(BUFF_E_INVALID, "a chunk holds at most 2^32-1 rows (n is stored as u32)");
  if (n && !x) return fail(BUFF_E_INVALID, "null input");
  const int prec = scale == 0 ? 0 : prec_of_scale(scale);
  const int dl = precision_map(prec);
  if (dl < 0) return fail(BUFF_E_REF_PANIC, "precision %d (scale %llu) is not in PRECISION_MAP", prec,
                          (unsigned long long)scale);
  const uint32_t dlen = (uint32_t)dl;
  const int32_t pexp = precision_exp_of_prec(prec);

  int64_t mn = INT64_MAX, mx = INT64_MIN;
  BUFF_TRY(quirk_minmax(ctx, x, n, pexp, dlen, &mn, &mx));
  const int64_t delta = (int64_t)((uint64_t)mx - (uint64_t)mn);
  const uint32_t F = width_of_delta(delta);
  // buff_slice_encode writes a constant column as one 0x80 plane (buff_slice.rs:372-379)
  if (F == 0 && n >= 1 && !(flags & BUFF_SEMANTIC) && !slice)
    return fail(BUFF_E_REF_PANIC, "all %llu values map to one fixed-point value (F = 0): the reference panics",
                (unsigned long long)n);
  if (F > 63) return fail(BUFF_E_UNSUPPORTED, "bit width %u exceeds 63", F);

  Chunk* ch = new (std::nothrow) Chunk();
  if (!ch) return fail(BUFF_E_OOM, "host allocation");
  ch->ctx = ctx;
  ch->n = n;
  ch->scale = scale;
  ch->flags = flags & (BUFF_FLIP | BUFF_SLICE);
  ch->dlen = dlen;
  ch->ilen = (uint32_t)((uint64_t)F - dlen);
  ch->F = F;
  ch->view.base = mn;
  ch->view.dlen = dlen;
  chunk_set_width(ch, F);
  int rc = chunk_alloc(ctx, ch);
  if (rc != BUFF_OK) {
    delete ch;
    return rc;
  }
  if (n > 0 && (ch->view.P || ch->view.r)) {
    ProfScope ps(ctx, "k_encode");
    hipError_t e = launch_encode(x, ch->view, pexp, ctx->stream, ctx->shape(Op::Encode, ch->view));
    if (e != hipSuccess) {
      (void)ctx->release(ch->mem);
      delete ch;
      return fail(BUFF_E_DEVICE, "k_encode launch: %s", hipGetErrorString(e));
    }
  }
  *out = ch;
  return BUFF_OK;
}

int chunk_serialize(Ctx* ctx, const Chunk* ch, uint8_t* dst, uint64_t* len) {
  if (!ctx || !ch || !len) return fail(BUFF_E_INVALID, "null argument");
  if (ch->flags & BUFF_RESIDUE) return residue_serialize(ctx, ch, dst, len);
  if (ch->flags & BUFF_MAJORITY) return major_serialize(ctx, ch, dst, len);
  if (ch->flags & BUFF_SLICE) return slice_serialize(ctx, ch, dst, len);
  const uint64_t total = stream_bytes_of(ch->n, ch->view.P, ch->view.r);
  const uint64_t need = *len;
  *len = total;
  if (!dst) return BUFF_OK;
  if (need < total) return fail(BUFF_E_SMALL_BUFFER, "stream needs %llu bytes", (unsigned long long)total);
  const uint64_t ubase = (uint64_t)ch->view.base;
  const uint32_t hdr[5] = {(uint32_t)ubase, (uint32_t)(ubase >> 32), (uint32_t)ch->n, ch->ilen, ch->dlen};
  memcpy(dst, hdr, 20);  // little-endian u32s, as BitPack::write(v, 32) lays them out
  const uint64_t body = total - kStreamHeader;
  if (body == 0) return BUFF_OK;
  BUFF_TRY(ctx->scratch(ctx->s_stream, body));
  const uint64_t tail_bytes = (ch->n * ch->view.r + 7) / 8;
  {
    ProfScope ps(ctx, "k_serialize");
    HIP_TRY(launch_serialize(ch->view, (ch->flags & BUFF_FLIP) ? 0x80 : 0, (uint8_t*)ctx->s_stream.ptr, tail_bytes,
                             ctx->stream));
  }
  HIP_TRY(hipMemcpyAsync(dst + kStreamHeader, ctx->s_stream.ptr, body, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return BUFF_OK;
}

int chunk_load(Ctx* ctx, const uint8_t* s, uint64_t len, uint64_t scale, uint32_t flags, Chunk** out) {
  if (!ctx || !s || !out) return fail(BUFF_E_INVALID, "null argument");
  if ((flags & BUFF_SLICE) && (flags & (BUFF_FLIP | BUFF_MAJORITY)))
    return fail(BUFF_E_INVALID, "BUFF_SLICE excludes BUFF_FLIP and BUFF_MAJORITY");
  if ((flags & BUFF_RESIDUE) && (flags & (BUFF_FLIP | BUFF_MAJORITY | BUFF_SLICE)))
    return fail(BUFF_E_INVALID, "BUFF_RESIDUE excludes BUFF_FLIP, BUFF_MAJORITY and BUFF_SLICE");
  if (flags & BUFF_RESIDUE) return residue_load(ctx, s, len, scale, flags, out);
  if (flags & BUFF_MAJORITY) return major_load(ctx, s, len, scale, flags, out);
  if (flags & BUFF_SLICE) return slice_load(ctx, s, len, scale, flags, out);
  if (len < kStreamHeader) return fail(BUFF_E_INVALID, "stream of %llu bytes has no 20-byte header",
                                       (unsigned long long)len);
  uint32_t hdr[5];
  memcpy(hdr, s, 20);
  const int64_t base = (int64_t)((uint64_t)hdr[0] | ((uint64_t)hdr[1] << 32));
  const uint64_t n = hdr[2];
  const uint32_t ilen = hdr[3], dlen = hdr[4];
  const uint32_t F = dlen + ilen;  // u32 wrap, as the readers compute `dlen + ilen`
  if (F > 63) return fail(BUFF_E_INVALID, "stored width dlen+ilen = %u is not a BUFF stream width", F);
  Chunk* ch = new (std::nothrow) Chunk();
  if (!ch) return fail(BUFF_E_OOM, "host allocation");
  ch->ctx = ctx;
  ch->n = n;
  ch->scale = scale;
  ch->flags = flags & BUFF_FLIP;
  ch->dlen = dlen;
  ch->ilen = ilen;
  ch->F = F;
  ch->view.base = base;
  ch->view.dlen = dlen;
  chunk_set_width(ch, F);
  const uint64_t expect = stream_bytes_of(n, ch->view.P, ch->view.r);
  if (len < expect) {
    delete ch;
    return fail(BUFF_E_INVALID, "stream is %llu bytes, its header needs %llu", (unsigned long long)len,
                (unsigned long long)expect);
  }
  int rc = chunk_alloc(ctx, ch);
  if (rc != BUFF_OK) {
    delete ch;
    return rc;
  }
  const uint64_t body = expect - kStreamHeader;
  if (n > 0 && (ch->view.P || ch->view.r)) {
    rc = ctx->scratch(ctx->s_stream, body);
    if (rc == BUFF_OK) {
      hipError_t e = hipMemcpyAsync(ctx->s_stream.ptr, s + kStreamHeader, body, hipMemcpyHostToDevice, ctx->stream);
      if (e == hipSuccess) {
        ProfScope ps(ctx, "k_load");
        e = launch_load(ch->view, (ch->flags & BUFF_FLIP) ? 0x80 : 0, (const uint8_t*)ctx->s_stream.ptr,
                        (n * ch->view.r + 7) / 8, ctx->stream);
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

// fetch_fixed_aligned of a predicate at the chunk's scale and dlen
// (prec_double.rs:190-205 with the operators' PrecisionBound).
int64_t fixed_of_pred(const Chunk* ch, double pred) {
  const int32_t pexp = precision_exp_of_prec(prec_of_scale(ch->scale));
  uint64_t u;
  memcpy(&u, &pred, 8);
  return fetch_fixed_aligned_bits(u, pexp, ch->dlen);
}

// Predicate -> per-plane target, or a constant answer. Returns true when the
// kernel must run; otherwise *fill is 0 (no rows) or 1 (all rows).
bool filter_plan(const Chunk* ch, int op, double pred, uint32_t flags, FilterTarget* tg, int* fill) {
  const ChunkView& v = ch->view;
  const int prec = prec_of_scale(ch->scale);
  const int32_t pexp = precision_exp_of_prec(prec);
  const int64_t fixed = fetch_fixed_aligned_bits([&] { uint64_t u; memcpy(&u, &pred, 8); return u; }(), pexp,
                                                 ch->dlen);
  memset(tg, 0, sizeof(*tg));
  if (fixed < v.base) {  // buff_simd.rs:4531-4534 / 4949-4952
    *fill = op == BUFF_OP_GT ? 1 : 0;
    return false;
  }
  const uint64_t t = (uint64_t)fixed - (uint64_t)v.base;
  const bool semantic = (flags & BUFF_SEMANTIC) != 0;
  if (v.F >= 8) {
    if (semantic && v.F < 64 && t >= (1ull << v.F)) {  // predicate above every stored offset
      *fill = 0;
      return false;
    }
    if (op == BUFF_OP_EQ && v.r == 0 && !semantic) {  // matches stay in rb1; res is empty
      *fill = 0;
      return false;
    }
    for (uint32_t p = 0; p < v.P; ++p) tg->digit[p] = (uint32_t)((t >> (v.F - 8 * (p + 1))) & 0xFF);
    tg->tail = v.r ? (t & ((1ull << v.r) - 1)) : 0;
    return true;
  }
  // F < 8: the whole value is the tail and compares against the full target
  if (v.F == 0) {
    *fill = (op == BUFF_OP_EQ && t == 0) ? 1 : 0;
    return false;
  }
  if (t >= (1ull << v.F)) {
    *fill = 0;
    return false;
  }
  tg->tail = t;
  return true;
}

int filter_device(Ctx* ctx, const Chunk* ch, int op, double pred, uint32_t flags, uint64_t* words,
                  unsigned long long* stats) {
  if (!ctx || !ch || !words) return fail(BUFF_E_INVALID, "null argument");
  const int rs = slice_filter_dispatch(ctx, ch, op, pred, flags, false, nullptr, words);
  if (rs >= 0) return rs;
  if (op != BUFF_OP_GT && op != BUFF_OP_EQ) return fail(BUFF_E_INVALID, "unknown filter op %d", op);
  if ((ch->flags & BUFF_RESIDUE) && !(flags & BUFF_SEMANTIC))
    return fail(BUFF_E_UNSUPPORTED, "the reference has no filter of a residue stream (BUFF_SEMANTIC answers it)");
  const uint64_t nwords = (ch->n + 63) / 64;
  if (nwords == 0) return BUFF_OK;
  if (ch->S > 0) return major_filter_device(ctx, ch, op, pred, flags, words);
  FilterTarget tg;
  int fill = 0;
  if (!filter_plan(ch, op, pred, flags, &tg, &fill)) {
    ProfScope ps(ctx, "k_fill_bitmap");
    HIP_TRY(launch_fill_bitmap(words, nwords, ch->n, fill, ctx->stream));
    return BUFF_OK;
  }
  ProfScope ps(ctx, op == BUFF_OP_GT ? "k_filter_gt" : "k_filter_eq");
  HIP_TRY(launch_filter(ch->view, op == BUFF_OP_GT ? 0 : 1, tg, words, nwords, stats, ctx->stream,
                        ctx->shape(op == BUFF_OP_GT ? Op::FilterGt : Op::FilterEq, ch->view)));
  return BUFF_OK;
}

int filter_host(Ctx* ctx, const Chunk* ch, int op, double pred, uint32_t flags, uint32_t* words, uint64_t nwords32,
                uint64_t* count) {
  if (!ctx || !ch) return fail(BUFF_E_INVALID, "null argument");
  const uint64_t need32 = (ch->n + 31) / 32;
  if (words && nwords32 < need32)
    return fail(BUFF_E_SMALL_BUFFER, "bitmap needs %llu u32 words", (unsigned long long)need32);
  const uint64_t nwords64 = (ch->n + 63) / 64;
  std::vector<uint64_t> host(nwords64);
  if (nwords64) {
    BUFF_TRY(ctx->scratch(ctx->s_bitmap, nwords64 * 8));
    BUFF_TRY(filter_device(ctx, ch, op, pred, flags, (uint64_t*)ctx->s_bitmap.ptr, nullptr));
    HIP_TRY(hipMemcpyAsync(host.data(), ctx->s_bitmap.ptr, nwords64 * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  uint64_t cnt = 0;
  for (uint64_t w : host) cnt += (uint64_t)__builtin_popcountll(w);
  if (count) *count = cnt;
  if (words && need32) memcpy(words, host.data(), need32 * 4);
  return BUFF_OK;
}

// ---------------------------------------------------------------------------
// Filter -> project on the rows of a condition bitmap (ceil(n/64) u64 words,
// the layout filter_device writes).

// The plan of a conditional filter. buff_range_filter_condition
// (buff_simd.rs:4683-4921) differs from buff_range_filter only where it
// returns early: a predicate below the column base gives no rows (4706-4709),
// and with F < 8 the target is its low byte (`fixed_target as u8`, 4718). The
// equality form has no reference counterpart: buff_filter's equality on the
// condition rows. buff_range_filter_majority_condition (buff_simd.rs:1279-1526)
// is the same on a buff-major chunk without sparse columns. fill: 0 = no rows,
// 1 = every condition row.
static bool cond_plan(const Chunk* ch, int op, double pred, uint32_t flags, FilterTarget* tg, int* fill) {
  const ChunkView& v = ch->view;
  if (op == BUFF_OP_GT && !(flags & BUFF_SEMANTIC)) {
    const int32_t pexp = precision_exp_of_prec(prec_of_scale(ch->scale));
    const int64_t fixed = fetch_fixed_aligned_bits([&] { uint64_t u; memcpy(&u, &pred, 8); return u; }(), pexp,
                                                   ch->dlen);
    if (fixed < v.base) {
      *fill = 0;
      return false;
    }
    if (v.F > 0 && v.F < 8) {
      const uint64_t t8 = ((uint64_t)fixed - (uint64_t)v.base) & 0xFF;
      if (t8 >= (1ull << v.F)) {
        *fill = 0;
        return false;
      }
      memset(tg, 0, sizeof(*tg));
      tg->tail = t8;
      return true;
    }
  }
  return filter_plan(ch, op, pred, flags, tg, fill);
}

int filter_cond_device(Ctx* ctx, const Chunk* ch, int op, double pred, uint32_t flags, const uint64_t* cond,
                       uint64_t* words) {
  if (!ctx || !ch) return fail(BUFF_E_INVALID, "null argument");
  const int rs = slice_filter_dispatch(ctx, ch, op, pred, flags, true, cond, words);
  if (rs >= 0) return rs;
  if (op != BUFF_OP_GT && op != BUFF_OP_EQ) return fail(BUFF_E_INVALID, "unknown filter op %d", op);
  const uint64_t nwords = (ch->n + 63) / 64;
  if (nwords == 0) return BUFF_OK;
  if (!words || !cond) return fail(BUFF_E_INVALID, "null bitmap");
  FilterTarget tg;
  int fill = 0;
  bool run = true;
  if (ch->flags & BUFF_RESIDUE) return fail(BUFF_E_UNSUPPORTED, "conditions and windows of residue chunks");
  if (ch->flags & BUFF_MAJORITY) {
    if (ch->F < 8) return fail(BUFF_E_UNSUPPORTED, "buff-major chunk with F < 8: its column-count byte holds tail bits");
    if (ch->S > 0) {
      // buff_range_filter_majority_condition (buff_simd.rs:1313-1315) skips every
      // column of a stream with sparse columns: no rows. The intended answer
      // (BUFF_SEMANTIC) and the equality form, which has no reference
      // counterpart, are the chunk's own filter on the condition rows.
      if (op == BUFF_OP_EQ || (flags & BUFF_SEMANTIC)) {
        BUFF_TRY(ctx->scratch(ctx->s_slice, nwords * 8));
        BUFF_TRY(major_filter_device(ctx, ch, op, pred, flags, (uint64_t*)ctx->s_slice.ptr));
        HIP_TRY(launch_bitmap_logic(words, (const uint64_t*)ctx->s_slice.ptr, nullptr, cond, nwords, ch->n, 0, 0,
                                    ctx->stream));
        return BUFF_OK;
      }
      run = false;
      fill = 0;
    }
  }
  if (!run || !cond_plan(ch, op, pred, flags, &tg, &fill)) {
    ProfScope ps(ctx, "k_cond_fill");
    HIP_TRY(launch_cond_fill(words, cond, nwords, ch->n, fill, ctx->stream));
    return BUFF_OK;
  }
  ProfScope ps(ctx, op == BUFF_OP_GT ? "k_filter_cond_gt" : "k_filter_cond_eq");
  HIP_TRY(launch_filter_cond(ch->view, op == BUFF_OP_GT ? 0 : 1, tg, cond, words, nwords, ctx->stream,
                             ctx->shape(op == BUFF_OP_GT ? Op::FilterGt : Op::FilterEq, ch->view)));
  return BUFF_OK;
}

// Host condition words (ceil(n/32) u32, BitVec<u32> order) -> ctx->s_condin.
static int upload_cond(Ctx* ctx, const Chunk* ch, const uint32_t* cond_words, const uint64_t** out) {
  const uint64_t nwords64 = (ch->n + 63) / 64, need32 = (ch->n + 31) / 32;
  BUFF_TRY(ctx->scratch(ctx->s_condin, nwords64 * 8));
  std::vector<uint64_t> host(nwords64, 0);
  if (need32) memcpy(host.data(), cond_words, need32 * 4);
  HIP_TRY(hipMemcpyAsync(ctx->s_condin.ptr, host.data(), nwords64 * 8, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));   // the host vector goes away
  *out = (const uint64_t*)ctx->s_condin.ptr;
  return BUFF_OK;
}

int filter_cond_host(Ctx* ctx, const Chunk* ch, int op, double pred, uint32_t flags, const uint32_t* cond_words,
                     uint32_t* words, uint64_t nwords32, uint64_t* count) {
  if (!ctx || !ch) return fail(BUFF_E_INVALID, "null argument");
  const uint64_t need32 = (ch->n + 31) / 32, nwords64 = (ch->n + 63) / 64;
  if (need32 && !cond_words) return fail(BUFF_E_INVALID, "null condition");
  if (words && nwords32 < need32)
    return fail(BUFF_E_SMALL_BUFFER, "bitmap needs %llu u32 words", (unsigned long long)need32);
  std::vector<uint64_t> host(nwords64);
  if (nwords64) {
    const uint64_t* cond = nullptr;
    BUFF_TRY(upload_cond(ctx, ch, cond_words, &cond));
    BUFF_TRY(ctx->scratch(ctx->s_bitmap, nwords64 * 8));
    BUFF_TRY(filter_cond_device(ctx, ch, op, pred, flags, cond, (uint64_t*)ctx->s_bitmap.ptr));
    HIP_TRY(hipMemcpyAsync(host.data(), ctx->s_bitmap.ptr, nwords64 * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  uint64_t cnt = 0;
  for (uint64_t w : host) cnt += (uint64_t)__builtin_popcountll(w);
  if (count) *count = cnt;
  if (words && need32) memcpy(words, host.data(), need32 * 4);
  return BUFF_OK;
}

// buff_decode_condition (buff_simd.rs:1528-1712): the values of the condition
// rows in row order. Default: the reference's tail placement (F >= 8 and
// F % 8 != 0: the tail bits OR-ed in shifted left by the tail width,
// 1693/1701) and its panic on an empty condition (`unwrap` of the first row).
// BUFF_SEMANTIC: the stored values; an empty condition gives no rows.
int materialize_cond_device(Ctx* ctx, const Chunk* ch, uint32_t flags, const uint64_t* cond, double* out,
                            uint64_t cap, uint64_t* rows) {
  if (!ctx || !ch) return fail(BUFF_E_INVALID, "null argument");
  if (ch->flags & BUFF_SLICE) {
    // buff_slice_decode_condition (buff_slice.rs:1266-1347): the stored values,
    // no rows for an empty condition, and the four-plane reader's panic
    if (!(flags & BUFF_SEMANTIC) && (ch->F == 0 || ch->F > 32))
      return fail(BUFF_E_REF_PANIC, "bit length greater than 32 is not supported yet (F = %u)", ch->F);
    flags |= BUFF_SEMANTIC;
  }
  const bool semantic = (flags & BUFF_SEMANTIC) != 0;
  if (ch->flags & BUFF_RESIDUE) return fail(BUFF_E_UNSUPPORTED, "conditions and windows of residue chunks");
  const bool major = (ch->flags & BUFF_MAJORITY) != 0;
  if (major) {
    if (ch->F < 8) return fail(BUFF_E_UNSUPPORTED, "buff-major chunk with F < 8: its column-count byte holds tail bits");
    if (ch->S > 0) {
      // buff_major_decode_condition (buff_simd.rs:1738-1740) decodes nothing
      // from a stream with sparse columns; BUFF_SEMANTIC: the stored values of
      // the condition rows (the whole column decoded, then gathered)
      if (!semantic) {
        if (rows) *rows = 0;
        return BUFF_OK;
      }
      uint64_t total = 0;
      if (ch->n) {
        if (!cond || (cap && !out)) return fail(BUFF_E_INVALID, "null argument");
        BUFF_TRY(ctx->scratch(ctx->s_condval, ch->n * 8));
        BUFF_TRY(major_materialize_device(ctx, ch, -1, BUFF_SEMANTIC, (double*)ctx->s_condval.ptr, nullptr));
        const uint64_t ntiles = (ch->n + kTileRows - 1) / kTileRows;
        BUFF_TRY(ctx->scratch(ctx->s_cond, cond_scratch_bytes(ntiles)));
        uint8_t* scr = (uint8_t*)ctx->s_cond.ptr;
        {
          ProfScope ps(ctx, "k_cond_decode");
          HIP_TRY(launch_cond_values((const double*)ctx->s_condval.ptr, ch->n, cond, out, cap, scr, ctx->stream));
        }
        BUFF_TRY(d2h_small(ctx, &total, scr + cond_total_offset(ntiles), 8));
        // the copy above waited for the gather: the decoded column goes back to
        // the context's block cache (reclaimable under the HBM budget)
        BUFF_TRY(ctx->release(ctx->s_condval.ptr));
        ctx->s_condval = Scratch{};
      }
      if (rows) *rows = total;
      if (total > cap)
        return fail(BUFF_E_SMALL_BUFFER, "%llu condition rows, room for %llu", (unsigned long long)total,
                    (unsigned long long)cap);
      return BUFF_OK;
    }
  }
  // buff_major_decode_condition places the tail bits correctly and reads an
  // empty condition as no rows
  const bool compat = !semantic && !major;
  uint64_t total = 0;
  if (ch->n) {
    if (!cond || (cap && !out)) return fail(BUFF_E_INVALID, "null argument");
    const uint64_t ntiles = ch->view.ntiles;
    BUFF_TRY(ctx->scratch(ctx->s_cond, cond_scratch_bytes(ntiles)));
    uint8_t* scr = (uint8_t*)ctx->s_cond.ptr;
    {
      ProfScope ps(ctx, "k_cond_decode");
      HIP_TRY(launch_cond_decode(ch->view, std::ldexp(1.0, -(int)ch->dlen), compat, cond, out, cap, scr,
                                 ctx->stream));
    }
    BUFF_TRY(d2h_small(ctx, &total, scr + cond_total_offset(ntiles), 8));
  }
  if (rows) *rows = total;
  if (total == 0 && compat) return fail(BUFF_E_REF_PANIC, "empty condition: the reference unwraps the first row");
  if (total > cap)
    return fail(BUFF_E_SMALL_BUFFER, "%llu condition rows, room for %llu", (unsigned long long)total,
                (unsigned long long)cap);
  return BUFF_OK;
}

int materialize_cond_host(Ctx* ctx, const Chunk* ch, uint32_t flags, const uint32_t* cond_words, double* out,
                          uint64_t cap, uint64_t* rows) {
  if (!ctx || !ch) return fail(BUFF_E_INVALID, "null argument");
  if (ch->n && !cond_words) return fail(BUFF_E_INVALID, "null condition");
  const uint64_t* cond = nullptr;
  if (ch->n) {
    BUFF_TRY(upload_cond(ctx, ch, cond_words, &cond));
    BUFF_TRY(ctx->scratch(ctx->s_values, ch->n * 8));
  }
  uint64_t total = 0;
  const int rc = materialize_cond_device(ctx, ch, flags, cond, (double*)ctx->s_values.ptr, ch->n, &total);
  if (rows) *rows = total;
  if (rc != BUFF_OK) return rc;
  if (total > cap)
    return fail(BUFF_E_SMALL_BUFFER, "%llu condition rows, room for %llu", (unsigned long long)total,
                (unsigned long long)cap);
  if (total) {
    if (!out) return fail(BUFF_E_INVALID, "null output");
    HIP_TRY(hipMemcpyAsync(out, ctx->s_values.ptr, total * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  return BUFF_OK;
}

// ---------------------------------------------------------------------------
// Windowed max: buff_max_range (buff_simd.rs:5096-5334).

// The reference's outcome on [s, e) in windows of W rows: the maxima of the
// full windows (its max_vec; a partial last window is dropped when F <= 8 and
// is a panic when F > 8, as is an empty range), and the argmax rows, which only
// its tail stage fills (F < 8, or F > 8 with a tail). BUFF_SEMANTIC: every
// window including a partial last one, and the argmax rows in every case.
int max_range_device(Ctx* ctx, const Chunk* ch, uint64_t s, uint64_t e, uint64_t W, uint32_t flags, double* vals,
                     uint64_t cap, uint64_t* nwin, uint64_t* argmax) {
  if (!ctx || !ch || !nwin) return fail(BUFF_E_INVALID, "null argument");
  if (W == 0 || s > e || e > ch->n)
    return fail(BUFF_E_INVALID, "window %llu over rows [%llu, %llu) of %llu", (unsigned long long)W,
                (unsigned long long)s, (unsigned long long)e, (unsigned long long)ch->n);
  if (ch->flags & BUFF_SLICE) {
    // buff_slice_max_range (buff_slice.rs:930-1098): every window and every
    // argmax row once its panics are ruled out
    BUFF_TRY(slice_max_range_check(ch, s, e, W, flags));
    flags |= BUFF_SEMANTIC;
  }
  const bool semantic = (flags & BUFF_SEMANTIC) != 0;
  const ChunkView& v = ch->view;
  const uint64_t L = e - s;
  if (ch->flags & BUFF_RESIDUE) return fail(BUFF_E_UNSUPPORTED, "conditions and windows of residue chunks");
  if (ch->flags & BUFF_MAJORITY) {
    // buff_max_majority_range (buff_simd.rs:2863-3114): with sparse columns it
    // logs an error and finds no window (2894-2896); without, it is
    // buff_max_range after the column-count byte, and the chunk's planes are the
    // byte_fixed layout.
    if (ch->F < 8) return fail(BUFF_E_UNSUPPORTED, "buff-major chunk with F < 8: its column-count byte holds tail bits");
    if (ch->S > 0) {
      if (semantic) return fail(BUFF_E_UNSUPPORTED, "windowed max of buff-major chunks with sparse columns");
      *nwin = 0;
      const uint64_t nwords = (ch->n + 63) / 64;
      if (argmax && nwords) HIP_TRY(launch_fill_bitmap(argmax, nwords, ch->n, 0, ctx->stream));
      return BUFF_OK;
    }
  }
  if (!semantic && v.F > 8 && (L == 0 || L % W))
    return fail(BUFF_E_REF_PANIC, "the reference asserts that windows of %llu rows tile %llu rows",
                (unsigned long long)W, (unsigned long long)L);
  const WmaxLayout lay = wmax_layout(s, e, W);
  const uint64_t k = semantic ? lay.nwin : L / W;
  const bool want = semantic || v.F < 8 || (v.F > 8 && v.r > 0);
  *nwin = k;
  if (k > cap)
    return fail(BUFF_E_SMALL_BUFFER, "%llu windows, room for %llu", (unsigned long long)k, (unsigned long long)cap);
  if (k && !vals) return fail(BUFF_E_INVALID, "null output");
  const uint64_t nwords = (ch->n + 63) / 64;
  if (lay.ntiles == 0) {
    if (argmax && nwords) HIP_TRY(launch_fill_bitmap(argmax, nwords, ch->n, 0, ctx->stream));
    return BUFF_OK;
  }
  BUFF_TRY(ctx->scratch(ctx->s_wmax, lay.bytes));
  uint8_t* scr = (uint8_t*)ctx->s_wmax.ptr;
  ProfScope ps(ctx, "k_wmax");
  HIP_TRY(launch_wmax(v, s, e, W, lay, scr, ctx->stream));
  if (argmax) HIP_TRY(launch_wmax_argmax(v, s, e, W, want, scr, lay, argmax, ctx->stream));
  HIP_TRY(launch_wmax_values(scr, lay, k, v.base, std::ldexp(1.0, -(int)ch->dlen), vals, ctx->stream));
  return BUFF_OK;
}

int max_range_host(Ctx* ctx, const Chunk* ch, uint64_t s, uint64_t e, uint64_t W, uint32_t flags, double* vals,
                   uint64_t cap, uint64_t* nwin, uint32_t* argmax_words, uint64_t nwords32) {
  if (!ctx || !ch || !nwin) return fail(BUFF_E_INVALID, "null argument");
  const uint64_t need32 = (ch->n + 31) / 32, nwords64 = (ch->n + 63) / 64;
  if (argmax_words && nwords32 < need32)
    return fail(BUFF_E_SMALL_BUFFER, "bitmap needs %llu u32 words", (unsigned long long)need32);
  if (W == 0 || s > e || e > ch->n)
    return fail(BUFF_E_INVALID, "window %llu over rows [%llu, %llu) of %llu", (unsigned long long)W,
                (unsigned long long)s, (unsigned long long)e, (unsigned long long)ch->n);
  const uint64_t nall = e > s ? (e - s + W - 1) / W : 0;
  BUFF_TRY(ctx->scratch(ctx->s_values, std::max<uint64_t>(nall, 1) * 8));
  if (argmax_words && nwords64) BUFF_TRY(ctx->scratch(ctx->s_bitmap, nwords64 * 8));
  BUFF_TRY(max_range_device(ctx, ch, s, e, W, flags, (double*)ctx->s_values.ptr, nall, nwin,
                            argmax_words && nwords64 ? (uint64_t*)ctx->s_bitmap.ptr : nullptr));
  if (*nwin > cap)
    return fail(BUFF_E_SMALL_BUFFER, "%llu windows, room for %llu", (unsigned long long)*nwin, (unsigned long long)cap);
  if (*nwin) {
    if (!vals) return fail(BUFF_E_INVALID, "null output");
    HIP_TRY(hipMemcpyAsync(vals, ctx->s_values.ptr, *nwin * 8, hipMemcpyDeviceToHost, ctx->stream));
  }
  std::vector<uint64_t> host(argmax_words ? nwords64 : 0);
  if (argmax_words && nwords64)
    HIP_TRY(hipMemcpyAsync(host.data(), ctx->s_bitmap.ptr, nwords64 * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (argmax_words && need32) memcpy(argmax_words, host.data(), need32 * 4);
  return BUFF_OK;
}

// split_double.rs:2127-2208 fold of the per-sub-column sums.
double reference_sum_fold(const Chunk* ch, const uint64_t* ps) {
  const uint32_t len = (uint32_t)ch->n, ilen = ch->ilen, dlen = ch->dlen;
  uint32_t remain = ilen + dlen;
  int32_t processed = 0;
  double sum_base = (double)len * (double)ch->view.base;
  double sum = 0.0;
  sum_base = sum_base / powi2((int32_t)dlen);
  uint32_t k = 0;
  while (remain >= 8) {
    remain -= 8;
    processed += 8;
    const double dec_scl = powi2((int32_t)((uint32_t)processed - ilen));
    sum = sum + (double)ps[k++] / dec_scl;
    if (remain == 0) return sum;   // returns before adding the base term
  }
  if (remain > 0) {
    processed += (int32_t)remain;
    const double dec_scl = powi2((int32_t)((uint32_t)processed - ilen));
    sum = sum + (double)ps[k] / dec_scl;
  }
  sum += sum_base;
  return sum;
}

int sum_op(Ctx* ctx, const Chunk* ch, double* ref_sum, double* exact_sum, uint64_t* plane_sums) {
  if (!ctx || !ch) return fail(BUFF_E_INVALID, "null argument");
  if (ch->flags & BUFF_RESIDUE) return residue_sum_op(ctx, ch, -1, 0, ref_sum, exact_sum, plane_sums);
  if (ch->S > 0) return major_sum(ctx, ch, 0, ref_sum, exact_sum, plane_sums);
  uint64_t ps[kMaxPlanes + 1] = {0};
  const ChunkView& v = ch->view;
  if (ch->n > 0 && (v.P || v.r)) {
    // one launch into a slot set ([k][slot] u64), folded as it is handed over
    const LaunchShape ls = ctx->shape(Op::Sum, v);
    unsigned long long *set = nullptr, *next = nullptr;
    BUFF_TRY(ctx->sum_slots(&set, &next));
    {
      ProfScope sc(ctx, "k_sum");
      HIP_TRY(launch_sum_slots(v, set, next, ctx->stream, ls));
      ctx->sum_slots_used();
    }
    Publish pub;
    BUFF_TRY(ctx->publish(set, kPublishSumSlots, v.P + 1, &pub));
    BUFF_TRY(ctx->published(pub, ps, v.P + 1));
  }
  if (plane_sums) memcpy(plane_sums, ps, sizeof(ps));
  if (ref_sum) *ref_sum = reference_sum_fold(ch, ps);
  if (exact_sum) {
    __int128 tot = (__int128)ch->n * (__int128)v.base;
    for (uint32_t p = 0; p < v.P; ++p) tot += (__int128)ps[p] << (v.F - 8 * (p + 1));
    tot += (__int128)ps[v.P];
    *exact_sum = (double)tot / powi2((int32_t)ch->dlen);
  }
  return BUFF_OK;
}

int max_op(Ctx* ctx, const Chunk* ch, uint32_t flags, double* value, uint64_t* max_off, uint64_t* argmax) {
  if (!ctx || !ch) return fail(BUFF_E_INVALID, "null argument");
  if (argmax) *argmax = 0;
  if ((ch->flags & BUFF_RESIDUE) && !(flags & BUFF_SEMANTIC))
    return fail(BUFF_E_UNSUPPORTED, "the reference has no max of a residue stream (BUFF_SEMANTIC answers it)");
  // buff-major chunks report the tail-stage ties (buff_simd.rs:2856); a chunk
  // without sparse columns runs the same stages over its planes
  if (ch->S > 0 || ((ch->flags & BUFF_MAJORITY) && argmax && ch->n > 0))
    return major_max(ctx, ch, flags, value, max_off, argmax);
  uint64_t best = 0;
  const ChunkView& v = ch->view;
  if (ch->n > 0 && (v.P || v.r)) {
    // [best u64 | pad to 512 | per-tile maxima]; the kernels clear best
    const uint64_t ms = max_scratch_bytes(v.ntiles);
    BUFF_TRY(ctx->scratch(ctx->s_best, 512 + ms));
    uint8_t* sb = (uint8_t*)ctx->s_best.ptr;
    unsigned int* slots = nullptr;
    uint32_t tag = 0;
    BUFF_TRY(ctx->max_slots(&slots, &tag));
    {
      ProfScope sc(ctx, "k_max");
      HIP_TRY(launch_max(v, (unsigned long long*)sb, sb + 512, slots, tag, ctx->stream, ctx->shape(Op::Max, v)));
    }
    Publish pub;
    BUFF_TRY(ctx->publish(sb, kPublishCopy, 1, &pub));
    BUFF_TRY(ctx->published(pub, &best, 1));
  }
  if (max_off) *max_off = best;
  // split_double.rs:4102: (max as i64 + base_int) as f64 / 2^dlen
  if (value) *value = (double)(int64_t)(best + (uint64_t)v.base) / powi2((int32_t)ch->dlen);
  return BUFF_OK;
}

// Reduced-precision read rule of split_double.rs:1642-1664 (u32 arithmetic).
int precision_bits(const Chunk* ch, int32_t precision, uint32_t* R) {
  const uint32_t ilen = ch->ilen, dlen = ch->dlen;
  if (precision == 0) {
    uint32_t remain = ilen;
    uint32_t byte_needed = ilen / 8;
    if (ilen % 8 > 0) {
      byte_needed += 1;
      remain = byte_needed * 8;
    }
    *R = remain;
    return BUFF_OK;
  }
  const int bn = precision_map(precision);
  if (bn < 0) return fail(BUFF_E_REF_PANIC, "precision %d is not in PRECISION_MAP", precision);
  const uint32_t bits_needed = (uint32_t)bn;
  if (!(dlen >= bits_needed)) return fail(BUFF_E_REF_PANIC, "requested %d decimals (%u bits) but dlen is %u",
                                          precision, bits_needed, dlen);
  uint32_t remain = ilen + bits_needed;
  const uint32_t fixed_byte = (dlen + ilen) / 8;
  uint32_t byte_needed = (ilen + bits_needed) / 8;
  const uint32_t extra_bits = (ilen + bits_needed) % 8;
  if (byte_needed < fixed_byte && extra_bits > 0) {
    byte_needed += 1;
    remain = byte_needed * 8;
  }
  *R = remain;
  return BUFF_OK;
}

// split_double.rs:2463-2626 fold of byte_residue_sum_with_precision over the
// sums it reads: one term per whole sub-column (returning without the base term
// when the R-bit read ends on a byte boundary), then the R % 8-bit fields.
double precision_sum_fold(const Chunk* ch, uint32_t R, const uint64_t* sums) {
  const uint32_t ilen = ch->ilen;
  uint32_t remain = R;
  int32_t processed = 0;
  double sum_base = (double)(uint32_t)ch->n * (double)ch->view.base;
  double sum = 0.0;
  sum_base = sum_base / powi2((int32_t)ch->dlen);
  uint32_t k = 0;
  while (remain >= 8) {
    remain -= 8;
    processed += 8;
    sum = sum + (double)sums[k++] / powi2((int32_t)((uint32_t)processed - ilen));
    if (remain == 0) return sum;
  }
  if (remain > 0) {
    processed += (int32_t)remain;
    sum = sum + (double)sums[k] / powi2((int32_t)((uint32_t)processed - ilen));
  }
  sum += sum_base;
  return sum;
}

// byte_residue_sum_with_precision (split_double.rs:2409-2629) on a plain chunk
// (the stream branch of ilen <= 15, 2576-2624). The R-bit read rule selects
// R / 8 whole sub-columns, summed by k_sum over a view of just those planes,
// and R % 8 = q bits per row: the whole tail when q equals the stored tail
// width (summed in the same launch), else k_tail_fields (the fields the
// reference reads from the packed tail, and the rows' top q tail bits).
// ref_sum replays the reference's fold; exact_sum is the exact sum of the
// values truncated to their top min(R, F) stored bits, rounded once. With
// BUFF_SEMANTIC, ref_sum = exact_sum and the reference's panics and ilen > 15
// branch do not apply (more decimals than stored: every bit).
int sum_precision_op(Ctx* ctx, const Chunk* ch, int32_t precision, uint32_t flags, double* ref_sum,
                     double* exact_sum, uint64_t* sums_out) {
  if (!ctx || !ch) return fail(BUFF_E_INVALID, "null argument");
  if (precision < 0) return fail(BUFF_E_INVALID, "precision %d < 0", precision);
  if (ch->flags & BUFF_RESIDUE) return residue_sum_op(ctx, ch, precision, flags, ref_sum, exact_sum, sums_out);
  if (ch->flags & (BUFF_MAJORITY | BUFF_SLICE))
    return fail(BUFF_E_UNSUPPORTED, "the reference has no reduced-precision sum of a %s stream",
                (ch->flags & BUFF_SLICE) ? "byte-slice" : "buff-major");
  const bool semantic = (flags & BUFF_SEMANTIC) != 0;
  const ChunkView& v = ch->view;
  uint32_t R = 0;
  int rc = precision_bits(ch, precision, &R);
  if (rc == BUFF_E_REF_PANIC && semantic) {
    R = v.F;
  } else if (rc != BUFF_OK) {
    return rc;
  }
  if (!semantic) {
    if (ch->ilen > 15)
      return fail(BUFF_E_UNSUPPORTED, "ilen %u > 15: the reference parses an outlier section here", ch->ilen);
    if (R / 8 > v.P) {
      if (kStreamHeader + (uint64_t)(R / 8) * ch->n > stream_bytes_of(ch->n, v.P, v.r))
        return fail(BUFF_E_REF_PANIC, "the reference reads %u sub-columns of a %u-plane stream", R / 8, v.P);
      return fail(BUFF_E_UNSUPPORTED, "the reference would read %u whole bytes of a %u-plane stream", R / 8, v.P);
    }
  }
  const uint32_t Rs = std::min(R, v.F);   // bits the exact (truncated) sum keeps
  // R % 8 != 0 only when R / 8 == P (R <= F): the q bits come from the tail.
  const uint32_t q = semantic ? Rs % 8 : R % 8;
  const uint32_t nplanes = std::min<uint32_t>((semantic ? Rs : R) / 8, v.P);
  uint64_t ps[kMaxPlanes + 1] = {0};
  uint64_t fields[2] = {0, 0};   // reference fields, exact top-q tail bits
  if (ch->n > 0) {
    ChunkView pv = v;
    pv.P = nplanes;
    pv.r = (q > 0 && q == v.r) ? v.r : 0;
    if (pv.P || pv.r) {
      const LaunchShape ls = ctx->shape(Op::Sum, pv);
      BUFF_TRY(ctx->scratch(ctx->s_sums, sizeof(ps) + sum_scratch_bytes(ls)));
      ProfScope sc(ctx, "k_sum");
      HIP_TRY(launch_sum(pv, (unsigned long long*)ctx->s_sums.ptr, ctx->stream, ls));
    }
    if (q > 0 && q < v.r) {
      BUFF_TRY(ctx->scratch(ctx->s_best, 256));
      HIP_TRY(hipMemsetAsync(ctx->s_best.ptr, 0, 16, ctx->stream));
      ProfScope sc(ctx, "k_tail_fields");
      HIP_TRY(launch_tail_fields(v, q, (unsigned long long*)ctx->s_best.ptr, ctx->stream));
    }
    if (pv.P || pv.r) BUFF_TRY(d2h_small(ctx, ps, ctx->s_sums.ptr, sizeof(ps)));
    if (q > 0 && q < v.r) BUFF_TRY(d2h_small(ctx, fields, ctx->s_best.ptr, sizeof(fields)));
  }
  // sums in read order: the whole sub-columns, then the q-bit part
  uint64_t sums[kMaxPlanes + 1] = {0};
  for (uint32_t k = 0; k < nplanes; ++k) sums[k] = ps[k];
  uint64_t top_q = 0;
  if (q > 0) {
    sums[nplanes] = (q == v.r) ? ps[nplanes] : fields[0];
    top_q = (q == v.r) ? ps[nplanes] : fields[1];
  }
  if (sums_out) memcpy(sums_out, sums, sizeof(sums));
  __int128 tot = (__int128)ch->n * (__int128)v.base;
  for (uint32_t k = 0; k < std::min<uint32_t>(Rs / 8, v.P); ++k) tot += (__int128)ps[k] << (v.F - 8 * (k + 1));
  if (Rs % 8) tot += (__int128)top_q << (v.F - Rs);
  const double exact = (double)tot / powi2((int32_t)ch->dlen);
  if (exact_sum) *exact_sum = exact;
  if (ref_sum) *ref_sum = semantic ? exact : precision_sum_fold(ch, R, sums);
  return BUFF_OK;
}

// The materialization recipe of a plain chunk: full precision
// (split_double.rs:1243-1350), the reference's reduced-precision read
// (1612-1893) or, with BUFF_SEMANTIC, the top R stored bits re-scaled.
int mat_plan(const Chunk* ch, int32_t precision, uint32_t flags, MatParams* out) {
  const ChunkView& v = ch->view;
  MatParams mp;
  memset(&mp, 0, sizeof(mp));
  mp.inv_scale = std::ldexp(1.0, -(int)ch->dlen);
  mp.mode = kMatFull;
  mp.planes = v.P;
  mp.tail_bits = v.r;
  const bool semantic = (flags & BUFF_SEMANTIC) != 0;
  if (precision >= 0) {
    uint32_t R = 0;
    int rc = precision_bits(ch, precision, &R);
    if (rc == BUFF_E_REF_PANIC && semantic) {
      R = v.F;  // more decimals than stored: everything
    } else if (rc != BUFF_OK) {
      return rc;
    }
    if (!semantic) {
      if (ch->ilen > 15 && !(ch->flags & BUFF_RESIDUE))
        return fail(BUFF_E_UNSUPPORTED, "ilen %u > 15: the reference parses an outlier section here", ch->ilen);
      if (R / 8 > v.P) {
        // The reference reads R/8 whole n-byte sub-columns: past the end of the
        // stream (a panic) unless the column is tiny.
        if (kStreamHeader + (uint64_t)(R / 8) * ch->n > stream_bytes_of(ch->n, v.P, v.r))
          return fail(BUFF_E_REF_PANIC, "the reference reads %u sub-columns of a %u-plane stream", R / 8, v.P);
        return fail(BUFF_E_UNSUPPORTED, "the reference would read %u whole bytes of a %u-plane stream", R / 8, v.P);
      }
      mp.mode = kMatRefPrecision;
      mp.planes = R / 8;
      mp.tail_bits = R % 8;
    } else if (R < v.F) {
      mp.mode = kMatSemantic;
      if (R <= 8 * v.P) {
        mp.planes = (R + 7) / 8;
        mp.tail_bits = 0;
        mp.drop = 8 * mp.planes - R;
      } else {
        mp.planes = v.P;
        mp.tail_bits = R - 8 * v.P;
        mp.drop = 0;
      }
      mp.shift_back = v.F - R;
    }
  }
  *out = mp;
  return BUFF_OK;
}

int materialize_device(Ctx* ctx, const Chunk* ch, int32_t precision, uint32_t flags, double* out) {
  if (!ctx || !ch || (!out && ch->n)) return fail(BUFF_E_INVALID, "null argument");
  if (ch->flags & BUFF_SLICE) {
    BUFF_TRY(slice_materialize_check(ch, precision, flags));
    precision = -1;
  }
  if ((ch->flags & BUFF_RESIDUE) && ch->S > 0 && !(flags & BUFF_SEMANTIC))
    return residue_materialize_device(ctx, ch, precision, out, nullptr);
  if (ch->flags & BUFF_MAJORITY) {
    if (precision >= 0 && !(flags & BUFF_SEMANTIC) && !(ch->flags & BUFF_RESIDUE))
      return fail(BUFF_E_UNSUPPORTED, "the reference has no reduced-precision decode of a buff-major stream");
    if (ch->S > 0) return major_materialize_device(ctx, ch, precision, flags, out, nullptr);
  }
  MatParams mp;
  BUFF_TRY(mat_plan(ch, precision, flags, &mp));
  if (ch->n == 0) return BUFF_OK;
  ProfScope sc(ctx, "k_materialize");
  HIP_TRY(launch_materialize(ch->view, mp, out, ctx->stream, ctx->shape(Op::Materialize, ch->view)));
  return BUFF_OK;
}

int materialize_host(Ctx* ctx, const Chunk* ch, int32_t precision, uint32_t flags, double* out) {
  if (!ctx || !ch) return fail(BUFF_E_INVALID, "null argument");
  if (ch->n == 0) return BUFF_OK;
  if (!out) return fail(BUFF_E_INVALID, "null output");
  BUFF_TRY(ctx->scratch(ctx->s_values, ch->n * 8));
  BUFF_TRY(materialize_device(ctx, ch, precision, flags, (double*)ctx->s_values.ptr));
  const uint64_t rows = ((ch->flags & BUFF_RESIDUE) && !(flags & BUFF_SEMANTIC)) ? residue_decoded_rows(ch, precision)
                        : (ch->flags & BUFF_MAJORITY) ? major_decoded_rows(ch, flags)
                                                      : ch->n;
  if (rows) HIP_TRY(hipMemcpyAsync(out, ctx->s_values.ptr, rows * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return BUFF_OK;
}

// ---------------------------------------------------------------------------
// process-wide default context for the reference-shaped entry points

static std::mutex g_default_mu;
static Ctx* g_default = nullptr;

int default_ctx(Ctx** out) {
  std::lock_guard<std::mutex> lk(g_default_mu);
  if (!g_default) {
    buff_ctx* c = nullptr;
    int rc = buff_ctx_create(0, 0, &c);
    if (rc != BUFF_OK) return rc;
    g_default = reinterpret_cast<Ctx*>(c);
  }
  *out = g_default;
  return BUFF_OK;
}

}  // namespace buffgpu

using namespace buffgpu;

// ===========================================================================
// C ABI

namespace buffgpu {
namespace {

// Kernel execution time of one launch sequence (dispatch-level events).
struct OpTimer {
  hipEvent_t a = nullptr, b = nullptr;
  ~OpTimer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  int init() {
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    return BUFF_OK;
  }
  template <typename F>
  int time(Ctx* ctx, F&& launch, float* us) {
    LaunchEvents& le = launch_events();
    le.start = a;
    le.stop = b;
    le.started = false;
    hipError_t e = launch();
    const bool started = le.started;
    le = LaunchEvents{};
    if (e != hipSuccess) return fail(BUFF_E_DEVICE, "tuning launch: %s", hipGetErrorString(e));
    if (!started) return fail(BUFF_E_DEVICE, "tuning launch issued no kernel");
    HIP_TRY(hipEventSynchronize(b));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, a, b));
    *us = ms * 1e3f;
    (void)ctx;
    return BUFF_OK;
  }
};

// Times every candidate shape of the range filter, sum, max and full
// materialization on this chunk (best of three runs each, the memory-side cache
// evicted before every run) and keeps the fastest per operation for chunks of
// the same size class and byte planes.
int tune_chunk(Ctx* ctx, const Chunk* ch, double pred, buff_tune_choice* out) {
  if (ch->S > 0 || (ch->flags & BUFF_MAJORITY))
    return fail(BUFF_E_UNSUPPORTED, "launch-shape tuning covers byte_fixed chunks");
  const ChunkView& v = ch->view;
  if (ch->n == 0 || (!v.P && !v.r)) return fail(BUFF_E_INVALID, "nothing to tune on an empty chunk");
  FilterTarget tg{};
  int fill = 0;
  if (!filter_plan(ch, BUFF_OP_GT, pred, 0, &tg, &fill))
    return fail(BUFF_E_INVALID, "the predicate needs no scan on this chunk (all or no rows qualify)");
  MatParams mp{};
  BUFF_TRY(mat_plan(ch, -1, 0, &mp));
  const uint64_t nwords = (ch->n + 63) / 64;
  const uint64_t evict = 1ull << 30;
  const LaunchShape probe{4, 0, 0};
  BUFF_TRY(ctx->scratch(ctx->s_bitmap, nwords * 8));
  BUFF_TRY(ctx->scratch(ctx->s_sums, 9 * 8 + sum_scratch_bytes(probe)));
  BUFF_TRY(ctx->scratch(ctx->s_best, 512 + max_scratch_bytes(v.ntiles)));
  BUFF_TRY(ctx->scratch(ctx->s_values, ch->n * 8));
  BUFF_TRY(ctx->scratch(ctx->s_sink, 256));
  void* flush = nullptr;
  BUFF_TRY(ctx->alloc(evict, &flush));
  OpTimer tm;
  int rc = tm.init();
  const Op ops[4] = {Op::FilterGt, Op::Sum, Op::Max, Op::Materialize};
  const uint32_t tiles_opts[4][3] = {{2, 4, 8}, {1, 2, 4}, {4, 8, 16}, {0, 0, 0}};
  for (int oi = 0; oi < 4 && rc == BUFF_OK; ++oi) {
    const Op op = ops[oi];
    // sums and maxima: every shape's result (handed over as the operations
    // do) must equal the first (default shape) run's
    const uint32_t nres = op == Op::Sum ? v.P + 1 : (op == Op::Max ? 1u : 0u);
    const void* res = nullptr;   // where the last run left its result
    uint64_t want[kMaxPlanes + 1], got[kMaxPlanes + 1];
    bool have_want = false;
    auto run = [&](const LaunchShape& ls) -> hipError_t {
      switch (op) {
        case Op::FilterGt:
          return launch_filter(v, 0, tg, (uint64_t*)ctx->s_bitmap.ptr, nwords, nullptr, ctx->stream, ls);
        case Op::Sum: {
          unsigned long long *set = nullptr, *next = nullptr;
          if (ctx->sum_slots(&set, &next) != BUFF_OK) return hipErrorUnknown;
          const hipError_t e = launch_sum_slots(v, set, next, ctx->stream, ls);
          if (e == hipSuccess) ctx->sum_slots_used();
          res = set;
          return e;
        }
        case Op::Max: {
          uint8_t* sb = (uint8_t*)ctx->s_best.ptr;
          unsigned int* slots = nullptr;
          uint32_t tag = 0;
          if (ctx->max_slots(&slots, &tag) != BUFF_OK) return hipErrorUnknown;
          res = sb;
          return launch_max(v, (unsigned long long*)sb, sb + 512, slots, tag, ctx->stream, ls);
        }
        default:
          return launch_materialize(v, mp, (double*)ctx->s_values.ptr, ctx->stream, ls);
      }
    };
    auto measure = [&](const LaunchShape& ls, float* best_us) -> int {
      *best_us = 1e30f;
      for (int rep = 0; rep < 4; ++rep) {   // first run warms code and descriptors
        HIP_TRY(launch_touch_bytes(flush, evict, (uint32_t*)ctx->s_sink.ptr, ctx->stream));
        float us = 0;
        BUFF_TRY(tm.time(ctx, [&] { return run(ls); }, &us));
        if (rep > 0) *best_us = std::min(*best_us, us);
        if (nres) {
          Publish pub;
          BUFF_TRY(ctx->publish(res, op == Op::Sum ? kPublishSumSlots : kPublishCopy, nres, &pub));
          BUFF_TRY(ctx->published(pub, have_want ? got : want, nres));
          if (have_want && memcmp(got, want, 8ull * nres) != 0)
            return fail(BUFF_E_DEVICE, "tuning: launch shape (%u waves, %u workgroups, %u tiles, form %u) changed "
                        "the %s", ls.waves_per_wg, ls.max_wg, ls.tiles_per_wave, ls.prefetch,
                        op == Op::Sum ? "sum" : "max");
          have_want = true;
        }
      }
      return BUFF_OK;
    };
    buff_tune_choice ch_out{};
    ch_out.op = (uint32_t)op;
    LaunchShape best{4, 0, 0};
    float best_us = 0;
    rc = measure(best, &best_us);
    ch_out.default_us = best_us;
    for (uint32_t waves : {2u, 4u}) {
      for (uint32_t cap : {0u, 1024u, 2048u, 4096u, 8192u, 16384u}) {
        for (uint32_t t : tiles_opts[oi]) {
          for (uint32_t pf = 0; pf < ((op == Op::FilterGt || op == Op::Max) ? 2u : 1u); ++pf) {
            if (rc != BUFF_OK) break;
            if (op == Op::Materialize && waves != 4) continue;   // its LDS staging assumes four waves
            const LaunchShape ls{waves, cap, t, pf};
            float us = 0;
            rc = measure(ls, &us);
            if (rc == BUFF_OK && us < best_us) {
              best_us = us;
              best = ls;
            }
          }
          if (op == Op::Materialize) break;
        }
      }
    }
    if (rc != BUFF_OK) break;
    ctx->tuned[tune_key(op, v.ntiles, v.P)] = best;
    if (op == Op::FilterGt) ctx->tuned[tune_key(Op::FilterEq, v.ntiles, v.P)] = best;
    ch_out.waves_per_wg = best.waves_per_wg;
    ch_out.max_wg = best.max_wg;
    ch_out.tiles_per_wave = best.tiles_per_wave;
    ch_out.prefetch = best.prefetch;
    ch_out.best_us = best_us;
    if (out) out[oi] = ch_out;
  }
  (void)hipStreamSynchronize(ctx->stream);
  (void)ctx->release(flush);
  return rc;
}

}  // namespace

}  // namespace buffgpu

extern "C" {

const char* buff_last_error(void) { return g_last_error.c_str(); }

int buff_ctx_create(int device, uint64_t hbm_budget_bytes, buff_ctx** out) {
  if (!out) return fail(BUFF_E_INVALID, "null argument");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) return fail(BUFF_E_DEVICE, "no HIP device: %s", hipGetErrorString(e));
  if (device < 0 || device >= ndev) return fail(BUFF_E_INVALID, "device %d of %d", device, ndev);
  HIP_TRY(hipSetDevice(device));
  Ctx* c = new (std::nothrow) Ctx();
  if (!c) return fail(BUFF_E_OOM, "host allocation");
  c->device = device;
  c->budget = hbm_budget_bytes;
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc(&c->pinned, Ctx::kPinnedBytes, hipHostMallocDefault);
  // the result mailbox, made here so that no operation allocates to read a result back
  void* mb = nullptr;
  if (e == hipSuccess) e = hipHostMalloc(&mb, kPublishWords * 8, hipHostMallocMapped | hipHostMallocCoherent);
  if (e == hipSuccess) {
    c->mbox = static_cast<uint64_t*>(mb);
    memset(mb, 0, kPublishWords * 8);
    void* d = nullptr;
    e = hipHostGetDevicePointer(&d, mb, 0);
    c->mbox_dev = static_cast<unsigned long long*>(d);
  }
  if (e != hipSuccess) {
    delete c;
    return fail(BUFF_E_DEVICE, "context setup: %s", hipGetErrorString(e));
  }
  *out = reinterpret_cast<buff_ctx*>(c);
  return BUFF_OK;
}

int buff_ctx_destroy(buff_ctx* ctx) {
  if (!ctx) return BUFF_OK;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  {
    std::lock_guard<std::mutex> lk(g_default_mu);
    if (c == g_default) g_default = nullptr;
  }
  (void)hipSetDevice(c->device);
  delete c;
  return BUFF_OK;
}

int buff_ctx_synchronize(buff_ctx* ctx) {
  if (!ctx) return fail(BUFF_E_INVALID, "null argument");
  HIP_TRY(hipStreamSynchronize(reinterpret_cast<Ctx*>(ctx)->stream));
  return BUFF_OK;
}

int buff_ctx_hbm_used(buff_ctx* ctx, uint64_t* bytes) {
  if (!ctx || !bytes) return fail(BUFF_E_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  std::lock_guard<std::mutex> lk(c->mu);
  *bytes = c->held;
  return BUFF_OK;
}

int buff_dev_alloc(buff_ctx* ctx, uint64_t bytes, void** dptr) {
  if (!ctx || !dptr) return fail(BUFF_E_INVALID, "null argument");
  return reinterpret_cast<Ctx*>(ctx)->alloc(bytes, dptr);
}

int buff_dev_free(buff_ctx* ctx, void* dptr) {
  if (!ctx) return fail(BUFF_E_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  HIP_TRY(hipStreamSynchronize(c->stream));
  return c->release(dptr);
}

int buff_memcpy_h2d(buff_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return fail(BUFF_E_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return BUFF_OK;
}

int buff_memcpy_d2h(buff_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return fail(BUFF_E_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return BUFF_OK;
}

int buff_dev_fill(buff_ctx* ctx, void* dptr, uint8_t value, uint64_t bytes) {
  if (!ctx || (bytes && !dptr)) return fail(BUFF_E_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  ProfScope ps(c, "k_fill_bytes");
  HIP_TRY(launch_fill_bytes(dptr, bytes, value, c->stream));
  return BUFF_OK;
}

int buff_dev_copy(buff_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return fail(BUFF_E_INVALID, "null argument");
  if (bytes % 16 || (uintptr_t)dst % 16 || (uintptr_t)src % 16)
    return fail(BUFF_E_INVALID, "copy of %llu bytes: sizes and addresses must be multiples of 16",
                (unsigned long long)bytes);
  const uintptr_t d = (uintptr_t)dst, sp = (uintptr_t)src;
  if (bytes && d < sp + bytes && sp < d + bytes) return fail(BUFF_E_INVALID, "overlapping copy");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  ProfScope ps(c, "k_copy_bytes");
  HIP_TRY(launch_copy_bytes(dst, src, bytes, c->stream));
  return BUFF_OK;
}

int buff_dev_read(buff_ctx* ctx, const void* src, uint64_t bytes) {
  if (!ctx || (bytes && !src)) return fail(BUFF_E_INVALID, "null argument");
  if (bytes % 16 || (uintptr_t)src % 16)
    return fail(BUFF_E_INVALID, "read of %llu bytes: size and address must be multiples of 16",
                (unsigned long long)bytes);
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  BUFF_TRY(c->scratch(c->s_sink, 256));
  ProfScope ps(c, "k_read_bytes");
  HIP_TRY(launch_read_bytes(src, bytes, (uint32_t*)c->s_sink.ptr, c->stream));
  return BUFF_OK;
}

int buff_dev_touch(buff_ctx* ctx, const void* dptr, uint64_t bytes) {
  if (!ctx || (bytes && !dptr)) return fail(BUFF_E_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  BUFF_TRY(c->scratch(c->s_sink, 256));
  ProfScope ps(c, "k_touch_bytes");
  HIP_TRY(launch_touch_bytes(dptr, bytes, (uint32_t*)c->s_sink.ptr, c->stream));
  return BUFF_OK;
}

int buff_tune_chunk(buff_ctx* ctx, const buff_chunk* chunk, double pred, buff_tune_choice* choices) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  const Chunk* ch = reinterpret_cast<const Chunk*>(chunk);
  if (!c || !ch) return fail(BUFF_E_INVALID, "null argument");
  if (ch->ctx != c) return fail(BUFF_E_INVALID, "the chunk belongs to another context");
  return tune_chunk(c, ch, pred, choices);
}

int buff_prof_enable(buff_ctx* ctx, int on) {
  if (!ctx) return fail(BUFF_E_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  BUFF_TRY(c->prof_collect());
  c->prof_on = on != 0;
  return BUFF_OK;
}

int buff_prof_reset(buff_ctx* ctx) {
  if (!ctx) return fail(BUFF_E_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  BUFF_TRY(c->prof_collect());
  c->prof_stats.clear();
  return BUFF_OK;
}

int buff_prof_query(buff_ctx* ctx, const char* kernel, uint64_t* launches, double* total_ms) {
  if (!ctx || !kernel) return fail(BUFF_E_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  BUFF_TRY(c->prof_collect());
  auto it = c->prof_stats.find(kernel);
  if (launches) *launches = it == c->prof_stats.end() ? 0 : it->second.first;
  if (total_ms) *total_ms = it == c->prof_stats.end() ? 0.0 : it->second.second;
  return BUFF_OK;
}

int buff_encode_device(buff_ctx* ctx, const double* x_dev, uint64_t n, uint64_t scale, uint32_t flags,
                       buff_chunk** out) {
  Chunk* ch = nullptr;
  int rc = encode_device(reinterpret_cast<Ctx*>(ctx), x_dev, n, scale, flags, &ch);
  if (rc == BUFF_OK) *out = reinterpret_cast<buff_chunk*>(ch);
  return rc;
}

int buff_encode(buff_ctx* ctx, const double* x_host, uint64_t n, uint64_t scale, uint32_t flags, buff_chunk** out) {
  if (!ctx || !out || (n && !x_host)) return fail(BUFF_E_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  void* dx = nullptr;
  if (n) {
    BUFF_TRY(c->alloc(n * 8, &dx));
    hipError_t e = hipMemcpyAsync(dx, x_host, n * 8, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) {
      (void)c->release(dx);
      return fail(BUFF_E_DEVICE, "upload: %s", hipGetErrorString(e));
    }
  }
  int rc = buff_encode_device(ctx, (const double*)dx, n, scale, flags, out);
  if (dx) {
    (void)hipStreamSynchronize(c->stream);
    c->release(dx);
  }
  return rc;
}

int buff_chunk_info_get(const buff_chunk* chunk, buff_chunk_info* info) {
  if (!chunk || !info) return fail(BUFF_E_INVALID, "null argument");
  const Chunk* ch = reinterpret_cast<const Chunk*>(chunk);
  info->base = ch->view.base;
  info->n = (uint32_t)ch->n;
  info->ilen = ch->ilen;
  info->dlen = ch->dlen;
  info->width = ch->F;
  info->planes = ch->view.P;
  info->tail_bits = ch->view.r;
  info->flags = ch->flags;
  info->scale = ch->scale;
  info->stream_bytes = stream_bytes_of(ch->n, ch->view.P, ch->view.r);
  if (ch->flags & BUFF_MAJORITY) {
    // the snappy-compressed bitmaps are only sized by buff_chunk_serialize(dst = NULL)
    info->stream_bytes = ch->S ? 0 : (ch->F >= 8 ? info->stream_bytes + 1 : info->stream_bytes);
  }
  if (ch->flags & BUFF_SLICE) info->stream_bytes = slice_stream_bytes(ch->n, ch->F);
  info->device_bytes = ch->device_bytes;
  info->sparse_cols = ch->S;
  info->reserved = 0;
  info->decoded_rows = (ch->flags & BUFF_RESIDUE)  ? residue_decoded_rows(ch, -1)
                       : (ch->flags & BUFF_MAJORITY) ? major_decoded_rows(ch, 0)
                                                     : ch->n;
  return BUFF_OK;
}

int buff_chunk_sparse_col(const buff_chunk* chunk, uint32_t j, uint32_t* major, uint64_t* outliers,
                          uint64_t* last_row) {
  if (!chunk) return fail(BUFF_E_INVALID, "null argument");
  const Chunk* ch = reinterpret_cast<const Chunk*>(chunk);
  if (j >= ch->S) return fail(BUFF_E_INVALID, "chunk has %u sparse columns", ch->S);
  if (major) *major = ch->sp[j].major;
  if (outliers) *outliers = ch->sp[j].n_out;
  if (last_row) *last_row = ch->sp[j].last;
  return BUFF_OK;
}

int buff_chunk_serialize(buff_ctx* ctx, const buff_chunk* chunk, uint8_t* dst, uint64_t* len) {
  return chunk_serialize(reinterpret_cast<Ctx*>(ctx), reinterpret_cast<const Chunk*>(chunk), dst, len);
}

int buff_chunk_load(buff_ctx* ctx, const uint8_t* stream, uint64_t len, uint64_t scale, uint32_t flags,
                    buff_chunk** out) {
  Chunk* ch = nullptr;
  int rc = chunk_load(reinterpret_cast<Ctx*>(ctx), stream, len, scale, flags, &ch);
  if (rc == BUFF_OK) *out = reinterpret_cast<buff_chunk*>(ch);
  return rc;
}

int buff_chunk_free(buff_chunk* chunk) {
  if (!chunk) return BUFF_OK;
  Chunk* ch = reinterpret_cast<Chunk*>(chunk);
  if (ch->borrowed) return fail(BUFF_E_INVALID, "this chunk belongs to a batch: free the batch");
  Ctx* c = ch->ctx;
  (void)hipStreamSynchronize(c->stream);
  major_release(ch);
  int rc = c->release(ch->mem);
  delete ch;
  return rc;
}

int buff_filter(buff_ctx* ctx, const buff_chunk* chunk, int op, double pred, uint32_t flags, uint32_t* words,
                uint64_t nwords, uint64_t* count) {
  return filter_host(reinterpret_cast<Ctx*>(ctx), reinterpret_cast<const Chunk*>(chunk), op, pred, flags, words,
                     nwords, count);
}

int buff_filter_device(buff_ctx* ctx, const buff_chunk* chunk, int op, double pred, uint32_t flags,
                       uint64_t* words_dev, uint64_t* stats_dev) {
  return filter_device(reinterpret_cast<Ctx*>(ctx), reinterpret_cast<const Chunk*>(chunk), op, pred, flags,
                       words_dev, (unsigned long long*)stats_dev);
}

int buff_filter_cond(buff_ctx* ctx, const buff_chunk* chunk, int op, double pred, uint32_t flags,
                     const uint32_t* cond_words, uint32_t* words, uint64_t nwords, uint64_t* count) {
  return filter_cond_host(reinterpret_cast<Ctx*>(ctx), reinterpret_cast<const Chunk*>(chunk), op, pred, flags,
                          cond_words, words, nwords, count);
}

int buff_filter_cond_device(buff_ctx* ctx, const buff_chunk* chunk, int op, double pred, uint32_t flags,
                            const uint64_t* cond_dev, uint64_t* words_dev) {
  return filter_cond_device(reinterpret_cast<Ctx*>(ctx), reinterpret_cast<const Chunk*>(chunk), op, pred, flags,
                            cond_dev, words_dev);
}

int buff_materialize_cond(buff_ctx* ctx, const buff_chunk* chunk, uint32_t flags, const uint32_t* cond_words,
                          double* out, uint64_t cap, uint64_t* rows) {
  return materialize_cond_host(reinterpret_cast<Ctx*>(ctx), reinterpret_cast<const Chunk*>(chunk), flags,
                               cond_words, out, cap, rows);
}

int buff_materialize_cond_device(buff_ctx* ctx, const buff_chunk* chunk, uint32_t flags, const uint64_t* cond_dev,
                                 double* out_dev, uint64_t cap, uint64_t* rows) {
  return materialize_cond_device(reinterpret_cast<Ctx*>(ctx), reinterpret_cast<const Chunk*>(chunk), flags,
                                 cond_dev, out_dev, cap, rows);
}

int buff_max_range(buff_ctx* ctx, const buff_chunk* chunk, uint64_t s, uint64_t e, uint64_t window, uint32_t flags,
                   double* max_values, uint64_t cap, uint64_t* nwindows, uint32_t* argmax_words, uint64_t nwords) {
  return max_range_host(reinterpret_cast<Ctx*>(ctx), reinterpret_cast<const Chunk*>(chunk), s, e, window, flags,
                        max_values, cap, nwindows, argmax_words, nwords);
}

int buff_max_range_device(buff_ctx* ctx, const buff_chunk* chunk, uint64_t s, uint64_t e, uint64_t window,
                          uint32_t flags, double* max_values_dev, uint64_t cap, uint64_t* nwindows,
                          uint64_t* argmax_dev) {
  return max_range_device(reinterpret_cast<Ctx*>(ctx), reinterpret_cast<const Chunk*>(chunk), s, e, window, flags,
                          max_values_dev, cap, nwindows, argmax_dev);
}

int buff_sum_ex(buff_ctx* ctx, const buff_chunk* chunk, uint32_t flags, double* ref_sum, double* exact_sum,
                uint64_t* plane_sums) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  const Chunk* ch = reinterpret_cast<const Chunk*>(chunk);
  if (!c || !ch) return fail(BUFF_E_INVALID, "null argument");
  if (ch->flags & BUFF_SLICE) return slice_sum(c, ch, flags, ref_sum, exact_sum, plane_sums);
  return sum_op(c, ch, ref_sum, exact_sum, plane_sums);
}

int buff_sum(buff_ctx* ctx, const buff_chunk* chunk, double* ref_sum, double* exact_sum, uint64_t* plane_sums) {
  return buff_sum_ex(ctx, chunk, 0, ref_sum, exact_sum, plane_sums);
}

int buff_sum_precision(buff_ctx* ctx, const buff_chunk* chunk, int32_t precision, uint32_t flags, double* ref_sum,
                       double* exact_sum, uint64_t* sums) {
  return sum_precision_op(reinterpret_cast<Ctx*>(ctx), reinterpret_cast<const Chunk*>(chunk), precision, flags,
                          ref_sum, exact_sum, sums);
}

int buff_max_ex(buff_ctx* ctx, const buff_chunk* chunk, uint32_t flags, double* max_value, uint64_t* max_off,
                uint64_t* argmax_count) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  const Chunk* ch = reinterpret_cast<const Chunk*>(chunk);
  if (!c || !ch) return fail(BUFF_E_INVALID, "null argument");
  if (ch->flags & BUFF_SLICE) return slice_max(c, ch, flags, max_value, max_off, argmax_count);
  return max_op(c, ch, flags, max_value, max_off, argmax_count);
}

int buff_max(buff_ctx* ctx, const buff_chunk* chunk, double* max_value, uint64_t* max_off) {
  return buff_max_ex(ctx, chunk, 0, max_value, max_off, nullptr);
}

int buff_materialize(buff_ctx* ctx, const buff_chunk* chunk, int32_t precision, uint32_t flags, double* out) {
  return materialize_host(reinterpret_cast<Ctx*>(ctx), reinterpret_cast<const Chunk*>(chunk), precision, flags,
                          out);
}

int buff_materialize_device(buff_ctx* ctx, const buff_chunk* chunk, int32_t precision, uint32_t flags,
                            double* out_dev) {
  return materialize_device(reinterpret_cast<Ctx*>(ctx), reinterpret_cast<const Chunk*>(chunk), precision, flags,
                            out_dev);
}

// ---- reference-shaped entry points ---------------------------------------------

static int sd_encode(const buff_split_double* sd, const double* data, uint64_t n, uint32_t flags, uint8_t* out,
                     uint64_t cap, uint64_t* out_len) {
  if (!sd || !out_len) return fail(BUFF_E_INVALID, "null argument");
  Ctx* c = nullptr;
  BUFF_TRY(default_ctx(&c));
  buff_chunk* ch = nullptr;
  BUFF_TRY(buff_encode(reinterpret_cast<buff_ctx*>(c), data, n, sd->scale, flags, &ch));
  uint64_t len = cap;
  int rc = buff_chunk_serialize(reinterpret_cast<buff_ctx*>(c), ch, out, &len);
  *out_len = len;
  (void)buff_chunk_free(ch);
  return rc;
}

static int sd_load(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, uint32_t flags, Ctx** c,
                   buff_chunk** ch) {
  if (!sd || !bytes) return fail(BUFF_E_INVALID, "null argument");
  BUFF_TRY(default_ctx(c));
  return buff_chunk_load(reinterpret_cast<buff_ctx*>(*c), bytes, len, sd->scale, flags, ch);
}

int buff_sd_byte_fixed_encode(const buff_split_double* sd, const double* data, uint64_t n, uint8_t* out, uint64_t cap,
                              uint64_t* out_len) {
  return sd_encode(sd, data, n, 0, out, cap, out_len);
}

int buff_sd_buff_simd256_encode(const buff_split_double* sd, const double* data, uint64_t n, uint8_t* out,
                                uint64_t cap, uint64_t* out_len) {
  return sd_encode(sd, data, n, BUFF_FLIP, out, cap, out_len);
}

static int sd_decode(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, uint32_t flags, int32_t prec,
                     double* out, uint64_t cap) {
  Ctx* c = nullptr;
  buff_chunk* ch = nullptr;
  BUFF_TRY(sd_load(sd, bytes, len, flags, &c, &ch));
  const uint64_t n = reinterpret_cast<Chunk*>(ch)->n;
  int rc = BUFF_OK;
  if (cap < n) rc = fail(BUFF_E_SMALL_BUFFER, "output needs %llu doubles", (unsigned long long)n);
  if (rc == BUFF_OK) rc = buff_materialize(reinterpret_cast<buff_ctx*>(c), ch, prec, 0, out);
  (void)buff_chunk_free(ch);
  return rc;
}

int buff_sd_byte_fixed_decode(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* out,
                              uint64_t cap) {
  return sd_decode(sd, bytes, len, 0, -1, out, cap);
}

int buff_sd_buff_simd256_decode(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* out,
                                uint64_t cap) {
  return sd_decode(sd, bytes, len, BUFF_FLIP, -1, out, cap);
}

// a residue stream decoded: values written to out (cap doubles), their count to n_out
static int sd_residue_decode(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, int32_t prec,
                             double* out, uint64_t cap, uint64_t* n_out) {
  Ctx* c = nullptr;
  buff_chunk* ch = nullptr;
  BUFF_TRY(sd_load(sd, bytes, len, BUFF_RESIDUE, &c, &ch));
  const Chunk* k = reinterpret_cast<Chunk*>(ch);
  const uint64_t rows = residue_decoded_rows(k, prec);
  int rc = BUFF_OK;
  if (cap < rows) rc = fail(BUFF_E_SMALL_BUFFER, "output needs %llu doubles", (unsigned long long)rows);
  if (rc == BUFF_OK) rc = buff_materialize(reinterpret_cast<buff_ctx*>(c), ch, prec, 0, out);
  if (rc == BUFF_OK && n_out) *n_out = rows;
  (void)buff_chunk_free(ch);
  return rc;
}

int buff_sd_byte_residue_encode(const buff_split_double* sd, const double* data, uint64_t n, uint8_t* out,
                                uint64_t cap, uint64_t* out_len) {
  return sd_encode(sd, data, n, BUFF_RESIDUE, out, cap, out_len);
}

int buff_sd_byte_residue_decode(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* out,
                                uint64_t cap, uint64_t* n_out) {
  return sd_residue_decode(sd, bytes, len, -1, out, cap, n_out);
}

int buff_sd_byte_residue_decode_with_precision(const buff_split_double* sd, const uint8_t* bytes, uint64_t len,
                                               uint32_t precision, double* out, uint64_t cap, uint64_t* n_out) {
  if (precision > 0x7FFFFFFFu) return fail(BUFF_E_REF_PANIC, "precision %u is not in PRECISION_MAP", precision);
  return sd_residue_decode(sd, bytes, len, (int32_t)precision, out, cap, n_out);
}

int buff_sd_byte_residue_sum(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* s) {
  Ctx* c = nullptr;
  buff_chunk* ch = nullptr;
  BUFF_TRY(sd_load(sd, bytes, len, BUFF_RESIDUE, &c, &ch));
  int rc = buff_sum(reinterpret_cast<buff_ctx*>(c), ch, s, nullptr, nullptr);
  (void)buff_chunk_free(ch);
  return rc;
}

static int sd_filter(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, uint32_t flags, int op,
                     double pred, uint32_t* words, uint64_t nwords, uint64_t* count) {
  Ctx* c = nullptr;
  buff_chunk* ch = nullptr;
  BUFF_TRY(sd_load(sd, bytes, len, flags, &c, &ch));
  int rc = buff_filter(reinterpret_cast<buff_ctx*>(c), ch, op, pred, 0, words, nwords, count);
  (void)buff_chunk_free(ch);
  return rc;
}

int buff_sd_buff_range_filter(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double pred,
                              uint32_t* words, uint64_t nwords, uint64_t* count) {
  return sd_filter(sd, bytes, len, 0, BUFF_OP_GT, pred, words, nwords, count);
}

int buff_sd_buff_equal_filter(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double pred,
                              uint32_t* words, uint64_t nwords, uint64_t* count) {
  return sd_filter(sd, bytes, len, 0, BUFF_OP_EQ, pred, words, nwords, count);
}

int buff_sd_buff_simd_range_filter(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double pred,
                                   uint32_t* words, uint64_t nwords, uint64_t* count) {
  return sd_filter(sd, bytes, len, BUFF_FLIP, BUFF_OP_GT, pred, words, nwords, count);
}

int buff_sd_buff_simd_equal_filter(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double pred,
                                   uint32_t* words, uint64_t nwords, uint64_t* count) {
  return sd_filter(sd, bytes, len, BUFF_FLIP, BUFF_OP_EQ, pred, words, nwords, count);
}

int buff_sd_byte_fixed_sum(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* s) {
  Ctx* c = nullptr;
  buff_chunk* ch = nullptr;
  BUFF_TRY(sd_load(sd, bytes, len, 0, &c, &ch));
  int rc = buff_sum(reinterpret_cast<buff_ctx*>(c), ch, s, nullptr, nullptr);
  (void)buff_chunk_free(ch);
  return rc;
}

int buff_sd_byte_residue_sum_with_precision(const buff_split_double* sd, const uint8_t* bytes, uint64_t len,
                                            uint32_t precision, double* s) {
  if (precision > 0x7FFFFFFFu) return fail(BUFF_E_REF_PANIC, "precision %u is not in PRECISION_MAP", precision);
  Ctx* c = nullptr;
  buff_chunk* ch = nullptr;
  BUFF_TRY(sd_load(sd, bytes, len, BUFF_RESIDUE, &c, &ch));
  int rc = buff_sum_precision(reinterpret_cast<buff_ctx*>(c), ch, (int32_t)precision, 0, s, nullptr, nullptr);
  (void)buff_chunk_free(ch);
  return rc;
}

int buff_sd_byte_fixed_max(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* max_value) {
  Ctx* c = nullptr;
  buff_chunk* ch = nullptr;
  BUFF_TRY(sd_load(sd, bytes, len, 0, &c, &ch));
  int rc = buff_max(reinterpret_cast<buff_ctx*>(c), ch, max_value, nullptr);
  (void)buff_chunk_free(ch);
  return rc;
}

// ---- buff-major (buff_simd.rs:317-2861) ----------------------------------------

int buff_sd_buff_encode_majority(const buff_split_double* sd, const double* data, uint64_t n, uint8_t* out,
                                 uint64_t cap, uint64_t* out_len) {
  return sd_encode(sd, data, n, BUFF_MAJORITY, out, cap, out_len);
}

int buff_sd_buff_decode_majority(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* out,
                                 uint64_t cap, uint64_t* n_out) {
  if (!n_out) return fail(BUFF_E_INVALID, "null argument");
  Ctx* c = nullptr;
  buff_chunk* ch = nullptr;
  BUFF_TRY(sd_load(sd, bytes, len, BUFF_MAJORITY, &c, &ch));
  const uint64_t rows = major_decoded_rows(reinterpret_cast<Chunk*>(ch), 0);
  int rc = BUFF_OK;
  if (cap < rows) rc = fail(BUFF_E_SMALL_BUFFER, "output needs %llu doubles", (unsigned long long)rows);
  if (rc == BUFF_OK) rc = buff_materialize(reinterpret_cast<buff_ctx*>(c), ch, -1, 0, out);
  *n_out = rc == BUFF_OK ? rows : 0;
  (void)buff_chunk_free(ch);
  return rc;
}

int buff_sd_buff_range_filter_majority(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double pred,
                                       uint32_t* words, uint64_t nwords, uint64_t* count) {
  return sd_filter(sd, bytes, len, BUFF_MAJORITY, BUFF_OP_GT, pred, words, nwords, count);
}

int buff_sd_buff_equal_filter_majority(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double pred,
                                       uint32_t* words, uint64_t nwords, uint64_t* count) {
  return sd_filter(sd, bytes, len, BUFF_MAJORITY, BUFF_OP_EQ, pred, words, nwords, count);
}

int buff_sd_buff_sum_majority(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* s) {
  Ctx* c = nullptr;
  buff_chunk* ch = nullptr;
  BUFF_TRY(sd_load(sd, bytes, len, BUFF_MAJORITY, &c, &ch));
  int rc = buff_sum(reinterpret_cast<buff_ctx*>(c), ch, s, nullptr, nullptr);
  (void)buff_chunk_free(ch);
  return rc;
}

int buff_sd_buff_max_majority(const buff_split_double* sd, const uint8_t* bytes, uint64_t len, double* max_value) {
  Ctx* c = nullptr;
  buff_chunk* ch = nullptr;
  BUFF_TRY(sd_load(sd, bytes, len, BUFF_MAJORITY, &c, &ch));
  int rc = buff_max(reinterpret_cast<buff_ctx*>(c), ch, max_value, nullptr);
  (void)buff_chunk_free(ch);
  return rc;
}

}  // extern "C"
