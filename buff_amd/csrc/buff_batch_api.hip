// Host side of batched execution (buff_batch.h): batch encode, query planning,
// the captured query graph, the batch autotuner and their C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "../../include/buffgpu.h"
#include "buff_batch.h"
#include "buff_common.h"
#include "buff_runtime.h"

namespace buffgpu {

namespace {

struct DevBuf {
  void* p = nullptr;
  uint64_t bytes = 0;
};

int ensure(Ctx* ctx, DevBuf& b, uint64_t bytes) {
  bytes = std::max<uint64_t>(bytes, 256);
  if (b.bytes >= bytes) return BUFF_OK;
  if (b.p) {
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    (void)ctx->release(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  const uint64_t want = round_up(bytes + bytes / 4, 4096);
  BUFF_TRY(ctx->alloc(want, &b.p));
  b.bytes = want;
  return BUFF_OK;
}

void drop(Ctx* ctx, DevBuf& b) {
  if (b.p) (void)ctx->release(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

}  // namespace

struct Batch {
  Ctx* ctx = nullptr;
  std::vector<Chunk*> chunks;
  DevBuf pool, views;
  // query machinery
  DevBuf hdr, q, items, acc, res, bitmaps, values, top, bbdev;
  BatchBuffers bb{};
  bool bb_valid = false;
  std::map<uint32_t, hipGraphExec_t> graphs;   // by persistent grid size
  std::vector<BQuery> hq;
  std::vector<BItem> hitems;
  std::vector<BResult> hres;
};

namespace {

constexpr uint64_t kEncodeItemTiles = 16;

// ---------------------------------------------------------------------------
// batch encode (byte_fixed_encode, split_double.rs:298-432, per chunk)

int batch_encode(Ctx* ctx, const double* x, const uint64_t* row_off, const uint64_t* scales, uint32_t nchunks,
                 uint32_t flags, Batch** out) {
  if (!ctx || !out || !row_off || !scales) return fail(BUFF_E_INVALID, "null argument");
  if (flags & ~BUFF_FLIP) return fail(BUFF_E_UNSUPPORTED, "batch encode takes BUFF_FLIP only");
  std::vector<BEncChunk> ck(nchunks);
  std::vector<BItem> items;
  for (uint32_t i = 0; i < nchunks; ++i) {
    if (row_off[i + 1] < row_off[i]) return fail(BUFF_E_INVALID, "row_off must not decrease (chunk %u)", i);
    const uint64_t n = row_off[i + 1] - row_off[i];
    if (n > 0xFFFFFFFFull) return fail(BUFF_E_INVALID, "chunk %u holds more than 2^32-1 rows", i);
    if (n && !x) return fail(BUFF_E_INVALID, "null input");
    const int prec = scales[i] == 0 ? 0 : prec_of_scale(scales[i]);
    const int dl = precision_map(prec);
    if (dl < 0)
      return fail(BUFF_E_REF_PANIC, "chunk %u: precision %d (scale %llu) is not in PRECISION_MAP", i, prec,
                  (unsigned long long)scales[i]);
    memset(&ck[i], 0, sizeof(BEncChunk));
    ck[i].x = x + row_off[i];
    ck[i].view.n = n;
    ck[i].view.ntiles = (n + kTileRows - 1) / kTileRows;
    ck[i].view.pitch = ck[i].view.ntiles * kTileRows;
    ck[i].view.dlen = (uint32_t)dl;
    ck[i].pexp = precision_exp_of_prec(prec);
    ck[i].item0 = (uint32_t)items.size();
    for (uint64_t c = 0; c < ck[i].view.ntiles; c += kEncodeItemTiles)
      items.push_back(BItem{i, 0, c, std::min<uint64_t>(c + kEncodeItemTiles, ck[i].view.ntiles)});
    ck[i].nitems = (uint32_t)items.size() - ck[i].item0;
  }
  const uint32_t nitems = (uint32_t)items.size();
  // scratch: [chunks | items | item partials | per-chunk min/max]
  const uint64_t o_items = round_up(sizeof(BEncChunk) * std::max<uint32_t>(nchunks, 1), 256);
  const uint64_t o_part = o_items + round_up(sizeof(BItem) * std::max<uint32_t>(nitems, 1), 256);
  const uint64_t o_mm = o_part + round_up(16ull * std::max<uint32_t>(nitems, 1), 256);
  const uint64_t total = o_mm + sizeof(BMinMax) * std::max<uint32_t>(nchunks, 1);
  BUFF_TRY(ctx->scratch(ctx->s_stream, total));
  uint8_t* sb = (uint8_t*)ctx->s_stream.ptr;
  BEncChunk* d_ck = (BEncChunk*)sb;
  BItem* d_items = (BItem*)(sb + o_items);
  std::vector<BMinMax> mm(nchunks);
  if (nchunks) {
    HIP_TRY(hipMemcpyAsync(d_ck, ck.data(), sizeof(BEncChunk) * nchunks, hipMemcpyHostToDevice, ctx->stream));
    if (nitems) HIP_TRY(hipMemcpyAsync(d_items, items.data(), sizeof(BItem) * nitems, hipMemcpyHostToDevice, ctx->stream));
    {
      ProfScope ps(ctx, "k_batch_minmax");
      HIP_TRY(launch_batch_minmax(d_ck, nchunks, d_items, nitems, (int64_t*)(sb + o_part), (BMinMax*)(sb + o_mm),
                                  std::max<uint32_t>(1, std::min<uint32_t>(nitems, 4096)), ctx->stream));
    }
    HIP_TRY(hipMemcpyAsync(mm.data(), sb + o_mm, sizeof(BMinMax) * nchunks, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  // widths on the host (glibc log2, as the reference's f64::log2 call)
  uint64_t pool_bytes = 0;
  std::vector<uint64_t> offs(nchunks);
  for (uint32_t i = 0; i < nchunks; ++i) {
    ChunkView& v = ck[i].view;
    const uint64_t n = v.n;
    int64_t mn = INT64_MAX, mx = INT64_MIN;
    if (n == 1) {
      mn = mm[i].min_all;
      mx = mm[i].first == INT64_MAX ? INT64_MAX : INT64_MIN;
    } else if (n > 1) {
      mn = mm[i].min_all;
      if (mm[i].max_rest >= mm[i].first) {
        mx = mm[i].max_rest;
      } else {
        // row 0 holds the unique maximum: the single-chunk replay of the `else if`
        BUFF_TRY(quirk_minmax(ctx, ck[i].x, n, ck[i].pexp, v.dlen, &mn, &mx));
      }
    }
    const uint32_t F = width_of_delta((int64_t)((uint64_t)mx - (uint64_t)mn));
    if (F == 0 && n >= 1)
      return fail(BUFF_E_REF_PANIC, "chunk %u: all %llu values map to one fixed-point value (F = 0)", i,
                  (unsigned long long)n);
    if (F > 63) return fail(BUFF_E_UNSUPPORTED, "chunk %u: bit width %u exceeds 63", i, F);
    v.base = mn;
    v.F = F;
    v.P = F >= 8 ? F / 8 : 0;
    v.r = F >= 8 ? F % 8 : F;
    offs[i] = pool_bytes;
    pool_bytes += round_up((uint64_t)v.P * v.pitch + v.ntiles * 64ull * v.r * 2ull, 256);
  }
  Batch* b = new (std::nothrow) Batch();
  if (!b) return fail(BUFF_E_OOM, "host allocation");
  b->ctx = ctx;
  int rc = ensure(ctx, b->pool, pool_bytes);
  if (rc == BUFF_OK) rc = ensure(ctx, b->views, sizeof(ChunkView) * std::max<uint32_t>(nchunks, 1));
  if (rc != BUFF_OK) {
    drop(ctx, b->pool);
    drop(ctx, b->views);
    delete b;
    return rc;
  }
  std::vector<ChunkView> views(nchunks);
  for (uint32_t i = 0; i < nchunks; ++i) {
    ChunkView& v = ck[i].view;
    v.planes = (uint8_t*)b->pool.p + offs[i];
    v.tail = (uint16_t*)(v.planes + (uint64_t)v.P * v.pitch);
    views[i] = v;
  }
  hipError_t e = hipSuccess;
  if (nchunks) {
    e = hipMemcpyAsync(d_ck, ck.data(), sizeof(BEncChunk) * nchunks, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
      ProfScope ps(ctx, "k_batch_encode");
      e = launch_batch_encode(d_ck, d_items, nitems, std::max<uint32_t>(1, std::min<uint32_t>(nitems, 4096)),
                              ctx->stream);
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->views.p, views.data(), sizeof(ChunkView) * nchunks, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);   // host arrays go out of scope
  }
  if (e != hipSuccess) {
    drop(ctx, b->pool);
    drop(ctx, b->views);
    delete b;
    return fail(BUFF_E_DEVICE, "batch encode: %s", hipGetErrorString(e));
  }
  b->chunks.resize(nchunks);
  for (uint32_t i = 0; i < nchunks; ++i) {
    Chunk* ch = new (std::nothrow) Chunk();
    if (!ch) {
      for (uint32_t k = 0; k < i; ++k) delete b->chunks[k];
      drop(ctx, b->pool);
      drop(ctx, b->views);
      delete b;
      return fail(BUFF_E_OOM, "host allocation");
    }
    ch->ctx = ctx;
    ch->n = views[i].n;
    ch->scale = scales[i];
    ch->flags = flags & BUFF_FLIP;
    ch->dlen = views[i].dlen;
    ch->F = views[i].F;
    ch->ilen = (uint32_t)((uint64_t)views[i].F - views[i].dlen);
    ch->view = views[i];
    ch->device_bytes = (uint64_t)views[i].P * views[i].pitch + views[i].ntiles * 128ull * views[i].r;
    ch->borrowed = true;
    b->chunks[i] = ch;
  }
  *out = b;
  return BUFF_OK;
}

// ---------------------------------------------------------------------------
// batched queries

int graph_for(Batch* b, uint32_t grid, const BatchTuning& tu, hipGraphExec_t* out) {
  auto it = b->graphs.find(grid * 4 + tu.first_planes);
  if (it != b->graphs.end()) {
    *out = it->second;
    return BUFF_OK;
  }
  Ctx* ctx = b->ctx;
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = launch_batch_queries((const BatchBuffers*)b->bbdev.p, tu, ctx->stream);
  hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(BUFF_E_DEVICE, "batch graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
  }
  hipGraphExec_t ex = nullptr;
  e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(BUFF_E_DEVICE, "batch graph instantiate: %s", hipGetErrorString(e));
  b->graphs[grid * 4 + tu.first_planes] = ex;
  *out = ex;
  return BUFF_OK;
}

int batch_query(Ctx* ctx, Batch* b, const buff_query* qs, uint32_t nq, buff_query_result* results,
                const BatchTuning& tu, float* graph_ms) {
  if (!ctx || !b || (nq && (!qs || !results))) return fail(BUFF_E_INVALID, "null argument");
  if (b->ctx != ctx) return fail(BUFF_E_INVALID, "the batch belongs to another context");
  const uint64_t T = std::max<uint32_t>(tu.tiles_per_item, 1);
  b->hq.assign(nq, BQuery{});
  b->hitems.clear();
  uint64_t bit_off = 0, val_off = 0, top_off = 0;
  for (uint32_t k = 0; k < nq; ++k) {
    const buff_query& in = qs[k];
    buff_query_result& r = results[k];
    memset(&r, 0, sizeof(r));
    BQuery& q = b->hq[k];
    memset(&q, 0, sizeof(q));
    q.op = kBqSum;   // a query that fails planning gets no items
    if (in.chunk >= b->chunks.size()) {
      r.status = fail(BUFF_E_INVALID, "query %u: chunk %u outside the batch", k, in.chunk);
      continue;
    }
    const Chunk* ch = b->chunks[in.chunk];
    q.chunk = in.chunk;
    const ChunkView& v = ch->view;
    bool tiles = ch->n > 0;
    switch (in.op) {
      case BUFF_Q_GT:
      case BUFF_Q_EQ: {
        q.op = in.op == BUFF_Q_GT ? kBqGt : kBqEq;
        int fill = 0;
        const bool run = filter_plan(ch, in.op == BUFF_Q_GT ? BUFF_OP_GT : BUFF_OP_EQ, in.pred, in.flags, &q.tg, &fill);
        q.fill = run ? 0 : (fill ? 2 : 1);
        q.out = bit_off;
        r.out_offset = bit_off;
        bit_off += (ch->n + 63) / 64;
        break;
      }
      case BUFF_Q_SUM:
        q.op = kBqSum;
        tiles = tiles && (v.P || v.r);
        break;
      case BUFF_Q_MAX:
        q.op = kBqMax;
        tiles = tiles && (v.P || v.r);
        q.out = top_off;
        top_off += v.ntiles;
        break;
      case BUFF_Q_MAT: {
        const int rc = mat_plan(ch, in.precision, in.flags, &q.mp);
        if (rc != BUFF_OK) {
          r.status = rc;
          q.op = kBqSum;
          tiles = false;
          break;
        }
        q.op = kBqMat;
        q.out = val_off;
        r.out_offset = val_off;
        val_off += ch->n;
        break;
      }
      default:
        r.status = fail(BUFF_E_INVALID, "query %u: unknown op %u", k, in.op);
        tiles = false;
    }
    q.item0 = (uint32_t)b->hitems.size();
    if (tiles)
      for (uint64_t c = 0; c < v.ntiles; c += T) b->hitems.push_back(BItem{k, 0, c, std::min<uint64_t>(c + T, v.ntiles)});
    q.nitems = (uint32_t)b->hitems.size() - q.item0;
  }
  const uint32_t nitems = (uint32_t)b->hitems.size();
  BUFF_TRY(ensure(ctx, b->hdr, sizeof(BHeader)));
  BUFF_TRY(ensure(ctx, b->q, sizeof(BQuery) * std::max<uint32_t>(nq, 1)));
  BUFF_TRY(ensure(ctx, b->items, sizeof(BItem) * std::max<uint32_t>(nitems, 1)));
  BUFF_TRY(ensure(ctx, b->acc, 8ull * (kMaxPlanes + 1) * std::max<uint32_t>(nitems, 1)));
  BUFF_TRY(ensure(ctx, b->res, sizeof(BResult) * std::max<uint32_t>(nq, 1)));
  BUFF_TRY(ensure(ctx, b->bitmaps, 8 * bit_off));
  BUFF_TRY(ensure(ctx, b->values, 8 * val_off));
  BUFF_TRY(ensure(ctx, b->top, top_off));
  BUFF_TRY(ensure(ctx, b->bbdev, sizeof(BatchBuffers)));
  BatchBuffers bb;
  bb.hdr = (const BHeader*)b->hdr.p;
  bb.views = (const ChunkView*)b->views.p;
  bb.q = (const BQuery*)b->q.p;
  bb.items = (const BItem*)b->items.p;
  bb.bitmaps = (uint64_t*)b->bitmaps.p;
  bb.values = (double*)b->values.p;
  bb.tile_top = (uint8_t*)b->top.p;
  bb.item_acc = (unsigned long long*)b->acc.p;
  bb.res = (BResult*)b->res.p;
  if (!b->bb_valid || memcmp(&bb, &b->bb, sizeof(bb)) != 0) {
    HIP_TRY(hipMemcpyAsync(b->bbdev.p, &bb, sizeof(bb), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    b->bb = bb;
    b->bb_valid = true;
  }
  if (nq == 0) return BUFF_OK;
  const BHeader h{nq, nitems, {0, 0}};
  HIP_TRY(hipMemcpyAsync(b->hdr.p, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(b->q.p, b->hq.data(), sizeof(BQuery) * nq, hipMemcpyHostToDevice, ctx->stream));
  if (nitems)
    HIP_TRY(hipMemcpyAsync(b->items.p, b->hitems.data(), sizeof(BItem) * nitems, hipMemcpyHostToDevice, ctx->stream));
  hipGraphExec_t ex = nullptr;
  BUFF_TRY(graph_for(b, tu.grid, tu, &ex));
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (graph_ms) {
    HIP_TRY(hipEventCreate(&ev[0]));
    HIP_TRY(hipEventCreate(&ev[1]));
    HIP_TRY(hipEventRecord(ev[0], ctx->stream));
  }
  {
    ProfScope ps(ctx, "k_batch_graph");
    HIP_TRY(hipGraphLaunch(ex, ctx->stream));
  }
  if (graph_ms) HIP_TRY(hipEventRecord(ev[1], ctx->stream));
  b->hres.resize(nq);
  HIP_TRY(hipMemcpyAsync(b->hres.data(), b->res.p, sizeof(BResult) * nq, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));   // hq / hitems are reused by the next call
  if (graph_ms) {
    (void)hipEventElapsedTime(graph_ms, ev[0], ev[1]);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
  }
  for (uint32_t k = 0; k < nq; ++k) {
    buff_query_result& r = results[k];
    if (r.status != BUFF_OK || qs[k].chunk >= b->chunks.size()) continue;
    const Chunk* ch = b->chunks[qs[k].chunk];
    const BResult& br = b->hres[k];
    switch (qs[k].op) {
      case BUFF_Q_GT:
      case BUFF_Q_EQ:
        r.count = br.v[0];
        break;
      case BUFF_Q_SUM: {
        uint64_t ps[kMaxPlanes + 1] = {0};
        for (uint32_t p = 0; p < ch->view.P; ++p) ps[p] = br.v[p];
        ps[ch->view.P] = br.v[kMaxPlanes];
        r.value = reference_sum_fold(ch, ps);
        __int128 tot = (__int128)ch->n * (__int128)ch->view.base;
        for (uint32_t p = 0; p < ch->view.P; ++p) tot += (__int128)ps[p] << (ch->view.F - 8 * (p + 1));
        tot += (__int128)ps[ch->view.P];
        r.exact = (double)tot / powi2((int32_t)ch->dlen);
        break;
      }
      case BUFF_Q_MAX:
        r.max_off = br.v[0];
        r.value = (double)(int64_t)(br.v[0] + (uint64_t)ch->view.base) / powi2((int32_t)ch->dlen);
        break;
      default:
        break;
    }
  }
  return BUFF_OK;
}

void batch_free(Batch* b) {
  Ctx* ctx = b->ctx;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  for (Chunk* ch : b->chunks) delete ch;
  for (DevBuf* d : {&b->pool, &b->views, &b->hdr, &b->q, &b->items, &b->acc, &b->res, &b->bitmaps, &b->values,
                    &b->top, &b->bbdev})
    drop(ctx, *d);
  delete b;
}

BatchTuning ctx_tuning(const Ctx* ctx) {
  BatchTuning t;
  t.grid = ctx->bt_grid;
  t.first_planes = ctx->bt_first;
  t.tiles_per_item = ctx->bt_tiles;
  t.pad = 0;
  return t;
}

// Times every candidate on the caller's query mix (best of three graph runs
// each) and keeps the fastest.
int batch_tune(Ctx* ctx, Batch* b, const buff_query* qs, uint32_t nq, BatchTuning* best_out) {
  std::vector<buff_query_result> tmp(nq);
  BatchTuning best = ctx_tuning(ctx);
  float best_ms = 1e30f;
  for (uint32_t grid : {1024u, 2048u, 4096u}) {
    for (uint32_t tiles : {4u, 16u, 64u}) {
      for (uint32_t first : {1u, 2u}) {
        BatchTuning t{grid, first, tiles, 0};
        float m = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
          float ms = 0;
          BUFF_TRY(batch_query(ctx, b, qs, nq, tmp.data(), t, &ms));
          m = std::min(m, ms);
        }
        if (m < best_ms) {
          best_ms = m;
          best = t;
        }
      }
    }
  }
  ctx->bt_grid = best.grid;
  ctx->bt_tiles = best.tiles_per_item;
  ctx->bt_first = best.first_planes;
  ctx->bt_tuned = true;
  *best_out = best;
  return BUFF_OK;
}

}  // namespace

}  // namespace buffgpu

using namespace buffgpu;

extern "C" {

int buff_batch_encode_device(buff_ctx* ctx, const double* x_dev, const uint64_t* row_off, const uint64_t* scales,
                             uint32_t nchunks, uint32_t flags, buff_batch** out) {
  Batch* b = nullptr;
  int rc = batch_encode(reinterpret_cast<Ctx*>(ctx), x_dev, row_off, scales, nchunks, flags, &b);
  if (rc == BUFF_OK) *out = reinterpret_cast<buff_batch*>(b);
  return rc;
}

int buff_batch_size(const buff_batch* batch, uint32_t* nchunks) {
  if (!batch || !nchunks) return fail(BUFF_E_INVALID, "null argument");
  *nchunks = (uint32_t)reinterpret_cast<const Batch*>(batch)->chunks.size();
  return BUFF_OK;
}

int buff_batch_chunk(buff_batch* batch, uint32_t i, buff_chunk** chunk) {
  if (!batch || !chunk) return fail(BUFF_E_INVALID, "null argument");
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (i >= b->chunks.size()) return fail(BUFF_E_INVALID, "chunk %u outside the batch", i);
  *chunk = reinterpret_cast<buff_chunk*>(b->chunks[i]);
  return BUFF_OK;
}

int buff_batch_free(buff_batch* batch) {
  if (!batch) return BUFF_OK;
  batch_free(reinterpret_cast<Batch*>(batch));
  return BUFF_OK;
}

int buff_batch_query(buff_ctx* ctx, buff_batch* batch, const buff_query* qs, uint32_t nq, buff_query_result* results) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(BUFF_E_INVALID, "null argument");
  return batch_query(c, reinterpret_cast<Batch*>(batch), qs, nq, results, ctx_tuning(c), nullptr);
}

int buff_batch_arena(buff_batch* batch, const uint64_t** bitmaps_dev, const double** values_dev) {
  if (!batch) return fail(BUFF_E_INVALID, "null argument");
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (bitmaps_dev) *bitmaps_dev = (const uint64_t*)b->bitmaps.p;
  if (values_dev) *values_dev = (const double*)b->values.p;
  return BUFF_OK;
}

int buff_batch_tune(buff_ctx* ctx, buff_batch* batch, const buff_query* qs, uint32_t nq, uint32_t* grid,
                    uint32_t* tiles_per_item, uint32_t* first_planes) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !batch) return fail(BUFF_E_INVALID, "null argument");
  BatchTuning t;
  BUFF_TRY(batch_tune(c, reinterpret_cast<Batch*>(batch), qs, nq, &t));
  if (grid) *grid = t.grid;
  if (tiles_per_item) *tiles_per_item = t.tiles_per_item;
  if (first_planes) *first_planes = t.first_planes;
  return BUFF_OK;
}

}  // extern "C"
