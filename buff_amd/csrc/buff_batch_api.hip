// Host side of batched execution (buff_batch.h): batch encode, query planning,
// the captured query graph, the batch autotuner and their C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
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

// Exactly `bytes` (no growth slack), for buffers sized up front.
int reserve(Ctx* ctx, DevBuf& b, uint64_t bytes) {
  bytes = round_up(std::max<uint64_t>(bytes, 256), 256);
  if (b.bytes >= bytes) return BUFF_OK;
  if (b.p) {
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    (void)ctx->release(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  BUFF_TRY(ctx->alloc(bytes, &b.p));
  b.bytes = bytes;
  return BUFF_OK;
}

void drop(Ctx* ctx, DevBuf& b) {
  if (b.p) (void)ctx->release(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

}  // namespace

// One in-flight query batch: descriptor / result buffers, the captured graph
// per tuning, its stream and completion event. A Batch uses one on the
// context's stream; a Server runs several on streams of its own.
struct QueryExec {
  Ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipEvent_t done = nullptr;
  DevBuf hdr, q, items, acc, res, top, arena, bbdev;
  BatchBuffers bb{};
  bool bb_valid = false;
  std::map<uint64_t, hipGraphExec_t> graphs;   // by (grid, parts, first_planes)
  std::vector<BQuery> hq;
  // pinned host staging [header | queries | results] so the copies are
  // asynchronous (a copy to or from pageable memory can hold the host thread
  // until the stream has drained, which would serialize a server's slots)
  uint8_t* pin = nullptr;       // host address
  uint8_t* pin_dev = nullptr;   // the same memory as the kernels address it
  uint32_t pin_cap = 0;
  std::vector<buff_query_result> res_host;
  BHeader h{};
  uint32_t nq = 0;
  // queries without a batched kernel (a max on a buff-major chunk with sparse
  // columns, every query on a residue chunk) run the single-chunk call on
  // `aux`, a context of the executor's own, when the batch completes
  // (host_ops; aux == nullptr: answered BUFF_E_UNSUPPORTED)
  Ctx* aux = nullptr;
  std::vector<uint32_t> host_ops;
};

struct Batch {
  Ctx* ctx = nullptr;
  std::vector<Chunk*> chunks;
  DevBuf pool, views;
  QueryExec ex;
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
      ProfScope ps(ctx, "k_batch_minmax", false);
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
      ProfScope ps(ctx, "k_batch_encode", false);
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

int exec_init(QueryExec& ex, Ctx* ctx, bool own_stream) {
  ex.ctx = ctx;
  if (own_stream) {
    HIP_TRY(hipStreamCreateWithFlags(&ex.stream, hipStreamNonBlocking));
    ex.own_stream = true;
  } else {
    ex.stream = ctx->stream;
  }
  HIP_TRY(hipEventCreateWithFlags(&ex.done, hipEventDisableTiming));
  return BUFF_OK;
}

int graph_for(QueryExec& ex, const BatchTuning& tu, uint32_t parts, hipGraphExec_t* out);

uint64_t pin_queries_off() { return kHostQueries; }
uint64_t pin_results_off(uint32_t cap) { return 256 + round_up(sizeof(BQuery) * (uint64_t)cap, 256); }

int pin_reserve(QueryExec& ex, uint32_t nq) {
  if (ex.pin && ex.pin_cap >= nq) return BUFF_OK;
  const uint32_t cap = std::max<uint32_t>(nq, 64);
  if (ex.pin) {
    HIP_TRY(hipStreamSynchronize(ex.stream));
    (void)hipHostFree(ex.pin);
    ex.pin = nullptr;
    ex.pin_dev = nullptr;
    ex.pin_cap = 0;
  }
  const uint64_t bytes = pin_results_off(cap) + sizeof(BResult) * (uint64_t)cap;
  // coherent (fine-grained): the kernels read and write it over the bus, uncached
  void* p = nullptr;
  hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocMapped | hipHostMallocCoherent);
  if (e != hipSuccess) return fail(BUFF_E_OOM, "pinned host staging (%llu bytes): %s", (unsigned long long)bytes,
                                   hipGetErrorString(e));
  void* d = nullptr;
  e = hipHostGetDevicePointer(&d, p, 0);
  if (e != hipSuccess) {
    (void)hipHostFree(p);
    return fail(BUFF_E_DEVICE, "pinned host staging device address: %s", hipGetErrorString(e));
  }
  ex.pin = static_cast<uint8_t*>(p);
  ex.pin_dev = static_cast<uint8_t*>(d);
  ex.pin_cap = cap;
  return BUFF_OK;
}

// Sizes a server slot's buffers and captures its graphs before it serves:
// device allocations and graph instantiation can wait for work already running
// on the device, which would serialize the slots. Larger batches still grow
// the buffers when they need to.
int exec_reserve(QueryExec& ex, uint32_t max_q, uint64_t arena_bytes, const BatchTuning& tu) {
  Ctx* ctx = ex.ctx;
  constexpr uint64_t kItems = 1u << 20;
  BUFF_TRY(reserve(ctx, ex.hdr, sizeof(BHeader)));
  BUFF_TRY(reserve(ctx, ex.q, sizeof(BQuery) * (uint64_t)max_q));
  BUFF_TRY(reserve(ctx, ex.res, sizeof(BResult) * (uint64_t)max_q));
  BUFF_TRY(reserve(ctx, ex.items, sizeof(BItem) * kItems));
  BUFF_TRY(reserve(ctx, ex.acc, 8ull * (kMaxPlanes + 1) * kItems));
  BUFF_TRY(reserve(ctx, ex.top, 64ull << 20));
  BUFF_TRY(reserve(ctx, ex.arena, arena_bytes));
  BUFF_TRY(reserve(ctx, ex.bbdev, sizeof(BatchBuffers)));
  BUFF_TRY(pin_reserve(ex, max_q));
  // every combination of kernels a batch can need (the scan always runs)
  for (uint32_t extra = 0; extra < 16; ++extra) {
    const uint32_t parts = kBqPartScan | ((extra & 1) ? kBqPartMax : 0u) | ((extra & 2) ? kBqPartMat : 0u) |
                           ((extra & 4) ? kBqPartMajor : 0u) | ((extra & 8) ? kBqPartMajorMat : 0u);
    hipGraphExec_t gx = nullptr;
    BUFF_TRY(graph_for(ex, tu, parts, &gx));
  }
  return BUFF_OK;
}

void exec_free(QueryExec& ex) {
  if (!ex.ctx) return;
  (void)hipStreamSynchronize(ex.stream);
  for (auto& kv : ex.graphs) (void)hipGraphExecDestroy(kv.second);
  ex.graphs.clear();
  for (DevBuf* d : {&ex.hdr, &ex.q, &ex.items, &ex.acc, &ex.res, &ex.top, &ex.arena, &ex.bbdev}) drop(ex.ctx, *d);
  if (ex.pin) (void)hipHostFree(ex.pin);
  ex.pin = nullptr;
  ex.pin_dev = nullptr;
  ex.pin_cap = 0;
  if (ex.done) (void)hipEventDestroy(ex.done);
  if (ex.own_stream) (void)hipStreamDestroy(ex.stream);
  ex.done = nullptr;
  ex.stream = nullptr;
}

int graph_for(QueryExec& ex, const BatchTuning& tu, uint32_t parts, hipGraphExec_t* out) {
  const uint64_t key = ((uint64_t)tu.grid << 32) | ((uint64_t)parts << 16) | tu.first_planes;
  auto it = ex.graphs.find(key);
  if (it != ex.graphs.end()) {
    *out = it->second;
    return BUFF_OK;
  }
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamBeginCapture(ex.stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = launch_batch_queries((const BatchBuffers*)ex.bbdev.p, tu, parts, ex.stream);
  hipError_t e2 = hipStreamEndCapture(ex.stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(BUFF_E_DEVICE, "batch graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
  }
  hipGraphExec_t gx = nullptr;
  e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(BUFF_E_DEVICE, "batch graph instantiate: %s", hipGetErrorString(e));
  ex.graphs[key] = gx;
  *out = gx;
  return BUFF_OK;
}

// Bytes a query needs in the arena when the caller gives no output address.
uint64_t arena_need(const Chunk* ch, const buff_query& in) {
  if (in.op == BUFF_Q_GT || in.op == BUFF_Q_EQ) return round_up(((ch->n + 63) / 64) * 8, 256);
  if (in.op == BUFF_Q_MAT) return round_up(ch->n * 8, 256);
  return 0;
}

// Plans and launches nq queries (asynchronously on ex.stream); outs[k] is the
// query's output address or 0 for the arena. Results come from exec_finish.
int exec_launch(QueryExec& ex, const std::vector<Chunk*>& chunks, const void* views_dev, const void* majors_dev,
                const buff_query* qs, uint32_t nq, const uint64_t* outs, const BatchTuning& tu) {
  Ctx* ctx = ex.ctx;
  const uint64_t Tmax = std::min<uint32_t>(std::max<uint32_t>(tu.tiles_per_item, 1), kMaxTilesPerItem);
  ex.hq.assign(nq, BQuery{});
  std::vector<uint64_t> qtiles(nq, 0);
  uint64_t total_tiles = 0;
  ex.res_host.assign(nq, buff_query_result{});
  ex.nq = nq;
  ex.host_ops.clear();
  uint64_t arena_off = 0, top_off = 0;
  std::vector<uint64_t> arena_at(nq, ~0ull);
  for (uint32_t k = 0; k < nq; ++k) {
    const buff_query& in = qs[k];
    buff_query_result& r = ex.res_host[k];
    BQuery& q = ex.hq[k];
    memset(&q, 0, sizeof(q));
    q.op = kBqSum;   // a query that fails planning gets no items
    if (in.chunk >= chunks.size()) {
      r.status = fail(BUFF_E_INVALID, "query %u: chunk %u outside the batch", k, in.chunk);
      continue;
    }
    const Chunk* ch = chunks[in.chunk];
    q.chunk = in.chunk;
    const ChunkView& v = ch->view;
    bool tiles = ch->n > 0;
    const uint64_t need = (outs && outs[k]) ? 0 : arena_need(ch, in);
    if (need) {
      arena_at[k] = arena_off;
      arena_off += need;
    }
    const bool residue = (ch->flags & BUFF_RESIDUE) != 0;
    // a residue chunk's sums and full-precision materializations run batched (its
    // resident form is buff-major's with majority byte 0, or byte_fixed's); its
    // other queries have no items and run the single-chunk call in exec_finish
    if (residue && in.op != BUFF_Q_SUM && !(in.op == BUFF_Q_MAT && in.precision < 0)) {
      if (ex.aux)
        ex.host_ops.push_back(k);
      else
        r.status = fail(BUFF_E_UNSUPPORTED, "query %u: this query of a residue chunk runs as its single-chunk call", k);
      continue;
    }
    const bool sparse = ch->S > 0;   // buff-major with sparse columns: k_bq_major / k_bq_major_mat
    const bool sem = (in.flags & BUFF_SEMANTIC) != 0;
    const uint32_t sparse_word = sparse ? (ch->S | (sem ? kBqSparseSem : 0u)) : 0u;
    switch (in.op) {
      case BUFF_Q_GT:
      case BUFF_Q_EQ: {
        q.op = in.op == BUFF_Q_GT ? kBqGt : kBqEq;
        const int op = in.op == BUFF_Q_GT ? BUFF_OP_GT : BUFF_OP_EQ;
        if (sparse) {
          MajorTarget mt;
          const int fill = major_filter_plan(ch, op, in.pred, in.flags, &mt);
          q.fill = fill < 0 ? 0 : (fill ? 2 : 1);
          q.sparse = fill < 0 ? sparse_word : 0u;   // a host-decided fill needs no sparse column
          for (uint32_t d = 0; d < kMaxPlanes; ++d) q.tg.digit[d] = mt.digit[d];
          q.tg.tail = mt.tail;
        } else {
          int fill = 0;
          const bool run = filter_plan(ch, op, in.pred, in.flags, &q.tg, &fill);
          q.fill = run ? 0 : (fill ? 2 : 1);
        }
        break;
      }
      case BUFF_Q_SUM:
        q.op = kBqSum;
        q.sparse = sparse_word;
        tiles = tiles && (sparse || v.P || v.r);
        break;
      case BUFF_Q_MAX:
        if (sparse) {
          if (ex.aux)
            ex.host_ops.push_back(k);   // no items: answered in exec_finish
          else
            r.status = fail(BUFF_E_UNSUPPORTED, "query %u: batched max of a buff-major chunk with sparse columns "
                            "(the single-chunk buff_max answers it)", k);
          tiles = false;
          break;
        }
        q.op = kBqMax;
        tiles = tiles && (v.P || v.r);
        q.top = top_off;
        top_off += v.ntiles;
        break;
      case BUFF_Q_MAT: {
        int rc = BUFF_OK;
        if ((ch->flags & BUFF_MAJORITY) && in.precision >= 0 && !sem)
          rc = fail(BUFF_E_UNSUPPORTED, "query %u: the reference has no reduced-precision decode of a buff-major "
                    "stream", k);
        if (rc == BUFF_OK && sparse) {
          if (!sem)
            for (uint32_t j = 0; j < ch->S && rc == BUFF_OK; ++j)
              if (ch->sp[j].n_out == 0)
                rc = fail(BUFF_E_REF_PANIC, "query %u: sparse column %u has no outliers (%s)", k, j,
                          residue ? "split_double.rs:1405" : "buff_simd.rs:683");
          memset(&q.mp, 0, sizeof(q.mp));
          // a residue decode (byte_residue_decode) is the true offsets
          q.mp.mode = (sem || residue) ? kMatSemantic : kMatFull;
          q.mp.inv_scale = std::ldexp(1.0, -(int)ch->dlen);
          if (rc == BUFF_OK) rc = major_precision_drop(ch, in.precision, &q.mp.drop);
          tiles = tiles && (residue && !sem ? residue_decoded_rows(ch, -1) : major_decoded_rows(ch, in.flags)) > 0;
          q.sparse = sparse_word;
        } else if (rc == BUFF_OK) {
          rc = mat_plan(ch, in.precision, in.flags, &q.mp);
        }
        if (rc != BUFF_OK) {
          r.status = rc;
          q.op = kBqSum;
          q.sparse = 0;
          tiles = false;
          break;
        }
        q.op = kBqMat;
        break;
      }
      default:
        r.status = fail(BUFF_E_INVALID, "query %u: unknown op %u", k, in.op);
        tiles = false;
    }
    qtiles[k] = tiles ? v.ntiles : 0;
    total_tiles += qtiles[k];
  }
  // Item length: the tuned one, shortened when the batch has too few tiles to
  // give every wave of the grid (4 per workgroup) an item.
  const uint64_t waves = 4ull * std::max<uint32_t>(tu.grid, 1);
  const uint64_t T = std::max<uint64_t>(1, std::min<uint64_t>(Tmax, (total_tiles + waves - 1) / waves));
  // item ranges: scan items first, materializations after them (their item0 is
  // rebased below); k_bq_items writes the items on the device
  uint64_t n_scan = 0, n_mat = 0;
  for (uint32_t k = 0; k < nq; ++k) {
    BQuery& q = ex.hq[k];
    uint64_t& next = q.op == kBqMat ? n_mat : n_scan;
    q.item0 = (uint32_t)next;
    q.nitems = (uint32_t)((qtiles[k] + T - 1) / T);
    next += q.nitems;
  }
  if (n_scan + n_mat > 0xFFFFFFFFull) return fail(BUFF_E_INVALID, "batch work items exceed 2^32");
  const uint32_t nscan = (uint32_t)n_scan;
  for (BQuery& q : ex.hq)
    if (q.op == kBqMat) q.item0 += nscan;
  const uint32_t nitems = (uint32_t)(n_scan + n_mat);
  BUFF_TRY(ensure(ctx, ex.hdr, sizeof(BHeader)));
  BUFF_TRY(ensure(ctx, ex.q, sizeof(BQuery) * std::max<uint32_t>(nq, 1)));
  BUFF_TRY(ensure(ctx, ex.items, sizeof(BItem) * std::max<uint32_t>(nitems, 1)));
  BUFF_TRY(ensure(ctx, ex.acc, 8ull * (kMaxPlanes + 1) * std::max<uint32_t>(nitems, 1)));
  BUFF_TRY(ensure(ctx, ex.res, sizeof(BResult) * std::max<uint32_t>(nq, 1)));
  BUFF_TRY(ensure(ctx, ex.top, top_off));
  BUFF_TRY(ensure(ctx, ex.arena, arena_off));
  BUFF_TRY(ensure(ctx, ex.bbdev, sizeof(BatchBuffers)));
  for (uint32_t k = 0; k < nq; ++k) {
    if (ex.res_host[k].status != BUFF_OK) continue;
    const uint64_t addr = (outs && outs[k]) ? outs[k] : (uint64_t)ex.arena.p + (arena_at[k] == ~0ull ? 0 : arena_at[k]);
    const bool has_out = qs[k].op == BUFF_Q_GT || qs[k].op == BUFF_Q_EQ || qs[k].op == BUFF_Q_MAT;
    ex.hq[k].out = addr;
    ex.res_host[k].out_offset = has_out ? addr : 0;
  }
  BUFF_TRY(pin_reserve(ex, nq));
  BatchBuffers bb;
  bb.hdr = (const BHeader*)ex.hdr.p;
  bb.views = (const ChunkView*)views_dev;
  bb.majors = (const MajorView*)majors_dev;
  bb.q = (const BQuery*)ex.q.p;
  bb.items = (BItem*)ex.items.p;
  bb.tile_top = (uint8_t*)ex.top.p;
  bb.item_acc = (unsigned long long*)ex.acc.p;
  bb.res = (BResult*)ex.res.p;
  bb.host_in = ex.pin_dev;
  bb.host_res = (BResult*)(ex.pin_dev + pin_results_off(ex.pin_cap));
  if (!ex.bb_valid || memcmp(&bb, &ex.bb, sizeof(bb)) != 0) {
    HIP_TRY(hipMemcpyAsync(ex.bbdev.p, &bb, sizeof(bb), hipMemcpyHostToDevice, ex.stream));
    HIP_TRY(hipStreamSynchronize(ex.stream));   // bb is a stack copy
    ex.bb = bb;
    ex.bb_valid = true;
  }
  if (nq == 0) return BUFF_OK;
  ex.h = BHeader{nq, nitems, nscan, (uint32_t)T};
  memcpy(ex.pin, &ex.h, sizeof(ex.h));
  memcpy(ex.pin + pin_queries_off(), ex.hq.data(), sizeof(BQuery) * nq);
  uint32_t parts = 0;
  for (const BQuery& q : ex.hq) {
    if (!q.nitems) continue;
    if (q.sparse) parts |= q.op == kBqMat ? kBqPartMajorMat : kBqPartMajor;
    else parts |= q.op == kBqMat ? kBqPartMat : kBqPartScan;
    if (q.op == kBqMax) parts |= kBqPartMax;
  }
  parts |= kBqPartScan;   // the fold also writes the results of queries without tiles
  hipGraphExec_t gx = nullptr;
  BUFF_TRY(graph_for(ex, tu, parts, &gx));
  HIP_TRY(hipGraphLaunch(gx, ex.stream));
  HIP_TRY(hipEventRecord(ex.done, ex.stream));
  return BUFF_OK;
}

// Waits for the batch and turns the device results into buff_query_result.
int exec_finish(QueryExec& ex, const std::vector<Chunk*>& chunks, const buff_query* qs) {
  HIP_TRY(hipEventSynchronize(ex.done));
  for (uint32_t k = 0; k < ex.nq; ++k) {
    buff_query_result& r = ex.res_host[k];
    if (r.status != BUFF_OK || qs[k].chunk >= chunks.size()) continue;
    const Chunk* ch = chunks[qs[k].chunk];
    const BResult& br = reinterpret_cast<const BResult*>(ex.pin + pin_results_off(ex.pin_cap))[k];
    switch (qs[k].op) {
      case BUFF_Q_GT:
      case BUFF_Q_EQ:
        r.count = br.v[0];
        break;
      case BUFF_Q_SUM: {
        uint64_t ps[kMaxPlanes + 1] = {0};
        for (uint32_t p = 0; p < ch->view.P; ++p) ps[p] = br.v[p];
        ps[ch->view.P] = br.v[kMaxPlanes];
        uint64_t outsum[kMaxSparse] = {0};   // dense planes, then the sparse columns' outlier sums (k_bq_major)
        for (uint32_t j = 0; j < ch->S; ++j) outsum[j] = br.v[ch->view.P + j];
        if (ch->flags & BUFF_RESIDUE) {
          r.status = residue_full_sum_fold(ch, outsum, ps, &r.value, &r.exact);
          break;
        }
        if (ch->S > 0) {
          major_sum_fold(ch, outsum, ps, &r.value, &r.exact, nullptr);
          break;
        }
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
  // the single-chunk calls, on the executor's own context
  for (uint32_t k : ex.host_ops) {
    buff_query_result& r = ex.res_host[k];
    if (r.status != BUFF_OK) continue;
    const buff_query& q = qs[k];
    const Chunk* ch = chunks[q.chunk];
    const uint64_t out_offset = r.out_offset;   // the output address planned for it
    r = buff_query_result{};
    r.out_offset = out_offset;
    switch (q.op) {
      case BUFF_Q_GT:
      case BUFF_Q_EQ: {
        uint64_t* words = reinterpret_cast<uint64_t*>(ex.hq[k].out);
        r.status = filter_device(ex.aux, ch, q.op == BUFF_Q_GT ? BUFF_OP_GT : BUFF_OP_EQ, q.pred, q.flags, words,
                                 nullptr);
        const uint64_t nw = (ch->n + 63) / 64;
        if (r.status == BUFF_OK && nw) {   // the count, from the bitmap
          std::vector<uint64_t> host(nw);
          hipError_t e = hipMemcpyAsync(host.data(), words, nw * 8, hipMemcpyDeviceToHost, ex.aux->stream);
          if (e == hipSuccess) e = hipStreamSynchronize(ex.aux->stream);
          if (e != hipSuccess) {
            r.status = fail(BUFF_E_DEVICE, "query %u: filter count: %s", k, hipGetErrorString(e));
            break;
          }
          for (uint64_t w : host) r.count += (uint64_t)__builtin_popcountll(w);
        }
        break;
      }
      case BUFF_Q_SUM:   // buff_sum_ex: the flags do not change a sum of these layouts
        r.status = sum_op(ex.aux, ch, &r.value, &r.exact, nullptr);
        break;
      case BUFF_Q_MAX:
        r.status = max_op(ex.aux, ch, q.flags, &r.value, &r.max_off, nullptr);
        break;
      case BUFF_Q_MAT:
        r.status = materialize_device(ex.aux, ch, q.precision, q.flags, reinterpret_cast<double*>(ex.hq[k].out));
        break;
      default:
        r.status = fail(BUFF_E_INVALID, "query %u: unknown op %u", k, q.op);
    }
  }
  if (!ex.host_ops.empty()) HIP_TRY(hipStreamSynchronize(ex.aux->stream));   // outputs written
  return BUFF_OK;
}

BatchTuning ctx_tuning(const Ctx* ctx) {
  BatchTuning t;
  t.grid = ctx->bt_grid;
  t.first_planes = ctx->bt_first;
  t.tiles_per_item = ctx->bt_tiles;
  t.pad = 0;
  return t;
}

// Batch queries: results for filters / materializations point into the batch's
// arena (out_offset = byte offset from the arena base).
int batch_query(Ctx* ctx, Batch* b, const buff_query* qs, uint32_t nq, buff_query_result* results,
                const BatchTuning& tu, float* graph_ms) {
  if (!ctx || !b || (nq && (!qs || !results))) return fail(BUFF_E_INVALID, "null argument");
  if (b->ctx != ctx) return fail(BUFF_E_INVALID, "the batch belongs to another context");
  if (!b->ex.ctx) BUFF_TRY(exec_init(b->ex, ctx, false));
  hipEvent_t ev0 = nullptr;
  if (graph_ms) {
    HIP_TRY(hipEventCreate(&ev0));
    HIP_TRY(hipEventRecord(ev0, b->ex.stream));
  }
  int rc;
  {
    ProfScope ps(ctx, "k_batch_graph", false);
    rc = exec_launch(b->ex, b->chunks, b->views.p, nullptr, qs, nq, nullptr, tu);
  }
  if (rc == BUFF_OK) rc = exec_finish(b->ex, b->chunks, qs);
  if (graph_ms) {
    if (rc == BUFF_OK) {
      hipEvent_t ev1;
      (void)hipEventCreate(&ev1);
      (void)hipEventRecord(ev1, b->ex.stream);
      (void)hipEventSynchronize(ev1);
      (void)hipEventElapsedTime(graph_ms, ev0, ev1);
      (void)hipEventDestroy(ev1);
    }
    (void)hipEventDestroy(ev0);
  }
  if (rc != BUFF_OK) return rc;
  for (uint32_t k = 0; k < nq; ++k) {
    results[k] = b->ex.res_host[k];
    if (results[k].status == BUFF_OK &&
        (qs[k].op == BUFF_Q_GT || qs[k].op == BUFF_Q_EQ || qs[k].op == BUFF_Q_MAT))
      results[k].out_offset -= (uint64_t)b->ex.arena.p;
    else
      results[k].out_offset = 0;
  }
  return BUFF_OK;
}

void batch_free(Batch* b) {
  Ctx* ctx = b->ctx;
  (void)hipStreamSynchronize(ctx->stream);
  exec_free(b->ex);
  for (Chunk* ch : b->chunks) delete ch;
  drop(ctx, b->pool);
  drop(ctx, b->views);
  delete b;
}

// Times every candidate on the caller's query mix (best of three runs each)
// and keeps the fastest.
int batch_tune(Ctx* ctx, Batch* b, const buff_query* qs, uint32_t nq, BatchTuning* best_out) {
  std::vector<buff_query_result> tmp(nq);
  BatchTuning best = ctx_tuning(ctx);
  float best_ms = 1e30f;
  for (uint32_t grid : {1024u, 2048u, 4096u, 8192u}) {
    for (uint32_t tiles : {16u, 64u, 256u}) {
      for (uint32_t first : {1u, kMaxPlanes}) {
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

// ---------------------------------------------------------------------------
// Server: resident chunks, a request queue and a worker thread doing continuous
// batching. While one batch runs on the GPU the worker plans the next from the
// requests that arrived meanwhile and launches it on another stream, so up to
// `slots` batches overlap on the device. A request joins a batch when a slot is
// free and the batch still has room: at most max_batch queries, and arena bytes
// for its output when the caller gave no output address.

struct Req {
  buff_query q;
  uint64_t out = 0;             // the caller's output address (0: the server arena)
  uint64_t arena_off = ~0ull;   // the arena range holding its output, until buff_server_release
  double t_submit = 0, t_done = 0;
  buff_query_result res{};
  bool done = false;
};

struct Server {
  Ctx* ctx = nullptr;
  Ctx* aux = nullptr;   // the worker's own context (maxima of buff-major chunks with sparse columns)
  buff_server_cfg cfg{};
  std::vector<Chunk*> chunks;
  DevBuf views, majors;   // per chunk: its (dense) view; the sparse columns of buff-major chunks
  bool views_dirty = false;
  std::vector<QueryExec> slots;
  std::vector<std::vector<uint64_t>> slot_tickets;
  std::vector<std::vector<buff_query>> slot_qs;
  std::vector<std::vector<Chunk*>> slot_chunks;   // the chunk list a slot's batch was planned on
  std::vector<bool> slot_busy;
  // Output memory of requests without an address: one device arena of
  // cfg.arena_bytes; a request is admitted when a range of it is free, and the
  // range stays held (its bitmap or values readable) until buff_server_release.
  DevBuf arena;
  std::map<uint64_t, uint64_t> arena_free;   // offset -> bytes, coalesced
  std::map<uint64_t, uint64_t> arena_held;   // offset -> bytes
  bool arena_blocked = false;                // the oldest pending request waits for arena memory
  uint64_t arena_peak = 0, arena_in_use = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> pending;
  std::unordered_map<uint64_t, Req> reqs;
  uint64_t next_ticket = 1;
  bool stop = false;
  int error = BUFF_OK;
  std::thread worker;
  uint64_t batches = 0, queries = 0, max_seen = 0;
};

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit range of the server arena (caller holds sv->mu).
bool arena_take(Server* sv, uint64_t need, uint64_t* off) {
  for (auto it = sv->arena_free.begin(); it != sv->arena_free.end(); ++it) {
    if (it->second < need) continue;
    const uint64_t o = it->first, sz = it->second;
    sv->arena_free.erase(it);
    if (sz > need) sv->arena_free[o + need] = sz - need;
    sv->arena_held[o] = need;
    sv->arena_in_use += need;
    sv->arena_peak = std::max(sv->arena_peak, sv->arena_in_use);
    *off = o;
    return true;
  }
  return false;
}

// Returns a held range to the free list, merged with its neighbours (caller holds sv->mu).
bool arena_give(Server* sv, uint64_t off) {
  auto h = sv->arena_held.find(off);
  if (h == sv->arena_held.end()) return false;
  uint64_t o = off, sz = h->second;
  sv->arena_held.erase(h);
  sv->arena_in_use -= sz;
  auto next = sv->arena_free.find(o + sz);
  if (next != sv->arena_free.end()) {
    sz += next->second;
    sv->arena_free.erase(next);
  }
  auto prev = sv->arena_free.lower_bound(o);
  if (prev != sv->arena_free.begin()) {
    --prev;
    if (prev->first + prev->second == o) {
      o = prev->first;
      sz += prev->second;
      sv->arena_free.erase(prev);
    }
  }
  sv->arena_free[o] = sz;
  sv->arena_blocked = false;
  return true;
}

void server_loop(Server* sv) {
  const BatchTuning tu = ctx_tuning(sv->ctx);
  const uint32_t nslots = (uint32_t)sv->slots.size();
  std::unique_lock<std::mutex> lk(sv->mu);
  uint32_t rr = 0;
  while (true) {
    bool progressed = false;
    // completions
    for (uint32_t s = 0; s < nslots; ++s) {
      if (!sv->slot_busy[s]) continue;
      if (hipEventQuery(sv->slots[s].done) != hipSuccess) continue;
      lk.unlock();
      int rc = exec_finish(sv->slots[s], sv->slot_chunks[s], sv->slot_qs[s].data());
      const double t = now_s();
      lk.lock();
      if (rc != BUFF_OK) sv->error = rc;
      for (uint32_t k = 0; k < sv->slot_tickets[s].size(); ++k) {
        Req& r = sv->reqs[sv->slot_tickets[s][k]];
        r.res = sv->slots[s].res_host[k];
        if (rc != BUFF_OK) r.res.status = rc;
        if (r.res.status != BUFF_OK && r.arena_off != ~0ull) {   // no output to hold
          (void)arena_give(sv, r.arena_off);
          r.arena_off = ~0ull;
          r.res.out_offset = 0;
        }
        r.t_done = t;
        r.done = true;
      }
      sv->slot_busy[s] = false;
      sv->cv_done.notify_all();
      progressed = true;
    }
    bool any_busy = false;
    for (uint32_t s = 0; s < nslots; ++s) any_busy |= sv->slot_busy[s];
    if (sv->stop && (sv->pending.empty() || (sv->arena_blocked && !any_busy))) {
      if (!any_busy) {
        // shutting down: requests still waiting for arena memory will not get it
        const double t = now_s();
        for (uint64_t id : sv->pending) {
          Req& r = sv->reqs[id];
          r.res.status = fail(BUFF_E_INVALID, "server destroyed while the request waited for arena memory");
          r.t_done = t;
          r.done = true;
        }
        sv->pending.clear();
        sv->cv_done.notify_all();
        break;
      }
    }
    // launches: a free slot takes what fits
    for (uint32_t i = 0; i < nslots && !sv->pending.empty() && !sv->arena_blocked; ++i) {
      const uint32_t s = (rr + i) % nslots;
      if (sv->slot_busy[s]) continue;
      std::vector<uint64_t>& tk = sv->slot_tickets[s];
      std::vector<buff_query>& qv = sv->slot_qs[s];
      std::vector<uint64_t> outs;
      tk.clear();
      qv.clear();
      while (!sv->pending.empty() && tk.size() < sv->cfg.max_batch) {
        const uint64_t id = sv->pending.front();
        Req& r = sv->reqs[id];
        const uint64_t need =
            (r.out || r.q.chunk >= sv->chunks.size()) ? 0 : arena_need(sv->chunks[r.q.chunk], r.q);
        if (need > sv->cfg.arena_bytes) {   // can never be admitted
          sv->pending.pop_front();
          r.res.status = fail(BUFF_E_OOM, "the request's output (%llu bytes) exceeds the server arena (%llu bytes)",
                              (unsigned long long)need, (unsigned long long)sv->cfg.arena_bytes);
          r.t_done = now_s();
          r.done = true;
          sv->cv_done.notify_all();
          continue;
        }
        uint64_t off = 0;
        if (need && !arena_take(sv, need, &off)) {   // waits until a held range is released
          sv->arena_blocked = true;
          break;
        }
        if (need) r.arena_off = off;
        sv->pending.pop_front();
        tk.push_back(id);
        qv.push_back(r.q);
        outs.push_back(r.out ? r.out : (need ? (uint64_t)sv->arena.p + off : 0));
      }
      if (tk.empty()) break;
      const bool vdirty = sv->views_dirty;
      sv->views_dirty = false;
      sv->slot_chunks[s] = sv->chunks;
      const std::vector<Chunk*>& chunks = sv->slot_chunks[s];
      lk.unlock();
      int rc = BUFF_OK;
      if (vdirty) {
        std::vector<ChunkView> v(chunks.size());
        std::vector<MajorView> mj(chunks.size());
        for (size_t k = 0; k < v.size(); ++k) {
          v[k] = chunks[k]->view;
          if (chunks[k]->S > 0) mj[k] = major_view(chunks[k]);
          else memset(&mj[k], 0, sizeof(MajorView));
        }
        for (uint32_t o = 0; o < nslots; ++o) (void)hipStreamSynchronize(sv->slots[o].stream);
        rc = ensure(sv->ctx, sv->views, sizeof(ChunkView) * std::max<size_t>(v.size(), 1));
        if (rc == BUFF_OK) rc = ensure(sv->ctx, sv->majors, sizeof(MajorView) * std::max<size_t>(mj.size(), 1));
        if (rc == BUFF_OK && !v.empty() &&
            (hipMemcpy(sv->views.p, v.data(), sizeof(ChunkView) * v.size(), hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(sv->majors.p, mj.data(), sizeof(MajorView) * mj.size(), hipMemcpyHostToDevice) != hipSuccess))
          rc = fail(BUFF_E_DEVICE, "server views upload");
      }
      if (rc == BUFF_OK)
        rc = exec_launch(sv->slots[s], chunks, sv->views.p, sv->majors.p, qv.data(), (uint32_t)qv.size(),
                         outs.data(), tu);
      lk.lock();
      if (rc != BUFF_OK) {
        sv->error = rc;
        const double t = now_s();
        for (uint64_t id : tk) {
          Req& r = sv->reqs[id];
          if (r.arena_off != ~0ull) (void)arena_give(sv, r.arena_off);
          r.arena_off = ~0ull;
          r.res.status = rc;
          r.t_done = t;
          r.done = true;
        }
        sv->cv_done.notify_all();
      } else {
        sv->slot_busy[s] = true;
        sv->batches += 1;
        sv->queries += tk.size();
        sv->max_seen = std::max<uint64_t>(sv->max_seen, tk.size());
      }
      progressed = true;
      rr = s + 1;
    }
    if (!progressed) {
      bool busy = false;
      for (uint32_t s = 0; s < nslots; ++s) busy |= sv->slot_busy[s];
      if (busy) {
        lk.unlock();
        std::this_thread::yield();   // poll the in-flight batches
        lk.lock();
      } else {
        sv->cv_work.wait(lk, [&] { return sv->stop || (!sv->pending.empty() && !sv->arena_blocked); });
      }
    }
  }
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
  if (bitmaps_dev) *bitmaps_dev = (const uint64_t*)b->ex.arena.p;
  if (values_dev) *values_dev = (const double*)b->ex.arena.p;
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

int buff_batch_set_tuning(buff_ctx* ctx, uint32_t grid, uint32_t tiles_per_item, uint32_t first_planes) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(BUFF_E_INVALID, "null argument");
  if (grid < 1 || grid > 65536) return fail(BUFF_E_INVALID, "grid %u outside [1, 65536]", grid);
  if (tiles_per_item < 1 || tiles_per_item > kMaxTilesPerItem)
    return fail(BUFF_E_INVALID, "tiles_per_item %u outside [1, %u]", tiles_per_item, kMaxTilesPerItem);
  if (first_planes != 1 && first_planes != kMaxPlanes)
    return fail(BUFF_E_INVALID, "first_planes %u is neither 1 nor %u", first_planes, kMaxPlanes);
  c->bt_grid = grid;
  c->bt_tiles = tiles_per_item;
  c->bt_first = first_planes;
  c->bt_tuned = true;
  return BUFF_OK;
}

int buff_server_create(buff_ctx* ctx, const buff_server_cfg* cfg, buff_server** out) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !out) return fail(BUFF_E_INVALID, "null argument");
  Server* sv = new (std::nothrow) Server();
  if (!sv) return fail(BUFF_E_OOM, "host allocation");
  sv->ctx = c;
  sv->cfg.max_batch = cfg && cfg->max_batch ? cfg->max_batch : 4096;
  sv->cfg.slots = cfg && cfg->slots ? std::min<uint32_t>(cfg->slots, 8) : 2;
  sv->cfg.arena_bytes = cfg && cfg->arena_bytes ? cfg->arena_bytes : (1ull << 30);
  sv->slots.resize(sv->cfg.slots);
  sv->slot_tickets.resize(sv->cfg.slots);
  sv->slot_qs.resize(sv->cfg.slots);
  sv->slot_chunks.resize(sv->cfg.slots);
  sv->slot_busy.assign(sv->cfg.slots, false);
  int rc = reserve(c, sv->arena, sv->cfg.arena_bytes);
  sv->arena_free[0] = sv->arena.bytes;
  buff_ctx* aux = nullptr;
  if (rc == BUFF_OK) rc = buff_ctx_create(c->device, 0, &aux);
  sv->aux = reinterpret_cast<Ctx*>(aux);
  for (auto& ex : sv->slots) {
    if (rc == BUFF_OK) rc = exec_init(ex, c, true);
    if (rc == BUFF_OK) rc = exec_reserve(ex, sv->cfg.max_batch, 0, ctx_tuning(c));
    ex.aux = sv->aux;
  }
  if (rc != BUFF_OK) {
    for (auto& e2 : sv->slots) exec_free(e2);
    drop(c, sv->arena);
    (void)buff_ctx_destroy(aux);
    delete sv;
    return rc;
  }
  sv->worker = std::thread(server_loop, sv);
  *out = reinterpret_cast<buff_server*>(sv);
  return BUFF_OK;
}

int buff_server_add_chunk(buff_server* server, buff_chunk* chunk, uint32_t* id) {
  Server* sv = reinterpret_cast<Server*>(server);
  if (!sv || !chunk) return fail(BUFF_E_INVALID, "null argument");
  Chunk* ch = reinterpret_cast<Chunk*>(chunk);
  if (ch->ctx != sv->ctx) return fail(BUFF_E_INVALID, "the chunk belongs to another context");
  if (ch->flags & BUFF_SLICE) return fail(BUFF_E_UNSUPPORTED, "served chunks use the byte_fixed, buff-major or residue layout");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (id) *id = (uint32_t)sv->chunks.size();
  sv->chunks.push_back(ch);
  sv->views_dirty = true;
  return BUFF_OK;
}

int buff_server_submit(buff_server* server, const buff_query* q, void* out_dev, uint64_t* ticket) {
  Server* sv = reinterpret_cast<Server*>(server);
  if (!sv || !q || !ticket) return fail(BUFF_E_INVALID, "null argument");
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) return fail(BUFF_E_INVALID, "server is shutting down");
    const uint64_t id = sv->next_ticket++;
    Req& r = sv->reqs[id];
    r.q = *q;
    r.out = (uint64_t)out_dev;
    r.t_submit = now_s();
    sv->pending.push_back(id);
    *ticket = id;
  }
  sv->cv_work.notify_one();
  return BUFF_OK;
}

int buff_server_poll(buff_server* server, uint64_t ticket, int* done) {
  Server* sv = reinterpret_cast<Server*>(server);
  if (!sv || !done) return fail(BUFF_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return fail(BUFF_E_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
  *done = it->second.done ? 1 : 0;
  return BUFF_OK;
}

int buff_server_release(buff_server* server, const void* out) {
  Server* sv = reinterpret_cast<Server*>(server);
  if (!sv) return fail(BUFF_E_INVALID, "null argument");
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    const uint64_t a = (uint64_t)out, base = (uint64_t)sv->arena.p;
    if (a < base || a >= base + sv->arena.bytes || !arena_give(sv, a - base))
      return fail(BUFF_E_INVALID, "%p is not a held output of this server's arena", out);
  }
  sv->cv_work.notify_one();
  return BUFF_OK;
}

int buff_server_arena_stats(buff_server* server, uint64_t* in_use, uint64_t* peak, uint64_t* capacity) {
  Server* sv = reinterpret_cast<Server*>(server);
  if (!sv) return fail(BUFF_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (in_use) *in_use = sv->arena_in_use;
  if (peak) *peak = sv->arena_peak;
  if (capacity) *capacity = sv->arena.bytes;
  return BUFF_OK;
}

int buff_server_wait(buff_server* server, uint64_t ticket, buff_query_result* res, double* latency_us) {
  Server* sv = reinterpret_cast<Server*>(server);
  if (!sv) return fail(BUFF_E_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  if (sv->reqs.find(ticket) == sv->reqs.end())
    return fail(BUFF_E_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
  // look the request up again after every wake-up: submissions from other
  // threads insert into the map meanwhile and may rehash it
  sv->cv_done.wait(lk, [&] { return sv->reqs.find(ticket)->second.done; });
  auto it = sv->reqs.find(ticket);
  if (res) *res = it->second.res;
  if (latency_us) *latency_us = (it->second.t_done - it->second.t_submit) * 1e6;
  sv->reqs.erase(it);
  return BUFF_OK;
}

int buff_server_run_trace(buff_server* server, const buff_query* qs, const double* arrival_s, void* const* outs,
                          uint32_t n, buff_query_result* res, double* latency_us) {
  Server* sv = reinterpret_cast<Server*>(server);
  if (!sv || (n && (!qs || !arrival_s))) return fail(BUFF_E_INVALID, "null argument");
  std::vector<uint64_t> tickets(n);
  uint32_t drained = 0;   // requests [0, drained) are collected
  // collects request `drained` (waiting for it when `block`), returning its arena
  // output range at once: outputs are not handed back, so the arena recycles
  // while later requests are still being submitted
  auto drain_one = [&](bool block, bool* took) -> int {
    *took = false;
    if (drained >= n) return BUFF_OK;
    if (!block) {
      int done = 0;
      BUFF_TRY(buff_server_poll(server, tickets[drained], &done));
      if (!done) return BUFF_OK;
    }
    const uint32_t i = drained;
    buff_query_result r{};
    BUFF_TRY(buff_server_wait(server, tickets[i], &r, latency_us ? &latency_us[i] : nullptr));
    if (r.status == BUFF_OK && !(outs && outs[i]) && (qs[i].op == BUFF_Q_GT || qs[i].op == BUFF_Q_EQ ||
                                                      qs[i].op == BUFF_Q_MAT)) {
      BUFF_TRY(buff_server_release(server, (const void*)r.out_offset));
      r.out_offset = 0;
    }
    if (res) res[i] = r;
    ++drained;
    *took = true;
    return BUFF_OK;
  };
  const double t0 = now_s();
  for (uint32_t i = 0; i < n; ++i) {
    while (now_s() - t0 < arrival_s[i]) {
      bool took = false;
      if (drained < i) BUFF_TRY(drain_one(false, &took));
      if (!took) std::this_thread::yield();
    }
    BUFF_TRY(buff_server_submit(server, &qs[i], outs ? outs[i] : nullptr, &tickets[i]));
  }
  while (drained < n) {
    bool took = false;
    BUFF_TRY(drain_one(true, &took));
  }
  return BUFF_OK;
}

int buff_trace_per_call(buff_ctx* ctx, buff_chunk* const* chunks, uint32_t nchunks, const buff_query* qs,
                        const double* arrival_s, void* const* outs, uint32_t n, buff_query_result* res,
                        double* latency_us) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || (n && (!qs || !arrival_s)) || (nchunks && !chunks)) return fail(BUFF_E_INVALID, "null argument");
  // output memory for requests without an address, sized before the replay starts
  uint64_t max_rows = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (qs[i].chunk >= nchunks) return fail(BUFF_E_INVALID, "request %u: chunk %u outside the list", i, qs[i].chunk);
    max_rows = std::max<uint64_t>(max_rows, reinterpret_cast<const Chunk*>(chunks[qs[i].chunk])->n);
  }
  BUFF_TRY(c->scratch(c->s_values, std::max<uint64_t>(max_rows * 8, 256)));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const double t0 = now_s();
  for (uint32_t i = 0; i < n; ++i) {
    while (now_s() - t0 < arrival_s[i]) std::this_thread::yield();
    const buff_query& q = qs[i];
    const Chunk* ch = reinterpret_cast<const Chunk*>(chunks[q.chunk]);
    void* out = (outs && outs[i]) ? outs[i] : c->s_values.ptr;
    buff_query_result r{};
    switch (q.op) {
      case BUFF_Q_GT:
      case BUFF_Q_EQ:
        r.status = filter_device(c, ch, q.op == BUFF_Q_GT ? BUFF_OP_GT : BUFF_OP_EQ, q.pred, q.flags,
                                 (uint64_t*)out, nullptr);
        r.out_offset = (outs && outs[i]) ? (uint64_t)out : 0;
        break;
      case BUFF_Q_SUM:
        r.status = sum_op(c, ch, &r.value, &r.exact, nullptr);
        break;
      case BUFF_Q_MAX:
        r.status = max_op(c, ch, q.flags, &r.value, &r.max_off, nullptr);
        break;
      case BUFF_Q_MAT:
        r.status = materialize_device(c, ch, q.precision, q.flags, (double*)out);
        r.out_offset = (outs && outs[i]) ? (uint64_t)out : 0;
        break;
      default:
        r.status = fail(BUFF_E_INVALID, "request %u: unknown op %u", i, q.op);
    }
    // a filter's bitmap or a materialization's values are on the device once
    // the stream has drained; a sum or max is complete when the call returns
    // (its result is on the host)
    if (r.status == BUFF_OK && q.op != BUFF_Q_SUM && q.op != BUFF_Q_MAX) {
      const hipError_t e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) return fail(BUFF_E_DEVICE, "per-call replay: %s", hipGetErrorString(e));
    }
    if (latency_us) latency_us[i] = (now_s() - t0 - arrival_s[i]) * 1e6;
    if (res) res[i] = r;
  }
  return BUFF_OK;
}

int buff_server_stats(buff_server* server, uint64_t* batches, uint64_t* queries, uint64_t* largest_batch) {
  Server* sv = reinterpret_cast<Server*>(server);
  if (!sv) return fail(BUFF_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (batches) *batches = sv->batches;
  if (queries) *queries = sv->queries;
  if (largest_batch) *largest_batch = sv->max_seen;
  return sv->error;
}

int buff_server_destroy(buff_server* server) {
  Server* sv = reinterpret_cast<Server*>(server);
  if (!sv) return BUFF_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv_work.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  for (auto& ex : sv->slots) exec_free(ex);
  drop(sv->ctx, sv->views);
  drop(sv->ctx, sv->majors);
  drop(sv->ctx, sv->arena);
  (void)buff_ctx_destroy(reinterpret_cast<buff_ctx*>(sv->aux));
  delete sv;
  return BUFF_OK;
}

}  // extern "C"
