// Host-side runtime objects behind the C ABI: context, chunk, scratch buffers,
// event timing. Internal to libbuffgpu.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "buff_common.h"
#include "buff_kernels.h"

namespace buffgpu {

int fail(int code, const char* fmt, ...);

inline uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

enum class Op : uint32_t { Encode = 0, FilterGt, FilterEq, Sum, Max, Materialize, kCount };

uint32_t tune_bucket(uint64_t ntiles);
inline uint64_t tune_key(Op op, uint64_t ntiles) { return ((uint64_t)op << 32) | tune_bucket(ntiles); }

struct Scratch {
  void* ptr = nullptr;
  uint64_t bytes = 0;
};

struct ProfRec {
  std::string name;
  hipEvent_t a, b;
};

struct Ctx {
  static constexpr uint64_t kPinnedBytes = 4096;

  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t budget = 0;  // 0 = unlimited
  uint64_t held = 0;    // bytes obtained from hipMalloc and not yet returned
  std::mutex mu;
  std::map<void*, uint64_t> live;
  std::multimap<uint64_t, void*> free_blocks;
  void* pinned = nullptr;

  Scratch s_partial, s_minmax, s_stream, s_bitmap, s_sums, s_best, s_values, s_sink;

  bool prof_on = false;
  std::vector<hipEvent_t> prof_pool;
  std::vector<ProfRec> prof_pending;
  std::map<std::string, std::pair<uint64_t, double>> prof_stats;

  std::map<uint64_t, LaunchShape> tuned;

  ~Ctx();
  int alloc(uint64_t bytes, void** out);
  int release(void* p);
  int scratch(Scratch& s, uint64_t bytes);
  void prof_begin(const char* name, hipEvent_t* ev);
  void prof_end(const char* name, hipEvent_t* ev);
  int prof_collect();
  LaunchShape shape(Op op, uint64_t ntiles) const;
};

struct ProfScope {
  Ctx* ctx;
  const char* name;
  hipEvent_t ev[2];
  ProfScope(Ctx* c, const char* n) : ctx(c), name(n) { ctx->prof_begin(name, ev); }
  ~ProfScope() { ctx->prof_end(name, ev); }
};

struct Chunk {
  Ctx* ctx = nullptr;
  uint64_t n = 0;
  uint64_t scale = 0;
  uint32_t flags = 0;  // BUFF_FLIP
  uint32_t dlen = 0;
  uint32_t ilen = 0;
  ChunkView view{};
  void* mem = nullptr;
  uint64_t device_bytes = 0;
};

}  // namespace buffgpu
