// buff_sum_noalloc: byte_fixed_sum (split_double.rs:2127-2208) on a resident
// plain chunk that allocates nothing. k_sum's per-workgroup rows and the nine
// result words live in a file-scope device array instead of context scratch,
// for callers that cannot allocate device memory (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).
//
// The array holds (kMaxSumRows + 1) rows of kMaxPlanes + 1 u64: the 9 results
// and the most rows k_sum ever writes (its grid is capped at kMaxSumRows
// workgroups whatever the column length), 589,896 bytes per GPU, held until
// the process ends. k_sum writes every row it later reads and k_sum_final every
// result word, so nothing has to start at zero. Every call on one GPU uses the
// same array: calls that overlap on one GPU (from several threads or contexts)
// race; each call returns only after its stream has finished with the array.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "../../include/buffgpu.h"
#include "buff_common.h"
#include "buff_kernels.h"
#include "buff_runtime.h"

namespace buffgpu {
namespace {

constexpr uint64_t kNoAllocScratchBytes = (uint64_t)(kMaxSumRows + 1) * (kMaxPlanes + 1) * 8;

// Device memory only: host code uses the per-device address from
// hipGetSymbolAddress below, never this name.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// The array's address on `device`, looked up once per device with that device
// current (each GPU holds its own copy at its own address).
int scratch_on(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return fail(BUFF_E_INVALID, "device %d", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != device) HIP_TRY(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return fail(BUFF_E_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_dev[device] = p;
  }
  *out = g_scratch_dev[device];
  return BUFF_OK;
}

int sum_noalloc(Ctx* ctx, const Chunk* ch, double* ref_sum, double* exact_sum, uint64_t* plane_sums) {
  if (!ctx || !ch) return fail(BUFF_E_INVALID, "null argument");
  if (ch->ctx != ctx) return fail(BUFF_E_INVALID, "the chunk belongs to another context");
  if (ch->flags & (BUFF_MAJORITY | BUFF_SLICE))
    return fail(BUFF_E_UNSUPPORTED, "buff_sum_noalloc takes plain (byte_fixed) chunks");
  void* dev = nullptr;
  BUFF_TRY(scratch_on(ctx->device, &dev));
  uint64_t ps[kMaxPlanes + 1] = {0};
  const ChunkView& v = ch->view;
  if (ch->n > 0 && (v.P || v.r)) {
    const LaunchShape ls = ctx->shape(Op::Sum, v);
    const uint64_t need = sum_scratch_bytes(ls);   // the 9 results, then at most kMaxSumRows rows
    if (need > kNoAllocScratchBytes)
      return fail(BUFF_E_UNSUPPORTED, "the sum needs %llu bytes of scratch, above the %llu of buff_sum_noalloc",
                  (unsigned long long)need, (unsigned long long)kNoAllocScratchBytes);
    HIP_TRY(launch_sum(v, (unsigned long long*)dev, ctx->stream, ls));
    BUFF_TRY(d2h_small(ctx, ps, dev, sizeof(ps)));   // through the context's mailbox, made at creation
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

}  // namespace
}  // namespace buffgpu

extern "C" int buff_sum_noalloc(buff_ctx* ctx, const buff_chunk* chunk, double* ref_sum, double* exact_sum,
                                uint64_t* plane_sums) {
  return buffgpu::sum_noalloc(reinterpret_cast<buffgpu::Ctx*>(ctx), reinterpret_cast<const buffgpu::Chunk*>(chunk),
                              ref_sum, exact_sum, plane_sums);
}
