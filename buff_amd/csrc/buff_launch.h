// Kernel launches that can carry timing events. While a profiled operation is
// armed (ProfScope on the calling thread), the first kernel it launches records
// the start event when it starts executing and every kernel records the stop
// event when it finishes (hipExtLaunchKernelGGL), so the measured interval is
// the kernels' own execution, first start to last end, without the dispatch
// latency in front of them. Unarmed, this is a plain launch.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>

namespace buffgpu {

// Workgroups of `kernel` (at this block size and dynamic LDS) the current
// device holds resident at once. Persistent grid-stride kernels launch at most
// this many: a larger grid leaves its last workgroups waiting for the first to
// finish, a tail in which part of the GPU idles.
template <typename K>
inline uint32_t resident_wgs(K kernel, uint32_t block, uint32_t shmem = 0) {
  static std::mutex mu;
  static std::map<std::tuple<const void*, uint32_t, uint32_t, int>, uint32_t> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const auto key = std::make_tuple(reinterpret_cast<const void*>(kernel), block, shmem, dev);
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, (int)block, (size_t)shmem) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
  const uint32_t v = (uint32_t)per_cu * (uint32_t)cus;
  std::lock_guard<std::mutex> lk(mu);
  cache[key] = v;
  return v;
}

struct LaunchEvents {
  hipEvent_t start = nullptr;
  hipEvent_t stop = nullptr;
  bool started = false;
};

// The calling thread's armed events (defined in buff_kernels.hip).
LaunchEvents& launch_events();

template <typename K, typename... Args>
inline void launch_k(K kernel, const dim3& grid, const dim3& block, uint32_t shmem, hipStream_t s, Args... args) {
  LaunchEvents& le = launch_events();
  if (le.stop) {
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, le.started ? nullptr : le.start, le.stop, 0, args...);
    le.started = true;
  } else {
    hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
  }
}

}  // namespace buffgpu
