// fp_pool.hip -- the per-device memory pool behind every solver allocation.
//
// A scene's fixed buffers are carved from one block (its arena); growing
// staging buffers (emit, grab) are blocks of their own.  Blocks come from a
// per-device cache in size classes of a quarter octave (at most 25% above the
// request) and go back to it when a scene is destroyed, so admitting and
// retiring scenes in a serving loop neither calls hipMalloc once the cache
// is warm nor calls hipFree (which waits for the whole device).  The cache
// is returned to the runtime by fp_pool_trim.
#include <map>
#include <mutex>
#include <unordered_map>

#include "fp_internal.h"

namespace {

struct DevicePool {
  std::multimap<size_t, void*> free_blocks;      // class bytes -> block
  std::unordered_map<void*, size_t> class_of;   // every block the pool made
  int64_t reserved = 0;                          // bytes of all blocks
  int64_t cached = 0;                            // bytes of free blocks
};

std::mutex g_pool_mutex;
std::map<int, DevicePool> g_pools;

}  // namespace

size_t fp_pool_class(size_t bytes) {
  if (bytes <= 4096) return 4096;
  int k = 63 - __builtin_clzll((unsigned long long)bytes);  // 2^k <= bytes < 2^(k+1)
  const size_t step = (size_t)1 << (k - 2);
  return (bytes + step - 1) / step * step;
}

hipError_t fp_pool_get(int device, size_t bytes, void** out, size_t* class_bytes) {
  const size_t cb = fp_pool_class(bytes);
  {
    std::lock_guard<std::mutex> lk(g_pool_mutex);
    DevicePool& p = g_pools[device];
    auto it = p.free_blocks.find(cb);
    if (it != p.free_blocks.end()) {
      *out = it->second;
      p.free_blocks.erase(it);
      p.cached -= (int64_t)cb;
      if (class_bytes) *class_bytes = cb;
      return hipSuccess;
    }
  }
  void* ptr = nullptr;
  hipError_t e = hipMalloc(&ptr, cb);
  if (e != hipSuccess) {
    // the cache may hold what the runtime lacks: give it back and retry once
    (void)hipGetLastError();
    if (fp_pool_trim_device(device) > 0) e = hipMalloc(&ptr, cb);
    if (e != hipSuccess) return e;
  }
  std::lock_guard<std::mutex> lk(g_pool_mutex);
  DevicePool& p = g_pools[device];
  p.class_of[ptr] = cb;
  p.reserved += (int64_t)cb;
  *out = ptr;
  if (class_bytes) *class_bytes = cb;
  return hipSuccess;
}

size_t fp_pool_put(int device, void* ptr) {
  if (!ptr) return 0;
  std::lock_guard<std::mutex> lk(g_pool_mutex);
  DevicePool& p = g_pools[device];
  auto it = p.class_of.find(ptr);
  if (it == p.class_of.end()) return 0;  // not a pool block
  p.free_blocks.emplace(it->second, ptr);
  p.cached += (int64_t)it->second;
  return it->second;
}

int64_t fp_pool_trim_device(int device) {
  std::vector<void*> blocks;
  int64_t freed = 0;
  {
    std::lock_guard<std::mutex> lk(g_pool_mutex);
    DevicePool& p = g_pools[device];
    for (auto& kv : p.free_blocks) {
      blocks.push_back(kv.second);
      p.class_of.erase(kv.second);
      freed += (int64_t)kv.first;
    }
    p.free_blocks.clear();
    p.reserved -= freed;
    p.cached = 0;
  }
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(device);
  for (void* b : blocks) (void)hipFree(b);
  (void)hipSetDevice(cur);
  return freed;
}

extern "C" {

int fp_pool_info(int32_t device, int64_t* reserved, int64_t* cached) {
  if (!reserved || !cached) return FP_EINVAL;
  if (device < 0 && hipGetDevice(&device) != hipSuccess) device = 0;
  std::lock_guard<std::mutex> lk(g_pool_mutex);
  auto it = g_pools.find(device);
  *reserved = it == g_pools.end() ? 0 : it->second.reserved;
  *cached = it == g_pools.end() ? 0 : it->second.cached;
  return FP_OK;
}

int fp_pool_trim(int32_t device) {
  if (device < 0 && hipGetDevice(&device) != hipSuccess) device = 0;
  (void)fp_pool_trim_device(device);
  return FP_OK;
}

}  // extern "C"
