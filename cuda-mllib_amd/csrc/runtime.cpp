// runtime.cpp -- device choice, allocation contract and error state.
//
// Reference: utilities.cu:23-94 (get_gpu, mallocBest, freeBest) and the
// error macros of utilities.h:25-34 / als/src/cuda/als.h:43-88, which
// exit() or throw.  Here nothing exits: failures are reported through
// return values and a thread-local message (cumf_last_error).
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "../../include/als.h"
#include "../../include/utilities.h"
#include "cumf_internal.h"

namespace cumf {

static thread_local std::string t_last_error;

void set_last_error(const std::string& msg) { t_last_error = msg; }
void clear_last_error() { t_last_error.clear(); }

}  // namespace cumf

namespace {

std::atomic<int> g_device_count{-1};
std::atomic<unsigned> g_next_device{0};
std::mutex g_init_lock;
// Test hook: when set, mallocBest skips device memory and takes the
// mapped-host fallback (CUMF_FORCE_HOST_ALLOC=1 or cumf_debug_force_host_alloc).
std::atomic<int> g_force_host{-1};

bool force_host_alloc() {
  int v = g_force_host.load();
  if (v < 0) {
    const char* e = std::getenv("CUMF_FORCE_HOST_ALLOC");
    v = (e && std::strcmp(e, "1") == 0) ? 1 : 0;
    g_force_host.store(v);
  }
  return v == 1;
}

}  // namespace

extern "C" const char* cumf_last_error(void) { return cumf::t_last_error.c_str(); }

extern "C" int cumf_debug_force_host_alloc(int on) {
  g_force_host.store(on ? 1 : 0);
  return 0;
}

// Round-robin over the visible devices.  The first call counts them (no
// device reset); every call then claims the next slot atomically and
// returns the slot it claimed.
extern "C" int get_gpu(void) {
  int count = g_device_count.load();
  if (count < 0) {
    std::lock_guard<std::mutex> lock(g_init_lock);
    count = g_device_count.load();
    if (count < 0) {
      int n = 0;
      if (hipGetDeviceCount(&n) != hipSuccess || n < 0) {
        (void)hipGetLastError();
        n = 0;
      }
      g_device_count.store(n);
      count = n;
    }
  }
  if (count <= 0) return -1;
  if (count == 1) return 0;
  return (int)(g_next_device.fetch_add(1u) % (unsigned)count);
}

hipError_t mallocBest(void** devPtr, size_t size) {
  if (!devPtr) return hipErrorInvalidValue;
  *devPtr = nullptr;
  if (!force_host_alloc()) {
    if (hipMalloc(devPtr, size) == hipSuccess) return hipSuccess;
    (void)hipGetLastError();
    *devPtr = nullptr;
  }
  void* host = nullptr;
  hipError_t e = hipHostMalloc(&host, size, hipHostMallocMapped | hipHostMallocWriteCombined);
  if (e != hipSuccess) return e;
  e = hipHostGetDevicePointer(devPtr, host, 0);
  if (e != hipSuccess) {
    (void)hipHostFree(host);
    *devPtr = nullptr;
  }
  return e;
}

hipError_t freeBest(void* devPtr) {
  if (!devPtr) return hipSuccess;
  hipPointerAttribute_t attr;
  std::memset(&attr, 0, sizeof(attr));
  hipError_t e = hipPointerGetAttributes(&attr, devPtr);
  if (e != hipSuccess) return e;
  if (attr.type == hipMemoryTypeDevice) return hipFree(devPtr);
  return hipHostFree(attr.hostPointer ? attr.hostPointer : devPtr);
}

extern "C" int cumf_malloc_best(void** devPtr, size_t size) { return (int)mallocBest(devPtr, size); }
extern "C" int cumf_free_best(void* devPtr) { return (int)freeBest(devPtr); }

extern "C" int cumf_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  CUMF_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return CUMF_OK;
}

extern "C" int cumf_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  CUMF_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return CUMF_OK;
}
