// noalloc.hip -- a half-sweep entry point that allocates nothing.
//
// cumf_half_sweep_device allocates its task plan, partial-Gram scratch and
// counters and reads the row pointers back to plan.  Callers that cannot
// allocate or synchronise (code under HIP graph capture, tight serving
// loops) use cumf_half_sweep_device_noalloc instead: same arithmetic (FP32
// Gram on MFMA, FP32 Cholesky), no host work beyond argument checks, no
// device allocation and no synchronisation.  A persistent grid of
// wavefronts takes whole rows from a counter; the counter and the
// failed-row count live in `scratch`, a file-scope device array of
// kScratchInts ints that each GPU holds for the life of the process.
//
// The array's name is a host-side stand-in: its device address comes from
// hipGetSymbolAddress, once per device, on the first call made with that
// device current, and only that address is handed to kernels.  Every call
// on a GPU shares that GPU's array: the counters are cleared at the start
// of each call (by a kernel on the caller's stream, so the call can be
// captured), and calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <climits>
#include <mutex>
#include <string>

#include "../../include/als.h"
#include "als_kernels.h"
#include "cumf_internal.h"

namespace {
constexpr int kScratchInts = 64;  // see scratch below
}  // namespace

// [0] next row, [1] rows with a non-positive pivot; 256 bytes per GPU.
static __device__ __align__(256) int scratch[kScratchInts];

namespace {
constexpr int kMaxDevices = 64;
constexpr int kMaxFusedF = 140;
constexpr int64_t kMaxRows = INT_MAX - (1 << 20);  // the row counter may overshoot by one per wavefront

__global__ void clear_scratch_kernel(int* p) {
  if (threadIdx.x < kScratchInts) p[threadIdx.x] = 0;
}

struct DeviceScratch {
  int* ptr = nullptr;
  int num_cus = 0;
};

std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

// The device address of `scratch` on the current device (resolved once per
// device) and its compute-unit count.
int device_scratch(DeviceScratch* out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    cumf::set_last_error("cumf_half_sweep_device_noalloc: no current device");
    return CUMF_ERR_NODEVICE;
  }
  std::lock_guard<std::mutex> g(g_mu);
  DeviceScratch& d = g_dev[dev];
  if (!d.ptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      cumf::set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return CUMF_ERR_HIP;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) {
      (void)hipGetLastError();
      cus = 256;
    }
    d.ptr = static_cast<int*>(p);
    d.num_cus = cus;
  }
  *out = d;
  return CUMF_OK;
}

}  // namespace

extern "C" int cumf_half_sweep_device_noalloc(hipStream_t stream, const int32_t* d_row_ptr, const int32_t* d_col,
                                              const float* d_val, const float* d_theta, float* d_x, int64_t m,
                                              int64_t n, int f, int64_t nnz, float lambda) {
  cumf::clear_last_error();
  if (f < 1 || f > kMaxFusedF || f % 4 != 0) {
    cumf::set_last_error("cumf_half_sweep_device_noalloc: f must be a multiple of 4 in [4, 140] (factor rows are "
                         "read in place, 16 bytes at a time); use cumf_half_sweep_device otherwise");
    return CUMF_ERR_INVALID;
  }
  if (m <= 0 || n <= 0 || nnz < 0 || nnz > INT32_MAX || !d_row_ptr || !d_theta || !d_x ||
      (nnz > 0 && (!d_col || !d_val)) || !(lambda >= 0.f)) {
    cumf::set_last_error("cumf_half_sweep_device_noalloc: bad arguments");
    return CUMF_ERR_INVALID;
  }
  if (m > kMaxRows) {
    cumf::set_last_error("cumf_half_sweep_device_noalloc: at most " + std::to_string(kMaxRows) +
                         " rows (the row counter in its 256-byte scratch array is an int)");
    return CUMF_ERR_INVALID;
  }
  DeviceScratch ds;
  CUMF_TRY(device_scratch(&ds));
  const int variant = cumf::sweep_nb(f) >= 9 ? 1 : 2;
  cumf::SweepArgs a{};
  a.row_ptr = d_row_ptr;
  a.col = d_col;
  a.val = d_val;
  a.theta = d_theta;
  a.x = d_x;
  a.ldf = f;
  a.tasks = nullptr;  // whole rows, taken in index order from the counter
  a.ntasks = m;
  a.fail_counter = ds.ptr + 1;
  a.f = f;
  a.lambda = lambda;
  a.persistent = ds.num_cus * 4 * variant;  // one wavefront per SIMD per register-budget slot
  a.task_counter = ds.ptr;
  hipLaunchKernelGGL(clear_scratch_kernel, dim3(1), dim3(kScratchInts), 0, stream, ds.ptr);
  CUMF_HIP_TRY(hipGetLastError());
  CUMF_HIP_TRY(variant == 1 ? cumf::w1::chol::launch_task_kernel(a, stream)
                            : cumf::w2::chol::launch_task_kernel(a, stream));
  return CUMF_OK;
}
