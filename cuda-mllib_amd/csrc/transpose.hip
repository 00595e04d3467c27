// transpose.hip -- CSR -> CSC (= CSR of R^T) on the device.
//
// The Theta half-sweep walks the ratings by item, so both orientations stay
// resident.  The reference's caller (Spark) builds each orientation on the
// host; here the transpose is a stable LSD radix sort of the ratings by
// column index (hipCUB), so rows stay ascending inside every column and the
// result is deterministic and identical to a host counting-sort transpose.
//
// Per rating, 4 bytes of key and 8 bytes of (row, value) payload are sorted;
// column pointers come from a binary search of the sorted keys.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "../../include/als.h"
#include "cumf_internal.h"

namespace {

// payload[k] = (row of k) << 32 | bits(val[k]); key[k] = col[k].
__global__ void pack_kernel(const int32_t* __restrict__ row_ptr, int64_t m, const int32_t* __restrict__ col,
                            const float* __restrict__ val, int64_t nnz, uint32_t* __restrict__ key,
                            uint64_t* __restrict__ payload) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * blockDim.x) {
    // Row u with row_ptr[u] <= k < row_ptr[u+1] (the last such u: empty rows
    // share pointers).
    int64_t lo = 0, hi = m;  // invariant: row_ptr[lo] <= k < row_ptr[hi]
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (row_ptr[mid] <= k) lo = mid; else hi = mid;
    }
    key[k] = (uint32_t)col[k];
    payload[k] = ((uint64_t)(uint32_t)lo << 32) | (uint64_t)__float_as_uint(val[k]);
  }
}

// col_ptr[j] = number of sorted keys < j, for j in [0, n].
__global__ void colptr_kernel(const uint32_t* __restrict__ key, int64_t nnz, int64_t n,
                              int32_t* __restrict__ col_ptr) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j <= n;
       j += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((int64_t)key[mid] < j) lo = mid + 1; else hi = mid;
    }
    col_ptr[j] = (int32_t)lo;
  }
}

__global__ void unpack_kernel(const uint64_t* __restrict__ payload, int64_t nnz, int32_t* __restrict__ row,
                              float* __restrict__ val) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t p = payload[k];
    row[k] = (int32_t)(p >> 32);
    val[k] = __uint_as_float((uint32_t)(p & 0xffffffffu));
  }
}

int bits_for(int64_t n) {
  int b = 1;
  while (b < 32 && ((int64_t)1 << b) < n) ++b;
  return b;
}

}  // namespace

// Transposes an m x n CSR (device arrays) into the CSC arrays col_ptr
// (n+1), row (nnz), val (nnz), also on the device.  Synchronises `stream`.
extern "C" int cumf_csr_to_csc_device(hipStream_t stream, int64_t m, int64_t n, int64_t nnz,
                                      const int32_t* d_row_ptr, const int32_t* d_col, const float* d_val,
                                      int32_t* d_col_ptr, int32_t* d_row, float* d_val_out) {
  cumf::clear_last_error();
  if (m <= 0 || n <= 0 || nnz < 0 || nnz > INT32_MAX || m > INT32_MAX || n > INT32_MAX || !d_row_ptr ||
      !d_col_ptr || (nnz > 0 && (!d_col || !d_val || !d_row || !d_val_out))) {
    cumf::set_last_error("cumf_csr_to_csc_device: bad arguments");
    return CUMF_ERR_INVALID;
  }
  if (nnz == 0) {
    CUMF_HIP_TRY(hipMemsetAsync(d_col_ptr, 0, sizeof(int32_t) * (n + 1), stream));
    CUMF_HIP_TRY(hipStreamSynchronize(stream));
    return CUMF_OK;
  }
  uint32_t *k0 = nullptr, *k1 = nullptr;
  uint64_t *p0 = nullptr, *p1 = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int rc = CUMF_OK;
  auto fail = [&](const char* what, hipError_t e) {
    cumf::set_last_error(std::string(what) + ": " + hipGetErrorString(e));
    rc = CUMF_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipMalloc(&k0, sizeof(uint32_t) * nnz)) != hipSuccess ||
      (e = hipMalloc(&k1, sizeof(uint32_t) * nnz)) != hipSuccess ||
      (e = hipMalloc(&p0, sizeof(uint64_t) * nnz)) != hipSuccess ||
      (e = hipMalloc(&p1, sizeof(uint64_t) * nnz)) != hipSuccess) {
    fail("transpose buffers", e);
  }
  const int bits = bits_for(n);
  if (rc == CUMF_OK) {
    hipLaunchKernelGGL(pack_kernel, dim3(4096), dim3(256), 0, stream, d_row_ptr, m, d_col, d_val, nnz, k0, p0);
    if ((e = hipGetLastError()) != hipSuccess) fail("pack_kernel", e);
  }
  if (rc == CUMF_OK &&
      (e = hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k0, k1, p0, p1, (int)nnz, 0, bits, stream)) !=
          hipSuccess)
    fail("radix sort sizing", e);
  if (rc == CUMF_OK && (e = hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 16)) != hipSuccess) fail("sort scratch", e);
  if (rc == CUMF_OK &&
      (e = hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k0, k1, p0, p1, (int)nnz, 0, bits, stream)) !=
          hipSuccess)
    fail("radix sort", e);
  if (rc == CUMF_OK) {
    hipLaunchKernelGGL(colptr_kernel, dim3(1024), dim3(256), 0, stream, k1, nnz, n, d_col_ptr);
    hipLaunchKernelGGL(unpack_kernel, dim3(4096), dim3(256), 0, stream, p1, nnz, d_row, d_val_out);
    if ((e = hipGetLastError()) != hipSuccess) fail("colptr/unpack", e);
  }
  if ((e = hipStreamSynchronize(stream)) != hipSuccess && rc == CUMF_OK) fail("transpose", e);
  (void)hipFree(k0);
  (void)hipFree(k1);
  (void)hipFree(p0);
  (void)hipFree(p1);
  (void)hipFree(tmp);
  return rc;
}
