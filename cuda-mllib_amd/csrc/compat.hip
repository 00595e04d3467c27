// compat.hip -- reproduction of the reference's half-sweep arithmetic,
// including its Gram write slip, for callers that need the reference's
// factors rather than the exact ALS update.
//
// Reference behaviour restated (als/src/cuda/als.cu):
//   * get_hermitian100 / get_hermitianT10 (:47-912): every Gram entry is one
//     thread's FP32 register, updated with one fused multiply-add per rating
//     in rating order (zero-padded windows add +0, which changes nothing);
//   * the slip at :275 / :718: in every upper 10x10 tile (10I, 10J), I <= J,
//     the element (10I+4, 10J) of the column-major tt receives G(10I, 10J);
//   * ridge (:462-466, :905-909): tt[i][i] += (n_u * lambda), the product
//     formed in FP32, added in FP64;
//   * updateX (:935-997): FP64 LU without pivoting, FP64 triangular solves,
//     the solution stored as FP32.
// The RHS of the reference comes from cusparseScsrmm2 (summation order not
// published); here it is an FP32 fused multiply-add per rating in rating
// order.  The LU's operation order differs from cuBLAS's (not published).
//
// One 512-thread work-group per row; the FP64 matrix lives in LDS
// (f*f*8 bytes, f <= 140).  This is a parity path, not a fast path.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "als_kernels.h"

namespace cumf {

constexpr int kCompatThreads = 512;
constexpr int kCompatWindow = 24;  // factor rows staged per window
constexpr int kCompatMaxQ = (140 * 140 + kCompatThreads - 1) / kCompatThreads;

__global__ __launch_bounds__(512) void compat_row_kernel(CompatArgs a) {
  extern __shared__ double Msh[];
  const int f = a.f;
  const int tid = threadIdx.x;
  const int64_t u = a.rows ? a.rows[blockIdx.x] : (int64_t)blockIdx.x;
  const int64_t s = a.row_ptr[u], e = a.row_ptr[u + 1];
  const int n = (int)(e - s);
  float* stage = reinterpret_cast<float*>(Msh);       // kCompatWindow x f
  float* rwin = stage + kCompatWindow * f;            // kCompatWindow

  // Entry q of this thread: e = tid + 512 q, (i, j) = (e / f, e % f).
  float acc[kCompatMaxQ];
  int ij[kCompatMaxQ];  // (i << 16) | j of entry q, or -1
#pragma unroll
  for (int q = 0; q < kCompatMaxQ; ++q) {
    acc[q] = 0.0f;
    const int ee = tid + kCompatThreads * q;
    ij[q] = ee < f * f ? (((ee / f) << 16) | (ee % f)) : -1;
  }
  float bacc = 0.0f;

  for (int64_t w = 0; w < n; w += kCompatWindow) {
    const int cnt = (int)((n - w) < kCompatWindow ? (n - w) : kCompatWindow);
    __syncthreads();
    for (int t = tid; t < cnt * f; t += kCompatThreads) {
      const int k = t / f, i = t - k * f;
      stage[t] = a.theta[(int64_t)a.col[s + w + k] * a.ldf + i];
    }
    if (tid < cnt) rwin[tid] = a.val[s + w + tid];
    __syncthreads();
    for (int k = 0; k < cnt; ++k) {
      const float* th = stage + k * f;
#pragma unroll
      for (int q = 0; q < kCompatMaxQ; ++q)
        if (ij[q] >= 0) acc[q] = fmaf(th[ij[q] >> 16], th[ij[q] & 0xffff], acc[q]);
      if (tid < f) bacc = fmaf(rwin[k], th[tid], bacc);
    }
  }
  __syncthreads();
  double* M = Msh;            // f x f row-major, M[i*f+j] = matrix element (i, j)
  double* bsh = Msh + f * f;  // f
#pragma unroll
  for (int q = 0; q < kCompatMaxQ; ++q) {
    const int ee = tid + kCompatThreads * q;
    if (ee < f * f) M[ee] = (double)acc[q];
  }
  if (tid < f) bsh[tid] = (double)bacc;
  __syncthreads();
  // The slip: element (10I+4, 10J) := G(10I, 10J).  Sources (rows = 0 mod
  // 10) are never targets (rows = 4 mod 10).
  const int nt = f / 10;
  for (int p = tid; p < nt * nt; p += kCompatThreads) {
    const int I = p / nt, J = p - I * nt;
    if (I <= J) M[(10 * I + 4) * f + 10 * J] = M[(10 * I) * f + 10 * J];
  }
  __syncthreads();
  if (tid < f) M[tid * f + tid] += (double)((float)n * a.lambda);
  __syncthreads();
  if (a.dump) {  // test hook: the system the solver sees
    for (int t = tid; t < f * f; t += kCompatThreads) a.dump[(int64_t)blockIdx.x * f * (f + 1) + t] = M[t];
    if (tid < f) a.dump[(int64_t)blockIdx.x * f * (f + 1) + f * f + tid] = bsh[tid];
    return;
  }
  float* xrow = a.x + u * a.ldf;
  if (n == 0) {
    for (int i = tid; i < a.ldf; i += kCompatThreads) xrow[i] = 0.0f;
    return;
  }
  // LU without pivoting (Doolittle, right-looking).
  __shared__ int bad;
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int k = 0; k < f; ++k) {
    const double piv = M[k * f + k];
    if (piv == 0.0) {
      if (tid == 0) bad = 1;
      break;  // uniform: every thread read the same pivot
    }
    for (int i = k + 1 + tid; i < f; i += kCompatThreads) M[i * f + k] /= piv;
    __syncthreads();
    const int rem = f - k - 1;
    for (int t = tid; t < rem * rem; t += kCompatThreads) {
      const int i = k + 1 + t / rem, j = k + 1 + t % rem;
      M[i * f + j] -= M[i * f + k] * M[k * f + j];
    }
    __syncthreads();
  }
  __syncthreads();
  if (bad) {
    for (int i = tid; i < a.ldf; i += kCompatThreads) xrow[i] = 0.0f;
    if (tid == 0) atomicAdd(a.fail_counter, 1);
    return;
  }
  // L y = b (unit lower), then U x = y; column-oriented.
  for (int k = 0; k < f; ++k) {
    const double yk = bsh[k];
    __syncthreads();
    for (int i = k + 1 + tid; i < f; i += kCompatThreads) bsh[i] -= M[i * f + k] * yk;
    __syncthreads();
  }
  for (int k = f - 1; k >= 0; --k) {
    if (tid == 0) bsh[k] /= M[k * f + k];
    __syncthreads();
    const double xk = bsh[k];
    for (int i = tid; i < k; i += kCompatThreads) bsh[i] -= M[i * f + k] * xk;
    __syncthreads();
  }
  for (int i = tid; i < a.ldf; i += kCompatThreads) xrow[i] = i < f ? (float)bsh[i] : 0.0f;
}

size_t compat_lds_bytes(int f) {
  const size_t m = sizeof(double) * (size_t)f * (f + 1);
  const size_t st = sizeof(float) * (size_t)kCompatWindow * (f + 1);
  return m > st ? m : st;
}

// Raises the kernel's dynamic-LDS limit to what f needs (once per device,
// outside any stream capture).
hipError_t prepare_compat(int f) {
  if (f < 10 || f > 140 || f % 10 != 0) return hipErrorInvalidValue;
  return hipFuncSetAttribute((const void*)compat_row_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)compat_lds_bytes(140));
}

hipError_t launch_compat(const CompatArgs& a, int64_t nrows, hipStream_t s) {
  if (a.f < 10 || a.f > 140 || a.f % 10 != 0) return hipErrorInvalidValue;
  if (nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(compat_row_kernel, dim3((unsigned)nrows), dim3(kCompatThreads), compat_lds_bytes(a.f), s, a);
  return hipGetLastError();
}

}  // namespace cumf
