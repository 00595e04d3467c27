// als_kernels.hip -- common gfx950 kernels of the ALS library: Gram dump
// (tests), held-out RMSE, BF16 factor copies, MFMA layout probes, the copy
// kernel used to calibrate byte counters, and the CG fallback-list reset.
// The half-sweep kernels are in als_sweep.hip; their building blocks (Gram
// accumulation, in-register Cholesky, CG) in als_device.h.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "als_device.h"

namespace cumf {

// Clears n counters (a kernel, not a memset, so that a captured sweep
// replays it reliably).
__global__ void zero_ints_kernel(int* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// Gram dump for tests: the full augmented (16NB)^2 matrix of each listed
// row (ridge included), row-major, upper tiles mirrored.
template <int NB, bool BF>
__global__ __launch_bounds__(64) void als_dump_kernel(SweepArgs a, const int32_t* rows, float* out) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int u = rows[blockIdx.x];
  f4 acc[num_tiles(NB)];
#pragma unroll
  for (int i = 0; i < num_tiles(NB); ++i) acc[i] = f4{0.f, 0.f, 0.f, 0.f};
  const int64_t s = a.row_ptr[u];
  const int cnt = a.row_ptr[u + 1] - a.row_ptr[u];
  if (cnt > 0) gram_any<NB, BF>(acc, a, s, cnt, lds, lane);
  add_ridge<NB>(acc, a.f, a.lambda * (float)cnt, lane);
  const int D = 16 * NB;
  float* o = out + (int64_t)blockIdx.x * D * D;
#pragma unroll
  for (int I = 0; I < NB; ++I)
#pragma unroll
    for (int J = I; J < NB; ++J)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * I + 4 * g + r, j = 16 * J + c;
        o[i * D + j] = acc[tile_index(I, J, NB)][r];
        if (I != J) o[j * D + i] = acc[tile_index(I, J, NB)][r];
      }
}

// Held-out squared error: 16 lanes per rating, FP64 dot and sum; one
// partial per block.
__global__ __launch_bounds__(256) void rmse_partial_kernel(const int32_t* __restrict__ u,
                                                           const int32_t* __restrict__ v,
                                                           const float* __restrict__ r, int64_t cnt,
                                                           const float* __restrict__ X,
                                                           const float* __restrict__ T, int64_t ldf,
                                                           int f, double* __restrict__ partial) {
  __shared__ double red[16];
  const int lane16 = threadIdx.x & 15;
  const int64_t grp = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  const int64_t ngrp = ((int64_t)gridDim.x * blockDim.x) >> 4;
  double se = 0.0;
  for (int64_t k = grp; k < cnt; k += ngrp) {
    const float* x = X + (int64_t)u[k] * ldf;
    const float* t = T + (int64_t)v[k] * ldf;
    double p = 0.0;
    for (int i = lane16; i < f; i += 16) p += (double)x[i] * (double)t[i];
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) p += __shfl_xor(p, o, 16);
    if (lane16 == 0) {
      const double d = (double)r[k] - p;
      se += d * d;
    }
  }
  // block reduction in a fixed order
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) se += __shfl_xor(se, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = se;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    partial[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(64) void sum_partials_kernel(const double* __restrict__ partial, int n,
                                                          double* __restrict__ out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 64) s += partial[i];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
  if (threadIdx.x == 0) *out = s;
}

// One v_mfma_f32_16x16x4_f32 with A[i][k] = 100 i + k, B[k][j] = 1000 k + j
// fed in the operand order the Gram kernel uses; writes each lane's four
// result registers.  Lets a test pin the lane layout of the builtin.
__global__ __launch_bounds__(64) void mfma_layout_kernel(float* out) {
  const int l = threadIdx.x;
  const float av = 100.0f * (float)(l & 15) + (float)(l >> 4);
  const float bv = 1000.0f * (float)(l >> 4) + (float)(l & 15);
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) out[l * 4 + r] = acc[r];
}

// FP32 rows (stride ldf) -> BF16 rows (stride ldb, zero padded).
__global__ __launch_bounds__(256) void to_bf16_kernel(const float* __restrict__ src, int64_t ldf,
                                                      __bf16* __restrict__ dst, int64_t ldb, int64_t rows, int f) {
  const int64_t total = rows * ldb;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = k / ldb, c = k - r * ldb;
    dst[k] = (__bf16)(c < f ? src[r * ldf + c] : 0.0f);
  }
}

// One v_mfma_f32_16x16x32_bf16 with A[i][k] = i + k/32... (small integers,
// exact in bf16) fed as lane l: A[l%16][8(l/16)+j]; writes the 64 x 4
// result registers.  Pins the operand and result layout of the builtin.
__global__ __launch_bounds__(64) void mfma_bf16_layout_kernel(float* out) {
  const int l = threadIdx.x;
  bf8 av, bv;
  for (int t = 0; t < 8; ++t) {
    const int k = 8 * (l >> 4) + t;
    av[t] = (__bf16)(float)((l & 15) + 2 * (k & 3));   // A[i][k] = i + 2 (k mod 4)
    bv[t] = (__bf16)(float)((k % 5) + 3 * (l & 15));   // B[k][j] = k mod 5 + 3 j
  }
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) out[l * 4 + r] = acc[r];
}

// Plain device copy used to calibrate the HBM byte counters.
__global__ __launch_bounds__(256) void copy_kernel(const f4* __restrict__ src, f4* __restrict__ dst,
                                                   int64_t n4) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
// 16-column blocks of the augmented tile grid: room for the FP32 layout
// (factor columns to ldf, rating at ldf) and the BF16 layout (factor
// columns to ldb = round_up(f,8), rating hi/lo at f, f+1).
int sweep_nb(int f) {
  const int ldf = (f + 3) & ~3, ldb = (f + 7) & ~7;
  int w = ldf + 1;
  if (ldb > w) w = ldb;
  if (f + 2 > w) w = f + 2;
  return (w + 15) / 16;
}

hipError_t launch_fallback_reset(int* count, int n, hipStream_t s) {
  if (n < 1 || n > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, s, count, n);
  return hipGetLastError();
}
template <int NB>
static hipError_t launch_dump_nb(const SweepArgs& a, const int32_t* rows, int64_t nrows, float* out, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  const size_t shmem = sizeof(float) * (size_t)lds_floats_nb(NB, a.precision != 0);
  if (a.precision != 0)
    hipLaunchKernelGGL((als_dump_kernel<NB, true>), dim3((unsigned)nrows), dim3(64), shmem, s, a, rows, out);
  else
    hipLaunchKernelGGL((als_dump_kernel<NB, false>), dim3((unsigned)nrows), dim3(64), shmem, s, a, rows, out);
  return hipGetLastError();
}

hipError_t launch_dump_kernel(const SweepArgs& a, const int32_t* rows, int64_t nrows, float* out,
                              hipStream_t s) {
  switch (sweep_nb(a.f)) {
    case 1: return launch_dump_nb<1>(a, rows, nrows, out, s);
    case 2: return launch_dump_nb<2>(a, rows, nrows, out, s);
    case 3: return launch_dump_nb<3>(a, rows, nrows, out, s);
    case 4: return launch_dump_nb<4>(a, rows, nrows, out, s);
    case 5: return launch_dump_nb<5>(a, rows, nrows, out, s);
    case 6: return launch_dump_nb<6>(a, rows, nrows, out, s);
    case 7: return launch_dump_nb<7>(a, rows, nrows, out, s);
    case 8: return launch_dump_nb<8>(a, rows, nrows, out, s);
    case 9: return launch_dump_nb<9>(a, rows, nrows, out, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_rmse(const int32_t* u, const int32_t* v, const float* r, int64_t cnt, const float* X,
                       const float* T, int64_t ldf, int f, double* partial, int nblocks, double* out,
                       hipStream_t s) {
  hipLaunchKernelGGL(rmse_partial_kernel, dim3(nblocks), dim3(256), 0, s, u, v, r, cnt, X, T, ldf, f,
                     partial);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(64), 0, s, partial, nblocks, out);
  return hipGetLastError();
}

hipError_t launch_mfma_layout(float* out, hipStream_t s) {
  hipLaunchKernelGGL(mfma_layout_kernel, dim3(1), dim3(64), 0, s, out);
  return hipGetLastError();
}

hipError_t launch_to_bf16(const float* src, int64_t ldf, void* dst, int64_t ldb, int64_t rows, int f,
                          hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(to_bf16_kernel, dim3(2048), dim3(256), 0, s, src, ldf, (__bf16*)dst, ldb, rows, f);
  return hipGetLastError();
}

hipError_t launch_mfma_bf16_layout(float* out, hipStream_t s) {
  hipLaunchKernelGGL(mfma_bf16_layout_kernel, dim3(1), dim3(64), 0, s, out);
  return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t s) {
  const int64_t n4 = bytes / 16;
  hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, s, (const f4*)src, (f4*)dst, n4);
  return hipGetLastError();
}

size_t partial_floats_per_slot(int f) {
  const int NB = sweep_nb(f);
  return (size_t)num_tiles(NB) * 256;
}

int dump_dim(int f) { return 16 * sweep_nb(f); }

}  // namespace cumf
