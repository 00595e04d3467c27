// als_large.hip -- the half-sweep for 140 < f <= 256 (FP32 Gram + FP32
// Cholesky), where the fused kernels' register-resident Gram tiles no
// longer fit a wavefront.
//
// One 256-thread work-group per row (reference semantics as the fused path:
// als/src/cuda/als.cu:47-51, 462-466 and 905-909 for the Gram and ridge,
// :1039-1054 for b, :935-997 for the solve):
//   * Gram: chunks of 16 ratings are staged in LDS as rows of the augmented
//     [y r] (width W = 8 ceil((f+1)/8), zero padded); each thread owns up to
//     three 8x8 blocks (I <= J) of the augmented upper triangle and
//     accumulates them in registers with FP32 fused multiply-adds.
//   * The blocks are written to LDS as the packed upper triangle of the
//     (f+1) x (f+1) augmented matrix (<= 33,153 floats, 130 KB at f = 256),
//     the ridge lambda n_u is added to the diagonal, and a blocked
//     right-looking Cholesky runs in place (panels of 8 pivots, then a
//     rank-8 update of the trailing rows in 4x4 register tiles); the
//     augmented column becomes z = U^{-T} b.
//   * Back substitution U x = z in wavefront 0, z in registers.
// Rows without ratings get x = 0; a non-positive pivot counts a failed row
// and gives x = 0.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "als_kernels.h"

namespace cumf {
namespace {

constexpr int kLT = 256;      // threads per row
constexpr int kLC = 16;       // ratings per staged chunk
constexpr int kMaxPairs = 3;  // 8x8 block pairs per thread: ceil(33*34/2 / 256)
constexpr int kLargeMaxF = 256;

__host__ __device__ constexpr int packed_floats(int f) { return (f + 1) * (f + 2) / 2; }
// Offset of row i in the packed upper triangle of an n x n matrix, minus i,
// so that element (i, j >= i) is at row_base(i, n) + j.
__device__ __forceinline__ int row_base(int i, int n) { return i * n - (i * (i - 1)) / 2 - i; }

__global__ __launch_bounds__(kLT) __attribute__((amdgpu_waves_per_eu(1))) void als_large_kernel(SweepArgs a) {
  extern __shared__ float lds[];
  const int t = threadIdx.x;
  const int64_t u = blockIdx.x;
  const int f = a.f, F1 = f + 1, FB = (F1 + 7) / 8, W = 8 * FB;
  const int64_t s0 = a.row_ptr[u];
  const int cnt = a.row_ptr[u + 1] - a.row_ptr[u];
  float* xrow = a.x + u * a.ldf;
  if (cnt <= 0) {
    for (int i = t; i < (int)a.ldf; i += kLT) xrow[i] = 0.0f;
    return;
  }

  // This thread's block pairs (I, J), I <= J < FB, numbered row by row.
  const int nbp = FB * (FB + 1) / 2;
  int pI[kMaxPairs], pJ[kMaxPairs];
#pragma unroll
  for (int q = 0; q < kMaxPairs; ++q) {
    int p = t + q * kLT;
    pI[q] = -1;
    pJ[q] = -1;
    if (p < nbp) {
      int I = 0;
      while (p >= FB - I) {
        p -= FB - I;
        ++I;
      }
      pI[q] = I;
      pJ[q] = I + p;
    }
  }
  float acc[kMaxPairs][8][8];
#pragma unroll
  for (int q = 0; q < kMaxPairs; ++q)
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
      for (int s = 0; s < 8; ++s) acc[q][r][s] = 0.0f;

  float* Y = lds;  // [kLC][W]
  for (int c0 = 0; c0 < cnt; c0 += kLC) {
    const int nk = min(kLC, cnt - c0);
    __syncthreads();  // the previous chunk has been consumed
    for (int e = t; e < kLC * W; e += kLT) {
      const int k = e / W, i = e - k * W;
      float v = 0.0f;
      if (k < nk) {
        if (i < f) v = a.theta[(int64_t)a.col[s0 + c0 + k] * a.ldf + i];
        else if (i == f) v = a.val[s0 + c0 + k];
      }
      Y[e] = v;
    }
    __syncthreads();
    for (int k = 0; k < nk; ++k) {
      const float* yk = Y + k * W;
#pragma unroll
      for (int q = 0; q < kMaxPairs; ++q) {
        if (pI[q] < 0) continue;
        const float4 a0 = *reinterpret_cast<const float4*>(yk + 8 * pI[q]);
        const float4 a1 = *reinterpret_cast<const float4*>(yk + 8 * pI[q] + 4);
        const float4 b0 = *reinterpret_cast<const float4*>(yk + 8 * pJ[q]);
        const float4 b1 = *reinterpret_cast<const float4*>(yk + 8 * pJ[q] + 4);
        const float av[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
        const float bv[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
          for (int s = 0; s < 8; ++s) acc[q][r][s] = fmaf(av[r], bv[s], acc[q][r][s]);
      }
    }
  }
  __syncthreads();  // staging done; the packed matrix reuses the space

  float* P = lds;  // packed upper triangle, F1 x F1
#pragma unroll
  for (int q = 0; q < kMaxPairs; ++q) {
    if (pI[q] < 0) continue;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int i = 8 * pI[q] + r;
      if (i >= F1) continue;
      const int rb = row_base(i, F1);
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const int j = 8 * pJ[q] + s;
        if (j >= i && j < F1) P[rb + j] = acc[q][r][s];
      }
    }
  }
  __syncthreads();
  const float ridge = a.lambda * (float)cnt;
  for (int i = t; i < f; i += kLT) P[row_base(i, F1) + i] += ridge;

  // Blocked right-looking Cholesky.  Panels of 8 pivots are factored row
  // by row, one thread per column: the pivot and the panel's part of row k
  // are read, a barrier, then each thread scales its element of row k and
  // updates the panel rows below it in its column.  The rows past the panel
  // then take one rank-8 update in 4x4 register tiles (a 16 x 16 thread
  // grid over the tiles), eight times fewer LDS accesses than updating
  // after every pivot.  The augmented column f rides along, so it ends as
  // z = U^{-T} b.
  const int ta = t >> 4, tc = t & 15;
  bool ok = true;
  for (int kb = 0; kb < f; kb += 8) {
    const int ke = min(kb + 8, f);
    for (int k = kb; k < ke; ++k) {
      __syncthreads();
      const int rbk = row_base(k, F1);
      const float d = P[rbk + k];
      if (!(d > 0.0f && d < __builtin_huge_valf())) {  // same value in every thread
        ok = false;
        break;
      }
      const float inv = rsqrtf(d);
      float uk[7];
#pragma unroll
      for (int q = 0; q < 7; ++q) {
        const int i = k + 1 + q;
        uk[q] = i < ke ? P[rbk + i] * inv : 0.0f;
      }
      __syncthreads();
      for (int j = k + t; j < F1; j += kLT) {
        const float v = P[rbk + j] * inv;
        P[rbk + j] = v;
#pragma unroll
        for (int q = 0; q < 7; ++q) {
          const int i = k + 1 + q;
          if (i < ke && i <= j) {
            const int ri = row_base(i, F1);
            P[ri + j] = fmaf(-uk[q], v, P[ri + j]);
          }
        }
      }
    }
    if (!ok || ke >= f) break;  // uniform
    __syncthreads();            // the panel rows are final
    const int T = (F1 - ke + 3) >> 2;
    for (int ti = ta; ti < T; ti += 16) {
      const int i0 = ke + 4 * ti;
      if (i0 >= f) break;
      for (int tj = ti + ((tc - ti) & 15); tj < T; tj += 16) {  // tj >= ti, tj = tc (mod 16)
        const int j0 = ke + 4 * tj;
        float acc4[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) acc4[r][c] = 0.0f;
        for (int k = kb; k < ke; ++k) {
          const int rbk = row_base(k, F1);
          float ui[4], uj[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            ui[r] = P[rbk + min(i0 + r, F1 - 1)];  // clamped reads; their results are not stored
            uj[r] = P[rbk + min(j0 + r, F1 - 1)];
          }
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc4[r][c] = fmaf(ui[r], uj[c], acc4[r][c]);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = i0 + r;
          if (i >= f) continue;
          const int ri = row_base(i, F1);
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const int j = j0 + c;
            if (j >= i && j < F1) P[ri + j] -= acc4[r][c];
          }
        }
      }
    }
  }
  __syncthreads();
  if (!ok) {
    if (t == 0) atomicAdd(a.fail_counter, 1);
    for (int i = t; i < (int)a.ldf; i += kLT) xrow[i] = 0.0f;
    return;
  }
  if (t >= 64) return;

  // Back substitution in wavefront 0: entry i of z and x in lane i % 64,
  // slot i / 64.
  const int lane = t;
  float z[kLargeMaxF / 64], xv[kLargeMaxF / 64];
#pragma unroll
  for (int q = 0; q < kLargeMaxF / 64; ++q) {
    const int i = lane + 64 * q;
    z[q] = i < f ? P[row_base(i, F1) + f] : 0.0f;
    xv[q] = 0.0f;
  }
  for (int k = f - 1; k >= 0; --k) {
    const int qk = k >> 6, lk = k & 63;
    float zk_src = z[0];
#pragma unroll
    for (int q = 1; q < kLargeMaxF / 64; ++q)
      if (qk == q) zk_src = z[q];
    const float zk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(zk_src), lk));
    const float xk = zk / P[row_base(k, F1) + k];
#pragma unroll
    for (int q = 0; q < kLargeMaxF / 64; ++q) {
      const int i = lane + 64 * q;
      if (i == k) xv[q] = xk;
      if (i < k) z[q] = fmaf(-P[row_base(i, F1) + k], xk, z[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < kLargeMaxF / 64; ++q) {
    const int i = lane + 64 * q;
    if (i < (int)a.ldf) xrow[i] = i < f ? xv[q] : 0.0f;
  }
}

}  // namespace

size_t large_lds_bytes(int f) {
  const int W = 8 * ((f + 8) / 8);  // staged row width, 8 ceil((f+1)/8)
  const int staged = kLC * W, packed = packed_floats(f);
  return sizeof(float) * (size_t)(packed > staged ? packed : staged);
}

hipError_t prepare_large_kernel(int f) {
  if (f <= 0 || f > kLargeMaxF) return hipErrorInvalidValue;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&als_large_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)large_lds_bytes(f));
}

hipError_t launch_large_kernel(const SweepArgs& a, int64_t rows, hipStream_t s) {
  if (a.f <= 0 || a.f > kLargeMaxF || rows < 0) return hipErrorInvalidValue;
  if (rows == 0) return hipSuccess;
  hipLaunchKernelGGL(als_large_kernel, dim3((unsigned)rows), dim3(kLT), large_lds_bytes(a.f), s, a);
  return hipGetLastError();
}

}  // namespace cumf
