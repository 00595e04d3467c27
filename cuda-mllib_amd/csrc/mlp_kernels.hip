// mlp_kernels.hip -- gfx950 kernels of the warm-start MLP.
//
// The network (f = rank): features [log n_u, diag(A_u)/n_u, b_u/n_u]
// (2f+1, zero-padded to K0 = round_up(2f+1, 32)) -> 256 -> 256 -> f,
// GELU (erf form) on the hidden layers, BF16 operands with FP32
// accumulation on v_mfma_f32_16x16x32_bf16, FP32 master weights updated
// by Adam.  Every product is one form, C = A B^T with A (M x K) and B
// (N x K) both row-major and K-contiguous, so that forward, backward-data
// and weight-gradient products share one kernel: the forward epilogue
// also writes the activations transposed (the weight gradient of the next
// layer reads them K-contiguous over the batch), and the backward epilogue
// writes the pre-activation gradient in both layouts.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "mlp_kernels.h"

namespace cumf {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int GT = 64;   // output tile (rows and columns) per work-group
constexpr int GK = 32;   // K step (one MFMA)
constexpr int GS = 40;   // LDS row stride in bf16 (80 B: 16-byte reads of 8 consecutive rows hit distinct banks)

__device__ __forceinline__ float gelu_f(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad_f(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

// 256 threads = 4 waves; wave w computes the 32 x 32 quadrant
// (32 (w >> 1), 32 (w & 1)) of the 64 x 64 tile as 2 x 2 MFMA tiles.
// Operands are staged row-major (K-contiguous) in LDS; lane l reads 8
// consecutive k of row l % 16 at k offset 8 (l / 16), which is the
// operand layout of v_mfma_f32_16x16x32_bf16 for both A and B^T.
// blockIdx.z selects a K slice (split-K); EPI_STORE then writes slice z
// at C + z * split_stride.
__global__ __launch_bounds__(256) void mlp_gemm_kernel(GemmArgs a) {
  __shared__ __align__(16) __bf16 As[GT * GS];
  __shared__ __align__(16) __bf16 Bs[GT * GS];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int kper = a.K / (int)gridDim.z;
  const int kbeg = (int)blockIdx.z * kper, kend = kbeg + kper;
  const int lr = tid >> 2, lc = (tid & 3) * 8;
  const bool am = m0 + lr < a.M, bn = n0 + lr < a.N;
  const __bf16* ap = a.A + (int64_t)(am ? m0 + lr : 0) * a.lda + lc;
  const __bf16* bp = a.B + (int64_t)(bn ? n0 + lr : 0) * a.ldb + lc;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  bf8 zero8;
#pragma unroll
  for (int t = 0; t < 8; ++t) zero8[t] = (__bf16)0.0f;
  bf8 av = am ? *reinterpret_cast<const bf8*>(ap + kbeg) : zero8;
  bf8 bv = bn ? *reinterpret_cast<const bf8*>(bp + kbeg) : zero8;
  for (int k0 = kbeg; k0 < kend; k0 += GK) {
    __syncthreads();
    *reinterpret_cast<bf8*>(As + lr * GS + lc) = av;
    *reinterpret_cast<bf8*>(Bs + lr * GS + lc) = bv;
    __syncthreads();
    if (k0 + GK < kend) {  // next slice's loads in flight during the MFMAs
      av = am ? *reinterpret_cast<const bf8*>(ap + k0 + GK) : zero8;
      bv = bn ? *reinterpret_cast<const bf8*>(bp + k0 + GK) : zero8;
    }
    bf8 ao[2], bo[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) ao[i] = *reinterpret_cast<const bf8*>(As + (wm + 16 * i + c) * GS + 8 * g);
#pragma unroll
    for (int j = 0; j < 2; ++j) bo[j] = *reinterpret_cast<const bf8*>(Bs + (wn + 16 * j + c) * GS + 8 * g);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ao[i], bo[j], acc[i][j], 0, 0, 0);
  }
  // Lane (g, c), register r of tile (i, j) holds C[wm+16i+4g+r][wn+16j+c].
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + 4 * g + r, n = n0 + wn + 16 * j + c;
        if (m >= a.M || n >= a.N) continue;
        const float v = acc[i][j][r];
        switch (a.epi) {
          case EPI_STORE:
            a.C[(int64_t)blockIdx.z * a.split_stride + (int64_t)m * a.ldc + n] = v;
            break;
          case EPI_FWD_GELU: {
            const float z = v + a.bias[n];
            if (a.C) a.C[(int64_t)m * a.ldc + n] = z;
            const __bf16 h = (__bf16)gelu_f(z);
            if (a.H) a.H[(int64_t)m * a.ldh + n] = h;
            if (a.Ht) a.Ht[(int64_t)n * a.ldht + m] = h;
            break;
          }
          case EPI_LINEAR:
            if (m < a.mlimit && n < a.nlimit) a.C[(int64_t)m * a.ldc + n] = v + a.bias[n];
            break;
          case EPI_BWD_GELU: {
            const __bf16 d = (__bf16)(v * gelu_grad_f(a.Z[(int64_t)m * a.ldz + n]));
            if (a.H) a.H[(int64_t)m * a.ldh + n] = d;
            if (a.Ht) a.Ht[(int64_t)n * a.ldht + m] = d;
            break;
          }
          default:
            break;
        }
      }
}

// dY = bf16(2 (Y - T) / (B f)) on the f real outputs (0 on padding), in
// both layouts; one FP64 partial sum of squared errors per block.
__global__ __launch_bounds__(256) void mlp_loss_kernel(const float* __restrict__ Y, int64_t ldy,
                                                       const float* __restrict__ T, int64_t ldt, int B, int f,
                                                       int n3, __bf16* __restrict__ dY, __bf16* __restrict__ dYt,
                                                       double* __restrict__ partial) {
  __shared__ double red[4];
  double se = 0.0;
  const float scale = 2.0f / ((float)B * (float)f);
  const int64_t total = (int64_t)B * n3;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(idx / n3), n = (int)(idx - (int64_t)b * n3);
    const float d = n < f ? Y[(int64_t)b * ldy + n] - T[(int64_t)b * ldt + n] : 0.0f;
    se += (double)d * (double)d;
    const __bf16 gv = (__bf16)(d * scale);
    dY[(int64_t)b * n3 + n] = gv;
    dYt[(int64_t)n * B + b] = gv;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) se += __shfl_xor(se, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = se;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// out[n] = sum_b Dt[n][b] (BF16 rows of length B), fixed order: lane l sums
// b = l, l + 64, ...; then a butterfly over the wave.
__global__ __launch_bounds__(64) void mlp_rowsum_kernel(const __bf16* __restrict__ Dt, int B, float* __restrict__ out) {
  const int n = blockIdx.x;
  float s = 0.0f;
  for (int b = threadIdx.x; b < B; b += 64) s += (float)Dt[(int64_t)n * B + b];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
  if (threadIdx.x == 0) out[n] = s;
}

// G[i] = sum_s W[s * stride + i], s in order.
__global__ __launch_bounds__(256) void mlp_reduce_kernel(const float* __restrict__ W, int64_t stride, int S,
                                                         int64_t count, float* __restrict__ G) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    float s = 0.0f;
    for (int k = 0; k < S; ++k) s += W[k * stride + i];
    G[i] = s;
  }
}

// Adam (Kingma & Ba 2015, Algorithm 1); c1 = 1/(1-b1^t), c2 = 1/(1-b2^t).
__global__ __launch_bounds__(256) void mlp_adam_kernel(float* __restrict__ P, const float* __restrict__ G,
                                                       float* __restrict__ Mm, float* __restrict__ V, int64_t count,
                                                       float lr, float b1, float b2, float eps, float c1, float c2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const float g = G[i];
    const float m = b1 * Mm[i] + (1.0f - b1) * g;
    const float v = b2 * V[i] + (1.0f - b2) * g * g;
    Mm[i] = m;
    V[i] = v;
    P[i] = P[i] - lr * (m * c1) / (sqrtf(v * c2) + eps);
  }
}

// BF16 copies of one layer's weights: Wt (N x K, as stored) and W (K x N).
__global__ __launch_bounds__(256) void mlp_publish_kernel(const float* __restrict__ Pw, int N, int K,
                                                          __bf16* __restrict__ Wt, __bf16* __restrict__ W) {
  const int64_t total = (int64_t)N * K;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int n = (int)(i / K), k = (int)(i - (int64_t)n * K);
    const __bf16 v = (__bf16)Pw[i];
    Wt[i] = v;
    W[(int64_t)k * N + n] = v;
  }
}

// Batch assembly: rows perm[0..B) of the feature / target sets into H0
// (B x K0), H0t (K0 x B) and T (B x f).
__global__ __launch_bounds__(256) void mlp_gather_kernel(const __bf16* __restrict__ feats, int K0,
                                                         const float* __restrict__ targets, int f,
                                                         const int64_t* __restrict__ perm, int B,
                                                         __bf16* __restrict__ H0, __bf16* __restrict__ H0t,
                                                         float* __restrict__ T) {
  const int64_t total = (int64_t)B * K0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(i / K0), k = (int)(i - (int64_t)b * K0);
    const int64_t s = perm[b];
    const __bf16 v = feats[s * K0 + k];
    H0[i] = v;
    H0t[(int64_t)k * B + b] = v;
    if (k < f) T[(int64_t)b * f + k] = targets[s * f + k];
  }
}

// FP32 -> BF16, any count.
__global__ __launch_bounds__(256) void mlp_to_bf16_kernel(const float* __restrict__ src, __bf16* __restrict__ dst,
                                                          int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (__bf16)src[i];
}

// Features of rows row0 .. row0+nrows-1 of a CSR side, one wavefront per
// row: [log n_u, diag/n_u + lambda, b/n_u, 0...] (all zero for an empty
// row) in BF16, from FP32 sums diag_i = sum theta_vi^2, b_i = sum r theta_vi.
// The wavefront works as four groups of 16 lanes; group q takes ratings
// q, q+4, q+8, ... of the row, lane c of a group holds columns 4c..4c+3
// (+64, +128) and moves them with 16-byte loads; the four group sums are
// added in group order at the end.
__global__ __launch_bounds__(64) void mlp_features_kernel(const int32_t* __restrict__ row_ptr,
                                                          const int32_t* __restrict__ col,
                                                          const float* __restrict__ val,
                                                          const float* __restrict__ theta, int64_t ldf, int f,
                                                          float lambda, int64_t row0, int K0,
                                                          __bf16* __restrict__ out) {
  __shared__ float sd[4][192], sb[4][192];
  const int lane = threadIdx.x, q = lane >> 4, c = lane & 15;
  const int64_t u = row0 + blockIdx.x;
  const int64_t s = row_ptr[u], e = row_ptr[u + 1];
  const int nq = (int)((ldf + 63) / 64);  // 64-column slices of a factor row (<= 3)
  f4 d[3], b[3];
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    d[t] = f4{0.f, 0.f, 0.f, 0.f};
    b[t] = f4{0.f, 0.f, 0.f, 0.f};
  }
  for (int64_t k = s + q; k < e; k += 4) {
    const float r = val[k];
    const f4* th = reinterpret_cast<const f4*>(theta + (int64_t)col[k] * ldf);
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      if (t < nq && 64 * t + 4 * c < ldf) {
        const f4 v = th[16 * t + c];
#pragma unroll
        for (int z = 0; z < 4; ++z) {
          d[t][z] = fmaf(v[z], v[z], d[t][z]);
          b[t][z] = fmaf(r, v[z], b[t][z]);
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 3; ++t)
#pragma unroll
    for (int z = 0; z < 4; ++z) {
      sd[q][64 * t + 4 * c + z] = d[t][z];
      sb[q][64 * t + 4 * c + z] = b[t][z];
    }
  __syncthreads();
  const int n = (int)(e - s);
  const float inv = n > 0 ? 1.0f / (float)n : 0.0f;
  __bf16* o = out + (int64_t)blockIdx.x * K0;
  for (int i = lane; i < K0; i += 64) {
    float vv = 0.0f;
    if (n > 0) {
      if (i == 0) {
        vv = logf((float)n);
      } else if (i <= f) {
        const int j = i - 1;
        vv = (((sd[0][j] + sd[1][j]) + sd[2][j]) + sd[3][j]) * inv + lambda;
      } else if (i <= 2 * f) {
        const int j = i - 1 - f;
        vv = (((sb[0][j] + sb[1][j]) + sb[2][j]) + sb[3][j]) * inv;
      }
    }
    o[i] = (__bf16)vv;
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_mlp_gemm(const GemmArgs& a, int splits, hipStream_t s) {
  if (a.M <= 0 || a.N <= 0) return hipSuccess;
  if (a.K <= 0 || splits <= 0 || a.K % (GK * splits) != 0 || a.lda % 8 || a.ldb % 8) return hipErrorInvalidValue;
  const dim3 grid((a.N + GT - 1) / GT, (a.M + GT - 1) / GT, splits);
  hipLaunchKernelGGL(mlp_gemm_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_mlp_loss(const float* Y, int64_t ldy, const float* T, int64_t ldt, int B, int f, int n3,
                           __bf16* dY, __bf16* dYt, double* partial, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(mlp_loss_kernel, dim3(nblocks), dim3(256), 0, s, Y, ldy, T, ldt, B, f, n3, dY, dYt, partial);
  return hipGetLastError();
}

hipError_t launch_mlp_rowsum(const __bf16* Dt, int N, int B, float* out, hipStream_t s) {
  hipLaunchKernelGGL(mlp_rowsum_kernel, dim3(N), dim3(64), 0, s, Dt, B, out);
  return hipGetLastError();
}

hipError_t launch_mlp_reduce(const float* W, int64_t stride, int S, int64_t count, float* G, hipStream_t s) {
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3(256), dim3(256), 0, s, W, stride, S, count, G);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(float* P, const float* G, float* M, float* V, int64_t count, float lr, float b1,
                           float b2, float eps, float c1, float c2, hipStream_t s) {
  hipLaunchKernelGGL(mlp_adam_kernel, dim3(256), dim3(256), 0, s, P, G, M, V, count, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t launch_mlp_publish(const float* Pw, int N, int K, __bf16* Wt, __bf16* W, hipStream_t s) {
  hipLaunchKernelGGL(mlp_publish_kernel, dim3(256), dim3(256), 0, s, Pw, N, K, Wt, W);
  return hipGetLastError();
}

hipError_t launch_mlp_gather(const __bf16* feats, int K0, const float* targets, int f, const int64_t* perm, int B,
                             __bf16* H0, __bf16* H0t, float* T, hipStream_t s) {
  hipLaunchKernelGGL(mlp_gather_kernel, dim3(512), dim3(256), 0, s, feats, K0, targets, f, perm, B, H0, H0t, T);
  return hipGetLastError();
}

hipError_t launch_mlp_to_bf16(const float* src, __bf16* dst, int64_t count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_to_bf16_kernel, dim3(512), dim3(256), 0, s, src, dst, count);
  return hipGetLastError();
}

hipError_t launch_mlp_features(const int32_t* row_ptr, const int32_t* col, const float* val, const float* theta,
                               int64_t ldf, int f, float lambda, int64_t row0, int64_t nrows, int K0, __bf16* out,
                               hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  if (f > 192 || ldf > 192 || ldf % 4) return hipErrorInvalidValue;
  hipLaunchKernelGGL(mlp_features_kernel, dim3((unsigned)nrows), dim3(64), 0, s, row_ptr, col, val, theta, ldf, f,
                     lambda, row0, K0, out);
  return hipGetLastError();
}

}  // namespace cumf
