// mlp_kernels.h -- internal interface of the warm-start MLP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cumf {

enum GemmEpi {
  EPI_STORE = 0,     // C (fp32) = A B^T; split-K slice z at C + z * split_stride
  EPI_FWD_GELU = 1,  // z = A B^T + bias: C = z (optional), H = bf16(gelu(z)), Ht = H^T (optional)
  EPI_LINEAR = 2,    // C = A B^T + bias on rows < mlimit, columns < nlimit
  EPI_BWD_GELU = 3,  // d = bf16((A B^T) * gelu'(Z)): H = d (optional), Ht = d^T (optional)
};

// C (M x N) = A (M x K, row stride lda) * B (N x K, row stride ldb)^T.
// K a multiple of 32 per split, lda and ldb multiples of 8 (16-byte rows).
struct GemmArgs {
  const __bf16* A;
  int64_t lda;
  const __bf16* B;
  int64_t ldb;
  int M, N, K;
  int epi;
  const float* bias;
  float* C;
  int64_t ldc;
  int64_t split_stride;
  __bf16* H;
  int64_t ldh;
  __bf16* Ht;
  int64_t ldht;
  const float* Z;
  int64_t ldz;
  int mlimit, nlimit;
};

hipError_t launch_mlp_gemm(const GemmArgs& a, int splits, hipStream_t s);
hipError_t launch_mlp_loss(const float* Y, int64_t ldy, const float* T, int64_t ldt, int B, int f, int n3,
                           __bf16* dY, __bf16* dYt, double* partial, int nblocks, hipStream_t s);
hipError_t launch_mlp_rowsum(const __bf16* Dt, int N, int B, float* out, hipStream_t s);
hipError_t launch_mlp_reduce(const float* W, int64_t stride, int S, int64_t count, float* G, hipStream_t s);
hipError_t launch_mlp_adam(float* P, const float* G, float* M, float* V, int64_t count, float lr, float b1,
                           float b2, float eps, float c1, float c2, hipStream_t s);
hipError_t launch_mlp_publish(const float* Pw, int N, int K, __bf16* Wt, __bf16* W, hipStream_t s);
hipError_t launch_mlp_gather(const __bf16* feats, int K0, const float* targets, int f, const int64_t* perm, int B,
                             __bf16* H0, __bf16* H0t, float* T, hipStream_t s);
hipError_t launch_mlp_to_bf16(const float* src, __bf16* dst, int64_t count, hipStream_t s);
hipError_t launch_mlp_features(const int32_t* row_ptr, const int32_t* col, const float* val, const float* theta,
                               int64_t ldf, int f, float lambda, int64_t row0, int64_t nrows, int K0, __bf16* out,
                               hipStream_t s);

}  // namespace cumf
