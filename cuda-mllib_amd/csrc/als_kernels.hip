// als_kernels.hip -- CDNA4 (gfx950) kernels of the ALS half-sweep.
//
// One half-sweep solves, for every row u of a CSR rating matrix,
//     (Y_u^T Y_u + lambda n_u I) x_u = Y_u^T r_u
// where Y_u stacks the n_u factor rows of the other side named by row u's
// column indices (reference: get_hermitian100 / get_hermitianT10 at
// als/src/cuda/als.cu:47-912, the cusparseScsrmm2 RHS at :1039-1054 and
// the batched LU solve updateX at :935-997).
//
// Design (one 64-lane wavefront per task):
//   * Ratings are consumed in chunks of KB=16.  Each chunk's factor rows are
//     gathered with 16-byte loads (a whole row is contiguous) and staged in
//     LDS as a KB x (16*NB) tile: columns [0, ldf) hold the factor row,
//     column ldf holds the rating, the rest are zero.
//   * The augmented Gram [Y r]^T [Y r] is accumulated with
//     v_mfma_f32_16x16x4_f32 over the NB(NB+1)/2 upper 16x16 tiles.  Tile
//     (I, J) gives A (I,J < ldf), b (column ldf) and r.r in one product.
//   * The accumulators stay in registers.  Their layout (lane l holds rows
//     16I + 4(l/16) + {0..3} of column 16J + l%16) is a 2-D block-cyclic
//     distribution, so a right-looking Cholesky runs on it in place: row k
//     of U is broadcast through a 16*NB-float LDS buffer and every lane
//     applies the rank-1 update to the elements it owns.  Running the
//     factorisation over the augmented column leaves z = U^{-T} b there.
//   * U and z are spilled to LDS (packed columns) for the back substitution
//     U x = z, one row index per lane.
//   * Rows with more than `split` ratings are cut into chunks whose partial
//     Grams go to a scratch buffer in register order; a second kernel sums
//     them in chunk order (deterministic) and runs the same solve.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "als_kernels.h"

namespace cumf {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int KB = 16;  // ratings per staged chunk

__host__ __device__ constexpr int num_tiles(int NB) { return NB * (NB + 1) / 2; }
__host__ __device__ constexpr int tile_index(int I, int J, int NB) {
  return I * NB - I * (I - 1) / 2 + (J - I);
}
// LDS row stride of the staged chunk: >= 16*NB and an odd multiple of 16
// floats, so the four 16-lane groups of a wavefront read four different
// 16-bank windows.
__host__ __device__ constexpr int stage_ld(int NB) { return 16 * NB + ((NB % 2 == 0) ? 16 : 0); }

template <int NB>
__host__ __device__ constexpr int lds_floats(int f) {
  return (KB * stage_ld(NB) > f * (f + 1) / 2 + 16 * NB) ? KB * stage_ld(NB)
                                                          : f * (f + 1) / 2 + 16 * NB;
}

__device__ __forceinline__ void wave_sync() {
  // Work-groups here are a single wavefront: this orders LDS traffic
  // between lanes (and is a compiler fence).
  __syncthreads();
}

__device__ __forceinline__ float readlane_f(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// ---------------------------------------------------------------------------
// Gram accumulation over a contiguous range of one row's ratings.
// ---------------------------------------------------------------------------
template <int NB>
__device__ __forceinline__ void gram_range(f4 (&acc)[num_tiles(NB)], const int32_t* __restrict__ col,
                                           const float* __restrict__ val,
                                           const float* __restrict__ theta, int64_t ldf,
                                           int64_t start, int count, float* __restrict__ stage,
                                           int lane) {
  constexpr int LD = stage_ld(NB);
  const int g = lane >> 4, c = lane & 15;
  const int Q = (int)(ldf >> 2);        // float4 per factor row
  const int slots = KB * Q;             // float4 per chunk
  const int bcol = (int)ldf;

  // Zero the staging tile once (padding columns stay zero afterwards).
  for (int i = lane; i < KB * LD; i += 64) stage[i] = 0.0f;

  // Per-lane (row, quad) of each load slot; NB slots per lane suffice
  // because ldf < 16*NB.
  int srow[NB], squad[NB];
#pragma unroll
  for (int it = 0; it < NB; ++it) {
    int s = it * 64 + lane;
    srow[it] = s < slots ? s / Q : KB;  // KB marks an idle slot
    squad[it] = s < slots ? s % Q : 0;
  }

  const int nch = (count + KB - 1) / KB;
  // Software pipeline: column indices/values two chunks ahead, factor rows
  // one chunk ahead.
  auto load_cv = [&](int ch, int& ci, float& cv) {
    int64_t k = (int64_t)ch * KB + lane;
    bool ok = lane < KB && ch < nch && k < count;
    ci = ok ? col[start + k] : -1;
    cv = ok ? val[start + k] : 0.0f;
  };
  auto load_rows = [&](int ci, f4 (&th)[NB]) {
#pragma unroll
    for (int it = 0; it < NB; ++it) {
      int src = __shfl(ci, srow[it] < KB ? srow[it] : 0);
      if (srow[it] < KB && src >= 0)
        th[it] = *reinterpret_cast<const f4*>(theta + (int64_t)src * ldf + 4 * squad[it]);
      else
        th[it] = f4{0.f, 0.f, 0.f, 0.f};
    }
  };

  int ci0, ci1;
  float cv0, cv1;
  f4 th[NB];
  load_cv(0, ci0, cv0);
  load_cv(1, ci1, cv1);
  load_rows(ci0, th);

  for (int ch = 0; ch < nch; ++ch) {
    wave_sync();  // previous chunk's LDS reads are done
#pragma unroll
    for (int it = 0; it < NB; ++it)
      if (srow[it] < KB) *reinterpret_cast<f4*>(stage + srow[it] * LD + 4 * squad[it]) = th[it];
    if (lane < KB) stage[lane * LD + bcol] = cv0;
    wave_sync();
    // Next chunk in flight while this one is multiplied.
    int ci2;
    float cv2;
    load_cv(ch + 2, ci2, cv2);
    load_rows(ci1, th);
    ci0 = ci1; cv0 = cv1; ci1 = ci2; cv1 = cv2;

    float a[KB / 4][NB];
#pragma unroll
    for (int s = 0; s < KB / 4; ++s)
#pragma unroll
      for (int I = 0; I < NB; ++I) a[s][I] = stage[(4 * s + g) * LD + 16 * I + c];
#pragma unroll
    for (int s = 0; s < KB / 4; ++s)
#pragma unroll
      for (int I = 0; I < NB; ++I)
#pragma unroll
        for (int J = I; J < NB; ++J)
          acc[tile_index(I, J, NB)] = __builtin_amdgcn_mfma_f32_16x16x4f32(
              a[s][I], a[s][J], acc[tile_index(I, J, NB)], 0, 0, 0);
  }
}

// Factorisation steps k in [16K, 16K+16) of block row K, then the next
// block row.  K and the register index kr are compile-time so every
// accumulator access is a fixed register; the 16-lane group kg that owns
// row k is a run-time loop.
template <int NB, int K>
__device__ __forceinline__ void chol_block(f4 (&acc)[num_tiles(NB)], int f, float* __restrict__ rowbuf,
                                           bool& ok, int lane) {
  const int g = lane >> 4, c = lane & 15;
  if (16 * K < f) {
#pragma unroll 1
    for (int kg = 0; kg < 4; ++kg) {
#pragma unroll
      for (int kr = 0; kr < 4; ++kr) {
        const int k = 16 * K + 4 * kg + kr;
        if (k < f) {
          if (g == kg) {
#pragma unroll
            for (int J = K; J < NB; ++J) rowbuf[16 * J + c] = acc[tile_index(K, J, NB)][kr];
          }
          wave_sync();
          const float d = rowbuf[k];
          float inv = 0.0f;
          if (d > 0.0f && d < __builtin_huge_valf()) inv = 1.0f / sqrtf(d);
          else ok = false;
          float colv[NB];
          float rowv[NB][4];
#pragma unroll
          for (int J = K; J < NB; ++J) colv[J] = rowbuf[16 * J + c] * inv;
#pragma unroll
          for (int I = K; I < NB; ++I) {
            const f4 v = *reinterpret_cast<const f4*>(rowbuf + 16 * I + 4 * g);
#pragma unroll
            for (int r = 0; r < 4; ++r) rowv[I][r] = (16 * I + 4 * g + r > k) ? v[r] * inv : 0.0f;
          }
          wave_sync();
          if (g == kg) {
#pragma unroll
            for (int J = K; J < NB; ++J) acc[tile_index(K, J, NB)][kr] = colv[J];
          }
#pragma unroll
          for (int I = K; I < NB; ++I)
#pragma unroll
            for (int J = I; J < NB; ++J)
#pragma unroll
              for (int r = 0; r < 4; ++r)
                acc[tile_index(I, J, NB)][r] =
                    fmaf(-rowv[I][r], colv[J], acc[tile_index(I, J, NB)][r]);
        }
      }
    }
  }
  if constexpr (K + 1 < NB) chol_block<NB, K + 1>(acc, f, rowbuf, ok, lane);
}

// ---------------------------------------------------------------------------
// In-register Cholesky of the augmented Gram + back substitution.
// Returns false if a pivot is not positive (x is then written as zeros).
// ---------------------------------------------------------------------------
template <int NB>
__device__ __forceinline__ bool chol_solve(f4 (&acc)[num_tiles(NB)], int f, int bcol,
                                           float* __restrict__ lds, float* __restrict__ xrow,
                                           int64_t ldf, int lane) {
  const int g = lane >> 4, c = lane & 15;
  bool ok = true;
  chol_block<NB, 0>(acc, f, lds, ok, lane);

  // Spill U (packed by columns: U[i][j] at j(j+1)/2 + i) and z.
  const int upk = f * (f + 1) / 2;
  float* ucol = lds;
  float* zbuf = lds + upk;
  wave_sync();
#pragma unroll
  for (int I = 0; I < NB; ++I)
#pragma unroll
    for (int J = I; J < NB; ++J) {
      const int j = 16 * J + c;
      const int cbase = j * (j + 1) / 2;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * I + 4 * g + r;
        const float v = acc[tile_index(I, J, NB)][r];
        if (j < f && i <= j) ucol[cbase + i] = v;
        if (j == bcol && i < f) zbuf[i] = v;
      }
    }
  wave_sync();

  // Back substitution, one row index per lane (f <= 192).
  float z[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int i = lane + 64 * q;
    z[q] = (i < f) ? zbuf[i] : 0.0f;
  }
  for (int k = f - 1; k >= 0; --k) {
    const int kq = k >> 6, kl = k & 63;
    const float zsel = kq == 0 ? z[0] : (kq == 1 ? z[1] : z[2]);
    const float zk = readlane_f(zsel, kl);
    const float xk = zk / ucol[k * (k + 1) / 2 + k];
    const float* uk = ucol + k * (k + 1) / 2;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int i = lane + 64 * q;
      if (i < k) z[q] = fmaf(-uk[i], xk, z[q]);
      if (i == k) z[q] = xk;
    }
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int i = lane + 64 * q;
    if (i < ldf) xrow[i] = (ok && i < f) ? z[q] : 0.0f;
  }
  return ok;
}

template <int NB>
__device__ __forceinline__ void add_ridge(f4 (&acc)[num_tiles(NB)], int f, float ridge, int lane) {
  const int g = lane >> 4, c = lane & 15;
#pragma unroll
  for (int I = 0; I < NB; ++I)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (c == 4 * g + r && 16 * I + c < f) acc[tile_index(I, I, NB)][r] += ridge;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
template <int NB>
__global__ __launch_bounds__(64) void als_task_kernel(SweepArgs a) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x;
  const Task t = a.tasks[blockIdx.x];
  f4 acc[num_tiles(NB)];
#pragma unroll
  for (int i = 0; i < num_tiles(NB); ++i) acc[i] = f4{0.f, 0.f, 0.f, 0.f};

  const int64_t start = (int64_t)a.row_ptr[t.row] + t.start;
  if (t.count > 0) gram_range<NB>(acc, a.col, a.val, a.theta, a.ldf, start, t.count, lds, lane);

  if (t.slot >= 0) {  // partial Gram of a heavy row
    float* dst = a.partial + (int64_t)t.slot * num_tiles(NB) * 256;
#pragma unroll
    for (int i = 0; i < num_tiles(NB); ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) dst[(i * 4 + r) * 64 + lane] = acc[i][r];
    return;
  }
  float* xrow = a.x + (int64_t)t.row * a.ldf;
  if (t.count == 0) {  // no ratings: x_u = 0
    for (int i = lane; i < a.ldf; i += 64) xrow[i] = 0.0f;
    return;
  }
  add_ridge<NB>(acc, a.f, a.lambda * (float)t.count, lane);
  wave_sync();
  if (!chol_solve<NB>(acc, a.f, (int)a.ldf, lds, xrow, a.ldf, lane) && lane == 0)
    atomicAdd(a.fail_counter, 1);
}

template <int NB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void als_finish_kernel(SweepArgs a) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x;
  const HeavyRow h = a.heavy[blockIdx.x];
  f4 acc[num_tiles(NB)];
#pragma unroll
  for (int i = 0; i < num_tiles(NB); ++i) acc[i] = f4{0.f, 0.f, 0.f, 0.f};
  for (int p = 0; p < h.nchunks; ++p) {
    const float* src = a.partial + (int64_t)(h.slot + p) * num_tiles(NB) * 256;
#pragma unroll
    for (int i = 0; i < num_tiles(NB); ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][r] += src[(i * 4 + r) * 64 + lane];
  }
  const int n_u = a.row_ptr[h.row + 1] - a.row_ptr[h.row];
  add_ridge<NB>(acc, a.f, a.lambda * (float)n_u, lane);
  float* xrow = a.x + (int64_t)h.row * a.ldf;
  if (!chol_solve<NB>(acc, a.f, (int)a.ldf, lds, xrow, a.ldf, lane) && lane == 0)
    atomicAdd(a.fail_counter, 1);
}

// Gram dump for tests: the full augmented (16NB)^2 matrix of each listed
// row (ridge included), row-major, upper tiles mirrored.
template <int NB>
__global__ __launch_bounds__(64) void als_dump_kernel(SweepArgs a, const int32_t* rows, float* out) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int u = rows[blockIdx.x];
  f4 acc[num_tiles(NB)];
#pragma unroll
  for (int i = 0; i < num_tiles(NB); ++i) acc[i] = f4{0.f, 0.f, 0.f, 0.f};
  const int64_t s = a.row_ptr[u];
  const int cnt = a.row_ptr[u + 1] - a.row_ptr[u];
  if (cnt > 0) gram_range<NB>(acc, a.col, a.val, a.theta, a.ldf, s, cnt, lds, lane);
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
int sweep_nb(int f) {
  const int ldf = (f + 3) & ~3;
  return (ldf + 1 + 15) / 16;
}

template <int NB>
static hipError_t launch_nb(const SweepArgs& a, hipStream_t s, int which, const int32_t* rows,
                            float* out, int64_t nrows) {
  const size_t shmem = sizeof(float) * (size_t)lds_floats<NB>(a.f);
  if (which == 0) {
    if (a.ntasks > 0)
      hipLaunchKernelGGL(als_task_kernel<NB>, dim3((unsigned)a.ntasks), dim3(64), shmem, s, a);
  } else if (which == 1) {
    if (a.nheavy > 0)
      hipLaunchKernelGGL(als_finish_kernel<NB>, dim3((unsigned)a.nheavy), dim3(64), shmem, s, a);
  } else {
    if (nrows > 0)
      hipLaunchKernelGGL(als_dump_kernel<NB>, dim3((unsigned)nrows), dim3(64), shmem, s, a, rows, out);
  }
  return hipGetLastError();
}

static hipError_t dispatch(const SweepArgs& a, hipStream_t s, int which, const int32_t* rows = nullptr,
                           float* out = nullptr, int64_t nrows = 0) {
  switch (sweep_nb(a.f)) {
    case 1: return launch_nb<1>(a, s, which, rows, out, nrows);
    case 2: return launch_nb<2>(a, s, which, rows, out, nrows);
    case 3: return launch_nb<3>(a, s, which, rows, out, nrows);
    case 4: return launch_nb<4>(a, s, which, rows, out, nrows);
    case 5: return launch_nb<5>(a, s, which, rows, out, nrows);
    case 6: return launch_nb<6>(a, s, which, rows, out, nrows);
    case 7: return launch_nb<7>(a, s, which, rows, out, nrows);
    case 8: return launch_nb<8>(a, s, which, rows, out, nrows);
    case 9: return launch_nb<9>(a, s, which, rows, out, nrows);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_task_kernel(const SweepArgs& a, hipStream_t s) { return dispatch(a, s, 0); }
hipError_t launch_finish_kernel(const SweepArgs& a, hipStream_t s) { return dispatch(a, s, 1); }
hipError_t launch_dump_kernel(const SweepArgs& a, const int32_t* rows, int64_t nrows, float* out,
                              hipStream_t s) {
  return dispatch(a, s, 2, rows, out, nrows);
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
