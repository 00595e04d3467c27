// als_large.hip -- the half-sweep for 140 < f <= 256 (FP32 Gram + FP32
// Cholesky), where the augmented Gram (up to 153 tiles of 16 x 16) no
// longer fits one wavefront's registers.
//
// One 256-thread work-group (four wavefronts, one per SIMD) per row, same
// arithmetic as the fused kernels (reference: als/src/cuda/als.cu:47-51,
// 462-466 and 905-909 for the Gram and ridge, :1039-1054 for b, :935-997
// for the solve):
//   * Gram on MFMA (v_mfma_f32_16x16x4_f32): the NT = NB(NB+1)/2 upper
//     tiles of the augmented [Y r]^T [Y r] are dealt round-robin to the
//     four wavefronts (tile t to wavefront t % 4), which keep them in
//     registers.  Chunks of 16 ratings are staged in LDS tile-major as in
//     the fused kernels (tile I of a chunk gathered by wavefront I % 4 with
//     global_load_lds_dwordx4; the rating column written by the wavefront
//     that gathers the last tile), double-buffered with one barrier per
//     chunk.
//   * The tiles then go to LDS (1 KB each, in accumulator order), and a
//     blocked right-looking Cholesky runs on them with two barriers per
//     16-column block K: the four wavefronts form the panel
//     U_KJ = W_K A_KJ on MFMA; then wavefront 0 updates the next diagonal
//     tile and factors it into W_{K+1} = U_{K+1,K+1}^{-T} (the
//     column-per-lane elimination of the fused kernels) while wavefronts
//     1-3 apply the rest of the trailing update A_IJ -= U_KI^T U_KJ on
//     MFMA, tile by tile.  The augmented column ends as z = U^{-T} b.
//   * Blocked back substitution x_K = W_K^T (z_K - sum_{J>K} U_KJ x_J) in
//     wavefront 0.
// Rows without ratings get x = 0; a pivot that is not positive and finite
// counts a failed row and gives x = 0.
//
// Rows of at most 95 ratings take the dual form instead (one wavefront per
// row, als_large_dual_kernel): with Y the n x f matrix of the row's factor
// rows and mu = lambda n, (Y^T Y + mu I)^{-1} Y^T r = Y^T (Y Y^T + mu I)^{-1} r,
// so z solves the n x n system (Y Y^T + mu I) z = r and x = Y^T z:
//   * K = Y Y^T in FP32 on v_mfma_f32_16x16x4_f32, straight from the factor
//     rows: for each 16-column slab, lane (g, c) loads columns 4g .. 4g+3 of
//     row 16I + c (16 bytes; the four lanes g of a row read 64 contiguous
//     bytes) for every row block I, and the four columns feed four MFMA
//     steps in which the same register is the A operand of block I and the
//     B operand of block J = I.  The next slab's loads are issued before the
//     current slab's MFMAs.
//   * r goes into column 16 NBD - 1 (past every live pivot), mu onto the
//     first n diagonal entries, and the register-tile Cholesky of the fused
//     kernels factors the NBD = (n + 16) / 16 <= 6 blocks and writes z to LDS.
//   * x = Y^T z in FP32: lane l forms columns 4l .. 4l+3, reading the n rows
//     again (one 16-byte load per lane and rating).
// At f = 256 that is at most 21 tiles of K against the 153 tiles of the
// primal Gram and a 6-block Cholesky against 17 blocks.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "als_device.h"

namespace cumf {
namespace {

constexpr int kLT = 256;  // threads per row: four wavefronts
constexpr int kLargeMaxF = 256;
// Row blocks of the work-group dual form: rows of 96 .. 255 ratings.
constexpr int kLargeDualMinBlocks = (kDualMaxRatingsHost + 1 + 16) / 16;
constexpr int kLargeDualMaxBlocks = kLargeDualWgMaxBlocks;
static_assert(kLargeDualMinBlocks == 7 && kLargeDualMaxBlocks == 16, "blocks 7 .. 16");

__host__ __device__ constexpr int qmax(int NB) { return (num_tiles(NB) + 3) / 4; }

template <typename F, int... qs>
__device__ __forceinline__ void static_for_impl(F&& fn, std::integer_sequence<int, qs...>) {
  (fn(std::integral_constant<int, qs>{}), ...);
}
// fn(std::integral_constant<int, q>) for q = 0 .. N-1, unrolled.
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& fn) {
  static_for_impl(fn, std::make_integer_sequence<int, N>{});
}
// LDS floats: the tiles, then the diagonal-tile scratch, the identity, the
// last block's W, z and x.  The Gram's staging (2 chunks of NB tiles and
// their ratings) overlays the tile area.
__host__ __device__ constexpr int large_lds_floats(int NB) { return num_tiles(NB) * 256 + 3 * 256 + 32 * NB; }
// The dual form also keeps the row's ratings and column indices (256 each).
__host__ __device__ constexpr int large_dual_wg_lds_floats(int NB) { return large_lds_floats(NB) + 512; }

// Tile t of the upper triangle in tile_index order -> (I, J).
__host__ __device__ constexpr int tile_row(int t, int NB) {
  int I = 0;
  while (t >= NB - I) {
    t -= NB - I;
    ++I;
  }
  return I;
}
__host__ __device__ constexpr int tile_col(int t, int NB) {
  int I = 0;
  while (t >= NB - I) {
    t -= NB - I;
    ++I;
  }
  return I + t;
}

// The Gram of wavefront W's tiles (t = W, W + 4, ...), left in acc, then
// the ridge and the store of the tiles to their LDS slots.
template <int NB, int W>
__device__ __forceinline__ void gram_tiles(const SweepArgs& a, int64_t start, int count, float* lds, int lane) {
  constexpr int NT = num_tiles(NB), Q = qmax(NB);
  constexpr int CF = NB * 256;                // floats per staged chunk
  constexpr bool WRITER = (NB - 1) % 4 == W;  // gathers the last tile, writes the ratings
  const int g = lane >> 4, c = lane & 15;
  const int sr = lane >> 2, sq = lane & 3;
  const int ldf = (int)a.ldf;
  const bool last_in = 16 * (NB - 1) + 4 * sq < ldf;
  const int rcol = ldf - 16 * (NB - 1);  // the rating's column in the last tile
  const int nch = (count + KB - 1) / KB;
  f4 acc[Q];

  auto load_ci = [&](int ch) -> int {
    const int k = ch * KB + sr;
    return (ch < nch && k < count) ? a.col[start + k] : -1;
  };
  auto load_r = [&](int ch) -> float {
    const int k = ch * KB + lane;
    return (lane < KB && ch < nch && k < count) ? a.val[start + k] : 0.0f;
  };
  auto issue = [&](int ch, int ci) {
    float* yb = lds + (ch & 1) * CF;
    const float* rp = ci >= 0 ? a.theta + (int64_t)ci * a.ldf + 4 * sq : kZeroRow;
#pragma unroll
    for (int I = W; I < NB - 1; I += 4)
      __builtin_amdgcn_global_load_lds(rp + 16 * I, (lds_void_t*)(yb + I * 256), 16, 0, 0);
    if constexpr (WRITER) {
      // (kZeroRow holds 256 floats: a row past the range reads the last
      // tile from its start)
      const float* rl = (last_in && ci >= 0) ? rp + 16 * (NB - 1) : kZeroRow;
      __builtin_amdgcn_global_load_lds(rl, (lds_void_t*)(yb + (NB - 1) * 256), 16, 0, 0);
    }
  };
  auto mfma_chunk = [&](const float* yb, auto first) {
#pragma unroll
    for (int s = 0; s < KB / 4; ++s) {
      const int row = 4 * s + g;
      static_for<Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value, t = 4 * q + W;
        if constexpr (t < NT) {
          constexpr int I = tile_row(t, NB), J = tile_col(t, NB);
          const float ai = yb[I * 256 + row * 16 + c], aj = yb[J * 256 + row * 16 + c];
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(
              ai, aj, (decltype(first)::value && s == 0) ? f4{0.f, 0.f, 0.f, 0.f} : acc[q], 0, 0, 0);
        }
      });
    }
  };

  float rv = WRITER ? load_r(0) : 0.0f;
  issue(0, load_ci(0));
  int ci_next = load_ci(1);
  for (int ch = 0; ch < nch; ++ch) {
    wait_vm_all();  // this wavefront's gathers of chunk ch (and its rating loads) are done
    const float* yb = lds + (ch & 1) * CF;
    if constexpr (WRITER) {
      if (lane < KB) lds[(ch & 1) * CF + (NB - 1) * 256 + lane * 16 + rcol] = rv;
    }
    __syncthreads();  // chunk ch staged by all; chunk ch-1 consumed by all
    if (ch + 1 < nch) issue(ch + 1, ci_next);
    ci_next = load_ci(ch + 2);
    if constexpr (WRITER) rv = load_r(ch + 1);
    if (ch == 0) mfma_chunk(yb, std::true_type{});
    else mfma_chunk(yb, std::false_type{});
  }
  wait_vm_all();
  __syncthreads();  // every wavefront is done with the staging area

  const float ridge = a.lambda * (float)count;
  static_for<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value, t = 4 * q + W;
    if constexpr (t < NT) {
      constexpr int I = tile_row(t, NB), J = tile_col(t, NB);
      if constexpr (I == J) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (c == 4 * g + r && 16 * I + c < a.f) acc[q][r] += ridge;
      }
      *reinterpret_cast<f4*>(lds + t * 256 + 4 * lane) = acc[q];
    }
  });
}

// Dual form on the work-group (rows of n <= 16 NB - 1 ratings):
// wavefront W's tiles of K = Y Y^T (t = W, W + 4, ...), over 16-column
// slabs of the factor rows listed in cs (lane (g, c) loads columns
// 4g .. 4g+3 of row 16I + c of every row block I, the next slab's loads
// issued before the current slab's MFMAs); then r into column 16 NB - 1,
// mu = lambda n on the first n diagonal entries, and the store of the
// tiles to their LDS slots.
template <int NB, int W>
__device__ __forceinline__ void dual_tiles(const SweepArgs& a, int n, const int* cs, const float* rs, float* lds,
                                           int lane) {
  constexpr int NT = num_tiles(NB), Q = qmax(NB);
  const int g = lane >> 4, c = lane & 15;
  const int ldf = (int)a.ldf;
  const int S = (ldf + 15) >> 4;  // 16-column slabs; columns past ldf are zero
  const float* rp[NB];
#pragma unroll
  for (int I = 0; I < NB; ++I) {
    const int r = 16 * I + c;
    rp[I] = r < n ? a.theta + (int64_t)cs[r] * ldf : kZeroRow;  // (kZeroRow holds 256 >= ldf floats)
  }
  auto load = [&](int sl, f4 (&v)[NB]) {
    const int k = 16 * sl + 4 * g;
#pragma unroll
    for (int I = 0; I < NB; ++I)
      v[I] = k < ldf ? *reinterpret_cast<const f4*>(rp[I] + k) : f4{0.f, 0.f, 0.f, 0.f};
  };
  f4 acc[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) acc[q] = f4{0.f, 0.f, 0.f, 0.f};
  auto mfmas = [&](const f4 (&v)[NB]) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      static_for<Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value, t = 4 * q + W;
        if constexpr (t < NT) {
          constexpr int I = tile_row(t, NB), J = tile_col(t, NB);
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(v[I][j], v[J][j], acc[q], 0, 0, 0);
        }
      });
  };
  f4 va[NB], vb[NB];
  load(0, va);
  for (int sl = 0; sl < S; sl += 2) {
    if (sl + 1 < S) load(sl + 1, vb);
    mfmas(va);
    if (sl + 1 < S) {
      if (sl + 2 < S) load(sl + 2, va);
      mfmas(vb);
    }
  }
  const float ridge = a.lambda * (float)n;
  static_for<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value, t = 4 * q + W;
    if constexpr (t < NT) {
      constexpr int I = tile_row(t, NB), J = tile_col(t, NB);
      if constexpr (J == NB - 1) {
        // r into column 16 NB - 1 (K is zero there: rows past n are zero)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (c == 15) acc[q][r] = rs[16 * I + 4 * g + r];
      }
      if constexpr (I == J) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (c == 4 * g + r && 16 * I + c < n) acc[q][r] += ridge;
      }
      *reinterpret_cast<f4*>(lds + t * 256 + 4 * lane) = acc[q];
    }
  });
}

// Diagonal tile K (in accumulator order in its slot) -> W_K = U_KK^{-T},
// column-major, into wdst; for the last block also U_KK back into the slot
// (its augmented columns hold z).  Wavefront 0 only.
template <int NB>
__device__ __forceinline__ void diag_tile(float* slot, float* wdst, float* tsc, const float* ident, int live,
                                          bool last, int lane) {
  const int g = lane >> 4, c = lane & 15;
  const f4 tv = *reinterpret_cast<const f4*>(slot + 4 * lane);
  *reinterpret_cast<f4*>(tsc + 16 * c + 4 * g) = tv;  // column-major T
  lds_order();
  const bool trow = (g & 1) == 0;
  const float* src = (trow ? tsc : ident) + 16 * c;
  float av[16];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const f4 v = *reinterpret_cast<const f4*>(src + 4 * q);
#pragma unroll
    for (int r = 0; r < 4; ++r) av[4 * q + r] = v[r];
  }
  diag_pivots<true, false>(av, live, NoHook{}, std::make_integer_sequence<int, 16>{});
  if (last) {
#pragma unroll
    for (int i = 1; i < 16; ++i) av[i] = (trow && i > c) ? 0.0f : av[i];
  }
  lds_order();
  float* dst = trow ? tsc + 16 * c : wdst + 16 * c;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    *reinterpret_cast<f4*>(dst + 4 * q) = f4{av[4 * q], av[4 * q + 1], av[4 * q + 2], av[4 * q + 3]};
  lds_order();
  if (last) *reinterpret_cast<f4*>(slot + 4 * lane) = *reinterpret_cast<const f4*>(tsc + 16 * c + 4 * g);
}

// DUAL: the tasks are rows of n ratings, 16 (NB - 1) <= n <= 16 NB - 1,
// solved in the dual form on NB row blocks (K = Y Y^T with r in column
// 16 NB - 1, then x = Y^T z).
template <int NB, bool DUAL>
__global__ __launch_bounds__(kLT) __attribute__((amdgpu_waves_per_eu(1))) void als_large_kernel(SweepArgs a) {
  extern __shared__ float lds[];
  constexpr int NT = num_tiles(NB);
  // (the wavefront index is uniform: branches on it are scalar)
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  // Whole rows: task blockIdx.x of a.tasks, or row blockIdx.x without a list.
  const int64_t u = a.tasks ? (int64_t)a.tasks[blockIdx.x].row : (int64_t)blockIdx.x;
  const int f = a.f;
  const int64_t s0 = a.tasks ? (int64_t)a.tasks[blockIdx.x].start : (int64_t)a.row_ptr[u];
  const int cnt = a.tasks ? a.tasks[blockIdx.x].count : a.row_ptr[u + 1] - a.row_ptr[u];
  float* xrow = a.x + u * a.ldf;
  if (cnt <= 0 || (DUAL && cnt > 16 * NB - 1)) {  // (the host launches only rows that fit: n + 1 <= 16 NB)
    for (int i = threadIdx.x; i < (int)a.ldf; i += kLT) xrow[i] = 0.0f;
    if (cnt > 0 && threadIdx.x == 0) atomicAdd(a.fail_counter, 1);
    return;
  }
  // Size of the system: f, or n in the dual form; its augmented column.
  const int fs = DUAL ? cnt : f;
  const int cb = DUAL ? 15 : (int)a.ldf - 16 * (NB - 1);
  float* tsc = lds + NT * 256;
  float* ident = tsc + 256;
  float* wlast = ident + 256;
  float* zb = wlast + 256;
  float* xs = zb + 16 * NB;
  float* rs = xs + 16 * NB;  // dual form: the ratings and their column indices
  int* cs = reinterpret_cast<int*>(rs + 256);
  if constexpr (DUAL) {
    for (int i = threadIdx.x; i < 256; i += kLT) {
      const bool in = i < cnt;
      rs[i] = in ? a.val[s0 + i] : 0.0f;
      cs[i] = in ? a.col[s0 + i] : 0;
    }
    __syncthreads();
  }
  if (wv == 0) {
    const f4 e = {4 * g == c ? 1.0f : 0.0f, 4 * g + 1 == c ? 1.0f : 0.0f, 4 * g + 2 == c ? 1.0f : 0.0f,
                  4 * g + 3 == c ? 1.0f : 0.0f};
    *reinterpret_cast<f4*>(ident + 16 * c + 4 * g) = e;
  }

  if constexpr (DUAL) {
    switch (wv) {
      case 0: dual_tiles<NB, 0>(a, cnt, cs, rs, lds, lane); break;
      case 1: dual_tiles<NB, 1>(a, cnt, cs, rs, lds, lane); break;
      case 2: dual_tiles<NB, 2>(a, cnt, cs, rs, lds, lane); break;
      default: dual_tiles<NB, 3>(a, cnt, cs, rs, lds, lane); break;
    }
  } else {
    switch (wv) {
      case 0: gram_tiles<NB, 0>(a, s0, cnt, lds, lane); break;
      case 1: gram_tiles<NB, 1>(a, s0, cnt, lds, lane); break;
      case 2: gram_tiles<NB, 2>(a, s0, cnt, lds, lane); break;
      default: gram_tiles<NB, 3>(a, s0, cnt, lds, lane); break;
    }
  }
  __syncthreads();  // all tiles in their slots

  auto slot = [&](int I, int J) { return lds + tile_index(I, J, NB) * 256; };
  // A_IJ -= U_KI^T U_KJ on the N tiles (I, J0) .. (I, J0+N-1), their MFMA
  // chains interleaved.
  auto trailing_run = [&](int K, int I, int J0, auto n) {
    constexpr int N = decltype(n)::value;
    const f4 uv = *reinterpret_cast<const f4*>(slot(K, I) + 4 * lane);
    const f4 nu = {-uv[0], -uv[1], -uv[2], -uv[3]};
    f4 v[N], d[N];
#pragma unroll
    for (int t = 0; t < N; ++t) {
      v[t] = *reinterpret_cast<const f4*>(slot(K, J0 + t) + 4 * lane);
      d[t] = *reinterpret_cast<const f4*>(slot(I, J0 + t) + 4 * lane);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int t = 0; t < N; ++t) d[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(nu[s], v[t][s], d[t], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < N; ++t) *reinterpret_cast<f4*>(slot(I, J0 + t) + 4 * lane) = d[t];
  };
  // U_KJ = W_K A_KJ on the N tiles J0, J0+4, ..., likewise; A operand
  // W_K[c][4g+s] from the column-major W.
  auto panel_run = [&](int K, int J0, auto n) {
    constexpr int N = decltype(n)::value;
    const float* w = slot(K, K);
    f4 wa;
#pragma unroll
    for (int s = 0; s < 4; ++s) wa[s] = w[16 * (4 * g + s) + c];
    f4 b[N], d[N];
#pragma unroll
    for (int t = 0; t < N; ++t) {
      b[t] = *reinterpret_cast<const f4*>(slot(K, J0 + 4 * t) + 4 * lane);
      d[t] = f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int t = 0; t < N; ++t) d[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(wa[s], b[t][s], d[t], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < N; ++t) *reinterpret_cast<f4*>(slot(K, J0 + 4 * t) + 4 * lane) = d[t];
  };
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I4 = std::integral_constant<int, 4>;
  // Block 0's diagonal tile (NB >= 7 here, and the system has at least 96
  // pivots: block 0 is neither the last nor short of pivots).
  if (wv == 0) diag_tile<NB>(slot(0, 0), slot(0, 0), tsc, ident, fs, false, lane);
  __syncthreads();  // W_0 is in LDS
  for (int K = 0; K + 1 < NB; ++K) {
    {
      // Panel: wavefront wv forms tiles J = K+1+wv, +4, ... (at most four:
      // NB <= 17).
      const int J0 = K + 1 + wv;
      const int n = J0 < NB ? (NB - J0 + 3) / 4 : 0;
      if (n == 4) panel_run(K, J0, I4{});
      else if (n == 3) panel_run(K, J0, I3{});
      else if (n == 2) panel_run(K, J0, I2{});
      else if (n == 1) panel_run(K, J0, I1{});
    }
    __syncthreads();  // the panel is final
    // Lookahead: wavefront 0 updates the next diagonal tile and factors it
    // while wavefronts 1-3 apply the rest of block K's trailing update:
    // each row I = K+1 .. NB-1 is cut into runs of four tiles (I, J),
    // J >= I, dealt round-robin; a full run's four MFMA chains are
    // interleaved.  Only the last block can hold no pivot of the f x f
    // system; its diagonal tile is then not factored (its x is zero).
    const int K1 = K + 1;
    if (wv == 0) {
      trailing_run(K, K1, K1, I1{});
      lds_order();
      if (16 * K1 < fs) {
        const bool last = K1 + 1 == NB;
        diag_tile<NB>(slot(K1, K1), last ? wlast : slot(K1, K1), tsc, ident, fs - 16 * K1, last, lane);
      }
    } else {
      // Runs are numbered row by row; this wavefront takes runs
      // q = wv - 1 (mod 3).
      int q0 = 0;  // number of the row's first run
      for (int I = K1; I < NB; ++I) {
        const int j_first = I == K1 ? K1 + 1 : I;
        const int nr = (NB - j_first + 3) / 4;
        for (int r = ((wv - 1 - q0) % 3 + 3) % 3; r < nr; r += 3) {
          const int J0 = j_first + 4 * r;
          const int n = NB - J0;
          if (n >= 4) trailing_run(K, I, J0, I4{});
          else if (n == 3) trailing_run(K, I, J0, I3{});
          else if (n == 2) trailing_run(K, I, J0, I2{});
          else trailing_run(K, I, J0, I1{});
        }
        q0 += nr;
      }
    }
    __syncthreads();  // block K's trailing update and W_{K+1} are in LDS
  }
  if (wv != 0) return;

  // z out of the augmented column (column cb of block NB-1).
  float* sink = tsc + 4 * lane;
  for (int I = 0; I < NB; ++I) {
    const f4 v = *reinterpret_cast<const f4*>(slot(I, NB - 1) + 4 * lane);
    *reinterpret_cast<f4*>(c == cb ? zb + 16 * I + 4 * g : sink) = v;
  }
  lds_order();
  bool fin = true;
  for (int K = NB - 1; K >= 0; --K) {
    const float* w = (K + 1 == NB ? wlast : slot(K, K)) + 16 * c;
    const f4 w0 = *reinterpret_cast<const f4*>(w), w1 = *reinterpret_cast<const f4*>(w + 4);
    const f4 w2 = *reinterpret_cast<const f4*>(w + 8), w3 = *reinterpret_cast<const f4*>(w + 12);
    const float zv = zb[16 * K + 4 * g + (c >> 2)];
    float y[4] = {0.f, 0.f, 0.f, 0.f};
    for (int J = K + 1; J < NB; ++J) {
      const f4 uv = *reinterpret_cast<const f4*>(slot(K, J) + 4 * lane);
      const float xj = xs[16 * J + c];
#pragma unroll
      for (int r = 0; r < 4; ++r) y[r] = fmaf(uv[r], xj, y[r]);
    }
    // t_i for i = 4g + (c >> 2), so t_j sits in lane 4j.
    const float t = zv - row_reduce_scatter(y[0], y[1], y[2], y[3], c);
    const f4 wq[4] = {w0, w1, w2, w3};
    float x = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; ++j) x = fmaf(wq[j >> 2][j & 3], readlane_f(t, 4 * j), x);
    x = 16 * K + c < fs ? x : 0.0f;  // a block past the system's size has no W_K and gives x_K = 0
    fin = fin && fabsf(x) < __builtin_huge_valf();
    if (g == 0) xs[16 * K + c] = x;
    lds_order();
  }
  // A pivot that was not positive and finite leaves NaN or infinities in x.
  const bool ok = !__any(!fin);
  if (!ok && lane == 0) atomicAdd(a.fail_counter, 1);
  if constexpr (DUAL) {
    // x = Y^T z, z in xs: lane l forms columns 4l .. 4l+3.
    const int ldf = (int)a.ldf;
    const int k = 4 * lane;
    if (k >= ldf) return;
    f4 xv = {0.f, 0.f, 0.f, 0.f};
    if (ok) {
      int i = 0;
      for (; i + 4 <= cnt; i += 4) {
        f4 y[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) y[q] = *reinterpret_cast<const f4*>(a.theta + (int64_t)cs[i + q] * ldf + k);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float zq = xs[i + q];
#pragma unroll
          for (int e = 0; e < 4; ++e) xv[e] = fmaf(zq, y[q][e], xv[e]);
        }
      }
      for (; i < cnt; ++i) {
        const f4 y = *reinterpret_cast<const f4*>(a.theta + (int64_t)cs[i] * ldf + k);
        const float zq = xs[i];
#pragma unroll
        for (int e = 0; e < 4; ++e) xv[e] = fmaf(zq, y[e], xv[e]);
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) xv[e] = k + e < f ? xv[e] : 0.0f;
    *reinterpret_cast<f4*>(xrow + k) = xv;
  } else {
    for (int i = lane; i < (int)a.ldf; i += 64) xrow[i] = (ok && i < f) ? xs[i] : 0.0f;
  }
}

// Dual form: row blocks of the largest system (95 ratings plus r's column),
// and the LDS floats of a row: the Cholesky's space for that many blocks,
// then z, the ratings and the ratings' column indices (96 each).
constexpr int kLargeDualBlocks = (kDualMaxRatings + 16) / 16;
static_assert(kLargeDualBlocks == 6, "six row blocks");
__host__ __device__ constexpr int large_dual_lds_floats() { return chol_lds_floats(kLargeDualBlocks) + 3 * 96; }

// K = Y Y^T for the n rows listed in cs (n + 1 <= 16 NBD), r and the ridge
// added, factored; z into zs.  False if a pivot was not positive and
// finite.
template <int NBD>
__device__ __forceinline__ bool large_dual_blocks(const SweepArgs& a, int n, const int* cs, const float* rs,
                                                  float* lds, float* zs, int lane) {
  const int g = lane >> 4, c = lane & 15;
  const int ldf = (int)a.ldf;
  const int S = (ldf + 15) >> 4;  // 16-column slabs; columns past ldf are zero
  const float* rp[NBD];
#pragma unroll
  for (int I = 0; I < NBD; ++I) {
    const int r = 16 * I + c;
    rp[I] = r < n ? a.theta + (int64_t)cs[r] * ldf : kZeroRow;  // (kZeroRow holds 256 >= ldf floats)
  }
  auto load = [&](int s, f4 (&v)[NBD]) {
    const int k = 16 * s + 4 * g;
#pragma unroll
    for (int I = 0; I < NBD; ++I)
      v[I] = k < ldf ? *reinterpret_cast<const f4*>(rp[I] + k) : f4{0.f, 0.f, 0.f, 0.f};
  };
  f4 acc[num_tiles(NBD)];
#pragma unroll
  for (int t = 0; t < num_tiles(NBD); ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
  // Four MFMA steps per slab: in step j, lane (g, c) supplies column
  // 16 s + 4 g + j of row 16 I + c, so every column enters each tile once.
  auto mfmas = [&](const f4 (&v)[NBD]) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int I = 0; I < NBD; ++I)
#pragma unroll
        for (int J = I; J < NBD; ++J)
          acc[tile_index(I, J, NBD)] =
              __builtin_amdgcn_mfma_f32_16x16x4f32(v[I][j], v[J][j], acc[tile_index(I, J, NBD)], 0, 0, 0);
  };
  f4 va[NBD], vb[NBD];
  load(0, va);
  for (int s = 0; s < S; s += 2) {
    if (s + 1 < S) load(s + 1, vb);
    mfmas(va);
    if (s + 1 < S) {
      if (s + 2 < S) load(s + 2, va);
      mfmas(vb);
    }
  }
  // r into column 16 NBD - 1 of the last block column (K is zero there:
  // rows past n are zero and 16 NBD - 1 >= n), mu on the diagonal.
#pragma unroll
  for (int I = 0; I < NBD; ++I)
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (c == 15) acc[tile_index(I, NBD - 1, NBD)][q] = rs[16 * I + 4 * g + q];
  add_ridge<NBD>(acc, n, a.lambda * (float)n, lane);
  return chol_solve<NBD>(acc, n, 16 * NBD - 1, lds, zs, 16 * NBD, lane);
}

// One wavefront per task (a whole row of at most 95 ratings).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void als_large_dual_kernel(SweepArgs a) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x;
  const Task t = a.tasks[blockIdx.x];
  const int ldf = (int)a.ldf;
  float* xrow = a.x + (int64_t)t.row * a.ldf;
  const int n = t.count;
  if (n <= 0 || n > kDualMaxRatings) {  // (the host launches only rows that fit)
    for (int i = lane; i < ldf; i += 64) xrow[i] = 0.0f;
    if (n > 0 && lane == 0) atomicAdd(a.fail_counter, 1);
    return;
  }
  float* zs = lds + chol_lds_floats(kLargeDualBlocks);
  float* rs = zs + 96;
  int* cs = reinterpret_cast<int*>(rs + 96);
  for (int i = lane; i < 96; i += 64) {
    const bool in = i < n;
    rs[i] = in ? a.val[t.start + i] : 0.0f;
    cs[i] = in ? a.col[t.start + i] : 0;
  }
  lds_order();
  const int nbd = (n + 16) >> 4;  // blocks of the n x n system plus r's column
  bool ok;
  if (nbd == 1) ok = large_dual_blocks<1>(a, n, cs, rs, lds, zs, lane);
  else if (nbd == 2) ok = large_dual_blocks<2>(a, n, cs, rs, lds, zs, lane);
  else if (nbd == 3) ok = large_dual_blocks<3>(a, n, cs, rs, lds, zs, lane);
  else if (nbd == 4) ok = large_dual_blocks<4>(a, n, cs, rs, lds, zs, lane);
  else if (nbd == 5) ok = large_dual_blocks<5>(a, n, cs, rs, lds, zs, lane);
  else ok = large_dual_blocks<6>(a, n, cs, rs, lds, zs, lane);
  lds_order();
  // x = Y^T z (chol_solve left z = 0 if it failed): columns 4 lane .. +3.
  const int k = 4 * lane;
  if (k >= ldf) return;
  f4 xv = {0.f, 0.f, 0.f, 0.f};
  int i = 0;
  for (; i + 4 <= n; i += 4) {
    f4 y[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) y[q] = *reinterpret_cast<const f4*>(a.theta + (int64_t)cs[i + q] * ldf + k);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float zq = zs[i + q];
#pragma unroll
      for (int e = 0; e < 4; ++e) xv[e] = fmaf(zq, y[q][e], xv[e]);
    }
  }
  for (; i < n; ++i) {
    const f4 y = *reinterpret_cast<const f4*>(a.theta + (int64_t)cs[i] * ldf + k);
    const float zq = zs[i];
#pragma unroll
    for (int e = 0; e < 4; ++e) xv[e] = fmaf(zq, y[e], xv[e]);
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) xv[e] = (ok && k + e < a.f) ? xv[e] : 0.0f;
  *reinterpret_cast<f4*>(xrow + k) = xv;
  if (!ok && lane == 0) atomicAdd(a.fail_counter, 1);
}

template <int NB, bool DUAL>
hipError_t set_lds_limit() {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&als_large_kernel<NB, DUAL>),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(sizeof(float) * (DUAL ? large_dual_wg_lds_floats(NB) : large_lds_floats(NB))));
}

template <int NB, bool DUAL>
void launch_nb(const SweepArgs& a, int64_t rows, hipStream_t s) {
  hipLaunchKernelGGL((als_large_kernel<NB, DUAL>), dim3((unsigned)rows), dim3(kLT),
                     sizeof(float) * (DUAL ? large_dual_wg_lds_floats(NB) : large_lds_floats(NB)), s, a);
}

}  // namespace

size_t large_lds_bytes(int f) {
  const int NB = sweep_nb(f);
  return sizeof(float) * (size_t)large_lds_floats(NB);
}

hipError_t prepare_large_kernel(int f) {
  if (f <= 0 || f > kLargeMaxF) return hipErrorInvalidValue;
  // the work-group dual form, every block count it takes
  for (hipError_t e : {set_lds_limit<kLargeDualMinBlocks, true>(), set_lds_limit<8, true>(),
                       set_lds_limit<9, true>(), set_lds_limit<10, true>(), set_lds_limit<11, true>(),
                       set_lds_limit<12, true>(), set_lds_limit<13, true>(), set_lds_limit<14, true>(),
                       set_lds_limit<15, true>(), set_lds_limit<kLargeDualMaxBlocks, true>()})
    if (e != hipSuccess) return e;
  switch (sweep_nb(f)) {
    case 10: return set_lds_limit<10, false>();
    case 11: return set_lds_limit<11, false>();
    case 12: return set_lds_limit<12, false>();
    case 13: return set_lds_limit<13, false>();
    case 14: return set_lds_limit<14, false>();
    case 15: return set_lds_limit<15, false>();
    case 16: return set_lds_limit<16, false>();
    case 17: return set_lds_limit<17, false>();
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_large_dual_kernel(const SweepArgs& a, int64_t ntasks, hipStream_t s) {
  if (a.f <= 0 || a.f > kLargeMaxF || a.ldf > kLargeMaxF || ntasks < 0 || (ntasks > 0 && !a.tasks))
    return hipErrorInvalidValue;
  if (ntasks == 0) return hipSuccess;
  hipLaunchKernelGGL(als_large_dual_kernel, dim3((unsigned)ntasks), dim3(64),
                     sizeof(float) * large_dual_lds_floats(), s, a);
  return hipGetLastError();
}

hipError_t launch_large_dual_wg_kernel(const SweepArgs& a, int64_t ntasks, int nbd, hipStream_t s) {
  if (a.f <= 0 || a.f > kLargeMaxF || a.ldf > kLargeMaxF || ntasks < 0 || (ntasks > 0 && !a.tasks))
    return hipErrorInvalidValue;
  if (ntasks == 0) return hipSuccess;
  switch (nbd) {
    case 7: launch_nb<7, true>(a, ntasks, s); break;
    case 8: launch_nb<8, true>(a, ntasks, s); break;
    case 9: launch_nb<9, true>(a, ntasks, s); break;
    case 10: launch_nb<10, true>(a, ntasks, s); break;
    case 11: launch_nb<11, true>(a, ntasks, s); break;
    case 12: launch_nb<12, true>(a, ntasks, s); break;
    case 13: launch_nb<13, true>(a, ntasks, s); break;
    case 14: launch_nb<14, true>(a, ntasks, s); break;
    case 15: launch_nb<15, true>(a, ntasks, s); break;
    case 16: launch_nb<16, true>(a, ntasks, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_large_kernel(const SweepArgs& a, int64_t rows, hipStream_t s) {
  if (a.f <= 0 || a.f > kLargeMaxF || rows < 0) return hipErrorInvalidValue;
  if (rows == 0) return hipSuccess;
  switch (sweep_nb(a.f)) {
    case 10: launch_nb<10, false>(a, rows, s); break;
    case 11: launch_nb<11, false>(a, rows, s); break;
    case 12: launch_nb<12, false>(a, rows, s); break;
    case 13: launch_nb<13, false>(a, rows, s); break;
    case 14: launch_nb<14, false>(a, rows, s); break;
    case 15: launch_nb<15, false>(a, rows, s); break;
    case 16: launch_nb<16, false>(a, rows, s); break;
    case 17: launch_nb<17, false>(a, rows, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace cumf
