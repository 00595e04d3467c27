// als_device.h -- device-side building blocks of the ALS half-sweep kernels
// (included by als_sweep.hip and als_kernels.hip).
//
// One half-sweep solves, for every row u of a CSR rating matrix,
//     (Y_u^T Y_u + lambda n_u I) x_u = Y_u^T r_u
// where Y_u stacks the n_u factor rows of the other side named by row u's
// column indices (reference: get_hermitian100 / get_hermitianT10 at
// als/src/cuda/als.cu:47-912, the cusparseScsrmm2 RHS at :1039-1054 and
// the batched LU solve updateX at :935-997).
//
// One 64-lane wavefront per task (a row, or a chunk of a long row):
//   * Gram: the factor rows of each chunk of ratings are gathered into LDS
//     and the augmented Gram [Y r]^T [Y r] is accumulated on MFMA over the
//     NB(NB+1)/2 upper 16x16 tiles (FP32: v_mfma_f32_16x16x4_f32; BF16:
//     v_mfma_f32_16x16x32_bf16 with the rating split into a BF16 high and
//     low column).  The extra column(s) carry b = Y^T r.
//   * The accumulators stay in registers (lane l holds rows
//     16I + 4(l/16) + {0..3} of column 16J + l%16 of tile (I, J)), and a
//     blocked right-looking Cholesky runs on the tiles in place: per
//     16-column block K the diagonal tile is factored on the VALU into
//     W_K = U_KK^{-T} (column-per-lane elimination with DPP broadcasts),
//     then the panel U_KJ = W_K A_KJ and the trailing update
//     A_IJ -= U_KI^T U_KJ run on MFMA, leaving z = U^{-T} b in the
//     augmented column; a blocked back substitution with the W_K (in LDS)
//     gives x.
//   * Or the mixed-precision solve (cg_solve): FP16 CG on the Jacobi-scaled
//     matrix in LDS inside FP32 refinement with FP64 acceptance residuals.
//   * Rows with more than `split` ratings are cut into chunks whose partial
//     Grams go to a scratch buffer in register order; a second kernel sums
//     them in chunk order (deterministic) and runs the same solve.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "als_kernels.h"

namespace cumf {


typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) bf4 lds_bf4_t;

constexpr int KB = 16;  // ratings per staged FP32 chunk
constexpr int KBB = 32;  // ratings per BF16 chunk (one MFMA k-step)
// Staging buffers of the BF16 Gram (NB KB each): three (gathers two chunks
// ahead) while eight wavefronts of them fit the CU's 160 KB of LDS, else two
// (one chunk ahead) -- at NB = 7 three would leave room for seven
// wavefronts, and the eighth is worth more than the deeper prefetch.
__host__ __device__ constexpr int bf16_stage_bufs(int NB) { return 8 * 3 * NB <= 160 ? 3 : 2; }


__host__ __device__ constexpr int num_tiles(int NB) { return NB * (NB + 1) / 2; }
__host__ __device__ constexpr int tile_index(int I, int J, int NB) {
  return I * NB - I * (I - 1) / 2 + (J - I);
}

// LDS floats per wave: the staging buffers of the Gram phase (FP32: two
// chunks of NB tiles of 16 x 16 floats plus ratings; BF16: bf16_stage_bufs
// chunks of NB tiles of 32 x 16 bf16) and the buffers of the Cholesky solve (W blocks,
// diagonal tile, identity, z, z2) share the same space.
__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }
__host__ __device__ constexpr int chol_lds_floats(int NB) { return NB * 256 + 512 + 32 * NB; }
__host__ __device__ constexpr int lds_floats_nb(int NB, bool bf) {
  return cmax(bf ? bf16_stage_bufs(NB) * NB * 256 : 2 * NB * 256 + 2 * KB, chol_lds_floats(NB));
}


__device__ __forceinline__ float readlane_f(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over the 16 lanes of each row group of four per-lane values y0..y3
// (rows 4g + 0..3): every lane of quad q of the group gets row 4g + q.
// DPP moves only: row_ror:8 pairs lane c with c ^ 8 (rows {0,1} stay in the
// low half, {2,3} in the high half), row_half_mirror pairs each lane with one
// lane of the other quad of its half (each quad keeps one row), two
// quad_perm steps sum the quad.
__device__ __forceinline__ float row_reduce_scatter(float y0, float y1, float y2, float y3, int c) {
  const bool h8 = (c & 8) != 0, h4 = (c & 4) != 0;
  float k0 = h8 ? y2 : y0, k1 = h8 ? y3 : y1;
  const float s0 = h8 ? y0 : y2, s1 = h8 ? y1 : y3;
  k0 += dpp_f<0x128>(s0);  // row_ror:8
  k1 += dpp_f<0x128>(s1);
  float k = h4 ? k1 : k0;
  k += dpp_f<0x141>(h4 ? k0 : k1);  // row_half_mirror
  k += dpp_f<0x4E>(k);              // quad_perm [2,3,0,1]
  k += dpp_f<0xB1>(k);              // quad_perm [1,0,3,2]
  return k;
}

typedef __attribute__((address_space(3))) void lds_void_t;

// Waits for every outstanding vector-memory operation of this wave,
// including LDS-DMA loads (global_load_lds), whose completion the compiler
// does not track; also a compiler memory barrier.
__device__ __forceinline__ void wait_vm_all() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
// Waits until at most N vector-memory operations of this wave are
// outstanding (they complete in issue order), also a compiler barrier.
template <int N>
__device__ __forceinline__ void wait_vm_recent() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}
// Orders LDS accesses of one wave (LDS executes a wave's DS instructions in
// issue order, so only the compiler must be kept from reordering).
__device__ __forceinline__ void lds_order() {
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

// Zeros that gathers read in place of factor rows past the end of a range
// (and of columns past the end of a factor row): 1 KB, never written.
static __device__ const float kZeroRow[256] = {};

// ---------------------------------------------------------------------------
// Gram of a contiguous range of one row's ratings (count >= 1): acc is set
// to the range's augmented Gram, not added to.
//
// Staging: two LDS buffers per chunk of KB = 16 ratings, each NB tiles of
// 16 rows x 16 floats (tile I holds columns 16I .. 16I+15 of the chunk's
// factor rows, row stride 64 bytes), plus the KB ratings.  Load I of a
// chunk fills tile I: lane l moves the 16 bytes of row l/4, columns
// 16I + 4(l%4) .. +3, with global_load_lds_dwordx4 to byte 16l of the tile,
// so every lane serves one fixed row of the chunk and loads that row's
// column index itself (no cross-lane exchange).  sweep_nb(f) gives
// 16 (NB-1) <= ldf < 16 NB: tiles 0 .. NB-2 lie inside the factor row, the
// last tile holds the rest of it, the rating column (ldf) and padding.  In
// the last tile, quads at or past ldf -- and every quad of a row past the
// end of the range -- come from kZeroRow; the rating column is fed to the
// MFMA from the rating buffer.  Gathers run one chunk ahead of the MFMAs.
// ---------------------------------------------------------------------------
template <int NB>
__device__ __forceinline__ void gram_range(f4 (&acc)[num_tiles(NB)], const int32_t* __restrict__ col,
                                           const float* __restrict__ val,
                                           const float* __restrict__ theta, int64_t ldf,
                                           int64_t start, int count, float* __restrict__ lds,
                                           int lane) {
  const int g = lane >> 4, c = lane & 15;
  const int sr = lane >> 2, sq = lane & 3;  // staging row and quad of this lane
  constexpr int CF = NB * 256;              // floats per staged chunk
  float* rb = lds + 2 * CF;                 // ratings, [2][KB]
  const bool last_in = 16 * (NB - 1) + 4 * sq < (int)ldf;
  const bool rating_lane = 16 * (NB - 1) + c == (int)ldf;
  const int nch = (count + KB - 1) / KB;

  // Column index of this lane's staging row in chunk ch (-1 past the end).
  auto load_ci = [&](int ch) -> int {
    const int k = ch * KB + sr;
    return (ch < nch && k < count) ? col[start + k] : -1;
  };
  auto issue = [&](int ch, int ci) {
    float* yb = lds + (ch & 1) * CF;
    const float* rp = ci >= 0 ? theta + (int64_t)ci * ldf + 4 * sq : kZeroRow;
#pragma unroll
    for (int I = 0; I < NB - 1; ++I)
      __builtin_amdgcn_global_load_lds(rp + 16 * I, (lds_void_t*)(yb + I * 256), 16, 0, 0);
    const float* rl = last_in ? rp + 16 * (NB - 1) : kZeroRow;
    __builtin_amdgcn_global_load_lds(rl, (lds_void_t*)(yb + (NB - 1) * 256), 16, 0, 0);
    const int k = ch * KB + lane;
    if (lane < KB && k < count)
      __builtin_amdgcn_global_load_lds(val + start + k, (lds_void_t*)(rb + (ch & 1) * KB), 4, 0, 0);
  };

  // The MFMAs of one staged chunk; the first chunk of the range starts the
  // accumulators from zero (an inline-constant accumulator operand).
  auto mfma_chunk = [&](const float* yb, const float* rr, int valid, auto first) {
#pragma unroll
    for (int s = 0; s < KB / 4; ++s) {
      const int row = 4 * s + g;
      float a[NB];
#pragma unroll
      for (int I = 0; I < NB; ++I) a[I] = yb[I * 256 + row * 16 + c];
      if (rating_lane) a[NB - 1] = row < valid ? rr[row] : 0.0f;
#pragma unroll
      for (int I = 0; I < NB; ++I)
#pragma unroll
        for (int J = I; J < NB; ++J)
          acc[tile_index(I, J, NB)] = __builtin_amdgcn_mfma_f32_16x16x4f32(
              a[I], a[J], (decltype(first)::value && s == 0) ? f4{0.f, 0.f, 0.f, 0.f} : acc[tile_index(I, J, NB)],
              0, 0, 0);
    }
  };

  issue(0, load_ci(0));
  int ci_next = load_ci(1);
  for (int ch = 0; ch < nch; ++ch) {
    wait_vm_all();  // chunk ch is in LDS; the column index of ch+1 is in ci_next
    if (ch + 1 < nch) issue(ch + 1, ci_next);
    ci_next = load_ci(ch + 2);
    const float* yb = lds + (ch & 1) * CF;
    const float* rr = rb + (ch & 1) * KB;
    const int valid = count - ch * KB;
    if (ch == 0) mfma_chunk(yb, rr, valid, std::true_type{});
    else mfma_chunk(yb, rr, valid, std::false_type{});
    lds_order();
  }
  wait_vm_all();
}

// ---------------------------------------------------------------------------
// BF16 Gram (precision mode): the same augmented product on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
//
// Factor rows come from a BF16 shadow copy (row stride ldb = round_up(f,8),
// zero padded).  A chunk of 32 ratings is staged in LDS as NB tiles of
// 32 rows x 16 bf16 (row stride 32 bytes).  Load I of a chunk fills tile I:
// lane l moves the 16 bytes of row l/2, columns 16I + 8(l%2) .. +7, with
// global_load_lds_dwordx4 to byte 16l of the tile; each lane loads its own
// row's column index.  sweep_nb(f) gives 16 (NB-1) <= ldb, so tiles
// 0 .. NB-2 lie inside the factor row; in the last tile, slots at or past
// ldb -- and every slot of a row past the end of the range -- come from
// kZeroRow.  The rating r is then written into columns f and f+1 as bf16(r)
// and bf16(r - bf16(r)), so b = Y^T r keeps ~16 significant bits (exact for
// integer ratings).
//
// MFMA operands come straight from the tiles with ds_read_b64_tr_b16:
// within a 16-lane group, lane j addresses 4 bf16 of row (8*kg + j/4),
// columns 4*(j%4) of tile I; the transpose read returns to lane i the
// element (i%4) of the segment of lane 4s + i/4, i.e. column 16I + i of
// rows 8kg + s, s = 0..3.  A second read 4 rows further gives the 8
// consecutive k values the 16x16x32 operand needs.
// ---------------------------------------------------------------------------

template <int NB>
__device__ __forceinline__ void gram_range_bf16(f4 (&acc)[num_tiles(NB)], const int32_t* __restrict__ col,
                                                const float* __restrict__ val,
                                                const __bf16* __restrict__ theta_b, int64_t ldb, int f,
                                                int64_t start, int count, __bf16* __restrict__ lds, int lane) {
  constexpr int TB = KBB * 16;   // bf16 per staged tile
  constexpr int CB = NB * TB;    // bf16 per staged chunk
  constexpr int NBUF = bf16_stage_bufs(NB);
  const int sr = lane >> 1, sh = lane & 1;  // staging row and half of this lane
  const bool last_in = 16 * (NB - 1) + 8 * sh < (int)ldb;
  const __bf16* zb = reinterpret_cast<const __bf16*>(kZeroRow);
  const int kg = lane >> 4, j = lane & 15;
  const int nch = (count + KBB - 1) / KBB;
  // this lane's transpose-read offset inside a tile (bf16 units)
  const int troff = (8 * kg + (j >> 2)) * 16 + 4 * (j & 3);
  // where the rating halves go: tile and column of f and f+1
  const int rt0 = (f >> 4) * TB + (f & 15), rt1 = ((f + 1) >> 4) * TB + ((f + 1) & 15);

  auto load_ci = [&](int ch) -> int {
    const int k = ch * KBB + sr;
    return (ch < nch && k < count) ? col[start + k] : -1;
  };
  auto load_r = [&](int ch) -> float {
    const int k = ch * KBB + lane;
    return (lane < KBB && ch < nch && k < count) ? val[start + k] : 0.0f;
  };
  auto issue = [&](int ch, int ci) {
    __bf16* yb = lds + (ch % NBUF) * CB;
    const __bf16* rp = ci >= 0 ? theta_b + (int64_t)ci * ldb + 8 * sh : zb;
#pragma unroll
    for (int I = 0; I < NB - 1; ++I)
      __builtin_amdgcn_global_load_lds(rp + 16 * I, (lds_void_t*)(yb + I * TB), 16, 0, 0);
    const __bf16* rl = last_in ? rp + 16 * (NB - 1) : zb;
    __builtin_amdgcn_global_load_lds(rl, (lds_void_t*)(yb + (NB - 1) * TB), 16, 0, 0);
  };

  auto chunk = [&](int ch, float r) {
    __bf16* yb = lds + (ch % NBUF) * CB;
    if (lane < KBB && ch * KBB + lane < count) {
      const __bf16 hi = (__bf16)r;
      const __bf16 lo = (__bf16)(r - (float)hi);
      yb[rt0 + 16 * lane] = hi;
      yb[rt1 + 16 * lane] = lo;
    }
    return yb;
  };
  // The MFMAs of one staged chunk; the first chunk of the range starts the
  // accumulators from zero (an inline-constant accumulator operand).
  auto mfmas = [&](const __bf16* yb, auto first) {
    lds_order();
    bf8 op[NB];
#pragma unroll
    for (int I = 0; I < NB; ++I) {
      const bf4 lo4 = __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf4_t*)(yb + I * TB + troff));
      const bf4 hi4 = __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf4_t*)(yb + I * TB + troff + 64));
      op[I] = __builtin_shufflevector(lo4, hi4, 0, 1, 2, 3, 4, 5, 6, 7);
    }
#pragma unroll
    for (int I = 0; I < NB; ++I)
#pragma unroll
      for (int J = I; J < NB; ++J)
        acc[tile_index(I, J, NB)] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            op[I], op[J], decltype(first)::value ? f4{0.f, 0.f, 0.f, 0.f} : acc[tile_index(I, J, NB)], 0, 0, 0);
    lds_order();
  };
  auto mfmas_any = [&](int ch, const __bf16* yb) {
    if (ch == 0) mfmas(yb, std::true_type{});
    else mfmas(yb, std::false_type{});
  };

  if constexpr (NBUF == 3) {
    // Gathers run two chunks ahead; column indices and ratings three
    // chunks ahead, in registers.
    int ci1 = load_ci(1), ci2 = load_ci(2);
    float rv = load_r(0), r1 = load_r(1), r2 = load_r(2);
    issue(0, load_ci(0));
    if (1 < nch) issue(1, ci1);
    int cn = load_ci(3);
    float rn = load_r(3);
    for (int ch = 0; ch < nch; ++ch) {
      // Chunk ch's gathers are older than the last NB vector-memory
      // operations whenever the two following chunks were issued (the next
      // chunk's NB gathers plus the index/rating loads came after them);
      // near the end of the range wait for everything.
      if (ch + 2 < nch) wait_vm_recent<NB>();
      else wait_vm_all();
      const __bf16* yb = chunk(ch, rv);
      if (ch + 2 < nch) issue(ch + 2, ci2);
      asm volatile("" ::: "memory");  // keep the next loads after the gathers
      rv = r1;
      r1 = r2;
      r2 = rn;
      ci2 = cn;
      cn = load_ci(ch + 4);
      rn = load_r(ch + 4);
      mfmas_any(ch, yb);
    }
  } else {
    // Gathers run one chunk ahead; column indices and ratings two chunks
    // ahead, in registers.
    int ci1 = load_ci(1);
    float rv = load_r(0), r1 = load_r(1);
    issue(0, load_ci(0));
    for (int ch = 0; ch < nch; ++ch) {
      wait_vm_all();  // chunk ch is in LDS; indices and ratings of ch+1 are in registers
      const __bf16* yb = chunk(ch, rv);
      if (ch + 1 < nch) issue(ch + 1, ci1);
      asm volatile("" ::: "memory");
      rv = r1;
      ci1 = load_ci(ch + 2);
      r1 = load_r(ch + 2);
      mfmas_any(ch, yb);
    }
  }
  wait_vm_all();
}

// ---------------------------------------------------------------------------
// Blocked Cholesky on the accumulator tiles, with the O(f^3) work on MFMA.
//
// Right-looking over the 16-row blocks K of the augmented Gram:
//   1. chol_diag: the diagonal tile T is factored by Gauss elimination of
//      [T | I] run column-wise -- lane c of the even 16-lane rows holds
//      column c of T, lane c of the odd rows column c of the identity, 16
//      registers each.  Pivot k scales row k by 1/sqrt(t_kk) and subtracts
//      u_ki * row k from every row i > k; the multipliers u_ki (lane i of
//      the T row) are copied into the identity rows with one permlane swap
//      and read by v_fmac_f32_dpp row_newbcast:i, one instruction per
//      multiplier.  The T half ends as U_KK (U_KK^T U_KK = T), the identity
//      half as W_K = U_KK^{-T}.  Pivots at or past f are not taken: those
//      rows of U and W are zero, so they contribute nothing below.
//   2. chol_panel: U_KJ = W_K A_KJ for J > K on v_mfma_f32_16x16x4f32.  In
//      the accumulator layout (lane (g, c), register s holds M[4g+s][c]) a
//      tile is the B operand of the four k-steps as it stands, with
//      k = 4g + s; the A operand of step s is W_K[c][4g+s], read from LDS.
//   3. chol_trailing: A_IJ -= U_KI^T U_KJ for K < I <= J: register s of tile
//      (K, I), negated, and of tile (K, J) are the A and B operands of step
//      s, so the update moves no data at all.
// The augmented column(s) ride along and end as z = U^{-T} b.  Each W_K
// stays in LDS (column-major) for the back substitution, which runs block
// by block from the last: x_K = W_K^T (z_K - sum_{J > K} U_KJ x_J).
//
// LDS (floats): W_K for every block (NB x 256), the diagonal tile being
// factored (256), a 16 x 16 identity (256), z and the second right-hand
// side's z (16 NB each).
// ---------------------------------------------------------------------------
// Elimination step of pivot K in the column-per-lane factorisation:
// a[i] -= u_i * a[K] for i = K+1..15, where the multiplier u_i is the value
// of u held by lane i of the lane's own 16-lane row, read (negated) with
// DPP row_newbcast:i -- one v_fmac_f32_dpp per multiplier.  The s_nop
// covers the two wait states a DPP read needs after the VALU write of u.
template <int K>
__device__ __forceinline__ void elim_dpp(float (&a)[16], float u) {
  if constexpr (K == 0) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, -%15, %16 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, -%15, %16 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, -%15, %16 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, -%15, %16 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %4, -%15, %16 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %5, -%15, %16 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %6, -%15, %16 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %7, -%15, %16 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %8, -%15, %16 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %9, -%15, %16 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %10, -%15, %16 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %11, -%15, %16 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %12, -%15, %16 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %13, -%15, %16 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %14, -%15, %16 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        : "+v"(a[1]), "+v"(a[2]), "+v"(a[3]), "+v"(a[4]), "+v"(a[5]), "+v"(a[6]), "+v"(a[7]), "+v"(a[8]), "+v"(a[9]), "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]), "+v"(a[14]), "+v"(a[15])
        : "v"(u), "v"(a[0]));
  }
  else if constexpr (K == 1) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, -%14, %15 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, -%14, %15 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, -%14, %15 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, -%14, %15 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %4, -%14, %15 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %5, -%14, %15 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %6, -%14, %15 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %7, -%14, %15 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %8, -%14, %15 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %9, -%14, %15 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %10, -%14, %15 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %11, -%14, %15 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %12, -%14, %15 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %13, -%14, %15 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        : "+v"(a[2]), "+v"(a[3]), "+v"(a[4]), "+v"(a[5]), "+v"(a[6]), "+v"(a[7]), "+v"(a[8]), "+v"(a[9]), "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]), "+v"(a[14]), "+v"(a[15])
        : "v"(u), "v"(a[1]));
  }
  else if constexpr (K == 2) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, -%13, %14 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, -%13, %14 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, -%13, %14 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, -%13, %14 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %4, -%13, %14 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %5, -%13, %14 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %6, -%13, %14 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %7, -%13, %14 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %8, -%13, %14 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %9, -%13, %14 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %10, -%13, %14 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %11, -%13, %14 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %12, -%13, %14 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        : "+v"(a[3]), "+v"(a[4]), "+v"(a[5]), "+v"(a[6]), "+v"(a[7]), "+v"(a[8]), "+v"(a[9]), "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]), "+v"(a[14]), "+v"(a[15])
        : "v"(u), "v"(a[2]));
  }
  else if constexpr (K == 3) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, -%12, %13 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, -%12, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, -%12, %13 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, -%12, %13 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %4, -%12, %13 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %5, -%12, %13 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %6, -%12, %13 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %7, -%12, %13 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %8, -%12, %13 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %9, -%12, %13 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %10, -%12, %13 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %11, -%12, %13 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        : "+v"(a[4]), "+v"(a[5]), "+v"(a[6]), "+v"(a[7]), "+v"(a[8]), "+v"(a[9]), "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]), "+v"(a[14]), "+v"(a[15])
        : "v"(u), "v"(a[3]));
  }
  else if constexpr (K == 4) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, -%11, %12 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, -%11, %12 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, -%11, %12 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, -%11, %12 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %4, -%11, %12 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %5, -%11, %12 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %6, -%11, %12 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %7, -%11, %12 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %8, -%11, %12 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %9, -%11, %12 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %10, -%11, %12 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        : "+v"(a[5]), "+v"(a[6]), "+v"(a[7]), "+v"(a[8]), "+v"(a[9]), "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]), "+v"(a[14]), "+v"(a[15])
        : "v"(u), "v"(a[4]));
  }
  else if constexpr (K == 5) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, -%10, %11 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, -%10, %11 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, -%10, %11 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, -%10, %11 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %4, -%10, %11 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %5, -%10, %11 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %6, -%10, %11 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %7, -%10, %11 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %8, -%10, %11 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %9, -%10, %11 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        : "+v"(a[6]), "+v"(a[7]), "+v"(a[8]), "+v"(a[9]), "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]), "+v"(a[14]), "+v"(a[15])
        : "v"(u), "v"(a[5]));
  }
  else if constexpr (K == 6) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, -%9, %10 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, -%9, %10 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, -%9, %10 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, -%9, %10 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %4, -%9, %10 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %5, -%9, %10 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %6, -%9, %10 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %7, -%9, %10 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %8, -%9, %10 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        : "+v"(a[7]), "+v"(a[8]), "+v"(a[9]), "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]), "+v"(a[14]), "+v"(a[15])
        : "v"(u), "v"(a[6]));
  }
  else if constexpr (K == 7) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, -%8, %9 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, -%8, %9 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, -%8, %9 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, -%8, %9 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %4, -%8, %9 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %5, -%8, %9 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %6, -%8, %9 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %7, -%8, %9 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        : "+v"(a[8]), "+v"(a[9]), "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]), "+v"(a[14]), "+v"(a[15])
        : "v"(u), "v"(a[7]));
  }
  else if constexpr (K == 8) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, -%7, %8 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, -%7, %8 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, -%7, %8 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, -%7, %8 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %4, -%7, %8 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %5, -%7, %8 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %6, -%7, %8 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        : "+v"(a[9]), "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]), "+v"(a[14]), "+v"(a[15])
        : "v"(u), "v"(a[8]));
  }
  else if constexpr (K == 9) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, -%6, %7 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, -%6, %7 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, -%6, %7 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, -%6, %7 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %4, -%6, %7 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %5, -%6, %7 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        : "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]), "+v"(a[14]), "+v"(a[15])
        : "v"(u), "v"(a[9]));
  }
  else if constexpr (K == 10) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, -%5, %6 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, -%5, %6 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, -%5, %6 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, -%5, %6 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %4, -%5, %6 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        : "+v"(a[11]), "+v"(a[12]), "+v"(a[13]), "+v"(a[14]), "+v"(a[15])
        : "v"(u), "v"(a[10]));
  }
  else if constexpr (K == 11) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, -%4, %5 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, -%4, %5 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, -%4, %5 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, -%4, %5 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        : "+v"(a[12]), "+v"(a[13]), "+v"(a[14]), "+v"(a[15])
        : "v"(u), "v"(a[11]));
  }
  else if constexpr (K == 12) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, -%3, %4 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, -%3, %4 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, -%3, %4 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        : "+v"(a[13]), "+v"(a[14]), "+v"(a[15])
        : "v"(u), "v"(a[12]));
  }
  else if constexpr (K == 13) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, -%2, %3 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, -%2, %3 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        : "+v"(a[14]), "+v"(a[15])
        : "v"(u), "v"(a[13]));
  }
  else if constexpr (K == 14) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, -%1, %2 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        : "+v"(a[15])
        : "v"(u), "v"(a[14]));
  }
}

// Pivot k of the column-per-lane factorisation (live = pivots of the block
// inside the f x f system).  A pivot at or past f scales its row by zero,
// and zero multipliers leave the other rows alone, so the step is
// branch-free and the compiler may interleave a block's VALU work with the
// previous block's trailing MFMAs; only the last block (GUARD: nothing
// follows it) skips the elimination of such pivots.  PARTIAL: the block may
// hold such pivots (without it every pivot is live and the select goes
// away).  A pivot that is not
// positive and finite turns its row into NaN or infinities, which reach
// the solution; chol_solve tests the solution instead of every pivot.
template <int k, bool PARTIAL, bool GUARD>
__device__ __forceinline__ void diag_pivot(float (&a)[16], int live) {
  const bool act = !PARTIAL || k < live;
  const float d = readlane_f(a[k], k);
  a[k] = PARTIAL ? (act ? a[k] * __builtin_amdgcn_rsqf(d) : 0.0f) : a[k] * __builtin_amdgcn_rsqf(d);
  if constexpr (k < 15) {
    if (!GUARD || act) {
      // Row k of U (u_kj in lane j of the T rows) copied into the identity
      // rows: v_permlane16_swap exchanges the odd rows of its first operand
      // with the even rows of its second, so the first result holds the T
      // rows' values in every row.
      const auto pr = __builtin_amdgcn_permlane16_swap(__float_as_int(a[k]), __float_as_int(a[k]), false, false);
      elim_dpp<k>(a, __int_as_float(pr[0]));
    }
  }
}
// The 16 pivots; after pivot k, hook(integral_constant<k>) issues work
// that does not depend on the elimination (see chol_step).
struct NoHook {
  template <typename T>
  __device__ __forceinline__ void operator()(T) const {}
};
template <bool PARTIAL, bool GUARD, typename H, int... ks>
__device__ __forceinline__ void diag_pivots(float (&a)[16], int live, const H& hook,
                                            std::integer_sequence<int, ks...>) {
  ((diag_pivot<ks, PARTIAL, GUARD>(a, live), hook(std::integral_constant<int, ks>{})), ...);
}

template <int NB, int K, typename H>
__device__ __forceinline__ void chol_diag(f4 (&acc)[num_tiles(NB)], int f, float* __restrict__ wcm,
                                          float* __restrict__ tsc, const float* __restrict__ ident, int lane,
                                          const H& hook) {
  const int g = lane >> 4, c = lane & 15;
  constexpr int TK = tile_index(K, K, NB);
  // Column-major copy of T: tsc[16 c + i] = T[i][c].
  *reinterpret_cast<f4*>(tsc + 16 * c + 4 * g) = acc[TK];
  lds_order();
  // Rows 0 and 2 of the wave (lanes 0..15, 32..47) hold the columns of T,
  // rows 1 and 3 the identity columns (rows 2, 3 repeat rows 0, 1), both
  // read from LDS.
  const bool trow = (g & 1) == 0;
  const float* src = (trow ? tsc : ident) + 16 * c;
  float a[16];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const f4 v = *reinterpret_cast<const f4*>(src + 4 * q);
#pragma unroll
    for (int r = 0; r < 4; ++r) a[4 * q + r] = v[r];
  }
  // Only the last two blocks can hold pivots at or past f (sweep_nb gives
  // 16 (NB - 1) <= f + 6, so f - 16 K >= 26 for K <= NB - 3).
  diag_pivots<K + 2 >= NB, K + 1 == NB>(a, f - 16 * K, hook, std::make_integer_sequence<int, 16>{});
  if constexpr (K + 1 == NB) {
    // The last block's U_KK goes back to the accumulators (its augmented
    // columns hold z), upper triangular: what elimination leaves below the
    // diagonal is rounding residue.  Earlier blocks' U_KK is not read again.
#pragma unroll
    for (int i = 1; i < 16; ++i) a[i] = (trow && i > c) ? 0.0f : a[i];
  }
  lds_order();
  // U_KK back into tsc (column-major), W_K into its slot; rows 2 and 3
  // write the same values as rows 0 and 1.
  float* dst = trow ? tsc + 16 * c : wcm + 256 * K + 16 * c;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    *reinterpret_cast<f4*>(dst + 4 * q) = f4{a[4 * q], a[4 * q + 1], a[4 * q + 2], a[4 * q + 3]};
  lds_order();
  if constexpr (K + 1 == NB) acc[TK] = *reinterpret_cast<const f4*>(tsc + 16 * c + 4 * g);
}

template <int NB, int K>
__device__ __forceinline__ void chol_panel(f4 (&acc)[num_tiles(NB)], const float* __restrict__ wcm, int lane) {
  const int g = lane >> 4, c = lane & 15;
  const float* w = wcm + 256 * K;
  f4 wa;
#pragma unroll
  for (int s = 0; s < 4; ++s) wa[s] = w[16 * (4 * g + s) + c];  // W_K[c][4g+s]
#pragma unroll
  for (int J = K + 1; J < NB; ++J) {
    const f4 b = acc[tile_index(K, J, NB)];
    f4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; ++s) d = __builtin_amdgcn_mfma_f32_16x16x4f32(wa[s], b[s], d, 0, 0, 0);
    acc[tile_index(K, J, NB)] = d;
  }
}

// Trailing update of block K, A_IJ -= U_KI^T U_KJ for K < I <= J, as a
// list of MFMAs: tile (K+1, K+1) first (the next diagonal factorisation
// waits for it), then the rest of row K+1 (the next panel reads it), then
// rows K+2 ..; four MFMAs per tile, m = 4 u + s for the u-th tile.
__host__ __device__ constexpr int trailing_tiles(int NB, int K) { return (NB - K - 1) * (NB - K) / 2; }
__host__ __device__ constexpr int trailing_tile_row(int NB, int K, int u) {
  int I = K + 1;
  while (u >= NB - I) {
    u -= NB - I;
    ++I;
  }
  return I;
}
__host__ __device__ constexpr int trailing_tile_col(int NB, int K, int u) {
  int I = K + 1;
  while (u >= NB - I) {
    u -= NB - I;
    ++I;
  }
  return I + u;
}
template <int NB, int K, int m>
__device__ __forceinline__ void trailing_mfma(f4 (&acc)[num_tiles(NB)]) {
  constexpr int u = m / 4, s = m % 4;
  constexpr int I = trailing_tile_row(NB, K, u), J = trailing_tile_col(NB, K, u);
  constexpr int tij = tile_index(I, J, NB), tki = tile_index(K, I, NB), tkj = tile_index(K, J, NB);
  acc[tij] = __builtin_amdgcn_mfma_f32_16x16x4f32(-acc[tki][s], acc[tkj][s], acc[tij], 0, 0, 0);
}
template <int NB, int K, int lo, int... ms>
__device__ __forceinline__ void trailing_mfmas_seq(f4 (&acc)[num_tiles(NB)], std::integer_sequence<int, ms...>) {
  (trailing_mfma<NB, K, lo + ms>(acc), ...);
}
// MFMAs lo .. hi-1 of block K's trailing update.
template <int NB, int K, int lo, int hi>
__device__ __forceinline__ void trailing_mfmas(f4 (&acc)[num_tiles(NB)]) {
  if constexpr (hi > lo) trailing_mfmas_seq<NB, K, lo>(acc, std::make_integer_sequence<int, hi - lo>{});
}

// Block K's diagonal tile is factored (W_K in LDS): its panel, then its
// trailing update and the next block, with one block of lookahead.  Tile
// (K+1, K+1) is updated first; of the rest, up to 32 MFMAs are issued
// between the 16 pivots of block K+1's diagonal factorisation (two after
// each), where the matrix pipe runs them beside the elimination's VALU
// work, and the remainder just before it.  Nothing the elimination reads
// is written by them: it reads only tile (K+1, K+1) and LDS.
template <int NB, int K>
__device__ __forceinline__ void chol_step(f4 (&acc)[num_tiles(NB)], int f, float* __restrict__ wcm,
                                          float* __restrict__ tsc, const float* __restrict__ ident, int lane) {
  if constexpr (K + 1 < NB) {
    chol_panel<NB, K>(acc, wcm, lane);
    constexpr int M = 4 * trailing_tiles(NB, K);  // MFMAs of the whole trailing update
    constexpr int R = M - 4 < 32 ? M - 4 : 32;     // interleaved with the next diagonal tile
    constexpr int P = M - R;                       // issued before it (tile (K+1, K+1) included)
    trailing_mfmas<NB, K, 0, P>(acc);
    // The last block may hold only the right-hand side column(s); then
    // nothing follows (R = 0).
    if (K + 2 < NB || 16 * (K + 1) < f) {
      const auto hook = [&](auto kc) {
        constexpr int k = decltype(kc)::value;
        __builtin_amdgcn_sched_barrier(0);
        trailing_mfmas<NB, K, P + k * R / 16, P + (k + 1) * R / 16>(acc);
        __builtin_amdgcn_sched_barrier(0);
      };
      chol_diag<NB, K + 1>(acc, f, wcm, tsc, ident, lane, hook);
    }
    chol_step<NB, K + 1>(acc, f, wcm, tsc, ident, lane);
  }
}

template <int NB>
__device__ __forceinline__ void chol_blocks(f4 (&acc)[num_tiles(NB)], int f, float* __restrict__ wcm,
                                            float* __restrict__ tsc, const float* __restrict__ ident, int lane) {
  chol_diag<NB, 0>(acc, f, wcm, tsc, ident, lane, NoHook{});
  chol_step<NB, 0>(acc, f, wcm, tsc, ident, lane);
}


template <int NB, int K>
__device__ __forceinline__ void chol_backsub(const f4 (&acc)[num_tiles(NB)], int f, const float* __restrict__ wcm,
                                             const float* __restrict__ zb, int lane, float (&xc)[NB]) {
  const int g = lane >> 4, c = lane & 15;
  {
    // Column c of W_K and this lane's z entry first (independent of x).
    // A block past f has W_K = 0 and gives x_K = 0.
    const float* w = wcm + 256 * K + 16 * c;
    const f4 w0 = *reinterpret_cast<const f4*>(w), w1 = *reinterpret_cast<const f4*>(w + 4);
    const f4 w2 = *reinterpret_cast<const f4*>(w + 8), w3 = *reinterpret_cast<const f4*>(w + 12);
    const float zv = zb[16 * K + 4 * g + (c >> 2)];
    float y[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int J = K + 1; J < NB; ++J)
#pragma unroll
      for (int r = 0; r < 4; ++r) y[r] = fmaf(acc[tile_index(K, J, NB)][r], xc[J], y[r]);
    // t_i for i = 4g + (c >> 2), so t_j sits in lane 4j.
    const float t = zv - row_reduce_scatter(y[0], y[1], y[2], y[3], c);
    const f4 wv[4] = {w0, w1, w2, w3};
    float x = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; ++j) x = fmaf(wv[j >> 2][j & 3], readlane_f(t, 4 * j), x);
    xc[K] = 16 * K + c < f ? x : 0.0f;
  }
  if constexpr (K > 0) chol_backsub<NB, K - 1>(acc, f, wcm, zb, lane, xc);
}

// bcol2 >= 0: a second right-hand-side column whose solution is added to
// the first (the BF16 low part of the ratings).  xb: optional BF16 copy of
// the solution (row of ldb entries, zero padded).  Returns false (and
// writes x = 0) if a pivot was not positive.
template <int NB>
__device__ __forceinline__ bool chol_solve(f4 (&acc)[num_tiles(NB)], int f, int bcol,
                                           float* __restrict__ lds, float* __restrict__ xrow,
                                           int64_t ldf, int lane, int bcol2 = -1,
                                           __bf16* __restrict__ xb = nullptr, int64_t ldb = 0) {
  const int g = lane >> 4, c = lane & 15;
  float* wcm = lds;
  float* tsc = lds + NB * 256;
  float* ident = tsc + 256;
  float* zb = ident + 256;
  float* zb2 = zb + 16 * NB;
  {
    // 16 x 16 identity: lane (g, c) writes rows 4g..4g+3 of column c.
    const f4 e = {4 * g == c ? 1.0f : 0.0f, 4 * g + 1 == c ? 1.0f : 0.0f, 4 * g + 2 == c ? 1.0f : 0.0f,
                  4 * g + 3 == c ? 1.0f : 0.0f};
    *reinterpret_cast<f4*>(ident + 16 * c + 4 * g) = e;
  }
  chol_blocks<NB>(acc, f, wcm, tsc, ident, lane);

  // z (and the second column's z) out of the augmented column(s), which
  // sweep_nb places in the last column block (bcol, bcol2 >= 16 (NB - 1)).
  // Lanes holding other columns store into the free diagonal-tile buffer.
  const int cb = bcol & 15, cb2 = bcol2 & 15;
  float* sink = tsc + 4 * lane;
#pragma unroll
  for (int I = 0; I < NB; ++I) {
    const f4 v = acc[tile_index(I, NB - 1, NB)];
    *reinterpret_cast<f4*>(c == cb ? zb + 16 * I + 4 * g : sink) = v;
    if (bcol2 >= 0) *reinterpret_cast<f4*>(c == cb2 ? zb2 + 16 * I + 4 * g : sink) = v;
  }
  lds_order();
  if (bcol2 >= 0) {
#pragma unroll
    for (int q = 0; q < (16 * NB + 63) / 64; ++q) {
      const int i = lane + 64 * q;
      if (i < 16 * NB) zb[i] += zb2[i];
    }
    lds_order();
  }
  float xc[NB];
#pragma unroll
  for (int J = 0; J < NB; ++J) xc[J] = 0.0f;
  chol_backsub<NB, NB - 1>(acc, f, wcm, zb, lane, xc);
  // A pivot that was not positive and finite leaves NaN or infinities in x.
  bool fin = true;
#pragma unroll
  for (int K = 0; K < NB; ++K) fin = fin && fabsf(xc[K]) < __builtin_huge_valf();
  const bool ok = !__any(!fin);
  if (g == 0) {
#pragma unroll
    for (int K = 0; K < NB; ++K) {
      const int i = 16 * K + c;
      const float xv = ok ? xc[K] : 0.0f;
      if (i < ldf) xrow[i] = xv;
      if (xb && i < ldb) xb[i] = (__bf16)xv;
    }
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
// Mixed-precision solve: conjugate gradients on the Jacobi-scaled matrix
// A^ = D^-1/2 A D^-1/2 held in LDS as FP16 (unit diagonal, |entries| <= 1,
// so FP16 holds it without range trouble), inside FP32 iterative
// refinement whose residual r = b - A x is formed from the FP32 Gram still
// in registers: in FP64 whenever it can accept the row, in FP32 for the
// start (it only drives the first round).  Each refinement round solves
// A^ d = D^-1/2 r / mu (mu = max |D^-1/2 r|, so the FP16 search directions
// stay O(1)) with a fixed number of CG iterations from d = 0 and adds
// mu D^-1/2 d to x.  The start is the row's current factor (the previous
// sweep's value), rescaled when far off.  A row whose relative residual
// ||b - A x|| / ||b|| has not reached the tolerance after the last round is
// solved again by the Cholesky path.
//
// LDS (after the Gram phase): A^ (f rows, stride cg_stride(f) halves), the
// FP16 direction p, x (16NB floats) and y = A x (16NB doubles).  Vectors
// live one entry per lane (entry i in lane i % 64, slot i / 64).
// ---------------------------------------------------------------------------
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

// Row stride of A^ in halves: f rounded up to 8, then to an odd multiple of
// 8, so that 16-byte row reads of consecutive lanes hit distinct banks.
__host__ __device__ constexpr int cg_stride(int f) {
  return (((f + 7) / 8) & 1) ? ((f + 7) & ~7) : ((f + 7) & ~7) + 8;
}
__host__ __device__ constexpr int cg_lds_floats(int NB, int f) {
  return (f + 1) * cg_stride(f) / 2 + 16 * NB + 32 * NB;
}
__host__ __device__ constexpr int rows_per_lane(int NB) { return (16 * NB + 63) / 64; }

// Sum / max over the wave, returned uniform: quad butterflies and two row
// rotations give every lane its 16-lane row total; the four row totals are
// combined in a fixed order from lanes 0, 16, 32, 48.
__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f<0x124>(v);  // row_ror:4
  v += dpp_f<0x128>(v);  // row_ror:8
  return (readlane_f(v, 0) + readlane_f(v, 16)) + (readlane_f(v, 32) + readlane_f(v, 48));
}
__device__ __forceinline__ float wave_max(float v) {
  v = fmaxf(v, dpp_f<0xB1>(v));
  v = fmaxf(v, dpp_f<0x4E>(v));
  v = fmaxf(v, dpp_f<0x124>(v));
  v = fmaxf(v, dpp_f<0x128>(v));
  return fmaxf(fmaxf(readlane_f(v, 0), readlane_f(v, 16)), fmaxf(readlane_f(v, 32), readlane_f(v, 48)));
}
__device__ __forceinline__ double readlane_d(double v, int lane) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}
// FP64 sum over the wave, uniform, in a fixed order (as wave_sum).
__device__ __forceinline__ double wave_sum_d(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x124>(v);
  v += dpp_d<0x128>(v);
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

// Every lane gets its value plus that of lane l ^ 16 (l ^ 32): the
// permlane swaps (two VALU moves, no LDS) exchange whole 16-lane (32-lane)
// rows; both lanes of a pair add the same two values.
__device__ __forceinline__ double swap_sum16(double v) {
  const int64_t b = __double_as_longlong(v);
  const auto lo = __builtin_amdgcn_permlane16_swap((int)(b & 0xffffffff), (int)(b & 0xffffffff), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((int)(b >> 32), (int)(b >> 32), false, false);
  const double a = __longlong_as_double(((int64_t)hi[0] << 32) | (uint32_t)lo[0]);
  const double o = __longlong_as_double(((int64_t)hi[1] << 32) | (uint32_t)lo[1]);
  return a + o;
}
__device__ __forceinline__ double swap_sum32(double v) {
  const int64_t b = __double_as_longlong(v);
  const auto lo = __builtin_amdgcn_permlane32_swap((int)(b & 0xffffffff), (int)(b & 0xffffffff), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((int)(b >> 32), (int)(b >> 32), false, false);
  const double a = __longlong_as_double(((int64_t)hi[0] << 32) | (uint32_t)lo[0]);
  const double o = __longlong_as_double(((int64_t)hi[1] << 32) | (uint32_t)lo[1]);
  return a + o;
}

// A register value the compiler must treat as changed at this point: keeps
// loop-invariant FP64 conversions of the Gram from being hoisted out of the
// refinement loop (which would hold 2x the Gram's registers live).
__device__ __forceinline__ f4 opaque(f4 v) {
  asm volatile("" : "+v"(v));
  return v;
}

// ys[0, 16NB) = A xs in FP64, A the Gram in registers (upper tiles; the
// diagonal tiles are full), xs zero beyond f.  Row direction: lane (g, c)
// of tile (I, J) adds A[16I+4g+r][16J+c] x[16J+c] to row 16I+4g+r; the 16
// lanes of a group are reduced with a reduce-scatter that leaves row
// 16I+4g+(c>>2) in lanes c with equal c>>2, all in DPP moves: row_ror:8
// pairs lane c with c ^ 8 (rows {0,1} stay in the low half, {2,3} in the
// high half), row_half_mirror pairs each lane with one lane of the other
// quad of its half (each quad keeps one row), and two quad_perm steps sum
// the quad.  Column direction (strictly upper tiles, the lower triangle by
// symmetry): row 16J+c collects A[16I+4g+r][16J+c] x[16I+4g+r] over r,
// I < J and the four groups (permlane swaps across the 16-lane rows).
template <int NB>
__device__ __forceinline__ void gram_matvec_f64(const f4 (&acc)[num_tiles(NB)], const float* __restrict__ xs,
                                                double* __restrict__ ys, int lane) {
  const int g = lane >> 4, c = lane & 15;
  double xc[NB];
#pragma unroll
  for (int J = 0; J < NB; ++J) xc[J] = (double)xs[16 * J + c];
  const bool h8 = (c & 8) != 0, h4b = (c & 4) != 0;
#pragma unroll
  for (int I = 0; I < NB; ++I) {
    double y0 = 0.0, y1 = 0.0, y2 = 0.0, y3 = 0.0;
#pragma unroll
    for (int J = I; J < NB; ++J) {
      const f4 v = opaque(acc[tile_index(I, J, NB)]);
      y0 = fma((double)v[0], xc[J], y0);
      y1 = fma((double)v[1], xc[J], y1);
      y2 = fma((double)v[2], xc[J], y2);
      y3 = fma((double)v[3], xc[J], y3);
    }
    double k0 = h8 ? y2 : y0, k1 = h8 ? y3 : y1;
    const double s0 = h8 ? y0 : y2, s1 = h8 ? y1 : y3;
    k0 += dpp_d<0x128>(s0);  // row_ror:8 (lane c ^ 8)
    k1 += dpp_d<0x128>(s1);
    double k = h4b ? k1 : k0;
    const double s = h4b ? k0 : k1;
    k += dpp_d<0x141>(s);    // row_half_mirror (lane 7 - c within the half)
    k += dpp_d<0x4E>(k);     // quad_perm [2,3,0,1]
    k += dpp_d<0xB1>(k);     // quad_perm [1,0,3,2]
    if ((c & 3) == 0) ys[16 * I + 4 * g + (c >> 2)] = k;
  }
  lds_order();
#pragma unroll
  for (int J = 1; J < NB; ++J) {
    double yc = 0.0;
#pragma unroll
    for (int I = 0; I < J; ++I) {
      const f4 xr = *reinterpret_cast<const f4*>(xs + 16 * I + 4 * g);
      const f4 v = opaque(acc[tile_index(I, J, NB)]);
#pragma unroll
      for (int r = 0; r < 4; ++r) yc = fma((double)v[r], (double)xr[r], yc);
    }
    yc = swap_sum32(swap_sum16(yc));
    if (g == 0) ys[16 * J + c] += yc;
  }
  lds_order();
}

// ysf[0, 16NB) = A xs with FP32 products and sums: the same tile walk and
// reductions as gram_matvec_f64, for the start's residual, which only
// drives the first CG round (no acceptance test reads it).
__device__ __forceinline__ float swap_sum16f(float v) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_int(v), __float_as_int(v), false, false);
  return __int_as_float(r[0]) + __int_as_float(r[1]);
}
__device__ __forceinline__ float swap_sum32f(float v) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_int(v), __float_as_int(v), false, false);
  return __int_as_float(r[0]) + __int_as_float(r[1]);
}
template <int NB>
__device__ __forceinline__ void gram_matvec_f32(const f4 (&acc)[num_tiles(NB)], const float* __restrict__ xs,
                                                float* __restrict__ ysf, int lane) {
  const int g = lane >> 4, c = lane & 15;
  float xc[NB];
#pragma unroll
  for (int J = 0; J < NB; ++J) xc[J] = xs[16 * J + c];
  const bool h8 = (c & 8) != 0, h4b = (c & 4) != 0;
#pragma unroll
  for (int I = 0; I < NB; ++I) {
    float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f, y3 = 0.0f;
#pragma unroll
    for (int J = I; J < NB; ++J) {
      const f4 v = acc[tile_index(I, J, NB)];
      y0 = fmaf(v[0], xc[J], y0);
      y1 = fmaf(v[1], xc[J], y1);
      y2 = fmaf(v[2], xc[J], y2);
      y3 = fmaf(v[3], xc[J], y3);
    }
    float k0 = h8 ? y2 : y0, k1 = h8 ? y3 : y1;
    const float s0 = h8 ? y0 : y2, s1 = h8 ? y1 : y3;
    k0 += dpp_f<0x128>(s0);
    k1 += dpp_f<0x128>(s1);
    float k = h4b ? k1 : k0;
    const float s = h4b ? k0 : k1;
    k += dpp_f<0x141>(s);
    k += dpp_f<0x4E>(k);
    k += dpp_f<0xB1>(k);
    if ((c & 3) == 0) ysf[16 * I + 4 * g + (c >> 2)] = k;
  }
  lds_order();
#pragma unroll
  for (int J = 1; J < NB; ++J) {
    float yc = 0.0f;
#pragma unroll
    for (int I = 0; I < J; ++I) {
      const f4 xr = *reinterpret_cast<const f4*>(xs + 16 * I + 4 * g);
      const f4 v = acc[tile_index(I, J, NB)];
#pragma unroll
      for (int r = 0; r < 4; ++r) yc = fmaf(v[r], xr[r], yc);
    }
    yc = swap_sum32f(swap_sum16f(yc));
    if (g == 0) ysf[16 * J + c] += yc;
  }
  lds_order();
}

// q = A^ p for the rows this lane owns (FP16 operands, FP32 accumulation).
// Rows past f are computed on row f-1 (same-address reads broadcast) and
// zeroed.
template <int NB>
__device__ __forceinline__ void cg_matvec(const _Float16* __restrict__ Ah, const _Float16* __restrict__ p16,
                                          int f, int sa, int nch, int lane, float (&q)[rows_per_lane(NB)]) {
  constexpr int RPL = rows_per_lane(NB);
  const _Float16* rowp[RPL];
  float q2[RPL];
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int i = lane + 64 * t;
    rowp[t] = Ah + (i < f ? i : f - 1) * sa;
    q[t] = 0.0f;
    q2[t] = 0.0f;
  }
#pragma unroll 4
  for (int k = 0; k < nch; ++k) {
    const h8 pk = *reinterpret_cast<const h8*>(p16 + 8 * k);
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
      const h8 av = *reinterpret_cast<const h8*>(rowp[t] + 8 * k);
      q[t] = __builtin_amdgcn_fdot2(h2{av[0], av[1]}, h2{pk[0], pk[1]}, q[t], false);
      q2[t] = __builtin_amdgcn_fdot2(h2{av[2], av[3]}, h2{pk[2], pk[3]}, q2[t], false);
      q[t] = __builtin_amdgcn_fdot2(h2{av[4], av[5]}, h2{pk[4], pk[5]}, q[t], false);
      q2[t] = __builtin_amdgcn_fdot2(h2{av[6], av[7]}, h2{pk[6], pk[7]}, q2[t], false);
    }
  }
#pragma unroll
  for (int t = 0; t < RPL; ++t) q[t] = (lane + 64 * t < f) ? q[t] + q2[t] : 0.0f;
}

struct CgParams {
  int iters;   // CG iterations per refinement round
  int rounds;  // refinement rounds
  float tol;   // accept x once ||b - A x|| <= tol ||b||
  int warm;    // start: 0 the row's current factor, 1 zero, 2 x0row, 3 the better of 0 and 2
};

// Squared FP64 residual norm of the x in xs (lane-per-row b in bv).
template <int NB>
__device__ __forceinline__ double residual_norm2(const f4 (&acc)[num_tiles(NB)], const float* xs, double* ys,
                                                 const float (&bv)[rows_per_lane(NB)], int f, int lane) {
  gram_matvec_f64<NB>(acc, xs, ys, lane);
  double rr = 0.0;
#pragma unroll
  for (int t = 0; t < rows_per_lane(NB); ++t) {
    const int i = lane + 64 * t;
    const double r = i < f ? (double)bv[t] - ys[i] : 0.0;
    rr += r * r;
  }
  return wave_sum_d(rr);
}

// Returns true (and writes x to xrow / xb) if x was accepted; false leaves
// the row to the Cholesky fallback kernel.
// x0row: the network's prediction for this row (warm 2 and 3).  trace
// (optional): initial and final relative residual, rounds run, fell back.
// xpre (optional): this lane's entries of the current factor, already
// loaded by the caller.
template <int NB>
__device__ __forceinline__ bool cg_solve(f4 (&acc)[num_tiles(NB)], int f, int bcol, float* __restrict__ lds,
                                         float* __restrict__ xrow, int64_t ldf, int lane, int bcol2,
                                         __bf16* __restrict__ xb, int64_t ldb, const CgParams& cp,
                                         const float* __restrict__ x0row, float* __restrict__ trace,
                                         const float* xpre = nullptr) {
  constexpr int RPL = rows_per_lane(NB);
  constexpr int D = 16 * NB;
  const int g = lane >> 4, c = lane & 15;
  const int sa = cg_stride(f);
  const int f8 = (f + 7) & ~7;
  _Float16* Ah = reinterpret_cast<_Float16*>(lds);
  _Float16* p16 = Ah + f * sa;
  float* xs = lds + (f + 1) * sa / 2;
  double* ys = reinterpret_cast<double*>(xs + D);
  float* ds = xs;                                  // diagonal, during set-up
  float* bs = reinterpret_cast<float*>(ys);        // b (and the BF16 low part at bs + D)
  float* ss = reinterpret_cast<float*>(ys);        // Jacobi scales, after b is read

  // Diagonal and right-hand side out of the register tiles.  Lane (g, c)
  // of diagonal tile I holds a_ii for i = 16I + c when c >> 2 == g (its
  // register c & 3); the right-hand side is the column bcol (bcol2) of the
  // block-column bcol >> 4, held by the lanes with c == bcol & 15.  Every
  // lane stores (rows past f are written too and never read); lanes that
  // hold nothing wanted write into a sink at the start of the matrix area
  // (16 bytes per lane where the area holds 1 KB, shared by lanes for small
  // f), which the FP16 matrix overwrites later.  No store is predicated, so
  // the set-up has no exec-mask branches.
  lds_order();
  const int jb = bcol >> 4, cb = bcol & 15;
  const int jb2 = bcol2 >= 0 ? bcol2 >> 4 : -1, cb2 = bcol2 & 15;
  const int span = min(256, (f * sa) >> 1);  // floats of the matrix area, capped (>= 4)
  float* sink = lds + ((4 * lane) & ((1 << (31 - __clz(span))) - 1));
#pragma unroll
  for (int I = 0; I < NB; ++I) {
    const f4 dv = acc[tile_index(I, I, NB)];
    const bool odd = (c & 1) != 0;
    const float dlo = odd ? dv[1] : dv[0], dhi = odd ? dv[3] : dv[2];
    const float dval = (c & 2) ? dhi : dlo;  // register c & 3, as selects
    *((c >> 2) == g ? ds + 16 * I + c : sink) = dval;
#pragma unroll
    for (int J = I; J < NB; ++J) {
      if (J == jb) *reinterpret_cast<f4*>(c == cb ? bs + 16 * I + 4 * g : sink) = acc[tile_index(I, J, NB)];
      if (J == jb2) *reinterpret_cast<f4*>(c == cb2 ? bs + D + 16 * I + 4 * g : sink) = acc[tile_index(I, J, NB)];
    }
  }
  lds_order();
  float sc[RPL], bv[RPL], xv[RPL], xm[RPL];
  bool bad = false, mbad = false;
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int i = lane + 64 * t;
    const bool in = i < f;
    const float d = in ? ds[i] : 1.0f;
    const bool good = d > 0.0f && d < __builtin_huge_valf();
    bad = bad || !good;
    sc[t] = (in && good) ? __builtin_amdgcn_rsqf(d) : 0.0f;
    bv[t] = in ? bs[i] + (bcol2 >= 0 ? bs[D + i] : 0.0f) : 0.0f;
    float x0 = (in && cp.warm != 1) ? (xpre ? xpre[t] : xrow[i]) : 0.0f;
    if (!(fabsf(x0) < __builtin_huge_valf())) x0 = 0.0f;
    xv[t] = x0;
    xm[t] = (in && cp.warm >= 2) ? x0row[i] : 0.0f;
    mbad = mbad || !(fabsf(xm[t]) < __builtin_huge_valf());
  }
  const bool net_bad = cp.warm >= 2 && __any(mbad);
  // A non-finite prediction is never used: with warm 2 the row goes to the
  // Cholesky solve, with warm 3 the current factor is the start.
  const bool any_bad = __any(bad) || (cp.warm == 2 && net_bad);
  if (cp.warm == 2 && !net_bad) {
#pragma unroll
    for (int t = 0; t < RPL; ++t) xv[t] = xm[t];
  }
  lds_order();
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int i = lane + 64 * t;
    if (i < D) {
      ss[i] = sc[t];
      xs[i] = xv[t];
    }
  }
  lds_order();
  bool accepted = false;
  float rel0 = __builtin_nanf(""), rel = __builtin_nanf("");
  int rounds_run = 0;
  if (!any_bad) {
    // A^ in FP16: tile (I, J) writes rows 16I+4g+r at column 16J+c and, for
    // I < J, row 16J+c at columns 16I+4g..+3 (one 8-byte store).  Entries
    // outside the f x f8 window go to the row after the matrix (the CG
    // direction vector, sa >= 8 halves, rewritten before every read), so
    // no store is predicated.  NB = sweep_nb(f) gives f >= 16 NB - 22, f8 >= 16 NB - 16:
    // row blocks up to NB-3 and column blocks up to NB-2 are always inside
    // (compile-time), and a transposed store's columns 16I+4g+3 <=
    // 16 NB - 17 always are.
    // The column scales stay in registers; each block row's scales are read
    // one block ahead, so no store waits on a scale read.
    float scc[NB];
#pragma unroll
    for (int J = 0; J < NB; ++J) scc[J] = ss[16 * J + c];
    f4 srn = *reinterpret_cast<const f4*>(ss + 4 * g);
#pragma unroll
    for (int I = 0; I < NB; ++I) {
      const f4 sr = srn;
      if (I + 1 < NB) srn = *reinterpret_cast<const f4*>(ss + 16 * (I + 1) + 4 * g);
#pragma unroll
      for (int J = I; J < NB; ++J) {
        const int j = 16 * J + c;
        const float scj = scc[J];
        const f4 v = acc[tile_index(I, J, NB)];
        h4 tv;
#pragma unroll
        for (int r = 0; r < 4; ++r) tv[r] = (_Float16)(v[r] * sr[r] * scj);
        _Float16* drow = Ah + (16 * I + 4 * g) * sa + j;
        const bool col_ok = J <= NB - 2 || j < f8;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const bool ok = col_ok && (I <= NB - 3 || 16 * I + 4 * g + r < f);
          *(ok ? drow + r * sa : p16 + (c & 7)) = tv[r];
        }
        if (I < J) {
          const bool ok = J <= NB - 3 || j < f;
          *reinterpret_cast<h4*>(ok ? Ah + j * sa + 16 * I + 4 * g : p16 + 16 * I + 4 * g) = tv;
        }
      }
    }
    double bb = 0.0;
#pragma unroll
    for (int t = 0; t < RPL; ++t) bb += (double)bv[t] * (double)bv[t];
    bb = wave_sum_d(bb);
    const double tol2 = (double)cp.tol * (double)cp.tol * bb;
    const int nch = f8 / 8;
    if (cp.warm == 3 && !net_bad) {
      // Start from the better of the current factor (in xs) and the prediction.
      const double rp = residual_norm2<NB>(acc, xs, ys, bv, f, lane);
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int i = lane + 64 * t;
        if (i < D) xs[i] = i < f ? xm[t] : 0.0f;
      }
      lds_order();
      const double rm = residual_norm2<NB>(acc, xs, ys, bv, f, lane);
      const bool use_net = rm < rp;
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int i = lane + 64 * t;
        if (use_net) xv[t] = xm[t];
        if (i < D) xs[i] = i < f ? xv[t] : 0.0f;
      }
      lds_order();
    }
    float* ysf = reinterpret_cast<float*>(ys);
    for (int round = 0;; ++round) {
      // Residual of the current x (xs holds it): the start's in FP32 (it
      // only drives the first CG round), every later one, which decides
      // acceptance, in FP64.
      if (round == 0) gram_matvec_f32<NB>(acc, xs, ysf, lane);
      else gram_matvec_f64<NB>(acc, xs, ys, lane);
      auto yat = [&](int i) -> double { return round == 0 ? (double)ysf[i] : ys[i]; };
      // A start far from the solution (the factor of the previous sweep
      // after the other side changed a lot) is first replaced by its
      // energy-optimal multiple s x0, s = <x0, b> / <x0, A x0>, whose
      // A-norm error is never above that of x0 or of zero.  The residual
      // b - s A x0 then comes from the same product.  Starts with
      // |s - 1| <= 1/8 are kept as they are.
      double s = 1.0;
      if (round == 0 && cp.warm != 1) {
        double xbd = 0.0, xad = 0.0;
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const int i = lane + 64 * t;
          if (i < f) {
            xbd += (double)xv[t] * (double)bv[t];
            xad += (double)xv[t] * yat(i);
          }
        }
        xbd = wave_sum_d(xbd);
        xad = wave_sum_d(xad);
        const double so = xbd / xad;
        if (xad > 0.0 && !(fabs(so - 1.0) <= 0.125)) s = so;
      }
      const bool scaled = s != 1.0;
      float rh[RPL];
      double rr = 0.0;
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int i = lane + 64 * t;
        const double r = i < f ? (double)bv[t] - s * yat(i) : 0.0;
        rr += r * r;
        rh[t] = (float)r * sc[t];
        if (scaled) xv[t] = (float)(s * (double)xv[t]);
      }
      rr = wave_sum_d(rr);
      if (round == 0) rel0 = (float)sqrt(rr / bb);
      rel = (float)sqrt(rr / bb);
      rounds_run = round;
      if (rr <= tol2 && round > 0) {
        accepted = true;
        break;
      }
      if (round == cp.rounds) break;
      float mu = 0.0f;
#pragma unroll
      for (int t = 0; t < RPL; ++t) mu = fmaxf(mu, fabsf(rh[t]));
      mu = wave_max(mu);
      if (!(mu > 0.0f && mu < __builtin_huge_valf())) break;
      const float imu = 1.0f / mu;
      float dv[RPL], rv[RPL], pv[RPL], q[RPL];
      float rho = 0.0f;
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        rv[t] = rh[t] * imu;
        pv[t] = rv[t];
        dv[t] = 0.0f;
        rho = fmaf(rv[t], rv[t], rho);
      }
      rho = wave_sum(rho);
      for (int it = 0; it < cp.iters; ++it) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const int i = lane + 64 * t;
          if (i < f8) p16[i] = (_Float16)pv[t];
        }
        lds_order();
        cg_matvec<NB>(Ah, p16, f, sa, nch, lane, q);
        float pq = 0.0f;
#pragma unroll
        for (int t = 0; t < RPL; ++t) pq = fmaf(pv[t], q[t], pq);
        pq = wave_sum(pq);
        if (!(pq > 0.0f) || !(rho > 0.0f)) break;
        const float alpha = rho * __builtin_amdgcn_rcpf(pq);  // step lengths need no IEEE division
        float rn = 0.0f;
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          dv[t] = fmaf(alpha, pv[t], dv[t]);
          rv[t] = fmaf(-alpha, q[t], rv[t]);
          rn = fmaf(rv[t], rv[t], rn);
        }
        rn = wave_sum(rn);
        const float beta = rn * __builtin_amdgcn_rcpf(rho);
        rho = rn;
#pragma unroll
        for (int t = 0; t < RPL; ++t) pv[t] = fmaf(beta, pv[t], rv[t]);
        lds_order();
      }
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int i = lane + 64 * t;
        xv[t] = fmaf(mu * sc[t], dv[t], xv[t]);
        if (i < D) xs[i] = i < f ? xv[t] : 0.0f;
      }
      lds_order();
    }
  }
  if (trace && lane == 0) {
    trace[0] = rel0;
    trace[1] = rel;
    trace[2] = (float)rounds_run;
    trace[3] = accepted ? 0.0f : 1.0f;
  }
  if (!accepted) return false;
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int i = lane + 64 * t;
    const float xo = i < f ? xv[t] : 0.0f;
    if (i < ldf) xrow[i] = xo;
    if (xb && i < ldb) xb[i] = (__bf16)xo;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
template <int NB, bool BF>
__device__ __forceinline__ void gram_any(f4 (&acc)[num_tiles(NB)], const SweepArgs& a, int64_t start, int count,
                                         float* lds, int lane) {
  if constexpr (BF)
    gram_range_bf16<NB>(acc, a.col, a.val, reinterpret_cast<const __bf16*>(a.theta_b), a.ldb, a.f,
                        start, count, reinterpret_cast<__bf16*>(lds), lane);
  else
    gram_range<NB>(acc, a.col, a.val, a.theta, a.ldf, start, count, lds, lane);
}

// Solves the system in registers for one row: Cholesky, or (CG) the
// mixed-precision refinement; a row CG does not accept is appended to the
// fallback list as `entry` (row >= 0, or -1 - index of a split row).
template <int NB, bool CG>
__device__ __forceinline__ void solve_row(f4 (&acc)[num_tiles(NB)], const SweepArgs& a, bool bf, float* lds,
                                          int64_t row, __bf16* xb, int lane, int entry,
                                          const float* xpre = nullptr) {
  float* xrow = a.x + row * a.ldf;
  const int bcol = bf ? a.f : (int)a.ldf;
  const int bcol2 = bf ? a.f + 1 : -1;
  if constexpr (CG) {
    const CgParams cp{a.cg_iters, a.cg_rounds, a.cg_tol, a.warm};
    const float* x0row = a.x0 ? a.x0 + row * a.ldf : nullptr;
    float* trace = a.cg_trace ? a.cg_trace + 4 * row : nullptr;
    const bool accepted =
        cg_solve<NB>(acc, a.f, bcol, lds, xrow, a.ldf, lane, bcol2, xb, a.ldb, cp, x0row, trace, xpre);
    if (!accepted && lane == 0) {
      a.fb_list[atomicAdd(a.fb_count, 1)] = entry;
      atomicAdd(a.cg_fallback, 1);
    }
  } else {
    const bool ok = chol_solve<NB>(acc, a.f, bcol, lds, xrow, a.ldf, lane, bcol2, xb, a.ldb);
    if (!ok && lane == 0) atomicAdd(a.fail_counter, 1);
  }
}

// ---------------------------------------------------------------------------
// Dual (kernel) form for rows with few ratings, BF16 factors.  With Y the
// n x f matrix of the row's gathered factor rows, r its ratings and
// mu = lambda n, the push-through identity
//     (Y^T Y + mu I)^{-1} Y^T r = Y^T (Y Y^T + mu I)^{-1} r
// gives the same x from an n x n system: z solves (Y Y^T + mu I) z = r and
// x = Y^T z.  For n <= kDualMaxRatings (95) that system has at most six
// 16-column blocks against NB for the f x f one (seven at f = 100), so the
// Cholesky costs a fraction of the primal one (the primal Gram's MFMA work
// is about the same as the dual's).
//   * Y is gathered as in gram_range_bf16: chunks of 32 rows, NB tiles of
//     32 rows x 16 bf16, all CH chunks (1 for rows of up to 31 ratings, 2 up
//     to 63, 3 up to 95) at once (rows past n are zeros).
//   * K = Y Y^T on v_mfma_f32_16x16x32_bf16: for k-step s and row block I,
//     lane (kg, i) supplies row 16I + i, columns 32s + 8kg .. +7 (one
//     16-byte LDS read; columns past the staged tiles read a zero block).
//     The same read is the A operand of block I and the B operand of block
//     J = I.
//   * r goes into column 16 NBD - 1 (past every live pivot), mu onto the
//     first n diagonal entries, and chol_solve factors the NBD blocks and
//     writes z to LDS.
//   * x = Y^T z in FP32 on the VALU: the rows are gathered again (the
//     Cholesky reused their LDS) and lane l forms columns 2l, 2l+1 from
//     them, rating by rating.
// ---------------------------------------------------------------------------
constexpr int kDualMaxRatings = kDualMaxRatingsHost;  // three staged chunks
static_assert(kDualMaxRatings == 3 * KBB - 1, "the dual form stages at most three chunks");

// Row blocks (of 16) of the largest dual system CH staged chunks allow:
// n <= 32 CH - 1 ratings plus r's column.
__host__ __device__ constexpr int dual_max_blocks(int CH) { return (32 * CH - 1 + 16) / 16; }
// LDS floats of the dual form: the CH staged chunks (the Cholesky's space
// once the Gram is formed), then z, the ratings and a zero block.
template <int NB, int CH>
__host__ __device__ constexpr int dual_zs_offset() {
  return cmax(CH * NB * 256, chol_lds_floats(dual_max_blocks(CH)));
}
template <int NB, int CH>
__host__ __device__ constexpr int dual_lds_floats() {
  return dual_zs_offset<NB, CH>() + 96 + 96 + 4;
}

template <int NB, int CH, int NBD>
__device__ __forceinline__ bool dual_blocks(const SweepArgs& a, int n, float* lds, float* zs, const float* rs,
                                            const __bf16* zero, int lane) {
  constexpr int TB = KBB * 16, CB = NB * TB;  // bf16 per staged tile / chunk
  constexpr int S = (NB + 1) / 2;             // 32-column k-steps over the staged tiles
  static_assert(NBD <= dual_max_blocks(CH), "rows past the staged chunks");
  const __bf16* yb = reinterpret_cast<const __bf16*>(lds);
  const int kg = lane >> 4, i = lane & 15;
  f4 acc[num_tiles(NBD)];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int t = 2 * s + (kg >> 1);  // tile of this lane's 8 columns
    bf8 op[NBD];
#pragma unroll
    for (int I = 0; I < NBD; ++I) {
      const int r = 16 * I + i;
      const __bf16* p = t < NB ? yb + (r >> 5) * CB + t * TB + (r & 31) * 16 + 8 * (kg & 1) : zero;
      op[I] = *reinterpret_cast<const bf8*>(p);
    }
#pragma unroll
    for (int I = 0; I < NBD; ++I)
#pragma unroll
      for (int J = I; J < NBD; ++J)
        acc[tile_index(I, J, NBD)] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            op[I], op[J], s == 0 ? f4{0.f, 0.f, 0.f, 0.f} : acc[tile_index(I, J, NBD)], 0, 0, 0);
  }
  // r into column 16 NBD - 1 of the last block column (K is zero there:
  // rows of Y past n are zero and 16 NBD - 1 >= n), mu on the diagonal.
  const int g = lane >> 4, c = lane & 15;
#pragma unroll
  for (int I = 0; I < NBD; ++I)
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (c == 15) acc[tile_index(I, NBD - 1, NBD)][q] = rs[16 * I + 4 * g + q];
  add_ridge<NBD>(acc, n, a.lambda * (float)n, lane);
  lds_order();  // the staged rows are dead: chol_solve reuses their space
  return chol_solve<NBD>(acc, n, 16 * NBD - 1, lds, zs, 16 * NBD, lane);
}

template <int NB, int CH>
__device__ __forceinline__ void dual_solve_row(const SweepArgs& a, const Task& t, float* lds, int lane) {
  constexpr int TB = KBB * 16, CB = NB * TB;
  float* xrow = a.x + (int64_t)t.row * a.ldf;
  __bf16* xb = reinterpret_cast<__bf16*>(a.x_b) + (int64_t)t.row * a.ldb;
  const int n = t.count;
  if (n <= 0 || n > KBB * CH - 1) {  // (the host launches only rows that fit)
    for (int i = lane; i < a.ldf; i += 64) xrow[i] = 0.0f;
    for (int i = lane; i < a.ldb; i += 64) xb[i] = (__bf16)0.0f;
    if (n > 0 && lane == 0) atomicAdd(a.fail_counter, 1);
    return;
  }
  const __bf16* theta_b = reinterpret_cast<const __bf16*>(a.theta_b);
  __bf16* yb = reinterpret_cast<__bf16*>(lds);
  float* zs = lds + dual_zs_offset<NB, CH>();
  float* rs = zs + 96;
  const __bf16* zero = reinterpret_cast<const __bf16*>(rs + 96);
  // Gather the CH chunks: lane l moves row l/2 of a chunk, 16 bytes of each
  // tile (rows past n and slots past ldb come from kZeroRow).  Done twice:
  // for the Gram, and again after the Cholesky (which reuses the space) for
  // x = Y^T z.
  const int sr = lane >> 1, sh = lane & 1;
  const bool last_in = 16 * (NB - 1) + 8 * sh < (int)a.ldb;
  const __bf16* zb = reinterpret_cast<const __bf16*>(kZeroRow);
  const __bf16* rp[CH];
#pragma unroll
  for (int ch = 0; ch < CH; ++ch) {
    const int k = ch * KBB + sr;
    rp[ch] = k < n ? theta_b + (int64_t)a.col[t.start + k] * a.ldb + 8 * sh : zb;
  }
  auto gather = [&] {
#pragma unroll
    for (int ch = 0; ch < CH; ++ch) {
      __bf16* dst = yb + ch * CB;
#pragma unroll
      for (int I = 0; I < NB - 1; ++I)
        __builtin_amdgcn_global_load_lds(rp[ch] + 16 * I, (lds_void_t*)(dst + I * TB), 16, 0, 0);
      __builtin_amdgcn_global_load_lds(last_in && rp[ch] != zb ? rp[ch] + 16 * (NB - 1) : zb,
                                       (lds_void_t*)(dst + (NB - 1) * TB), 16, 0, 0);
    }
  };
  gather();
  for (int i = lane; i < 96; i += 64) rs[i] = i < n ? a.val[t.start + i] : 0.0f;
  if (lane < 4) rs[96 + lane] = 0.0f;  // the zero block
  wait_vm_all();
  lds_order();
  bool ok;
  const int nbd = (n + 16) >> 4;  // blocks of the n x n system plus r's column
  if (nbd == 1) ok = dual_blocks<NB, CH, 1>(a, n, lds, zs, rs, zero, lane);
  else if constexpr (CH == 1) ok = dual_blocks<NB, CH, 2>(a, n, lds, zs, rs, zero, lane);
  else if (nbd == 2) ok = dual_blocks<NB, CH, 2>(a, n, lds, zs, rs, zero, lane);
  else if (nbd == 3) ok = dual_blocks<NB, CH, 3>(a, n, lds, zs, rs, zero, lane);
  else if constexpr (CH == 2) ok = dual_blocks<NB, CH, 4>(a, n, lds, zs, rs, zero, lane);
  else if (nbd == 4) ok = dual_blocks<NB, CH, 4>(a, n, lds, zs, rs, zero, lane);
  else if (nbd == 5) ok = dual_blocks<NB, CH, 5>(a, n, lds, zs, rs, zero, lane);
  else ok = dual_blocks<NB, CH, 6>(a, n, lds, zs, rs, zero, lane);
  lds_order();
  gather();
  wait_vm_all();
  lds_order();
  // x = Y^T z from the staged rows (chol_solve left z = 0 if it failed).
  // Lane l forms columns c0 = 2l + 128 p and c0 + 1 (tile c0 / 16) of the
  // staged width; the BF16 row's padding past it is zero.
  for (int i = 16 * NB + lane; i < (int)a.ldb; i += 64) xb[i] = (__bf16)0.0f;
  for (int c0 = 2 * lane; c0 < 16 * NB; c0 += 128) {
    const __bf16* col = yb + (c0 >> 4) * TB + (c0 & 15);
    float x0 = 0.0f, x1 = 0.0f;
#pragma unroll 8
    for (int k = 0; k < n; ++k) {
      const uint32_t w = *reinterpret_cast<const uint32_t*>(col + (k >> 5) * CB + (k & 31) * 16);
      const float zk = zs[k];
      x0 = fmaf(zk, __uint_as_float(w << 16), x0);
      x1 = fmaf(zk, __uint_as_float(w & 0xffff0000u), x1);
    }
    x0 = c0 < a.f ? x0 : 0.0f;
    x1 = c0 + 1 < a.f ? x1 : 0.0f;
    if (c0 < (int)a.ldf) xrow[c0] = x0;
    if (c0 + 1 < (int)a.ldf) xrow[c0 + 1] = x1;
    if (c0 < (int)a.ldb) xb[c0] = (__bf16)x0;
    if (c0 + 1 < (int)a.ldb) xb[c0 + 1] = (__bf16)x1;
  }
  if (!ok && lane == 0) atomicAdd(a.fail_counter, 1);
}

// ---------------------------------------------------------------------------
// The dual form of whole rows without LDS staging: the MFMA operands come
// straight from the BF16 factor rows.  For 32-column step s and row block
// I, lane (kg, i) loads the 16 bytes of row 16I + i, columns 32s + 8kg ..
// +7 -- the operand the staged form reads from LDS -- two steps at a time
// (rows past n and columns past ldb read zeros).  x = Y^T z: lane l forms
// columns 2l + 128p and the next one, each rating's row read whole by the
// wavefront (4 bytes a lane).  LDS holds only the Cholesky's space, z, the
// ratings and their column indices; CH bounds the row length (32 CH - 1)
// and with it the blocks and the registers.  Whole rows only (the host
// launches it where no task is a split row's chunk).
// ---------------------------------------------------------------------------
template <int CH>
__host__ __device__ constexpr int dual_g_lds_floats() {
  return chol_lds_floats(dual_max_blocks(CH)) + 3 * 96;
}

template <int NB, int NBD>
__device__ __forceinline__ bool dual_blocks_g(const SweepArgs& a, int n, const int* cs, const float* rs,
                                              float* lds, float* zs, int lane) {
  constexpr int S = (NB + 1) / 2;  // 32-column steps over the 16 NB columns
  const int ldb = (int)a.ldb;
  const __bf16* theta_b = reinterpret_cast<const __bf16*>(a.theta_b);
  const __bf16* zb = reinterpret_cast<const __bf16*>(kZeroRow);  // (512 zero bf16 >= ldb)
  const int kg = lane >> 4, i = lane & 15;
  const __bf16* rp[NBD];
#pragma unroll
  for (int I = 0; I < NBD; ++I) {
    const int r = 16 * I + i;
    rp[I] = r < n ? theta_b + (int64_t)cs[r] * ldb : zb;
  }
  f4 acc[num_tiles(NBD)];
#pragma unroll
  for (int t = 0; t < num_tiles(NBD); ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s0 = 0; s0 < S; s0 += 2) {
    bf8 op[2][NBD];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int I = 0; I < NBD; ++I) {
        const int col = 32 * (s0 + h) + 8 * kg;
        const __bf16* p = (s0 + h < S && col < ldb) ? rp[I] + col : zb;
        op[h][I] = *reinterpret_cast<const bf8*>(p);
      }
#pragma unroll
    for (int h = 0; h < 2; ++h)
      if (s0 + h < S) {
#pragma unroll
        for (int I = 0; I < NBD; ++I)
#pragma unroll
          for (int J = I; J < NBD; ++J)
            acc[tile_index(I, J, NBD)] =
                __builtin_amdgcn_mfma_f32_16x16x32_bf16(op[h][I], op[h][J], acc[tile_index(I, J, NBD)], 0, 0, 0);
      }
  }
  // r into column 16 NBD - 1 of the last block column (K is zero there:
  // rows of Y past n are zero and 16 NBD - 1 >= n), mu on the diagonal.
  const int g = lane >> 4, c = lane & 15;
#pragma unroll
  for (int I = 0; I < NBD; ++I)
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (c == 15) acc[tile_index(I, NBD - 1, NBD)][q] = rs[16 * I + 4 * g + q];
  add_ridge<NBD>(acc, n, a.lambda * (float)n, lane);
  return chol_solve<NBD>(acc, n, 16 * NBD - 1, lds, zs, 16 * NBD, lane);
}

template <int NB, int CH>
__device__ __forceinline__ void dual_solve_row_g(const SweepArgs& a, const Task& t, float* lds, int lane) {
  constexpr int MB = dual_max_blocks(CH);
  float* xrow = a.x + (int64_t)t.row * a.ldf;
  __bf16* xb = reinterpret_cast<__bf16*>(a.x_b) + (int64_t)t.row * a.ldb;
  const int n = t.count;
  if (n <= 0 || n > KBB * CH - 1 || t.slot >= 0) {  // (the host launches only whole rows that fit)
    for (int i = lane; i < a.ldf; i += 64) xrow[i] = 0.0f;
    for (int i = lane; i < a.ldb; i += 64) xb[i] = (__bf16)0.0f;
    if (n > 0 && lane == 0) atomicAdd(a.fail_counter, 1);
    return;
  }
  float* zs = lds + chol_lds_floats(MB);
  float* rs = zs + 96;
  int* cs = reinterpret_cast<int*>(rs + 96);
  for (int i = lane; i < 96; i += 64) {
    const bool in = i < n;
    rs[i] = in ? a.val[t.start + i] : 0.0f;
    cs[i] = in ? a.col[t.start + i] : 0;
  }
  lds_order();
  bool ok;
  const int nbd = (n + 16) >> 4;  // blocks of the n x n system plus r's column
  if (nbd == 1) ok = dual_blocks_g<NB, 1>(a, n, cs, rs, lds, zs, lane);
  else if constexpr (MB == 2) ok = dual_blocks_g<NB, 2>(a, n, cs, rs, lds, zs, lane);
  else if (nbd == 2) ok = dual_blocks_g<NB, 2>(a, n, cs, rs, lds, zs, lane);
  else if (nbd == 3) ok = dual_blocks_g<NB, 3>(a, n, cs, rs, lds, zs, lane);
  else if constexpr (MB == 4) ok = dual_blocks_g<NB, 4>(a, n, cs, rs, lds, zs, lane);
  else if (nbd == 4) ok = dual_blocks_g<NB, 4>(a, n, cs, rs, lds, zs, lane);
  else if (nbd == 5) ok = dual_blocks_g<NB, 5>(a, n, cs, rs, lds, zs, lane);
  else ok = dual_blocks_g<NB, 6>(a, n, cs, rs, lds, zs, lane);
  lds_order();
  // x = Y^T z (chol_solve left z = 0 if it failed); the BF16 row's padding
  // columns of Y are zero, so are x's there.
  const __bf16* theta_b = reinterpret_cast<const __bf16*>(a.theta_b);
  const int ldb = (int)a.ldb;
  for (int c0 = 2 * lane; c0 < ldb; c0 += 128) {
    float x0 = 0.0f, x1 = 0.0f;
#pragma unroll 8
    for (int k = 0; k < n; ++k) {
      const uint32_t w = *reinterpret_cast<const uint32_t*>(theta_b + (int64_t)cs[k] * ldb + c0);
      const float zk = zs[k];
      x0 = fmaf(zk, __uint_as_float(w << 16), x0);
      x1 = fmaf(zk, __uint_as_float(w & 0xffff0000u), x1);
    }
    x0 = c0 < a.f ? x0 : 0.0f;
    x1 = c0 + 1 < a.f ? x1 : 0.0f;
    if (c0 < (int)a.ldf) xrow[c0] = x0;
    if (c0 + 1 < (int)a.ldf) xrow[c0 + 1] = x1;
    xb[c0] = (__bf16)x0;
    xb[c0 + 1] = (__bf16)x1;
  }
  for (int i = ldb + lane; i < (int)a.ldf; i += 64) xrow[i] = 0.0f;  // (ldf <= ldb: none)
  if (!ok && lane == 0) atomicAdd(a.fail_counter, 1);
}

}  // namespace cumf
