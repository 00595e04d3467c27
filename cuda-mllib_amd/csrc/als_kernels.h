// als_kernels.h -- internal interface between the host driver and the
// gfx950 kernels of the ALS half-sweep.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cumf {

// One unit of work for a wavefront: `count` ratings of row `row`, starting
// at rating index `start` of the CSR (absolute, so that no row-pointer load
// precedes the first gather; rating indices fit int32 like the CSR's own).  slot < 0: the task covers the
// whole row and solves it; slot >= 0: it writes a partial Gram to
// partial[slot].
struct Task {
  int32_t row;
  int32_t start;
  int32_t count;
  int32_t slot;
};

// A row whose Gram is the sum of `nchunks` partials starting at `slot`.
struct HeavyRow {
  int32_t row;
  int32_t slot;
  int32_t nchunks;
  int32_t pad;
};

struct SweepArgs {
  const int32_t* row_ptr;  // CSR of the side being solved (m+1)
  const int32_t* col;      // column indices into theta (nnz)
  const float* val;        // ratings (nnz)
  const float* theta;      // factors of the other side, row stride ldf
  float* x;                // factors being solved, row stride ldf
  int64_t ldf;             // factor row stride in floats (multiple of 4, >= f)
  const Task* tasks;
  int64_t ntasks;
  const HeavyRow* heavy;
  int64_t nheavy;
  float* partial;          // partial Grams, partial_floats_per_slot(f) each
  int* fail_counter;       // rows whose factorisation met a non-positive pivot
  int f;
  float lambda;
  // precision != 0: Gram from BF16 shadow factors (theta_b, row stride ldb,
  // zero padded); solutions are also written to x_b in BF16.
  int precision;
  int64_t ldb;
  const uint16_t* theta_b;
  uint16_t* x_b;
  // solver != 0: FP16 conjugate gradients inside FP32 refinement with an
  // FP64 residual, warm-started from x; Cholesky for rows that do not reach
  // ||b - A x|| <= cg_tol ||b|| within cg_rounds rounds of cg_iters steps.
  int solver;
  int cg_iters;
  int cg_rounds;
  float cg_tol;
  int* cg_fallback;        // rows that fell back to Cholesky
  // CG start: 0 the current factor, 1 zero, 2 x0 (row stride ldf), 3 the
  // better of the current factor and x0 by FP64 residual.
  int warm;
  const float* x0;
  float* cg_trace;         // optional, 4 floats per row: initial / final relative residual, rounds, fell back
  // Rows the CG pass hands to the Cholesky fallback kernel: entry = row, or
  // -1 - index into heavy for a split row; fb_count = list length.
  int* fb_list;
  int* fb_count;
  // persistent > 0: launch that many wavefronts, each taking tasks from
  // *task_counter (zero before the launch) until ntasks are taken.
  int persistent;
  int* task_counter;
  // occupancy > 0 (one-task-per-group launches): at most that many
  // wavefronts per CU, enforced through the LDS each group requests.
  int occupancy;
  // Dual-form launches: 32-rating chunks staged per row, 1 (rows of at
  // most 31 ratings), 2 (at most 63) or 3 (at most 95); dual_partial != 0
  // if a task of the launch is a split row's chunk (its partial Gram is
  // formed: two or three chunks only).
  int dual_chunks;
  int dual_partial;
};

// Longest row the dual form takes (three staged chunks of 32 ratings).
constexpr int kDualMaxRatingsHost = 95;

// LDS of one CU, and the range of occupancy caps a launch accepts.
constexpr int kLdsBytesPerCU = 160 * 1024;
constexpr int kMinOccupancy = 3, kMaxOccupancy = 16;
// LDS a group requests so that at most n groups fit one CU: more than a
// (n+1)-th of the CU's LDS, in 512-byte steps.
inline size_t occupancy_lds_bytes(int n) { return (size_t)(kLdsBytesPerCU / (n + 1) / 512 + 1) * 512; }

// The reference-reproducing path (compat.hip): one work-group per row.
struct CompatArgs {
  const int32_t* row_ptr;
  const int32_t* col;
  const float* val;
  const float* theta;      // row stride ldf
  float* x;                // row stride ldf
  int64_t ldf;
  const int32_t* rows;     // optional list of rows (NULL: rows 0..nrows-1)
  double* dump;            // optional: write [M | b] (f*(f+1) doubles per row) and stop
  int* fail_counter;       // rows with a zero pivot
  int f;
  float lambda;
};
hipError_t prepare_compat(int f);
hipError_t launch_compat(const CompatArgs& a, int64_t nrows, hipStream_t s);

int sweep_nb(int f);
size_t partial_floats_per_slot(int f);
int dump_dim(int f);
// Clears n <= 64 ints (CG fallback-list length, persistent task counters).
hipError_t launch_fallback_reset(int* count, int n, hipStream_t s);
// Half-sweep kernels (als_sweep.hip), one set per register budget:
// w2 = two wavefronts per SIMD, w1 = one.
// Launchers of the sweep kernels, per register budget (w1, w2) and solver
// family (chol: Cholesky task/finish kernels and the CG path's fallback;
// cg: CG task/finish kernels; chol also holds the dual-form task kernel
// for rows of at most 95 ratings, BF16 factors only).
#define CUMF_SWEEP_LAUNCHERS                                                \
  namespace chol {                                                          \
  hipError_t launch_task_kernel(const SweepArgs& a, hipStream_t s);         \
  hipError_t launch_finish_kernel(const SweepArgs& a, hipStream_t s);       \
  hipError_t launch_fallback_kernel(const SweepArgs& a, hipStream_t s);     \
  hipError_t launch_dual_kernel(const SweepArgs& a, hipStream_t s);         \
  }                                                                         \
  namespace cg {                                                            \
  hipError_t launch_task_kernel(const SweepArgs& a, hipStream_t s);         \
  hipError_t launch_finish_kernel(const SweepArgs& a, hipStream_t s);       \
  }
namespace w1 { CUMF_SWEEP_LAUNCHERS }
namespace w2 { CUMF_SWEEP_LAUNCHERS }
#undef CUMF_SWEEP_LAUNCHERS
// Half-sweep for 140 < f <= 256 (FP32 Gram + FP32 Cholesky, one 256-thread
// work-group per row: the whole-row tasks a.tasks[0 .. rows), or rows
// 0 .. rows-1 of row_ptr when a.tasks is null); large_lds_bytes(f) of LDS.
// prepare_large_kernel(f) raises the kernel's dynamic LDS limit to
// large_lds_bytes(f); call it outside any stream capture before launching.
hipError_t launch_large_kernel(const SweepArgs& a, int64_t rows, hipStream_t s);
hipError_t prepare_large_kernel(int f);
size_t large_lds_bytes(int f);
// The dual form for 140 < f <= 256: one wavefront per whole-row task
// a.tasks[0 .. ntasks), each of at most kDualMaxRatingsHost ratings (FP32).
hipError_t launch_large_dual_kernel(const SweepArgs& a, int64_t ntasks, hipStream_t s);
// The dual form on the work-group for 140 < f <= 256: whole-row tasks
// a.tasks[0 .. ntasks) of more than kDualMaxRatingsHost ratings whose
// systems have nbd = (n + 16) / 16 row blocks (7 .. kLargeDualWgMaxBlocks).
// prepare_large_kernel raises the LDS limits of these kernels too.
constexpr int kLargeDualWgMaxBlocks = 16;
hipError_t launch_large_dual_wg_kernel(const SweepArgs& a, int64_t ntasks, int nbd, hipStream_t s);
hipError_t launch_dump_kernel(const SweepArgs& a, const int32_t* rows, int64_t nrows, float* out,
                              hipStream_t s);
hipError_t launch_rmse(const int32_t* u, const int32_t* v, const float* r, int64_t cnt, const float* X,
                       const float* T, int64_t ldf, int f, double* partial, int nblocks, double* out,
                       hipStream_t s);
hipError_t launch_mfma_layout(float* out, hipStream_t s);
hipError_t launch_mfma_bf16_layout(float* out, hipStream_t s);
hipError_t launch_to_bf16(const float* src, int64_t ldf, void* dst, int64_t ldb, int64_t rows, int f,
                          hipStream_t s);
hipError_t launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t s);

}  // namespace cumf
