/*
 * als.h -- C ABI of libGPUALS.so, the MI355X (gfx950) ALS update.
 *
 * Two layers:
 *   1. doALSWithCSR: the drop-in for the reference's entry point of the
 *      same name (IBMSparkGPU/CUDA-MLlib als/src/cuda/als.h:96-103,
 *      definition als/src/cuda/als.cu:1000-1115).  Same name, C linkage,
 *      parameter order, types and ownership; the stream is a hipStream_t
 *      (an opaque pointer, like cudaStream_t).
 *   2. cumf_*: a status-returning context API that keeps ratings and both
 *      factor matrices resident in device memory and runs whole ALS
 *      iterations, the held-out RMSE and per-kernel timing.  The reference
 *      has no counterpart (its sweep loop and RMSE live in Spark); these
 *      are what its Spark caller does around doALSWithCSR.
 *
 * No HIP or CUDA header is needed to include this file.
 */
#ifndef CUMF_ALS_H_
#define CUMF_ALS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same definition as <hip/hip_runtime_api.h>; repeating a typedef is legal. */
typedef struct ihipStream_t* hipStream_t;

/* ------------------------------------------------------------------------
 * 1. Drop-in entry point.
 *
 * One half-sweep: for every row u of the m x n CSR matrix R (0-based,
 * int32 row pointers and column indices, fp32 values, all in DEVICE memory
 * obtained from mallocBest), solve
 *     (Theta_u^T Theta_u + lambda * n_u * I) x_u = Theta_u^T r_u
 * where Theta_u are the rows of thetaTHost (HOST, n x f, row j at
 * thetaTHost[j*f]) named by row u's column indices; write x_u to
 * XTHost[u*f] (HOST, m x f).
 *
 * Matches als/src/cuda/als.cu:1000-1115:
 *   - returns after the result is in XTHost (synchronous);
 *   - CONSUMES the three CSR arrays: they are released with freeBest
 *     (als/src/cuda/als.cu:1058, 1101-1102), so they must come from
 *     mallocBest and must not be used by the caller afterwards;
 *   - X_BATCH (>= 1) bounded the reference's temporary memory; the fused
 *     kernels here need no per-row f x f buffer, so it is only validated.
 * Differences (documented improvements):
 *   - any 1 <= f <= 256 (the reference needs f % 10 == 0, als/README.md:63);
 *     f > 140 runs a work-group-per-row FP32 Gram + Cholesky kernel, and
 *     the FP32 dual form for lighter rows;
 *   - rows with no ratings get x_u = 0 (the reference divides by zero);
 *   - errors never exit(): the call returns early and cumf_last_error()
 *     describes the failure (the CSR arrays are still released).
 * ---------------------------------------------------------------------- */
void doALSWithCSR(hipStream_t cuda_stream, int* csrRowIndex, int* csrColIndex, float* csrVal,
                  float* thetaTHost, float* XTHost, const int m, const int n, const int f,
                  const long nnz, const float lambda, const int X_BATCH);

/* ------------------------------------------------------------------------
 * 2. Context API.
 * ---------------------------------------------------------------------- */
#define CUMF_OK 0
#define CUMF_ERR_INVALID 1      /* bad argument */
#define CUMF_ERR_HIP 2          /* HIP runtime error */
#define CUMF_ERR_NOMEM 3        /* allocation failed */
#define CUMF_ERR_NODEVICE 4     /* no usable GPU */
#define CUMF_ERR_STATE 5        /* call out of order (e.g. no ratings yet) */
#define CUMF_ERR_UNSUPPORTED 6  /* option combination not built */

#define CUMF_OPTIONS_VERSION 1

#define CUMF_PREC_FP32 0        /* FP32 Gram on MFMA, FP32 Cholesky */
#define CUMF_PREC_BF16_GRAM 1   /* Gram on BF16 MFMA (FP32 accumulation) from BF16
                                   copies of the factors, b from a BF16 hi/lo split
                                   of the ratings, FP32 Cholesky */

#define CUMF_SOLVER_CHOLESKY 0  /* FP32 Cholesky of every row's system */
#define CUMF_SOLVER_CG 1        /* FP16 conjugate gradients on the Jacobi-scaled
                                   system inside FP32 iterative refinement, every
                                   residual b - A x that can accept a row formed
                                   in FP64 (the start's, which only drives the
                                   first round, in FP32), started from the row's
                                   current factor; rows whose relative residual
                                   stays above cg_tol fall back to the Cholesky
                                   solve */
#define CUMF_SOLVER_DUAL 2      /* tile tables only, degree buckets 0, 1 and 2 (at
                                   most 95 ratings): the same x from the n x n
                                   system (Y Y^T + lambda n I) z = r, x = Y^T z,
                                   with Y the row's n gathered factor rows (K
                                   on BF16 MFMA from the BF16 copies, or on FP32
                                   MFMA from the factors on a CUMF_PREC_FP32
                                   handle; FP32 Cholesky of the smaller system).
                                   On an FP32 handle a launch holding a split
                                   row's chunk takes the Cholesky kernel */

#define CUMF_WARM_PREV 0        /* CG starts from the row's current factor */
#define CUMF_WARM_ZERO 1        /* CG starts from zero */
#define CUMF_WARM_MLP 2         /* X half-sweeps start from the attached network's
                                   prediction (cumf_mlp.h); Theta half-sweeps
                                   from the current factor */
#define CUMF_WARM_BEST 3        /* the better of PREV and MLP by FP64 residual */

typedef struct cumf_options {
  int32_t version;      /* CUMF_OPTIONS_VERSION */
  int32_t precision;    /* CUMF_PREC_* */
  int32_t split;        /* max ratings per wavefront task; longer rows are
                           split and their partial Grams reduced (0: 16384) */
  int32_t timing;       /* 1: time every kernel launch with HIP events */
  int32_t use_graph;    /* 1: each half-sweep (X, Theta) is replayed from a
                           hipGraph captured at its first run; per-kernel
                           times are then replaced by per-replay times */
  int32_t compat_hermitian_bug; /* 1: reproduce the reference's arithmetic and
                           its Gram write slip (als/src/cuda/als.cu:275, 718):
                           sequential FP32 Gram, element (10I+4, 10J) :=
                           G(10I, 10J), FP64 LU without pivoting.  Needs
                           f % 10 == 0; slow path, for parity with the
                           reference's factors.  The legacy doALSWithCSR
                           takes it from CUMF_COMPAT_HERMITIAN_BUG=1. */
  int32_t solver;       /* CUMF_SOLVER_* */
  int32_t cg_iters;     /* CG steps per refinement round (0: 6) */
  int32_t cg_rounds;    /* refinement rounds, at most (0: 6) */
  float cg_tol;         /* accept ||b - A x|| <= cg_tol ||b||, the residual
                           evaluated in FP64 on the Gram the GPU formed
                           (0: 1e-5) */
  int32_t warm_start;   /* CUMF_WARM_* (CG only) */
  int32_t warm_sweeps;  /* CUMF_WARM_MLP / CUMF_WARM_BEST: the network starts
                           only the first warm_sweeps X half-sweeps after
                           cumf_set_factors (the transient ones, where the
                           previous factor is a poor start), later ones start
                           from the current factor (0: every X half-sweep) */
  int32_t reserved[4];
} cumf_options;

typedef struct cumf_stats {
  double task_ms;          /* fused Gram + solve kernel (all launches) */
  int64_t task_launches;
  double finish_ms;        /* partial-Gram reduction + solve kernel */
  int64_t finish_launches;
  double rmse_ms;          /* held-out RMSE kernels */
  int64_t rmse_launches;
  double sweep_ms;         /* whole half-sweeps, first to last kernel */
  int64_t sweeps;
  int64_t failed_rows;     /* rows with a non-positive pivot (x_u set to 0) */
  int64_t gram_flops;      /* nnz f(f+1) (the f(f+1)/2 unique Gram entries, one
                              multiply-add each, per rating) summed over the
                              half-sweeps run */
  double iter_ms;          /* graph replays (use_graph), one per half-sweep */
  int64_t iters;           /* number of graph replays */
  int64_t cg_fallback_rows; /* CUMF_SOLVER_CG: rows solved again by Cholesky */
  double warm_ms;          /* features + network predictions (CUMF_WARM_MLP/BEST) */
  int64_t warm_launches;
  double fallback_ms;      /* Cholesky re-solve of the rows CG did not accept */
  int64_t fallback_launches;
} cumf_stats;

typedef struct cumf_handle cumf_handle;

int cumf_default_options(cumf_options* opt);
/* device < 0: pick with get_gpu().  f in [1, 256], lambda >= 0.  For f > 140
 * only CUMF_PREC_FP32 with CUMF_SOLVER_CHOLESKY is built (no compat mode, no
 * autotuner); other options give CUMF_ERR_UNSUPPORTED. */
int cumf_create(int device, int f, float lambda, const cumf_options* opt, cumf_handle** out);
int cumf_destroy(cumf_handle* h);
/* Host arrays; copied to the device.  The CSC (= CSR of R^T) is optional:
 * when csc_col_ptr is NULL it is built from the CSR on the device
 * (cumf_csr_to_csc_device). */
int cumf_set_ratings(cumf_handle* h, int64_t m, int64_t n, int64_t nnz, const int32_t* csr_row_ptr,
                     const int32_t* csr_col, const float* csr_val, const int32_t* csc_col_ptr,
                     const int32_t* csc_row, const float* csc_val);
/* Held-out ratings (COO, host arrays) used by cumf_heldout_rmse. */
int cumf_set_heldout(cumf_handle* h, int64_t count, const int32_t* u, const int32_t* v,
                     const float* r);
/* Host factors, row-major m x f and n x f; either may be NULL. */
int cumf_set_factors(cumf_handle* h, const float* X, const float* theta);
int cumf_get_factors(cumf_handle* h, float* X, float* theta);
/* Half-sweeps, asynchronous on the handle's stream. */
int cumf_update_x(cumf_handle* h);      /* X from Theta over the CSR */
int cumf_update_theta(cumf_handle* h);  /* Theta from X over the CSC */
/* iters x (update_x, update_theta). */
int cumf_iterate(cumf_handle* h, int iters);
/* sqrt(mean((r - x_u . theta_v)^2)) over the held-out set, FP64; blocks. */
int cumf_heldout_rmse(cumf_handle* h, double* rmse);
int cumf_synchronize(cumf_handle* h);
/* Synchronises, then reports accumulated statistics. */
int cumf_get_stats(cumf_handle* h, cumf_stats* stats);
int cumf_reset_stats(cumf_handle* h);
hipStream_t cumf_get_stream(cumf_handle* h);
/* Thread-local description of the last failure on this thread ("" if none). */
const char* cumf_last_error(void);

/* Device-resident half-sweep with caller-owned device buffers (nothing
 * consumed, nothing copied): x (m x f) from theta (n x f) over the CSR.
 * Asynchronous on `stream`; the workspace it needs is allocated and freed
 * inside the call, which therefore synchronises `stream` before returning. */
int cumf_half_sweep_device(hipStream_t stream, const int32_t* d_row_ptr, const int32_t* d_col,
                           const float* d_val, const float* d_theta, float* d_x, int64_t m,
                           int64_t n, int f, int64_t nnz, float lambda, const cumf_options* opt);
/* The same half-sweep for callers that cannot allocate or synchronise
 * (code under HIP graph capture, tight serving loops); no counterpart in
 * the reference.  FP32 Gram + FP32 Cholesky like cumf_half_sweep_device,
 * but: no device allocation, no host readback and no synchronisation (the
 * work is enqueued on `stream`); whole rows taken by a persistent grid from
 * a counter, so a very long row is solved by one wavefront (no split); the
 * CSR must be consistent (row_ptr[0] = 0, row_ptr[m] = nnz, columns < n),
 * which is not checked on the host.  Limits: f a multiple of 4 in [4, 140]
 * (factor rows are read in place, row stride f), m <= INT_MAX - 2^20.  Its
 * only scratch memory is a 256-byte device array per GPU (the row counter
 * and the failed-row count), held until the process ends and shared by
 * every call on that GPU: calls that overlap on one GPU race.  Rows with a
 * non-positive pivot get x_u = 0. */
int cumf_half_sweep_device_noalloc(hipStream_t stream, const int32_t* d_row_ptr, const int32_t* d_col,
                                   const float* d_val, const float* d_theta, float* d_x, int64_t m, int64_t n,
                                   int f, int64_t nnz, float lambda);

/* ------------------------------------------------------------------------
 * Rating-matrix ingestion.
 * ---------------------------------------------------------------------- */
/* The reference's raw binary CSR files (loadCSRSparseMatrixBin,
 * als/src/cuda/als.cu:914-933): int32 row_ptr[m+1], int32 col[nnz],
 * fp32 val[nnz], no header.  Every file must exist and have exactly that
 * size; returns CUMF_OK or CUMF_ERR_INVALID (see cumf_last_error). */
int cumf_load_csr_bin(const char* dataFile, const char* rowFile, const char* colFile, float* data,
                      int* row, int* col, int m, long nnz);
int cumf_save_csr_bin(const char* dataFile, const char* rowFile, const char* colFile,
                      const float* data, const int* row, const int* col, int m, long nnz);
/* Seeded held-out split of a CSR rating matrix (host arrays), for ingested
 * files that come without a test set; no counterpart in the reference (its
 * Spark caller splits before the GPU sees the data).  Each rating is held
 * out with probability `fraction`, decided by a hash of (seed, rating
 * index); every row and every column with ratings keeps at least one
 * training rating.  Writes the training CSR (train_row_ptr[m+1],
 * train_col/train_val) and the held-out COO (test_u/v/r), order preserved;
 * every output array except train_row_ptr must hold nnz entries.
 * *train_nnz + *test_nnz == nnz. */
int cumf_split_holdout(int64_t m, int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col,
                       const float* val, double fraction, uint64_t seed, int32_t* train_row_ptr,
                       int32_t* train_col, float* train_val, int64_t* train_nnz, int32_t* test_u,
                       int32_t* test_v, float* test_r, int64_t* test_nnz);
/* CSR (m x n, device arrays) -> CSC (col_ptr n+1, row nnz, val nnz, device
 * arrays) by a stable radix sort on the device: rows stay ascending inside
 * each column, so the result equals a host counting-sort transpose.  No
 * counterpart in the reference (its caller transposes on the host).
 * Synchronises `stream`. */
int cumf_csr_to_csc_device(hipStream_t stream, int64_t m, int64_t n, int64_t nnz,
                           const int32_t* d_row_ptr, const int32_t* d_col, const float* d_val,
                           int32_t* d_col_ptr, int32_t* d_row, float* d_val_out);

/* ------------------------------------------------------------------------
 * Batching form: degree buckets and the tile table (no counterpart in the
 * reference, which launches one block per row).
 *
 * Tasks (rows, or split chunks of long rows) fall into 8 degree buckets by
 * rating count ([0,32), [32,64), [64,96), [96,256), [256,1024),
 * [1024,4096), [4096,16384), [16384,...)).  A tile
 * table gives, per side (0 = X, 1 = Theta): the split size, the register
 * budget (1 or 2 wavefronts per SIMD) of the split-row finish and of the CG
 * fallback kernels, and per bucket the register budget, whether the bucket
 * runs as a persistent grid of N wavefronts taking tasks from a counter, its
 * solver (the handle's, Cholesky, the dual form in buckets 0-2, or -- on a
 * CG handle -- CG), and a cap on
 * the wavefronts per CU of a one-task-per-group launch (3..16; 0 = none),
 * enforced through the LDS each group requests.  Without
 * a table every half-sweep is one launch of all tasks,
 * with two wavefronts per SIMD up to f = 124 and one above; a side whose
 * non-empty buckets share one budget and no persistent grid is also one
 * launch.
 * ---------------------------------------------------------------------- */
int cumf_load_tile_table(cumf_handle* h, const char* path);
int cumf_save_tile_table(cumf_handle* h, const char* path);
/* Autotuner: on the loaded ratings, times each split size per side, then
 * every launch shape (register budget, persistent grid or occupancy cap)
 * and solver of every non-empty bucket (min over reps runs), and keeps the
 * fastest; the per-bucket shapes are kept only if a
 * whole half-sweep with them beats one launch of every task with the
 * default shape (with the handle's solver, or with one solver for the whole
 * side).  The factors are restored afterwards, also on failure.  Writes the
 * table to path unless NULL. */
int cumf_autotune(cumf_handle* h, int reps, const char* path);

/* ------------------------------------------------------------------------
 * Test and measurement hooks (no counterpart in the reference).
 * ---------------------------------------------------------------------- */
/* Augmented Gram of the listed rows of the X side (side 0, over the CSR)
 * or Theta side (side 1, over the CSC), ridge included, as dense D x D
 * row-major FP32 (D = cumf_debug_gram_dim(f)); A is the leading f x f
 * block, b sits in column round_up(f,4).  out: host, nrows * D * D. */
int cumf_debug_gram_dim(int f);
int cumf_debug_row_gram(cumf_handle* h, int side, const int32_t* rows, int64_t nrows, float* out);
/* The host task planner alone (no device): the wavefront tasks of a CSR
 * with the given split size, 4 int32 each (row, first rating, count, partial
 * slot or -1), longest first, and the split rows, 4 int32 each (row, first
 * slot, chunks, 0).  *ntasks / *nheavy receive the counts even when the
 * capacities are too small (then CUMF_ERR_INVALID). */
int cumf_debug_plan(const int32_t* row_ptr, int64_t m, int split, int f, int32_t* tasks, int64_t task_cap,
                    int64_t* ntasks, int32_t* heavy, int64_t heavy_cap, int64_t* nheavy);
/* Runs one v_mfma_f32_16x16x4_f32 with A[i][k] = 100 i + k, B[k][j] =
 * 1000 k + j and returns the 64 x 4 result registers (host out[256]). */
int cumf_debug_mfma_layout(int device, float* out);
/* Same for v_mfma_f32_16x16x32_bf16 with A[i][k] = i + 2 (k mod 4),
 * B[k][j] = (k mod 5) + 3 j, k < 32, operands fed as the BF16 Gram feeds them. */
int cumf_debug_mfma_bf16_layout(int device, float* out);
/* Times `iters` launches of a plain copy kernel over `bytes` bytes (for
 * calibrating HBM byte counters); ms_per_launch receives the mean. */
int cumf_debug_copy(int device, int64_t bytes, int iters, double* ms_per_launch);
/* compat_hermitian_bug path: the system each listed row's solver sees,
 * [M | b] as f*(f+1) doubles per row (M row-major, slip and ridge applied).
 * side 0 = X (CSR), 1 = Theta (CSC).  out: host. */
int cumf_debug_compat_system(cumf_handle* h, int side, const int32_t* rows, int64_t nrows,
                             double* out);
/* Runs one half-sweep of side 0 (X) or 1 (Theta) as cumf_update_x /
 * cumf_update_theta do, recording for each row solved by CUMF_SOLVER_CG:
 * out[4 u .. 4 u + 3] = relative residual ||b - A x|| / ||b|| (FP64, NaN
 * if CG did not start) of the start (after its rescaling, if any) and of
 * the result, refinement rounds run, 1 if the row fell back to Cholesky.
 * out: host, rows x 4. */
int cumf_debug_cg_trace(cumf_handle* h, int side, float* out);
/* 0: above f = 140, every row takes the work-group-per-row primal kernel;
 * 1 (the default): rows of at most 95 ratings take the dual form on one
 * wavefront, and rows whose dual system has at most two 16-column blocks
 * fewer than the f x f one take it on a work-group.  Process wide; returns
 * the previous setting. */
int cumf_debug_large_dual(int on);
/* 1: mallocBest skips device memory and takes its mapped-host fallback
 * (also set by the environment variable CUMF_FORCE_HOST_ALLOC=1). */
int cumf_debug_force_host_alloc(int on);
/* Host <-> device copies for callers without a HIP binding (blocking). */
int cumf_memcpy_h2d(void* dst, const void* src, size_t bytes);
int cumf_memcpy_d2h(void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
/* Same name, C++ linkage and parameters as the reference
 * (als/src/cuda/als.cu:914-916); status via cumf_last_error(). */
void loadCSRSparseMatrixBin(const char* dataFile, const char* rowFile, const char* colFile,
                            float* data, int* row, int* col, const int m, const long nnz);
#endif
#endif /* CUMF_ALS_H_ */
