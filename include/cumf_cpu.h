/* cumf_cpu.h -- host (CPU) ALS half-sweep: the "cpu" mode of the
 * reference's runALS.sh (als/runALS.sh:24-28, which runs Spark's JVM ALS)
 * without Spark, so the standalone driver can run the same workload on
 * the CPU and on the GPU.  libcumf_cpu.so, host code only (OpenMP).
 *
 * Arithmetic: for every row u of the CSR,
 *     x_u = (Y_u^T Y_u + fl32(lambda n_u) I)^-1 Y_u^T r_u
 * (the ALS-WR ridge of als/src/cuda/als.cu:905-909), the normal equations
 * accumulated in FP64 and solved by a Cholesky factorisation, the result
 * stored as FP32; rows without ratings get x_u = 0.
 */
#ifndef CUMF_CPU_H
#define CUMF_CPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* theta: n x f row-major; x: m x f row-major (written).  threads <= 0: all
 * OpenMP threads.  Returns 0; -1 for bad arguments; -2 if some system was
 * not positive definite (those rows are set to 0). */
int cumf_cpu_half_sweep(int64_t m, int64_t n, int f, const int32_t* row_ptr, const int32_t* col,
                        const float* val, const float* theta, float lambda, float* x, int threads);

/* CSR (m x n) -> CSC by a counting sort: col_ptr[n+1], row[nnz], val[nnz];
 * rows ascending inside each column.  Returns 0 or -1. */
int cumf_cpu_transpose(int64_t m, int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col,
                       const float* val, int32_t* col_ptr, int32_t* row, float* cval);

/* Held-out RMSE sqrt(mean((r - x_u . theta_v)^2)) accumulated in FP64. */
double cumf_cpu_rmse(int64_t count, const int32_t* u, const int32_t* v, const float* r, int f,
                     const float* x, const float* theta, int threads);

#ifdef __cplusplus
}
#endif
#endif
