/*
 * als_oracle.c -- FP64 CPU restatement of the cuMF ALS normal-equation update.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: the parity tests, the
 * smoke check and the benchmark's CPU-baseline leg call it; the product
 * library (libGPUALS.so) never links, loads or calls it.
 *
 * What it restates (IBMSparkGPU/CUDA-MLlib, paths relative to the repo root):
 *   - A_u = sum_{v in Omega_u} theta_v theta_v^T          als/src/cuda/als.cu:47-51
 *   - n_u = csrRowIndex[u+1] - csrRowIndex[u]              als/src/cuda/als.cu:58-64
 *   - A_u[i][i] += lambda * n_u  (ALS-WR weighting; the
 *     product lambda*n_u is formed in FP32 there)           als/src/cuda/als.cu:462-466, 905-909
 *   - b_u = sum_{v in Omega_u} r_uv theta_v                 als/src/cuda/als.cu:1039-1054
 *     (cusparseScsrmm2 of R with Theta, then a transpose)
 *   - x_u = A_u^{-1} b_u, FP64 solve, stored as FP32         als/src/cuda/als.cu:935-997
 *   - one call = one half-sweep over every row of a CSR     als/src/cuda/als.cu:1000-1004
 *   - factor rows are contiguous: theta_j = thetaT[j*f..]   als/src/cuda/als.cu:100-101, 532
 *   - compat mode: the Gram write slip in both Gram kernels
 *     (the element (10I+4, 10J) of every upper 10x10 tile
 *     receives G(10I, 10J)); the reference then solves the
 *     non-symmetric result by LU without pivoting           als/src/cuda/als.cu:275, 718, 952-955
 *
 * The reference has no CPU ALS, no sweep loop and no RMSE (those live in
 * Spark).  The sweep order, held-out RMSE and objective below are this
 * oracle's definitions: one iteration = X half-sweep from Theta over
 * CSR(R), then Theta half-sweep from X over CSR(R^T).
 *
 * Everything is computed in FP64.  Rows with n_u = 0 give x_u = 0 (the
 * reference divides by zero there, als/src/cuda/als.cu:905-909 + :954).
 *
 * Parity pins (each function is checked by tests/test_oracle.py):
 *   oracle_row_system   brute force numpy Y^T Y, closed-form trace, f=1 and
 *                       unit-vector special cases, symmetry/SPD invariants
 *   oracle_chol_solve   scipy.linalg.cho_solve, FP64 residual
 *   oracle_lu_solve     scipy.linalg.lu_solve on a diagonally dominant
 *                       non-symmetric matrix (no pivoting needed)
 *   oracle_half_sweep   gradient of the per-row objective vanishes; numpy
 *                       lstsq on the stacked regularised system
 *   compat mode         exactly 55 changed entries at f=100 vs clean mode,
 *                       at the positions the reference writes
 *   oracle_compat_system  FP32 fused multiply-adds in rating order, checked
 *                       against a float64 sum (FP32-rounding bound) and the
 *                       clean FP64 system away from the 55 slipped entries
 *   oracle_rmse         closed form on hand-made factors
 *   oracle_objective    monotone decrease across exact half-sweeps
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_OK 0
#define ORACLE_EINVAL -1
#define ORACLE_ENOTSPD -2
#define ORACLE_EZEROPIV -3
#define ORACLE_ENOMEM -4

/* Ratings of one row are consumed in chunks of this many factor rows: the
 * chunk is transposed into a contiguous f x CHUNK buffer so that every Gram
 * entry is a unit-stride dot product. */
#define CHUNK 256

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Applies the reference's Gram write slip to a full row-major f x f matrix
 * holding the correct Gram G (before regularisation).  In the reference the
 * thread owning the upper 10x10 tile (tile_x=10I, tile_y=10J), I <= J,
 * stores temp0 (= G(10I,10J)) where temp40 (= G(10I+4,10J)) belongs; the
 * solver reads tt column-major, so the matrix element (10I+4, 10J) is the
 * one overwritten.  Mirror writes (I < J) are correct.
 * als/src/cuda/als.cu:275 (get_hermitianT10), :718 (get_hermitian100). */
static void apply_compat_slip(int f, double* A) {
  int nt = f / 10;
  for (int I = 0; I < nt; ++I)
    for (int J = I; J < nt; ++J)
      A[(size_t)(10 * I + 4) * f + 10 * J] = A[(size_t)(10 * I) * f + 10 * J];
}

/* Forms the regularised normal equations of row u.
 * A: f*f row-major (full, symmetric unless compat), b: f.
 * theta: factor rows of the other side, row stride ldt (>= f), FP64.
 * work: CHUNK*f doubles of scratch. */
static void row_system_d(const int32_t* row_ptr, const int32_t* col, const float* val,
                         const double* theta, int64_t ldt, int f, float lambda, int64_t u,
                         int compat, double* A, double* b, double* work) {
  int64_t s = row_ptr[u], e = row_ptr[u + 1];
  memset(A, 0, sizeof(double) * (size_t)f * f);
  memset(b, 0, sizeof(double) * (size_t)f);
  for (int64_t c0 = s; c0 < e; c0 += CHUNK) {
    int64_t cn = e - c0 < CHUNK ? e - c0 : CHUNK;
    /* work[i*CHUNK + k] = theta_{col[c0+k]}[i] */
    for (int64_t k = 0; k < cn; ++k) {
      const double* t = theta + (int64_t)col[c0 + k] * ldt;
      double r = (double)val[c0 + k];
      for (int i = 0; i < f; ++i) {
        work[(size_t)i * CHUNK + k] = t[i];
        b[i] += r * t[i];
      }
    }
    for (int i = 0; i < f; ++i) {
      const double* wi = work + (size_t)i * CHUNK;
      for (int j = i; j < f; ++j) {
        const double* wj = work + (size_t)j * CHUNK;
        double acc = 0.0;
        for (int64_t k = 0; k < cn; ++k) acc += wi[k] * wj[k];
        A[(size_t)i * f + j] += acc;
      }
    }
  }
  for (int i = 0; i < f; ++i)
    for (int j = i + 1; j < f; ++j) A[(size_t)j * f + i] = A[(size_t)i * f + j];
  if (compat) apply_compat_slip(f, A);
  /* lambda * n_u is an FP32 product in the reference (int * float). */
  double reg = (double)(lambda * (float)(e - s));
  for (int i = 0; i < f; ++i) A[(size_t)i * f + i] += reg;
}

/* Cholesky A = L L^T in place (lower triangle), then solves L L^T x = b;
 * b is overwritten with x.  Golub & Van Loan, Alg. 4.2.1 (gaxpy form). */
int oracle_chol_solve(int f, double* A, double* b) {
  for (int j = 0; j < f; ++j) {
    double d = A[(size_t)j * f + j];
    for (int k = 0; k < j; ++k) d -= A[(size_t)j * f + k] * A[(size_t)j * f + k];
    if (!(d > 0.0)) return ORACLE_ENOTSPD;
    d = sqrt(d);
    A[(size_t)j * f + j] = d;
    for (int i = j + 1; i < f; ++i) {
      double s = A[(size_t)i * f + j];
      for (int k = 0; k < j; ++k) s -= A[(size_t)i * f + k] * A[(size_t)j * f + k];
      A[(size_t)i * f + j] = s / d;
    }
  }
  for (int i = 0; i < f; ++i) { /* L z = b */
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= A[(size_t)i * f + k] * b[k];
    b[i] = s / A[(size_t)i * f + i];
  }
  for (int i = f - 1; i >= 0; --i) { /* L^T x = z */
    double s = b[i];
    for (int k = i + 1; k < f; ++k) s -= A[(size_t)k * f + i] * b[k];
    b[i] = s / A[(size_t)i * f + i];
  }
  return ORACLE_OK;
}

/* LU without pivoting (Doolittle) in place, then L U x = b; b := x.
 * The reference calls cublasDgetrfBatched with a NULL pivot array and
 * cublasDgetrsBatched (als/src/cuda/als.cu:952-955, 974-976). */
int oracle_lu_solve(int f, double* A, double* b) {
  for (int k = 0; k < f; ++k) {
    double p = A[(size_t)k * f + k];
    if (p == 0.0) return ORACLE_EZEROPIV;
    for (int i = k + 1; i < f; ++i) {
      double l = A[(size_t)i * f + k] / p;
      A[(size_t)i * f + k] = l;
      for (int j = k + 1; j < f; ++j) A[(size_t)i * f + j] -= l * A[(size_t)k * f + j];
    }
  }
  for (int i = 0; i < f; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= A[(size_t)i * f + k] * b[k];
    b[i] = s;
  }
  for (int i = f - 1; i >= 0; --i) {
    double s = b[i];
    for (int k = i + 1; k < f; ++k) s -= A[(size_t)i * f + k] * b[k];
    b[i] = s / A[(size_t)i * f + i];
  }
  return ORACLE_OK;
}

/* Public: the regularised system of one row, FP32 factor input. */
int oracle_row_system(const int32_t* row_ptr, const int32_t* col, const float* val,
                      const float* theta, int64_t n, int f, float lambda, int64_t u,
                      int compat, double* A, double* b) {
  if (f <= 0 || (compat && f % 10 != 0)) return ORACLE_EINVAL;
  int64_t s = row_ptr[u], e = row_ptr[u + 1];
  double* th = (double*)malloc(sizeof(double) * (size_t)n * f);
  double* work = (double*)malloc(sizeof(double) * (size_t)CHUNK * f);
  if (!th || !work) { free(th); free(work); return ORACLE_ENOMEM; }
  for (int64_t k = s; k < e; ++k) /* only the rows this system touches */
    for (int i = 0; i < f; ++i) th[(int64_t)col[k] * f + i] = theta[(int64_t)col[k] * f + i];
  row_system_d(row_ptr, col, val, th, f, f, lambda, u, compat, A, b, work);
  free(th);
  free(work);
  return ORACLE_OK;
}

/* One half-sweep in FP64: for every row u of the CSR (m rows over n
 * columns), x_u = A_u^{-1} b_u with A_u, b_u from the factor rows theta
 * (n x f, FP64).  X: m x f FP64 output.  Rows with n_u = 0 get x_u = 0.
 * status (optional, m ints): 0 ok, else the solver's error code. */
int oracle_half_sweep_d(int64_t m, int64_t n, int f, const int32_t* row_ptr,
                        const int32_t* col, const float* val, const double* theta,
                        float lambda, int compat, double* X, int32_t* status, int nthreads) {
  (void)n;
  if (f <= 0 || (compat && f % 10 != 0)) return ORACLE_EINVAL;
  set_threads(nthreads);
  int err = ORACLE_OK;
#pragma omp parallel
  {
    double* A = (double*)malloc(sizeof(double) * (size_t)f * f);
    double* work = (double*)malloc(sizeof(double) * (size_t)CHUNK * f);
    if (!A || !work) {
#pragma omp atomic write
      err = ORACLE_ENOMEM;
    } else {
#pragma omp for schedule(dynamic, 16)
      for (int64_t u = 0; u < m; ++u) {
        double* x = X + u * f;
        int rc = ORACLE_OK;
        if (row_ptr[u + 1] == row_ptr[u]) {
          memset(x, 0, sizeof(double) * (size_t)f);
        } else {
          row_system_d(row_ptr, col, val, theta, f, f, lambda, u, compat, A, x, work);
          rc = compat ? oracle_lu_solve(f, A, x) : oracle_chol_solve(f, A, x);
        }
        if (status) status[u] = rc;
      }
    }
    free(A);
    free(work);
  }
  return err;
}

/* Same half-sweep with FP32 factors in and out (what one doALSWithCSR call
 * consumes and produces); internal arithmetic is FP64. */
int oracle_half_sweep_f(int64_t m, int64_t n, int f, const int32_t* row_ptr, const int32_t* col,
                        const float* val, const float* theta, float lambda, int compat,
                        float* X, int32_t* status, int nthreads) {
  double* th = (double*)malloc(sizeof(double) * (size_t)n * f);
  double* xd = (double*)malloc(sizeof(double) * (size_t)m * f);
  if (!th || !xd) { free(th); free(xd); return ORACLE_ENOMEM; }
  for (int64_t i = 0; i < n * f; ++i) th[i] = (double)theta[i];
  int rc = oracle_half_sweep_d(m, n, f, row_ptr, col, val, th, lambda, compat, xd, status, nthreads);
  for (int64_t i = 0; i < m * f; ++i) X[i] = (float)xd[i];
  free(th);
  free(xd);
  return rc;
}

/* Held-out RMSE: sqrt(mean((r - x_u . theta_v)^2)), FP64 throughout.
 * No clipping of predictions (explicit-feedback evaluation). */
double oracle_rmse_d(int64_t cnt, const int32_t* u, const int32_t* v, const float* r,
                     int f, const double* X, const double* T, int nthreads) {
  if (cnt <= 0) return 0.0;
  set_threads(nthreads);
  double se = 0.0;
#pragma omp parallel for reduction(+ : se) schedule(static)
  for (int64_t k = 0; k < cnt; ++k) {
    const double* x = X + (int64_t)u[k] * f;
    const double* t = T + (int64_t)v[k] * f;
    double p = 0.0;
    for (int i = 0; i < f; ++i) p += x[i] * t[i];
    double d = (double)r[k] - p;
    se += d * d;
  }
  return sqrt(se / (double)cnt);
}

double oracle_rmse_f(int64_t cnt, const int32_t* u, const int32_t* v, const float* r, int f,
                     const float* X, const float* T, int nthreads) {
  if (cnt <= 0) return 0.0;
  set_threads(nthreads);
  double se = 0.0;
#pragma omp parallel for reduction(+ : se) schedule(static)
  for (int64_t k = 0; k < cnt; ++k) {
    const float* x = X + (int64_t)u[k] * f;
    const float* t = T + (int64_t)v[k] * f;
    double p = 0.0;
    for (int i = 0; i < f; ++i) p += (double)x[i] * (double)t[i];
    double d = (double)r[k] - p;
    se += d * d;
  }
  return sqrt(se / (double)cnt);
}

/* ALS-WR objective over the training CSR:
 *   sum (r_uv - x_u.theta_v)^2 + lambda (sum_u n_u |x_u|^2 + sum_v n_v |theta_v|^2)
 * n_v is taken from the CSR column counts. */
double oracle_objective_d(int64_t m, int64_t n, int f, const int32_t* row_ptr, const int32_t* col,
                          const float* val, const double* X, const double* T, float lambda,
                          int nthreads) {
  set_threads(nthreads);
  double loss = 0.0, regx = 0.0;
#pragma omp parallel for reduction(+ : loss, regx) schedule(dynamic, 64)
  for (int64_t u = 0; u < m; ++u) {
    const double* x = X + u * f;
    for (int64_t k = row_ptr[u]; k < row_ptr[u + 1]; ++k) {
      const double* t = T + (int64_t)col[k] * f;
      double p = 0.0;
      for (int i = 0; i < f; ++i) p += x[i] * t[i];
      double d = (double)val[k] - p;
      loss += d * d;
    }
    double nx = 0.0;
    for (int i = 0; i < f; ++i) nx += x[i] * x[i];
    regx += (double)(row_ptr[u + 1] - row_ptr[u]) * nx;
  }
  int64_t* cnt = (int64_t*)calloc((size_t)n, sizeof(int64_t));
  if (!cnt) return NAN;
  for (int64_t k = 0; k < row_ptr[m]; ++k) cnt[col[k]]++;
  double regt = 0.0;
  for (int64_t v = 0; v < n; ++v) {
    double nt = 0.0;
    for (int i = 0; i < f; ++i) nt += T[v * f + i] * T[v * f + i];
    regt += (double)cnt[v] * nt;
  }
  free(cnt);
  return loss + (double)lambda * (regx + regt);
}

/* Full ALS in FP64.  One iteration = X half-sweep over CSR(R) from Theta,
 * then Theta half-sweep over CSR(R^T) from X.  Theta0 (n x f, FP32) seeds
 * the first X sweep.  X, T: FP64 outputs.  If ho_cnt > 0, rmse_hist[it]
 * receives the held-out RMSE after iteration it. */
int oracle_als_train(int64_t m, int64_t n, int f, const int32_t* rp, const int32_t* ci,
                     const float* rv, const int32_t* cp, const int32_t* ri, const float* cv,
                     const float* theta0, float lambda, int iters, int compat, int64_t ho_cnt,
                     const int32_t* ho_u, const int32_t* ho_v, const float* ho_r, double* X,
                     double* T, double* rmse_hist, int nthreads) {
  for (int64_t i = 0; i < n * f; ++i) T[i] = (double)theta0[i];
  for (int it = 0; it < iters; ++it) {
    int rc = oracle_half_sweep_d(m, n, f, rp, ci, rv, T, lambda, compat, X, NULL, nthreads);
    if (rc) return rc;
    rc = oracle_half_sweep_d(n, m, f, cp, ri, cv, X, lambda, compat, T, NULL, nthreads);
    if (rc) return rc;
    if (rmse_hist && ho_cnt > 0) rmse_hist[it] = oracle_rmse_d(ho_cnt, ho_u, ho_v, ho_r, f, X, T, nthreads);
  }
  return ORACLE_OK;
}

/* The system the reference's solver sees for row u, restated with the
 * reference's own arithmetic (als/src/cuda/als.cu:111-224 for
 * get_hermitianT10, :554-667 for get_hermitian100): each Gram entry is an
 * FP32 register updated by one fused multiply-add per rating, in rating
 * order; stored to FP64; the slip of :275/:718 applied; the ridge
 * (float)(n_u * lambda) added in FP64 (:462-466, :905-909).  b is formed
 * the same way (FP32 fused multiply-adds in rating order; the reference's
 * cusparseScsrmm2 order is not published).
 * M: f*f row-major doubles, b: f doubles.  Requires f % 10 == 0. */
int oracle_compat_system(const int32_t* row_ptr, const int32_t* col, const float* val,
                         const float* theta, int f, float lambda, int64_t u, double* M, double* b) {
  if (f <= 0 || f % 10 != 0) return ORACLE_EINVAL;
  float* G = (float*)calloc((size_t)f * f, sizeof(float));
  float* bf = (float*)calloc((size_t)f, sizeof(float));
  if (!G || !bf) { free(G); free(bf); return ORACLE_ENOMEM; }
  const int64_t s = row_ptr[u], e = row_ptr[u + 1];
  for (int64_t k = s; k < e; ++k) {
    const float* t = theta + (int64_t)col[k] * f;
    for (int i = 0; i < f; ++i) {
      for (int j = 0; j < f; ++j) G[(size_t)i * f + j] = fmaf(t[i], t[j], G[(size_t)i * f + j]);
      bf[i] = fmaf(val[k], t[i], bf[i]);
    }
  }
  for (size_t i = 0; i < (size_t)f * f; ++i) M[i] = (double)G[i];
  apply_compat_slip(f, M);
  const double reg = (double)((float)(e - s) * lambda);
  for (int i = 0; i < f; ++i) {
    M[(size_t)i * f + i] += reg;
    b[i] = (double)bf[i];
  }
  free(G);
  free(bf);
  return ORACLE_OK;
}
