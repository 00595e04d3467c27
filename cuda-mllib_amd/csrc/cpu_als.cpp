// cpu_als.cpp -- host ALS half-sweep, RMSE and CSR transpose (libcumf_cpu.so).
//
// The "cpu" leg of the standalone driver (als/runALS.sh:24-28 runs Spark's
// JVM ALS there).  Per row: the gathered factor rows are packed into a
// contiguous block of up to kChunk rows, the upper triangle of Y^T Y and
// Y^T r are accumulated from the block in FP64, the ridge fl32(lambda n_u)
// is added (als/src/cuda/als.cu:905-909), and the system is solved by an
// in-place Cholesky factorisation A = L L^T with forward and back
// substitution.  Rows are independent and spread over OpenMP threads.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/cumf_cpu.h"

namespace {

constexpr int kChunk = 64;  // factor rows packed per accumulation step

// Solves A x = b for SPD A (f x f, row-major, upper triangle filled);
// returns false if a pivot is not positive.
bool cholesky_solve(int f, std::vector<double>& A, std::vector<double>& b) {
  // L overwrites the lower triangle (L[i][j] at A[i*f + j], j <= i), read
  // from the upper triangle by symmetry.
  for (int j = 0; j < f; ++j) {
    double d = A[(size_t)j * f + j];
    for (int k = 0; k < j; ++k) d -= A[(size_t)j * f + k] * A[(size_t)j * f + k];
    if (!(d > 0.0)) return false;
    const double ljj = std::sqrt(d);
    A[(size_t)j * f + j] = ljj;
    for (int i = j + 1; i < f; ++i) {
      double s = A[(size_t)j * f + i];  // a_ij from the upper triangle
      for (int k = 0; k < j; ++k) s -= A[(size_t)i * f + k] * A[(size_t)j * f + k];
      A[(size_t)i * f + j] = s / ljj;
    }
  }
  for (int i = 0; i < f; ++i) {  // L y = b
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= A[(size_t)i * f + k] * b[k];
    b[i] = s / A[(size_t)i * f + i];
  }
  for (int i = f - 1; i >= 0; --i) {  // L^T x = y
    double s = b[i];
    for (int k = i + 1; k < f; ++k) s -= A[(size_t)k * f + i] * b[k];
    b[i] = s / A[(size_t)i * f + i];
  }
  return true;
}

}  // namespace

extern "C" int cumf_cpu_half_sweep(int64_t m, int64_t n, int f, const int32_t* row_ptr, const int32_t* col,
                                   const float* val, const float* theta, float lambda, float* x, int threads) {
  if (m < 0 || n < 0 || f < 1 || !row_ptr || !theta || !x || (row_ptr[m] > 0 && (!col || !val)) || !(lambda >= 0.f))
    return -1;
  for (int64_t u = 0; u < m; ++u)
    if (row_ptr[u + 1] < row_ptr[u]) return -1;
  for (int64_t k = 0; k < row_ptr[m]; ++k)
    if (col[k] < 0 || col[k] >= n) return -1;
  int bad = 0;
  const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt) reduction(+ : bad)
  {
    std::vector<double> A((size_t)f * f), b(f), Y((size_t)kChunk * f), r(kChunk);
#pragma omp for schedule(dynamic, 64)
    for (int64_t u = 0; u < m; ++u) {
      const int64_t s = row_ptr[u], e = row_ptr[u + 1];
      float* xu = x + u * f;
      if (s == e) {
        std::fill(xu, xu + f, 0.0f);
        continue;
      }
      std::fill(A.begin(), A.end(), 0.0);
      std::fill(b.begin(), b.end(), 0.0);
      for (int64_t c0 = s; c0 < e; c0 += kChunk) {
        const int cnt = (int)std::min<int64_t>(kChunk, e - c0);
        for (int t = 0; t < cnt; ++t) {
          const float* th = theta + (int64_t)col[c0 + t] * f;
          for (int j = 0; j < f; ++j) Y[(size_t)t * f + j] = th[j];
          r[t] = val[c0 + t];
        }
        for (int i = 0; i < f; ++i) {
          double* Ai = A.data() + (size_t)i * f;
          for (int t = 0; t < cnt; ++t) {
            const double* yt = Y.data() + (size_t)t * f;
            const double yi = yt[i];
            for (int j = i; j < f; ++j) Ai[j] += yi * yt[j];
            b[i] += yi * r[t];
          }
        }
      }
      const double ridge = (double)(lambda * (float)(e - s));
      for (int i = 0; i < f; ++i) A[(size_t)i * f + i] += ridge;
      if (cholesky_solve(f, A, b)) {
        for (int j = 0; j < f; ++j) xu[j] = (float)b[j];
      } else {
        std::fill(xu, xu + f, 0.0f);
        bad += 1;
      }
    }
  }
  return bad ? -2 : 0;
}

extern "C" int cumf_cpu_transpose(int64_t m, int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col,
                                  const float* val, int32_t* col_ptr, int32_t* row, float* cval) {
  if (m < 0 || n < 0 || nnz < 0 || !row_ptr || !col_ptr || row_ptr[m] != nnz ||
      (nnz > 0 && (!col || !val || !row || !cval)))
    return -1;
  std::vector<int64_t> next((size_t)n + 1, 0);
  for (int64_t k = 0; k < nnz; ++k) {
    if (col[k] < 0 || col[k] >= n) return -1;
    ++next[(size_t)col[k] + 1];
  }
  for (int64_t c = 0; c < n; ++c) next[c + 1] += next[c];
  for (int64_t c = 0; c <= n; ++c) col_ptr[c] = (int32_t)next[c];
  for (int64_t u = 0; u < m; ++u)
    for (int64_t k = row_ptr[u]; k < row_ptr[u + 1]; ++k) {
      const int64_t p = next[col[k]]++;
      row[p] = (int32_t)u;
      cval[p] = val[k];
    }
  return 0;
}

extern "C" double cumf_cpu_rmse(int64_t count, const int32_t* u, const int32_t* v, const float* r, int f,
                                const float* x, const float* theta, int threads) {
  if (count <= 0 || !u || !v || !r || !x || !theta || f < 1) return 0.0;
  double se = 0.0;
  const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for num_threads(nt) reduction(+ : se) schedule(static)
  for (int64_t k = 0; k < count; ++k) {
    const float* xu = x + (int64_t)u[k] * f;
    const float* tv = theta + (int64_t)v[k] * f;
    double p = 0.0;
    for (int j = 0; j < f; ++j) p += (double)xu[j] * (double)tv[j];
    const double d = (double)r[k] - p;
    se += d * d;
  }
  return std::sqrt(se / (double)count);
}
