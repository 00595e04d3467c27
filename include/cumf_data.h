/*
 * cumf_data.h -- seeded synthetic rating matrices (libcumf_data.so).
 *
 * No counterpart in the reference: IBMSparkGPU/CUDA-MLlib takes its ratings
 * from Spark (als/runALS.sh:33-37) and its only loader reads raw CSR files
 * (loadCSRSparseMatrixBin, als/src/cuda/als.cu:914-933).  This generator
 * feeds the benchmark, the tests and the standalone driver with matrices of
 * a stated shape.  Host-only C ABI; no GPU calls.
 */
#ifndef CUMF_DATA_H_
#define CUMF_DATA_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cumf_gen_params {
  int64_t m, n;          /* rows (users), columns (items) */
  int64_t nnz_total;     /* distinct (row, column) pairs drawn */
  int64_t n_holdout;     /* of which this many are held out */
  double row_s, col_s;   /* Zipf-Mandelbrot exponents */
  int64_t row_cap;       /* largest row degree */
  int64_t col_cap;       /* largest column weight, in expected ratings */
  int32_t rank;          /* latent rank */
  double latent_var;     /* per-entry variance of the latent factors */
  double noise_std;      /* rating noise standard deviation */
  double mu;             /* rating offset */
  int32_t round_clip;    /* 1: clip(round(x), 1, 5) */
  uint64_t seed;
} cumf_gen_params;

typedef struct cumf_gen_info_t {
  int64_t m, n, nnz, n_test;
  double q_row, q_col;   /* solved Zipf-Mandelbrot offsets */
  int64_t max_row_deg, min_row_deg, max_col_deg, min_col_deg;
  int64_t hist[5];       /* training ratings equal to 1..5 */
} cumf_gen_info_t;

typedef struct cumf_dataset cumf_dataset;

/* Named workloads: "zipf_10k_2k_500k", "netflix_480k_17770_99m",
 * "zipf_4m_500k_1b".  Returns 0 or -1 for an unknown name. */
int cumf_gen_preset(const char* name, cumf_gen_params* params);
int cumf_gen_create(const cumf_gen_params* params, cumf_dataset** out);
int cumf_gen_info(const cumf_dataset* ds, cumf_gen_info_t* info);
/* Copies any non-NULL output: CSR (m+1, nnz, nnz), CSC (n+1, nnz, nnz),
 * held-out COO (n_test each). */
int cumf_gen_copy(const cumf_dataset* ds, int32_t* csr_row_ptr, int32_t* csr_col, float* csr_val,
                  int32_t* csc_col_ptr, int32_t* csc_row, float* csc_val, int32_t* test_u,
                  int32_t* test_v, float* test_r);
void cumf_gen_free(cumf_dataset* ds);
/* rows x f factors, each row N(0,1)^f scaled to unit L2 norm. */
int cumf_gen_init_factors(int64_t rows, int f, uint64_t seed, float* out);

#ifdef __cplusplus
}
#endif
#endif
