/* cumf_mlp.h -- learned warm start for the mixed-precision solver.
 *
 * No counterpart in the reference (als/src/cuda/als.cu solves every row
 * from scratch); this is the library's learned part.  A small MLP maps the
 * features of row u of an X half-sweep,
 *     [log n_u, diag(A_u) / n_u, b_u / n_u]          (2f + 1 values)
 * to a prediction of x_u, used as the starting point of the CG solve
 * (CUMF_SOLVER_CG with cumf_options.warm_start = CUMF_WARM_MLP or
 * CUMF_WARM_BEST).  Layers: (2f+1) -> 256 -> 256 -> f, GELU on the hidden
 * layers, BF16 weights and activations with FP32 accumulation on the
 * matrix cores, FP32 master weights trained by Adam on the GPU.  Training
 * data are (features, x_u*) pairs harvested from the plain (Cholesky)
 * solver's own X half-sweeps.
 *
 * Parameter layout (FP32, flat): W1 (256 x K0), c1 (256), W2 (256 x 256),
 * c2 (256), W3 (N3 x 256), c3 (N3) with K0 = round_up(2f+1, 32) and
 * N3 = round_up(f, 32); W rows are output units, padded rows/columns zero.
 */
#ifndef CUMF_MLP_H_
#define CUMF_MLP_H_

#include <stdint.h>

#include "als.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cumf_mlp cumf_mlp;

/* device < 0: current device.  f in [1, 140].  Weights drawn from a
 * splitmix64 stream seeded with `seed` (uniform in +-1/sqrt(fan_in)). */
int cumf_mlp_create(int device, int f, uint64_t seed, cumf_mlp** out);
int cumf_mlp_destroy(cumf_mlp* mlp);
/* dims[0..3] = K0, 256, 256, N3; *count = number of FP32 parameters. */
int cumf_mlp_dims(const cumf_mlp* mlp, int32_t dims[4], int64_t* count);
int cumf_mlp_get_params(cumf_mlp* mlp, float* params);
/* Also clears the Adam moments and step count. */
int cumf_mlp_set_params(cumf_mlp* mlp, const float* params);

/* Host-side entry points (tests, small data).  feats: rows x K0 FP32
 * (rounded to BF16 on upload); targets and out: rows x f. */
int cumf_mlp_predict_host(cumf_mlp* mlp, const float* feats, int64_t rows, float* out);
/* Loss (mean over rows x f of the squared error) and its gradient with
 * respect to every parameter (flat layout above); weights unchanged.
 * rows must be a positive multiple of 256. */
int cumf_mlp_gradients_host(cumf_mlp* mlp, const float* feats, const float* targets, int64_t rows,
                            float* grads, double* loss);
/* One Adam step (lr, beta1 0.9, beta2 0.999, eps 1e-8) on the given batch. */
int cumf_mlp_step_host(cumf_mlp* mlp, const float* feats, const float* targets, int64_t rows, float lr,
                       double* loss);

/* Harvest: runs `iters` iterations of the handle's ALS with the Cholesky
 * solver; before each X half-sweep records the features of every row, after
 * it the solution, appending (features, x_u*) to the MLP's training set. */
int cumf_mlp_harvest(cumf_handle* h, cumf_mlp* mlp, int iters);
int64_t cumf_mlp_harvested(const cumf_mlp* mlp);
int cumf_mlp_clear_harvest(cumf_mlp* mlp);
/* Trains on the harvested set: `epochs` passes in shuffled batches of
 * `batch` rows (multiple of 256), Adam at learning rate lr; *loss = mean
 * training loss of the last epoch. */
int cumf_mlp_train(cumf_mlp* mlp, int epochs, int batch, float lr, uint64_t seed, double* loss);

/* Attaches (or, with NULL, detaches) a network whose predictions start the
 * CG solves of the handle's X half-sweeps when warm_start is CUMF_WARM_MLP
 * or CUMF_WARM_BEST.  The handle does not own it. */
int cumf_set_mlp(cumf_handle* h, cumf_mlp* mlp);

/* Test hook: the FP32 values of the BF16 features of the listed rows of
 * the X side (side 0, from theta) or Theta side (side 1, from X):
 * out is nrows x K0. */
int cumf_debug_features(cumf_handle* h, int side, const int32_t* rows, int64_t nrows, float* out);

#ifdef __cplusplus
}
#endif

#endif /* CUMF_MLP_H_ */
