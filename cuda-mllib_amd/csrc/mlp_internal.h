// mlp_internal.h -- what the ALS driver uses of the warm-start MLP.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct cumf_mlp;

namespace cumf {

int mlp_f(const cumf_mlp* m);
int mlp_k0(const cumf_mlp* m);
int mlp_device(const cumf_mlp* m);
// Allocates the prediction workspace (before any stream capture).
int mlp_prepare_predict(cumf_mlp* m);
// Predictions for `rows` rows of BF16 features (stride K0) into x (row
// stride ldx, f columns), on stream s.
int mlp_predict_device(cumf_mlp* m, const __bf16* feats, int64_t rows, float* x, int64_t ldx, hipStream_t s);
// Appends room for `rows` harvested pairs; returns where to write their
// features (rows x K0 BF16) and targets (rows x f FP32).
int mlp_harvest_reserve(cumf_mlp* m, int64_t rows, __bf16** feats, float** targets);

}  // namespace cumf
