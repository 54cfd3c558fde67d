// Cost model MLP shared definitions (mlp.hip) and its use by the cell path.
#pragma once

namespace rnng {

constexpr int MLP_I = 4, MLP_H = 64, MLP_O = 2;
constexpr int MLP_NPARAMS = MLP_H * MLP_I + MLP_H + MLP_H * MLP_H + MLP_H + MLP_O * MLP_H + MLP_O;   // 4610
constexpr float MLP_UNDER = 4.0f;   // loss weight of an estimate below its target

// Inference kernel on device arrays X[B][4] -> Y[B][2].
int mlp_infer_device(const float* P, const float* X, int B, float* Y);
// Device parameters of the cost model set by rnng_set_cost_model (nullptr if none).
const float* cost_model_device();

}  // namespace rnng
