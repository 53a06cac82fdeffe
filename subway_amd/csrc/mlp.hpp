// mlp.hpp -- the learned scheduler's network: 12 features -> 64 -> 64 -> 3
// heads, ReLU hidden layers, FP32 weights and accumulation, trained with the
// mean squared error of the heads a sample marks valid, and Adam.
//
// Parameter vector (MLP_PARAMS floats):
//   [0, 12)        feature mean         (inputs are standardised first)
//   [12, 24)       feature scale (1/std)
//   W1 12x64, b1 64, W2 64x64, b2 64, W3 64x3, b3 3   (row-major, in x out)
// Heads: 0 = log2 of the next iteration's active-edge fraction / 32, 1 =
// whole-range density threshold, 2 = log2 of the async inner-round cap / 16.
#ifndef SUBWAY_MLP_HPP
#define SUBWAY_MLP_HPP

#include <hip/hip_runtime.h>

#include "subway/globals.hpp"

namespace subway {

constexpr int MLP_IN = 12;
constexpr int MLP_H = 64;
constexpr int MLP_OUT = 3;
constexpr int MLP_NORM = 2 * MLP_IN;
constexpr int MLP_W1 = MLP_NORM;
constexpr int MLP_B1 = MLP_W1 + MLP_IN * MLP_H;
constexpr int MLP_W2 = MLP_B1 + MLP_H;
constexpr int MLP_B2 = MLP_W2 + MLP_H * MLP_H;
constexpr int MLP_W3 = MLP_B2 + MLP_H;
constexpr int MLP_B3 = MLP_W3 + MLP_H * MLP_OUT;
constexpr int MLP_PARAMS = MLP_B3 + MLP_OUT;   // 5211
constexpr int MLP_WEIGHTS = MLP_PARAMS - MLP_NORM;

// Forward for n samples (x: n x 12, out: n x 3), device pointers.
void launch_mlp_forward(const float *params, const float *x, ull n, float *out, hipStream_t st);
// Loss and gradient over n samples (gradient of the 5187 weights, the
// normalisation entries get 0); grad and loss are overwritten.
void launch_mlp_grad(const float *params, const float *x, const float *y, const float *mask, ull n, float *grad,
                     float *loss, hipStream_t st);
// One Adam step on the weights (normalisation entries untouched).
void launch_mlp_adam(float *params, const float *grad, float *m, float *v, int step, float lr, hipStream_t st);

}  // namespace subway

#endif  // SUBWAY_MLP_HPP
