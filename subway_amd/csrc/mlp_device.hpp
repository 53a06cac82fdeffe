// mlp_device.hpp -- device side of the scheduler network (layout in
// mlp.hpp): the forward pass of one sample on one wave64, lane j = hidden
// unit j of both layers.  Included by device translation units only.
#ifndef SUBWAY_MLP_DEVICE_HPP
#define SUBWAY_MLP_DEVICE_HPP

#include <hip/hip_runtime.h>

#include "mlp.hpp"

namespace subway {

struct Fwd {
    float xn[MLP_IN];
    float z1, h1, z2, h2;
    float out[MLP_OUT];
};

// Forward of one sample on one wave; lane = hidden unit.  xs: raw features
// (standardised here with the parameter block's mean and scale).
__device__ __forceinline__ void forward_wave(const float *__restrict__ p, const float *xs, Fwd &f)
{
    const int j = (int)__lane_id();
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) f.xn[i] = (xs[i] - p[i]) * p[MLP_IN + i];
    float a = p[MLP_B1 + j];
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) a += f.xn[i] * p[MLP_W1 + i * MLP_H + j];
    f.z1 = a;
    f.h1 = a > 0.f ? a : 0.f;
    float b = p[MLP_B2 + j];
    for (int k = 0; k < MLP_H; k++) b += __shfl(f.h1, k, 64) * p[MLP_W2 + k * MLP_H + j];
    f.z2 = b;
    f.h2 = b > 0.f ? b : 0.f;
#pragma unroll
    for (int m = 0; m < MLP_OUT; m++) {
        float c = f.h2 * p[MLP_W3 + j * MLP_OUT + m];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        f.out[m] = c + p[MLP_B3 + m];
    }
}

}  // namespace subway

#endif  // SUBWAY_MLP_DEVICE_HPP
