// mlp.hip -- forward, backward and Adam kernels of the learned scheduler's
// 12 -> 64 -> 64 -> 3 network (layout in mlp.hpp).  One wave64 per block;
// lane j owns hidden unit j of both layers, so every weight row read is a
// coalesced 256-byte access and the hidden activations move between lanes by
// shuffles.  Gradients are accumulated per lane over the block's samples and
// added to the global gradient once per block.
#include "mlp.hpp"
#include "mlp_device.hpp"

#include <hip/hip_runtime.h>

#include "subway/gpu_error_check.hpp"

namespace subway {

namespace {

constexpr int SPB = 32;  // samples per block

__global__ __launch_bounds__(64) void mlp_forward_kernel(const float *p, const float *x, ull n, float *out)
{
    for (ull s = (ull)blockIdx.x * SPB; s < n && s < (ull)(blockIdx.x + 1) * SPB; s++) {
        Fwd f;
        forward_wave(p, x + s * MLP_IN, f);
        if (__lane_id() < MLP_OUT) out[s * MLP_OUT + __lane_id()] = f.out[__lane_id()];
    }
}

__global__ __launch_bounds__(64) void mlp_grad_kernel(const float *p, const float *x, const float *y,
                                                      const float *mask, ull n, float *grad, float *loss)
{
    const int j = (int)__lane_id();
    float gW1[MLP_IN], gW2[MLP_H], gW3[MLP_OUT];
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, lsum = 0.f;
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) gW1[i] = 0.f;
    for (int k = 0; k < MLP_H; k++) gW2[k] = 0.f;
#pragma unroll
    for (int m = 0; m < MLP_OUT; m++) gW3[m] = 0.f;
    const float invN = 1.0f / (float)n;
    for (ull s = (ull)blockIdx.x * SPB; s < n && s < (ull)(blockIdx.x + 1) * SPB; s++) {
        Fwd f;
        forward_wave(p, x + s * MLP_IN, f);
        float dout[MLP_OUT];
#pragma unroll
        for (int m = 0; m < MLP_OUT; m++) {
            const float mk = mask[s * MLP_OUT + m];
            const float e = f.out[m] - y[s * MLP_OUT + m];
            dout[m] = 2.f * mk * e * invN;
            lsum += mk * e * e * invN;
        }
        // layer 3
        float dh2 = 0.f;
#pragma unroll
        for (int m = 0; m < MLP_OUT; m++) {
            gW3[m] += f.h2 * dout[m];
            dh2 += p[MLP_W3 + j * MLP_OUT + m] * dout[m];
        }
        if (j < MLP_OUT) gb3 += dout[j];
        // layer 2
        const float dz2 = f.z2 > 0.f ? dh2 : 0.f;
        gb2 += dz2;
        for (int k = 0; k < MLP_H; k++) gW2[k] += __shfl(f.h1, k, 64) * dz2;
        // d h1 for lane k = sum_j W2[k][j] dz2_j
        float dh1 = 0.f;
        for (int jj = 0; jj < MLP_H; jj++) dh1 += p[MLP_W2 + j * MLP_H + jj] * __shfl(dz2, jj, 64);
        const float dz1 = f.z1 > 0.f ? dh1 : 0.f;
        gb1 += dz1;
#pragma unroll
        for (int i = 0; i < MLP_IN; i++) gW1[i] += f.xn[i] * dz1;
    }
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) atomicAdd(&grad[MLP_W1 + i * MLP_H + j], gW1[i]);
    atomicAdd(&grad[MLP_B1 + j], gb1);
    for (int k = 0; k < MLP_H; k++) atomicAdd(&grad[MLP_W2 + k * MLP_H + j], gW2[k]);
    atomicAdd(&grad[MLP_B2 + j], gb2);
#pragma unroll
    for (int m = 0; m < MLP_OUT; m++) atomicAdd(&grad[MLP_W3 + j * MLP_OUT + m], gW3[m]);
    if (j < MLP_OUT) atomicAdd(&grad[MLP_B3 + j], gb3);
    if (j == 0) atomicAdd(loss, lsum);
}

__global__ void mlp_zero_kernel(float *grad, float *loss)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < MLP_PARAMS; i += gridDim.x * blockDim.x) grad[i] = 0.f;
    if (blockIdx.x == 0 && threadIdx.x == 0) *loss = 0.f;
}

__global__ void mlp_adam_kernel(float *p, const float *g, float *m, float *v, float lr, float c1, float c2)
{
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int i = MLP_NORM + blockIdx.x * blockDim.x + threadIdx.x; i < MLP_PARAMS; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

}  // namespace

void launch_mlp_forward(const float *params, const float *x, ull n, float *out, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n + SPB - 1) / SPB)), dim3(64), 0, st, params, x, n, out);
    SUBWAY_CHECK(hipGetLastError());
}

void launch_mlp_grad(const float *params, const float *x, const float *y, const float *mask, ull n, float *grad,
                     float *loss, hipStream_t st)
{
    hipLaunchKernelGGL(mlp_zero_kernel, dim3(8), dim3(256), 0, st, grad, loss);
    SUBWAY_CHECK(hipGetLastError());
    if (!n) return;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3((unsigned)((n + SPB - 1) / SPB)), dim3(64), 0, st, params, x, y, mask, n,
                       grad, loss);
    SUBWAY_CHECK(hipGetLastError());
}

void launch_mlp_adam(float *params, const float *grad, float *m, float *v, int step, float lr, hipStream_t st)
{
    const float c1 = 1.f - powf(0.9f, (float)step), c2 = 1.f - powf(0.999f, (float)step);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3(8), dim3(256), 0, st, params, grad, m, v, lr, c1, c2);
    SUBWAY_CHECK(hipGetLastError());
}

}  // namespace subway
