// kpt_mlp.hip — training and test kernels of the learned pitch-angle closure.
//
// Training is plain FP32 on the GPU: one lane per sample for the forward pass and
// the backward pass through the activations (weights broadcast from LDS), one
// lane per parameter for the weight gradients (a fixed-order sum over the batch,
// so a gradient is reproducible bit for bit), and an Adam update per parameter.
// The samples are exact slab-QLT evaluations drawn from a Philox stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

#include "kpt_device_common.h"
#include "kpt_internal.h"
#include "kpt_mlp_device.h"

namespace kpt {
namespace {

constexpr int kBlk = 256;
constexpr int kZ = 3 * kMlpH;  // stored pre-activations per sample

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}
// Hard-swish z u, u = clamp(z / 6 + 1/2, 0, 1); derivative u + z / 6 = 2u - 1/2
// inside the ramp, 0 below it and 1 above it.
__device__ __forceinline__ float hswish_derivative(float z) {
    const float u = __builtin_fmaf(z, 1.0f / 6.0f, 0.5f);
    return u <= 0.0f ? 0.0f : (u >= 1.0f ? 1.0f : 2.0f * u - 0.5f);
}
// Activation and derivative of the training kernels.
template <int ACT>
__device__ __forceinline__ float act_f(float z) {
    if constexpr (ACT == kActHardSwish) return hard_swish(z);
    else return silu(z);
}
template <int ACT>
__device__ __forceinline__ float act_df(float z) {
    if constexpr (ACT == kActHardSwish) return hswish_derivative(z);
    else return dsilu(z);
}

// Training set: |mu| half uniform on [mu_lo, 1), half log-uniform; ln xi and
// gamma_k uniform over the hull. Counter (index lo, index hi, 'MLP', 0).
__global__ void __launch_bounds__(kBlk) gen_kernel(float* __restrict__ feats, float* __restrict__ targets,
                                                   double* __restrict__ raw_lnr, uint64_t n, uint32_t k0,
                                                   uint32_t k1, uint64_t first, kpt_mlp_hull h, MlpScale m) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    if (i >= n) return;
    const uint64_t c = first + i;
    uint32_t r[4];
    philox4x32_10(static_cast<uint32_t>(c), static_cast<uint32_t>(c >> 32), 0x004D4C50u, 0u, k0, k1, r);
    const double u1 = u01(r[1]);
    const double mu = (r[0] & 1u) ? h.mu_lo + (1.0 - h.mu_lo) * u1 : exp(log(h.mu_lo) * (1.0 - u1));
    const double lxi = h.log_xi_lo + (h.log_xi_hi - h.log_xi_lo) * u01(r[2]);
    const double gk = h.gamma_lo + (h.gamma_hi - h.gamma_lo) * u01(r[3]);
    double lnr, y1;
    mlp_exact_target(mu, lxi, gk, h.h0, lnr, y1);
    feats[3 * i + 0] = (static_cast<float>(log(mu)) - m.c0) * m.s0;
    feats[3 * i + 1] = (static_cast<float>(lxi) - m.c1) * m.s1;
    feats[3 * i + 2] = (static_cast<float>(gk) - m.c2) * m.s2;
    targets[2 * i + 0] = static_cast<float>(lnr);
    targets[2 * i + 1] = static_cast<float>(y1);
    if (raw_lnr) raw_lnr[i] = lnr;
}

__global__ void __launch_bounds__(kBlk) normalize_kernel(float* __restrict__ targets, uint64_t n, float mean,
                                                         float inv_scale) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    if (i < n) targets[2 * i] = (targets[2 * i] - mean) * inv_scale;
}

template <int ACT>
__global__ void __launch_bounds__(kBlk) fwd_kernel(const float* __restrict__ params, const float* __restrict__ x,
                                                   uint64_t n, float* __restrict__ z, float* __restrict__ yhat) {
    __shared__ float w[kMlpParams];
    for (int i = threadIdx.x; i < kMlpParams; i += kBlk) w[i] = params[i];
    __syncthreads();
    const uint64_t b = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    if (b >= n) return;
    float h[kMlpH], g[kMlpH];
    const float x0 = x[3 * b], x1 = x[3 * b + 1], x2 = x[3 * b + 2];
    float* zb = z ? z + b * kZ : nullptr;
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) {
        const float v = w[kOffB1 + j] + w[kOffW1 + 3 * j] * x0 + w[kOffW1 + 3 * j + 1] * x1 + w[kOffW1 + 3 * j + 2] * x2;
        if (zb) zb[j] = v;
        h[j] = act_f<ACT>(v);
    }
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) {
        float v = w[kOffB2 + j];
#pragma unroll
        for (int k = 0; k < kMlpH; ++k) v += w[kOffW2 + kMlpH * j + k] * h[k];
        if (zb) zb[kMlpH + j] = v;
        g[j] = act_f<ACT>(v);
    }
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) {
        float v = w[kOffB3 + j];
#pragma unroll
        for (int k = 0; k < kMlpH; ++k) v += w[kOffW3 + kMlpH * j + k] * g[k];
        if (zb) zb[2 * kMlpH + j] = v;
        h[j] = act_f<ACT>(v);
    }
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o) {
        float v = w[kOffB4 + o];
#pragma unroll
        for (int k = 0; k < kMlpH; ++k) v += w[kOffW4 + kMlpH * o + k] * h[k];
        yhat[kMlpOut * b + o] = v;
    }
}

// Loss L = 1/(2n) sum_b sum_o (yhat - y)^2; dL/dyhat = (yhat - y) / n.
template <int ACT>
__global__ void __launch_bounds__(kBlk) bwd_data_kernel(const float* __restrict__ params, const float* __restrict__ z,
                                                        const float* __restrict__ yhat, const float* __restrict__ y,
                                                        uint64_t n, float* __restrict__ dz, float* __restrict__ dy,
                                                        float* __restrict__ loss_partials) {
    __shared__ float w[kMlpParams];
    __shared__ float red[kBlk / 64];
    for (int i = threadIdx.x; i < kMlpParams; i += kBlk) w[i] = params[i];
    __syncthreads();
    const uint64_t b = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    float loss = 0.0f;
    if (b < n) {
        const float inv_n = 1.0f / static_cast<float>(n);
        const float e0 = yhat[2 * b] - y[2 * b], e1 = yhat[2 * b + 1] - y[2 * b + 1];
        loss = 0.5f * (e0 * e0 + e1 * e1) * inv_n;
        const float d0 = e0 * inv_n, d1 = e1 * inv_n;
        dy[2 * b] = d0;
        dy[2 * b + 1] = d1;
        const float* zb = z + b * kZ;
        float* db = dz + b * kZ;
        float d3[kMlpH], d2[kMlpH];
#pragma unroll
        for (int k = 0; k < kMlpH; ++k)
            d3[k] = (w[kOffW4 + k] * d0 + w[kOffW4 + kMlpH + k] * d1) * act_df<ACT>(zb[2 * kMlpH + k]);
#pragma unroll
        for (int k = 0; k < kMlpH; ++k) {
            db[2 * kMlpH + k] = d3[k];
            float v = 0.0f;
#pragma unroll
            for (int j = 0; j < kMlpH; ++j) v += w[kOffW3 + kMlpH * j + k] * d3[j];
            d2[k] = v * act_df<ACT>(zb[kMlpH + k]);
        }
#pragma unroll
        for (int k = 0; k < kMlpH; ++k) {
            db[kMlpH + k] = d2[k];
            float v = 0.0f;
#pragma unroll
            for (int j = 0; j < kMlpH; ++j) v += w[kOffW2 + kMlpH * j + k] * d2[j];
            db[k] = v * act_df<ACT>(zb[k]);
        }
    }
    // fixed-order block reduction of the loss
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) loss += __shfl_xor(loss, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = loss;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int i = 0; i < kBlk / 64; ++i) t += red[i];
        loss_partials[blockIdx.x] = t;
    }
}

// One lane per parameter and batch chunk: partial[chunk][q] = sum over the
// chunk's samples in sample order; grad = sum over chunks in chunk order.
constexpr uint64_t kGradChunk = 128;

template <int ACT>
__global__ void __launch_bounds__(kBlk) bwd_weight_kernel(const float* __restrict__ x, const float* __restrict__ z,
                                                          const float* __restrict__ dz, const float* __restrict__ dy,
                                                          uint64_t n_total, float* __restrict__ partial) {
    const int q = blockIdx.x * kBlk + threadIdx.x;
    if (q >= kMlpParams) return;
    const uint64_t b0 = blockIdx.y * kGradChunk;
    const uint64_t n = b0 + kGradChunk < n_total ? b0 + kGradChunk : n_total;
    x += 3 * b0;
    z += b0 * kZ;
    dz += b0 * kZ;
    dy += 2 * b0;
    const uint64_t m = n - b0;
    float acc = 0.0f;
    if (q < kOffB1) {  // W1[i][j]
        const int i = q / kMlpIn, j = q % kMlpIn;
        for (uint64_t b = 0; b < m; ++b) acc += dz[b * kZ + i] * x[3 * b + j];
    } else if (q < kOffW2) {
        const int i = q - kOffB1;
        for (uint64_t b = 0; b < m; ++b) acc += dz[b * kZ + i];
    } else if (q < kOffB2) {
        const int i = (q - kOffW2) / kMlpH, j = (q - kOffW2) % kMlpH;
        for (uint64_t b = 0; b < m; ++b) acc += dz[b * kZ + kMlpH + i] * act_f<ACT>(z[b * kZ + j]);
    } else if (q < kOffW3) {
        const int i = q - kOffB2;
        for (uint64_t b = 0; b < m; ++b) acc += dz[b * kZ + kMlpH + i];
    } else if (q < kOffB3) {
        const int i = (q - kOffW3) / kMlpH, j = (q - kOffW3) % kMlpH;
        for (uint64_t b = 0; b < m; ++b) acc += dz[b * kZ + 2 * kMlpH + i] * act_f<ACT>(z[b * kZ + kMlpH + j]);
    } else if (q < kOffW4) {
        const int i = q - kOffB3;
        for (uint64_t b = 0; b < m; ++b) acc += dz[b * kZ + 2 * kMlpH + i];
    } else if (q < kOffB4) {
        const int i = (q - kOffW4) / kMlpH, j = (q - kOffW4) % kMlpH;
        for (uint64_t b = 0; b < m; ++b) acc += dy[2 * b + i] * act_f<ACT>(z[b * kZ + 2 * kMlpH + j]);
    } else {
        const int i = q - kOffB4;
        for (uint64_t b = 0; b < m; ++b) acc += dy[2 * b + i];
    }
    partial[blockIdx.y * static_cast<uint64_t>(kMlpParams) + q] = acc;
}

__global__ void __launch_bounds__(kBlk) grad_reduce_kernel(const float* __restrict__ partial, uint32_t chunks,
                                                           float* __restrict__ grad) {
    const int q = blockIdx.x * kBlk + threadIdx.x;
    if (q >= kMlpParams) return;
    float acc = 0.0f;
    for (uint32_t c = 0; c < chunks; ++c) acc += partial[c * static_cast<uint64_t>(kMlpParams) + q];
    grad[q] = acc;
}

__global__ void __launch_bounds__(kBlk) adam_kernel(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                                    const float* __restrict__ g, float lr, float b1c, float b2c) {
    const int q = blockIdx.x * kBlk + threadIdx.x;
    if (q >= kMlpParams) return;
    const float gq = g[q];
    const float mq = 0.9f * m[q] + 0.1f * gq;
    const float vq = 0.999f * v[q] + 0.001f * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] -= lr * (mq / b1c) / (sqrtf(vq / b2c) + 1e-8f);
}

// BF16 MFMA evaluation of n samples, one per lane; lanes past n evaluate zeros.
template <int ACT>
__global__ void __launch_bounds__(kBlk) infer_kernel(const void* __restrict__ image, const float* __restrict__ x,
                                                     uint64_t n, float* __restrict__ out) {
    __shared__ uint4 lds[kMlpLdsBytes / 16];
    mlp_load_lds(image, lds);
    __syncthreads();
    const MlpLds im = mlp_lds_view(lds);
    const uint64_t b = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
    if (b < n) {
        f0 = x[3 * b];
        f1 = x[3 * b + 1];
        f2 = x[3 * b + 2];
    }
    const float* b4 = reinterpret_cast<const float*>(static_cast<const char*>(image) + kMlpLdsBytes);
    float o0, o1;
    mlp_eval_wave<ACT>(im, f0, f1, f2, b4[0], b4[1], o0, o1);
    if (b < n) {
        out[2 * b] = o0;
        out[2 * b + 1] = o1;
    }
}

inline unsigned blocks(uint64_t n) { return static_cast<unsigned>((n + kBlk - 1) / kBlk); }

uint16_t to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return static_cast<uint16_t>(u >> 16);  // inf / nan
    u += 0x7FFFu + ((u >> 16) & 1u);                                               // round to nearest even
    return static_cast<uint16_t>(u >> 16);
}

}  // namespace

hipError_t launch_mlp_generate(float* feats, float* targets, double* raw_lnr, uint64_t n, uint64_t seed,
                               uint64_t first, const kpt_mlp_hull& h, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(gen_kernel, dim3(blocks(n)), dim3(kBlk), 0, stream, feats, targets, raw_lnr, n,
                       static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32), first, h, mlp_scale(h));
    return hipGetLastError();
}

hipError_t launch_mlp_normalize(float* targets, uint64_t n, float mean, float inv_scale, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(normalize_kernel, dim3(blocks(n)), dim3(kBlk), 0, stream, targets, n, mean, inv_scale);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(const float* params, int act, const float* x, uint64_t n, float* z, float* yhat,
                              hipStream_t stream) {
    if (!mlp_activation_ok(act)) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    if (act == kActHardSwish)
        hipLaunchKernelGGL(fwd_kernel<kActHardSwish>, dim3(blocks(n)), dim3(kBlk), 0, stream, params, x, n, z, yhat);
    else
        hipLaunchKernelGGL(fwd_kernel<kActSilu>, dim3(blocks(n)), dim3(kBlk), 0, stream, params, x, n, z, yhat);
    return hipGetLastError();
}

size_t mlp_loss_partials(uint64_t n) { return blocks(n); }
size_t mlp_grad_partials(uint64_t n) { return ((n + kGradChunk - 1) / kGradChunk) * kMlpParams; }

hipError_t launch_mlp_backward(const float* params, int act, const float* x, const float* z, const float* yhat,
                               const float* y, uint64_t n, float* dz, float* dy, float* grad_partials, float* grad,
                               float* loss_partials, hipStream_t stream) {
    if (!mlp_activation_ok(act)) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    const uint32_t chunks = static_cast<uint32_t>((n + kGradChunk - 1) / kGradChunk);
    const dim3 data_grid(blocks(n)), weight_grid(blocks(kMlpParams), chunks);
    if (act == kActHardSwish) {
        hipLaunchKernelGGL(bwd_data_kernel<kActHardSwish>, data_grid, dim3(kBlk), 0, stream, params, z, yhat, y, n, dz,
                           dy, loss_partials);
        hipLaunchKernelGGL(bwd_weight_kernel<kActHardSwish>, weight_grid, dim3(kBlk), 0, stream, x, z, dz, dy, n,
                           grad_partials);
    } else {
        hipLaunchKernelGGL(bwd_data_kernel<kActSilu>, data_grid, dim3(kBlk), 0, stream, params, z, yhat, y, n, dz, dy,
                           loss_partials);
        hipLaunchKernelGGL(bwd_weight_kernel<kActSilu>, weight_grid, dim3(kBlk), 0, stream, x, z, dz, dy, n,
                           grad_partials);
    }
    hipLaunchKernelGGL(grad_reduce_kernel, dim3(blocks(kMlpParams)), dim3(kBlk), 0, stream, grad_partials, chunks,
                       grad);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float* params, float* m, float* v, const float* grad, float lr, float b1c, float b2c,
                           hipStream_t stream) {
    hipLaunchKernelGGL(adam_kernel, dim3(blocks(kMlpParams)), dim3(kBlk), 0, stream, params, m, v, grad, lr, b1c,
                       b2c);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const void* image, int act, const float* x, uint64_t n, float* out,
                            hipStream_t stream) {
    if (!mlp_activation_ok(act)) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    if (act == kActHardSwish)
        hipLaunchKernelGGL(infer_kernel<kActHardSwish>, dim3(blocks(n)), dim3(kBlk), 0, stream, image, x, n, out);
    else
        hipLaunchKernelGGL(infer_kernel<kActSilu>, dim3(blocks(n)), dim3(kBlk), 0, stream, image, x, n, out);
    return hipGetLastError();
}

// Image = kFrags BF16 fragments | kBiasTiles FP32 accumulator initialisers |
// b4[2] (FP32, read by the kernels from global memory).
//
// A hard-swish network is stored as trained. A SiLU network is stored with the
// SiLU's exp2 scale folded in: every
// hidden pre-activation is stored as z' = -log2(e) z, so that
// silu(z) = -ln(2) z' / (1 + 2^z'). Layer 1 (weights and bias column) and the
// biases of layers 2 and 3 are scaled by -log2(e); the hidden activations are
// kept as h' = z' / (1 + 2^z') = -log2(e) silu(z), so the weights of layers 2 and
// 3 are unchanged (the factors cancel) and those of layer 4 are scaled by
// -ln(2). Each activation then costs one multiply fewer. Scales are applied in
// FP32 before the BF16 rounding.
void mlp_build_image(const float* p, int act, void* image) {
    const bool fold = act == kActSilu;
    const float kPre = fold ? -1.4426950408889634f : 1.0f;  // -log2(e)
    const float kOut = fold ? -0.6931471805599453f : 1.0f;  // -ln(2)
    uint16_t* frag = static_cast<uint16_t*>(image);
    float* bias = reinterpret_cast<float*>(static_cast<char*>(image) + kMlpFragBytes);
    auto put = [&](int f, int lane, int i, float v) { frag[(static_cast<size_t>(f) * 64 + lane) * 8 + i] = to_bf16(v); };
    for (int lane = 0; lane < 64; ++lane) {
        const int half = lane / 32, row = lane % 32;
        for (int i = 0; i < 8; ++i) {
            const int k = 8 * half + i;
            for (int t = 0; t < 2; ++t) {
                const int m = 32 * t + row;
                float v = 0.0f;
                if (k < kMlpIn) v = p[kOffW1 + kMlpIn * m + k];
                else if (k == kMlpIn) v = p[kOffB1 + m];
                put(kFragA1 + t, lane, i, v * kPre);
                for (int s = 0; s < 4; ++s) {
                    const int feat = mlp_k_feature(s, k);
                    put(kFragA2 + 4 * t + s, lane, i, p[kOffW2 + kMlpH * m + feat]);
                    put(kFragA3 + 4 * t + s, lane, i, p[kOffW3 + kMlpH * m + feat]);
                }
            }
            for (int s = 0; s < 4; ++s) {
                const int feat = mlp_k_feature(s, k);
                put(kFragA4 + s, lane, i, row < kMlpOut ? p[kOffW4 + kMlpH * row + feat] * kOut : 0.0f);
            }
        }
        // accumulator j of a lane sits in the lane's 16-byte slot of quarter j / 4
        for (int t = 0; t < 2; ++t)
            for (int j = 0; j < 16; ++j) {
                const int m = mlp_acc_row(t, j, half);
                bias[mlp_bias_index(0 * 2 + t, lane, j)] = p[kOffB2 + m] * kPre;
                bias[mlp_bias_index(1 * 2 + t, lane, j)] = p[kOffB3 + m] * kPre;
            }
    }
    float* b4 = reinterpret_cast<float*>(static_cast<char*>(image) + kMlpLdsBytes);
    b4[0] = p[kOffB4];
    b4[1] = p[kOffB4 + 1];
}

}  // namespace kpt
