// kpt_mlp_device.h — wavefront-level BF16 MFMA evaluation of the pitch-angle
// closure network (layout and reparametrisation in kpt_mlp.h).
#pragma once

#include <hip/hip_runtime.h>

#include "kpt_mlp.h"

namespace kpt {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// The network image as the kernels hold it in LDS: kFrags BF16 fragments and
// kBiasTiles accumulator initialisers, each 64 lanes wide.
struct MlpLds {
    const bf16x8* frag;
    const float4* bias;  // tile t, quarter q, lane l at bias[(4 t + q) 64 + l]
};

__device__ __forceinline__ MlpLds mlp_lds_view(const void* lds) {
    MlpLds v;
    v.frag = static_cast<const bf16x8*>(lds);
    v.bias = reinterpret_cast<const float4*>(static_cast<const char*>(lds) + kMlpFragBytes);
    return v;
}

// One bias tile of a lane as its accumulator initialiser.
__device__ __forceinline__ f32x16 mlp_bias_tile(const MlpLds& im, int tile, int lane) {
    f32x16 r;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 v = im.bias[(tile * 4 + q) * 64 + lane];
        r[4 * q + 0] = v.x;
        r[4 * q + 1] = v.y;
        r[4 * q + 2] = v.z;
        r[4 * q + 3] = v.w;
    }
    return r;
}

// Cooperative copy of the image (global -> LDS) by a whole block; the caller
// synchronises the block afterwards.
__device__ __forceinline__ void mlp_load_lds(const void* image, void* lds) {
    const uint4* src = static_cast<const uint4*>(image);
    uint4* dst = static_cast<uint4*>(lds);
    for (unsigned i = threadIdx.x; i < kMlpLdsBytes / 16; i += blockDim.x) dst[i] = src[i];
}

// SiLU on the image's scaled pre-activation z' = -log2(e) z (see
// mlp_build_image): returns z' / (1 + 2^z') = -log2(e) silu(z), with the native
// exp2 and reciprocal.
__device__ __forceinline__ float silu_scaled(float zs) {
    return zs * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(zs));
}

// Hard-swish z clamp(z / 6 + 1/2, 0, 1): one FMA with a [0, 1] clamp and a multiply.
__device__ __forceinline__ float hard_swish(float z) {
    return z * __builtin_amdgcn_fmed3f(__builtin_fmaf(z, 1.0f / 6.0f, 0.5f), 0.0f, 1.0f);
}

template <int ACT>
__device__ __forceinline__ float mlp_activation(float z) {
    if constexpr (ACT == kActHardSwish) return hard_swish(z);
    else return silu_scaled(z);
}

// Accumulators of one tile -> the B operands of K steps 2t and 2t + 1.
template <int ACT>
__device__ __forceinline__ void act_pack(const f32x16& acc, bf16x8& k0, bf16x8& k1) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        k0[i] = static_cast<__bf16>(mlp_activation<ACT>(acc[i]));
        k1[i] = static_cast<__bf16>(mlp_activation<ACT>(acc[8 + i]));
    }
}

#define KPT_MFMA(a, b, c) __builtin_amdgcn_mfma_f32_32x32x16_bf16((a), (b), (c), 0, 0, 0)

// One 32-sample column block through the network. Output rows 0 and 1 land in
// registers 0 and 1 of the lanes of half 0.
template <int ACT>
__device__ __forceinline__ void mlp_block(const MlpLds& im, int lane, const bf16x8& xin, float b40, float b41,
                                          float& o0, float& o1) {
    const f32x16 zero = {};
    f32x16 a0 = KPT_MFMA(im.frag[(kFragA1 + 0) * 64 + lane], xin, zero);
    f32x16 a1 = KPT_MFMA(im.frag[(kFragA1 + 1) * 64 + lane], xin, zero);
    bf16x8 h0, h1, h2, h3;
    act_pack<ACT>(a0, h0, h1);
    act_pack<ACT>(a1, h2, h3);
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
        const int fa = layer == 0 ? kFragA2 : kFragA3;
        a0 = mlp_bias_tile(im, 2 * layer + 0, lane);
        a1 = mlp_bias_tile(im, 2 * layer + 1, lane);
        a0 = KPT_MFMA(im.frag[(fa + 0) * 64 + lane], h0, a0);
        a1 = KPT_MFMA(im.frag[(fa + 4) * 64 + lane], h0, a1);
        a0 = KPT_MFMA(im.frag[(fa + 1) * 64 + lane], h1, a0);
        a1 = KPT_MFMA(im.frag[(fa + 5) * 64 + lane], h1, a1);
        a0 = KPT_MFMA(im.frag[(fa + 2) * 64 + lane], h2, a0);
        a1 = KPT_MFMA(im.frag[(fa + 6) * 64 + lane], h2, a1);
        a0 = KPT_MFMA(im.frag[(fa + 3) * 64 + lane], h3, a0);
        a1 = KPT_MFMA(im.frag[(fa + 7) * 64 + lane], h3, a1);
        act_pack<ACT>(a0, h0, h1);
        act_pack<ACT>(a1, h2, h3);
    }
    f32x16 y = {};
    if (lane < 32) {
        y[0] = b40;
        y[1] = b41;
    }
    y = KPT_MFMA(im.frag[(kFragA4 + 0) * 64 + lane], h0, y);
    y = KPT_MFMA(im.frag[(kFragA4 + 1) * 64 + lane], h1, y);
    y = KPT_MFMA(im.frag[(kFragA4 + 2) * 64 + lane], h2, y);
    y = KPT_MFMA(im.frag[(kFragA4 + 3) * 64 + lane], h3, y);
    o0 = y[0];
    o1 = y[1];
}

// The 64 samples of a wavefront, one per lane (features f0..f2 already scaled).
// Every lane of the wavefront must execute this call: the MFMAs and the swaps of
// the two lane halves involve all 64 lanes.
template <int ACT>
__device__ __forceinline__ void mlp_eval_wave(const MlpLds& im, float f0, float f1, float f2, float b40, float b41,
                                              float& o0, float& o1) {
    int lane = static_cast<int>(threadIdx.x & 63u);
    // An opaque lane index keeps the fragment and bias reads from LDS at each
    // evaluation: hoisted out of a caller's loop they would pin ~150 registers.
    asm volatile("" : "+v"(lane));
    const float g0 = __shfl_xor(f0, 32), g1 = __shfl_xor(f1, 32), g2 = __shfl_xor(f2, 32);
    bf16x8 x0 = {}, x1 = {};
    if (lane < 32) {
        x0[0] = static_cast<__bf16>(f0);
        x0[1] = static_cast<__bf16>(f1);
        x0[2] = static_cast<__bf16>(f2);
        x0[3] = static_cast<__bf16>(1.0f);
        x1[0] = static_cast<__bf16>(g0);
        x1[1] = static_cast<__bf16>(g1);
        x1[2] = static_cast<__bf16>(g2);
        x1[3] = static_cast<__bf16>(1.0f);
    }
    float p0, p1, q0, q1;
    mlp_block<ACT>(im, lane, x0, b40, b41, p0, p1);  // samples of lanes 0..31
    // One block's working set at a time: letting the scheduler interleave the two
    // blocks (one block's MFMAs beside the other's activations) measured 8% slower
    // in the push, from the registers the second working set takes.
    __builtin_amdgcn_sched_barrier(0);
    mlp_block<ACT>(im, lane, x1, b40, b41, q0, q1);  // samples of lanes 32..63, results in lanes 0..31
    const float r0 = __shfl_xor(q0, 32), r1 = __shfl_xor(q1, 32);
    o0 = lane < 32 ? p0 : r0;
    o1 = lane < 32 ? p1 : r1;
}

#undef KPT_MFMA

// Scaled features and hull membership of one closure evaluation; l2mu returns
// log2|mu|.
__device__ __forceinline__ bool mlp_features(const MlpScale& m, float mu_abs, float log_xi, float gk, float& f0,
                                             float& f1, float& f2, float& l2mu) {
    l2mu = __builtin_amdgcn_logf(mu_abs);
    const float lmu = l2mu * 0.6931471805599453f;
    f0 = (lmu - m.c0) * m.s0;
    f1 = (log_xi - m.c1) * m.s1;
    f2 = (gk - m.c2) * m.s2;
    return mu_abs >= m.mu_lo && mu_abs < 1.0f && log_xi >= m.lxi_lo && log_xi <= m.lxi_hi && gk >= m.g_lo &&
           gk <= m.g_hi;
}

// The pitch-cosine part of the hull test (xi and gamma_k fixed within a sub-step).
__device__ __forceinline__ bool mlp_mu_in_hull(const MlpScale& m, float mu_abs) {
    return mu_abs >= m.mu_lo && mu_abs < 1.0f;
}

}  // namespace kpt
