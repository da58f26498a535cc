// kpt_mlp.h — learned closure of the slab-QLT pitch-angle coefficient.
//
// LegencyModel::evaluate_pitch_angle_scattering (include/LegencyModel.hpp:
// 539-689) gives
//   D_mumu = (pi/4) Omega sigma2 C_gamma (1 - mu^2) R(|mu|, xi, gamma_k),
//   R = xi mu_h^(gamma_k - 1) / (mu_h^gamma_k + xi^gamma_k),  mu_h = |mu| + h0,
// with Omega = |q| |B| / gamma, xi = Omega Lc / v. The prefactor and the kinematic
// (1 - mu^2) are exact; the network stands in for the resonance function R:
//   inputs  f0 = ln|mu|, f1 = ln xi, f2 = gamma_k (affinely scaled to [-1, 1]),
//   outputs y0 = (ln R - out_mean) / out_scale,  y1 = |mu| d ln R / d|mu|,
// three hidden layers of 64 units (SiLU, or hard-swish z clamp(z / 6 + 1/2, 0, 1),
// which needs no transcendental). The five physical inputs of the closure
// (mu, ln p, ln sigma2, gamma_k, ln Lc) reduce exactly to (|mu|, ln xi, gamma_k):
// sigma2 is a prefactor and p, Lc and |B| enter only through xi.
// Then D = P (1 - mu^2) R and dD/dmu = P R ((1 - mu^2) sgn(mu) y1 / |mu| - 2 mu).
//
// Inside the push the network runs as BF16 MFMA (v_mfma_f32_32x32x16_bf16) with
// FP32 accumulation, one wavefront evaluating its 64 lanes' samples as two
// 32-sample column blocks. The MFMA lane layout (A: lane l holds
// A[l % 32][8 (l / 32) + i], B: B[8 (l / 32) + i][l % 32], D register j:
// row 8 (j / 4) + 4 (l / 32) + j % 4 of column l % 32) leaves each lane holding,
// for its column, exactly one K slice of the next layer's input when the next
// layer's weight columns are permuted by mlp_k_feature(); so hidden layers need
// no cross-lane traffic.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/kpt.h"

namespace kpt {

constexpr int kMlpIn = KPT_MLP_INPUTS;    // 3
constexpr int kMlpH = KPT_MLP_HIDDEN;     // 64
constexpr int kMlpOut = KPT_MLP_OUTPUTS;  // 2
// Flat FP32 parameter layout: W1[64][3] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[2][64] b4[2]
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpH * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpH;
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;
constexpr int kOffW3 = kOffB2 + kMlpH;
constexpr int kOffB3 = kOffW3 + kMlpH * kMlpH;
constexpr int kOffW4 = kOffB3 + kMlpH;
constexpr int kOffB4 = kOffW4 + kMlpOut * kMlpH;
constexpr int kMlpParams = kOffB4 + kMlpOut;
static_assert(kMlpParams == KPT_MLP_PARAMS, "parameter count");

// Hidden-layer activations.
constexpr int kActSilu = KPT_MLP_SILU;
constexpr int kActHardSwish = KPT_MLP_HARD_SWISH;
inline bool mlp_activation_ok(int32_t act) { return act == kActSilu || act == kActHardSwish; }

// MFMA fragments: 16 bytes (8 bf16) per lane, 64 lanes per fragment.
//   A1[t]      t = 0, 1         (layer 1, K = features, bias at k = 3)
//   A2[t][s]   s = 0..3 K steps (layer 2), A3[t][s], A4[s] (output rows 0, 1)
constexpr int kFragA1 = 0;
constexpr int kFragA2 = 2;
constexpr int kFragA3 = 10;
constexpr int kFragA4 = 18;
constexpr int kFrags = 22;
// Accumulator initialisers (biases of layers 2 and 3): 16 floats per lane per tile,
// stored as four quarters of 64 lanes x 16 bytes, so that a wavefront reads each
// quarter as one 16-byte load per lane at consecutive addresses (a lane-major
// 64-byte record per lane put every fourth lane on the same LDS banks).
constexpr int kBiasTiles = 4;  // B2[t], B3[t]
__host__ __device__ constexpr size_t mlp_bias_index(int tile, int lane, int j) {
    return ((static_cast<size_t>(tile) * 4 + static_cast<size_t>(j / 4)) * 64 + static_cast<size_t>(lane)) * 4 +
           static_cast<size_t>(j % 4);
}
constexpr size_t kMlpFragBytes = static_cast<size_t>(kFrags) * 64 * 16;
constexpr size_t kMlpBiasBytes = static_cast<size_t>(kBiasTiles) * 64 * 16 * 4;
constexpr size_t kMlpLdsBytes = kMlpFragBytes + kMlpBiasBytes;
// Device image = LDS part followed by the two output biases (FP32) and padding.
constexpr size_t kMlpImageBytes = kMlpLdsBytes + 16;

// Feature of K index k (0..15) in K step s (0..3): the row held in accumulator
// register j = 8 (s % 2) + k % 8 of tile s / 2 by lanes of half k / 8.
__host__ __device__ inline int mlp_k_feature(int s, int k) {
    return 32 * (s / 2) + 8 * (2 * (s % 2) + (k % 8) / 4) + 4 * (k / 8) + (k % 4);
}
// Row of accumulator register j of tile t in a lane of half h.
__host__ __device__ inline int mlp_acc_row(int t, int j, int h) { return 32 * t + 8 * (j / 4) + 4 * h + (j % 4); }

// Affine feature scaling and hull of the trained network.
struct MlpScale {
    float c0, s0, c1, s1, c2, s2;  // f = (x - c) * s
    float mu_lo, lxi_lo, lxi_hi, g_lo, g_hi;
    float out_mean, out_scale;
};

inline MlpScale mlp_scale(const kpt_mlp_hull& h) {
    MlpScale m;
    const double lmu = log(h.mu_lo);
    m.c0 = static_cast<float>(0.5 * lmu);
    m.s0 = static_cast<float>(-2.0 / lmu);
    m.c1 = static_cast<float>(0.5 * (h.log_xi_lo + h.log_xi_hi));
    m.s1 = static_cast<float>(2.0 / (h.log_xi_hi - h.log_xi_lo));
    m.c2 = static_cast<float>(0.5 * (h.gamma_lo + h.gamma_hi));
    m.s2 = static_cast<float>(2.0 / (h.gamma_hi - h.gamma_lo));
    m.mu_lo = static_cast<float>(h.mu_lo);
    m.lxi_lo = static_cast<float>(h.log_xi_lo);
    m.lxi_hi = static_cast<float>(h.log_xi_hi);
    m.g_lo = static_cast<float>(h.gamma_lo);
    m.g_hi = static_cast<float>(h.gamma_hi);
    m.out_mean = static_cast<float>(h.out_mean);
    m.out_scale = static_cast<float>(h.out_scale);
    return m;
}

// Exact targets (FP64): ln R and |mu| d ln R / d|mu| for |mu| > 0.
__host__ __device__ inline void mlp_exact_target(double mu_abs, double log_xi, double gk, double h0, double& ln_r,
                                                 double& y1) {
    const double mh = mu_abs + h0;
    const double lmh = log(mh);
    const double mg = exp(gk * lmh);
    const double xg = exp(gk * log_xi);
    ln_r = log_xi + (gk - 1.0) * lmh - log(mg + xg);
    y1 = mu_abs * ((gk - 1.0) / mh - gk * (mg / mh) / (mg + xg));
}

hipError_t launch_mlp_generate(float* feats, float* targets, double* raw_lnr, uint64_t n, uint64_t seed,
                               uint64_t first, const kpt_mlp_hull& h, hipStream_t stream);
hipError_t launch_mlp_normalize(float* targets, uint64_t n, float mean, float inv_scale, hipStream_t stream);
hipError_t launch_mlp_forward(const float* params, int act, const float* x, uint64_t n, float* z, float* yhat,
                              hipStream_t stream);
hipError_t launch_mlp_backward(const float* params, int act, const float* x, const float* z, const float* yhat,
                               const float* y, uint64_t n, float* dz, float* dy, float* grad_partials, float* grad,
                               float* loss_partials, hipStream_t stream);
hipError_t launch_mlp_adam(float* params, float* m, float* v, const float* grad, float lr, float b1c, float b2c,
                           hipStream_t stream);
hipError_t launch_mlp_infer(const void* lds_image, int act, const float* x, uint64_t n, float* out,
                            hipStream_t stream);
size_t mlp_loss_partials(uint64_t n);
size_t mlp_grad_partials(uint64_t n);  // floats of per-chunk gradient partial sums
// Host: BF16 MFMA fragments and bias initialisers from FP32 parameters of a
// network with activation act.
void mlp_build_image(const float* params, int act, void* image /* kMlpImageBytes */);

}  // namespace kpt
