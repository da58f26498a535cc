// kpt_device_common.h — device helpers shared by the push and the MLP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace kpt {

// Philox4x32-10 (Salmon et al., SC'11) in registers: 10 rounds of two 32x32->64
// multiplies, Weyl-incremented key.
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c0;
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// FocusTransportSolver::reflected_pitch_angle_cosine (include/FocusSolver.hpp:
// 454-470): mirror at +-upper. The mirror loop is bounded (16 passes); a value
// still outside is folded by the closed-form triangle wave of period 4 upper,
// so every lane leaves this function after a fixed amount of work.
__device__ __forceinline__ double reflect_mu(double mu, double upper) {
    const double lower = -upper;
    for (int k = 0; k < 16 && (mu > upper || mu < lower); ++k) {
        if (mu > upper) mu = 2.0 * upper - mu;
        if (mu < lower) mu = 2.0 * lower - mu;
    }
    if (mu > upper || mu < lower) {
        const double period = 4.0 * upper;
        double m = mu + upper;
        m -= period * floor(m / period);
        if (m > 2.0 * upper) m = period - m;
        mu = m - upper;
    }
    return mu;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// (r + 1/2) / 2^32: open-interval U(0, 1) from 32 bits.
__device__ __forceinline__ double u01(uint32_t r) {
    return (static_cast<double>(r) + 0.5) * (1.0 / 4294967296.0);
}

}  // namespace kpt
