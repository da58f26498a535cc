// kpt_kernels.hip — gfx950 kernels of the Parker / focused-transport SDE push.
//
// Layout in HBM
//  - particles: structure of arrays (x, y, p, t, step_dt, weight, mu: f64;
//    status, split_level: i32; id: u64), one lane per particle, so every
//    per-particle load and store of a wavefront is a coalesced 512 B (f64) line set;
//  - coefficient frames: node-major, `stride` components per node (see
//    kpt_internal.h); the four bilinear corners of a sample are two pairs of
//    adjacent nodes, each pair one contiguous 2 x 128 B run for FP64 Parker.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "kpt_internal.h"

namespace kpt {
namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) in registers: 10 rounds of two 32x32->64
// multiplies, Weyl-incremented key.
// ---------------------------------------------------------------------------
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

// (r + 1/2) / 2^32: open-interval U(0, 1) from 32 bits.
__device__ __forceinline__ double u01(uint32_t r) {
    return (static_cast<double>(r) + 0.5) * (1.0 / 4294967296.0);
}

constexpr double kSqrt3 = 1.7320508075688772935;

__device__ __forceinline__ double sqr(double v) { return v * v; }

// wrap_periodic_coordinate (src/main.cpp:641-665).
__device__ __forceinline__ double wrap_periodic(double v, double lo, double hi, double width) {
    double w = v;
    if (w >= hi) {
        w -= width;
        if (w >= hi) w -= floor((w - lo) / width) * width;
    } else if (w < lo) {
        w += width;
        if (w < lo) w -= floor((w - lo) / width) * width;
    }
    if (w >= hi) w -= width;
    if (w < lo) w += width;
    return w;
}

// MappedCoordinateGrid::wrap_coordinate (include/MappedCoordinateGrid.hpp:295-315).
__device__ __forceinline__ double grid_wrap(double x, double lo, double hi, double width,
                                            double inv_width) {
    double w = x - floor((x - lo) * inv_width) * width;
    if (w >= hi) w -= width;
    if (w < lo) w += width;
    return w;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ---------------------------------------------------------------------------
// Coefficient packing: make_parker_reconnection_coefficients /
// make_reconnection_coefficients (src/main.cpp:839-957) on the periodic grid.
// One lane per output node; derivatives are central differences with the
// cell-centre spacing of AnalyticCoordinateGrid (src/main.cpp:763-769).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wrap_index(int i, int n) {
    int w = i % n;
    return w < 0 ? w + n : w;
}

__global__ void __launch_bounds__(kBlock) pack_kernel(const double* __restrict__ raw, GridDesc g,
                                                      int model, uint64_t cplane,
                                                      double* __restrict__ out64,
                                                      float* __restrict__ out32) {
    const int nxp = g.nx + 2, nyp = g.ny + 2;
    const int64_t node = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x;
    if (node >= static_cast<int64_t>(nxp) * nyp) return;
    const int i = static_cast<int>(node % nxp) - 1;
    const int j = static_cast<int>(node / nxp) - 1;
    const int64_t plane = static_cast<int64_t>(g.nx) * g.ny;

    auto val = [&](int comp, int ii, int jj) -> double {
        return raw[comp * plane + static_cast<int64_t>(wrap_index(jj, g.ny)) * g.nx +
                   wrap_index(ii, g.nx)];
    };
    auto bmag = [&](int ii, int jj) -> double {
        const double bx = val(0, ii, jj), by = val(1, ii, jj), bz = val(2, ii, jj);
        return sqrt(bx * bx + by * by + bz * bz);
    };
    auto cell = [&](int d, double k) -> double {
        return 0.5 * ((g.start[d] + k * g.step[d]) + (g.start[d] + (k + 1.0) * g.step[d]));
    };
    const double hx = g.nx > 1 ? cell(0, i + 1) - cell(0, i - 1) : 1.0;
    const double hy = g.ny > 1 ? cell(1, j + 1) - cell(1, j - 1) : 1.0;
    // comp: 0..2 B, 3..5 V
    auto ddx = [&](int comp) -> double {
        return g.nx > 1 ? (val(comp, i + 1, j) - val(comp, i - 1, j)) / hx : 0.0;
    };
    auto ddy = [&](int comp) -> double {
        return g.ny > 1 ? (val(comp, i, j + 1) - val(comp, i, j - 1)) / hy : 0.0;
    };
    const double bx = val(0, i, j), by = val(1, i, j), bz = val(2, i, j);
    double c[24];
#pragma unroll
    for (int k = 0; k < 24; ++k) c[k] = 0.0;
    const double dbm_dx = g.nx > 1 ? (bmag(i + 1, j) - bmag(i - 1, j)) / hx : 0.0;
    const double dbm_dy = g.ny > 1 ? (bmag(i, j + 1) - bmag(i, j - 1)) / hy : 0.0;
    if (model == KPT_MODEL_PARKER16) {
        c[0] = val(3, i, j);
        c[1] = val(4, i, j);
        c[2] = bx;
        c[3] = by;
        c[4] = bz;
        c[5] = sqrt(bx * bx + by * by + bz * bz);
        c[6] = ddx(3);
        c[7] = ddy(4);
        c[8] = ddx(0);
        c[9] = ddy(0);
        c[10] = ddx(1);
        c[11] = ddy(1);
        c[12] = ddx(2);
        c[13] = ddy(2);
        c[14] = dbm_dx;
        c[15] = dbm_dy;
    } else {
        c[0] = val(3, i, j);
        c[1] = val(4, i, j);
        c[2] = val(5, i, j);
        c[3] = bx;
        c[4] = by;
        c[5] = bz;
        c[6] = sqrt(bx * bx + by * by + bz * bz);
        c[7] = ddx(3);
        c[8] = ddy(3);
        c[9] = ddx(4);
        c[10] = ddy(4);
        c[11] = ddx(5);
        c[12] = ddy(5);
        c[13] = ddx(0);
        c[14] = ddy(0);
        c[15] = ddx(1);
        c[16] = ddy(1);
        c[17] = ddx(2);
        c[18] = ddy(2);
        c[19] = dbm_dx;
        c[20] = dbm_dy;
    }
    for (int k = 0; k < model; ++k) {
        out64[k * cplane + node] = c[k];
        out32[k * cplane + node] = static_cast<float>(c[k]);
    }
}

// ---------------------------------------------------------------------------
// Bilinear, time-blended sample of C components at one particle.
// Corner order and weights follow LinearPositionInterpolator::sample
// (include/FieldInterpolator.hpp:363-378): (lo,lo), (hi,lo), (lo,hi), (hi,hi).
// ---------------------------------------------------------------------------
// Each corner value is its own load from component plane k; the two corners of a
// row, (i, j) and (i + 1, j), are adjacent in the plane.
template <int C, typename FT>
__device__ __forceinline__ void sample_corners(const FT* __restrict__ f, int64_t n00, int nxp,
                                               uint64_t plane, FT w00, FT w10, FT w01, FT w11,
                                               FT* out) {
#pragma unroll
    for (int k = 0; k < C; ++k) {
        const FT* q = f + k * plane + n00;
        out[k] = w00 * q[0] + w10 * q[1] + w01 * q[nxp] + w11 * q[nxp + 1];
    }
}

// Node index and weights of the stencil along one axis for a sampling coordinate
// q (already shifted by interpolation_shift): the cell-centred inverse map of
// include/AnalyticCoordinateGrid.hpp:205-210 after the periodic wrap of
// include/FieldInterpolator.hpp:465-480. Returns false on a NaN coordinate
// (the reference aborts there; the push marks the particle instead).
__device__ __forceinline__ bool axis_stencil(double q, const GridDesc& g, int d, int n, int& i0,
                                             double& w) {
    const double qw = grid_wrap(q, g.phys_lo[d], g.phys_hi[d], g.width[d], g.inv_width[d]);
    const double logical = (qw - g.start[d]) * g.inv_step[d] - 0.5;
    if (!(logical == logical)) return false;
    double fl = floor(logical);
    // Post-wrap `logical` lies in [-0.5, n - 0.5); clamping keeps the gather in
    // the stored node range [-1, n] for any input.
    fl = fl < -1.0 ? -1.0 : (fl > static_cast<double>(n - 1) ? static_cast<double>(n - 1) : fl);
    i0 = static_cast<int>(fl);
    w = logical - fl;
    w = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
    return true;
}

// ---------------------------------------------------------------------------
// The push: one lane advances one particle through the whole interval
// (move_parker_particles_one_interval, src/main.cpp:1113-1325, and the shared
// mover move_particles_one_interval, src/main.cpp:1364-1753).
// MODEL: 16 or 21 packed components. FOCUSED: reduced focused branch.
// FT: storage / interpolation type of the coefficient frames.
// ---------------------------------------------------------------------------
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

// Slab-QLT D_mumu and dD_mumu/dmu (include/LegencyModel.hpp:557-643) for one
// trial mu, with omega, xi = omega Lc / v and xi^gamma_k fixed for the sub-step.
// `pref` = (pi / 4) omega sigma2 gamma_k sin(pi / gamma_k) / pi.
__device__ __forceinline__ void qlt_scattering(double mu, double xi, double xi_g, double pref, double gk,
                                               double h0, double& D, double& dD) {
    const double one_minus_mu2 = 1.0 - mu * mu;
    D = 0.0;
    dD = 0.0;
    if (one_minus_mu2 <= 0.0) return;
    const double mu_h = fabs(mu) + h0;
    const double pm1 = pow(mu_h, gk - 1.0);  // mu_h^(gamma_k - 1)
    const double den = pm1 * mu_h + xi_g;    // mu_h^gamma_k + xi^gamma_k
    D = pref * one_minus_mu2 * xi * pm1 / den;
    if (D == 0.0 || mu == 0.0) return;
    const double sign_mu = mu > 0.0 ? 1.0 : -1.0;
    const double shape = sign_mu * ((gk - 1.0) / mu_h - gk * pm1 / den);
    dD = D * (-2.0 * mu / one_minus_mu2 + shape);
}

template <int MODEL, bool FOCUSED, typename FT, int CLOSURE>
__global__ void __launch_bounds__(kBlock) push_kernel(PushArgs a) {
    constexpr bool QLT = FOCUSED && CLOSURE != KPT_CLOSURE_REDUCED;
    constexpr int CP = MODEL;  // components read per corner
    const int64_t idx = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x;
    uint64_t my_steps = 0;
    uint32_t my_fail = 0, my_active = 0;

    if (idx < static_cast<int64_t>(a.count) && a.P.status[idx] == 0) {
        my_active = 1;
        const kpt_device_settings& s = a.s;
        const FT* __restrict__ flo = static_cast<const FT*>(a.field_lo);
        const FT* __restrict__ fhi = static_cast<const FT*>(a.field_hi);
        const int nxp = a.g.nx + 2;
        const double interval_start = a.t0;
        const double interval_end = a.t_end;
        double x = a.P.x[idx], y = a.P.y[idx], p = a.P.p[idx], t = a.P.t[idx];
        double dt = a.P.step_dt[idx];
        double diagnostic_dt = dt;
        double pitch_mu = 0.0;
        if (MODEL == 21) {
            const double m = a.P.mu[idx];
            pitch_mu = QLT ? reflect_mu(m, s.mu_max) : (m > s.mu_max ? s.mu_max : (m < -s.mu_max ? -s.mu_max : m));
        }
        const uint64_t id = a.P.id[idx];
        // Sub-steps already taken in this interval (earlier chunks): the step
        // guard and the Philox sub-step counter continue from there.
        int step_guard = a.first_chunk ? 0 : static_cast<int>(a.P.substep[idx]);
        int status = 0;
        const double t_stop = a.t_stop;

        while (t < t_stop && step_guard < s.max_particle_steps_per_interval && status == 0) {
            ++step_guard;
            x = wrap_periodic(x, s.x_lower, s.x_upper, s.x_width);
            y = wrap_periodic(y, s.y_lower, s.y_upper, s.y_width);

            int i0, j0;
            double wx, wy;
            if (!axis_stencil(x + s.interpolation_shift_x, a.g, 0, a.g.nx, i0, wx) ||
                !axis_stencil(y + s.interpolation_shift_y, a.g, 1, a.g.ny, j0, wy)) {
                status = 2;
                break;
            }
            const int64_t n00 = static_cast<int64_t>(j0 + 1) * nxp + (i0 + 1);
            const FT w00 = static_cast<FT>((1.0 - wx) * (1.0 - wy));
            const FT w10 = static_cast<FT>(wx * (1.0 - wy));
            const FT w01 = static_cast<FT>((1.0 - wx) * wy);
            const FT w11 = static_cast<FT>(wx * wy);

            double c[CP];
            {
                FT lo[CP];
                sample_corners<CP, FT>(flo, n00, nxp, a.plane, w00, w10, w01, w11, lo);
                if (s.time_interpolation != 0) {
                    FT hi[CP];
                    sample_corners<CP, FT>(fhi, n00, nxp, a.plane, w00, w10, w01, w11, hi);
                    const double rt = (t - interval_start) / s.mhd_dt;
                    const double uw = rt < 0.0 ? 0.0 : (rt > 1.0 ? 1.0 : rt);
                    const double lw = 1.0 - uw;
#pragma unroll
                    for (int k = 0; k < CP; ++k)
                        c[k] = lw * static_cast<double>(lo[k]) + uw * static_cast<double>(hi[k]);
                } else {
#pragma unroll
                    for (int k = 0; k < CP; ++k) c[k] = static_cast<double>(lo[k]);
                }
            }

            double vx, vy, bx, by, bz, b, dvx_dx, dvx_dy = 0.0, dvy_dx = 0.0, dvy_dy,
                                              dvz_dx = 0.0, dvz_dy = 0.0, dbx_dx, dbx_dy, dby_dx,
                                              dby_dy, dbz_dx, dbz_dy, db_dx, db_dy;
            if constexpr (MODEL == 16) {
                vx = c[0]; vy = c[1]; bx = c[2]; by = c[3]; bz = c[4]; b = c[5];
                dvx_dx = c[6]; dvy_dy = c[7]; dbx_dx = c[8]; dbx_dy = c[9];
                dby_dx = c[10]; dby_dy = c[11]; dbz_dx = c[12]; dbz_dy = c[13];
                db_dx = c[14]; db_dy = c[15];
            } else {
                vx = c[0]; vy = c[1]; bx = c[3]; by = c[4]; bz = c[5]; b = c[6];
                dvx_dx = c[7]; dvx_dy = c[8]; dvy_dx = c[9]; dvy_dy = c[10];
                dvz_dx = c[11]; dvz_dy = c[12]; dbx_dx = c[13]; dbx_dy = c[14];
                dby_dx = c[15]; dby_dy = c[16]; dbz_dx = c[17]; dbz_dy = c[18];
                db_dx = c[19]; db_dy = c[20];
            }

            if (!isfinite(p) || p <= 0.0 || !isfinite(b) || b <= 0.0) {
                status = 2;
                break;
            }
            const double ib = 1.0 / b;
            const double ib2 = ib * ib;
            const double ib3 = ib2 * ib;

            // kappa model (src/main.cpp:1177-1199)
            double knorm = 1.0;
            if (s.magnetic_dependency == 1) knorm *= pow(b, s.gamma_turb - 2.0);
            if (s.momentum_dependency == 1) knorm *= pow(p / s.p0, s.pindex);
            const double kpara = s.kpara0 * knorm;
            const double kperp = kpara * s.kperp_over_kpara;
            if (!isfinite(kpara) || kpara <= 0.0 || !isfinite(kperp) || kperp < 0.0) {
                status = 2;
                break;
            }
            const double skpara = sqrt(2.0 * kpara);
            const double skperp = sqrt(2.0 * kperp);
            const double skpara_perp = MODEL == 16 ? sqrt(2.0 * (kpara - kperp))
                                                   : (kpara > kperp ? sqrt(2.0 * (kpara - kperp)) : 0.0);
            const double kpp = FOCUSED ? -kperp : kpara - kperp;

            double dkdx = 0.0, dkdy = 0.0;
            if (s.magnetic_dependency == 1) {
                const double me = s.gamma_turb - 2.0;
                dkdx = db_dx * ib * me;
                dkdy = db_dy * ib * me;
            }
            // divergence of the diffusion tensor (src/main.cpp:1210-1223)
            const double dkxx_dx = kperp * dkdx + kpp * dkdx * bx * bx * ib2 +
                                   2.0 * kpp * bx * (dbx_dx * b - bx * db_dx) * ib3;
            const double dkyy_dy = kperp * dkdy + kpp * dkdy * by * by * ib2 +
                                   2.0 * kpp * by * (dby_dy * b - by * db_dy) * ib3;
            const double dkxy_dx = kpp * dkdx * bx * by * ib2 +
                                   kpp * ((dbx_dx * by + bx * dby_dx) * ib2 - 2.0 * bx * by * db_dx * ib3);
            const double dkxy_dy = kpp * dkdy * bx * by * ib2 +
                                   kpp * ((dbx_dy * by + bx * dby_dy) * ib2 - 2.0 * bx * by * db_dy * ib3);
            const double drift_den =
                sqrt(sqr(s.drift_param1 * s.p0 / p) + sqr(s.drift_param2 * s.p0 * s.p0 / (p * p)));

            double dx_dt, dy_dt, dp_dt, dmu_dt = 0.0, duu = 0.0;
            // focusing-rate pieces and QLT constants reused by the Picard passes
            double f_speed = 0.0, f_divv = 0.0, f_bb = 0.0, f_bv = 0.0, f_divb = 0.0;
            double q_xi = 0.0, q_xig = 0.0, q_pref = 0.0;
            if (FOCUSED) {
                // reduced focused branch (src/main.cpp:1499-1589)
                const double mu2 = pitch_mu * pitch_mu;
                const double muf1 = 0.5 * (1.0 - mu2);
                const double muf2 = 0.5 * (3.0 * mu2 - 1.0);
                const double kx = bx * dbx_dx + by * dbx_dy;
                const double ky = bx * dby_dx + by * dby_dy;
                const double kz = bx * dbz_dx + by * dbz_dy;
                const double bdot_curvb = bx * dbz_dy - by * dbz_dx + bz * (dby_dx - dbx_dy);
                const double vdp = 1.0 / s.charge / drift_den;
                const double vdx = vdp * (muf1 * (-bz * db_dy) * ib2 + mu2 * (by * kz - bz * ky) * ib3 +
                                          muf1 * bx * bdot_curvb * ib3);
                const double vdy = vdp * (muf1 * (bz * db_dx) * ib2 + mu2 * (bz * kx - bx * kz) * ib3 +
                                          muf1 * by * bdot_curvb * ib3);
                // QLT: relativistic speed p / gamma (m = c = 1)
                const double gamma_p = QLT ? sqrt(1.0 + p * p) : 1.0;
                const double speed = QLT ? p / gamma_p : s.particle_v0 * p / s.p0;
                if (!isfinite(speed) || speed <= 0.0) {
                    status = 2;
                    break;
                }
                const double vpb = speed * pitch_mu * ib;
                dx_dt = vx + vdx + vpb * bx + dkxx_dx + dkxy_dy;
                dy_dt = vy + vdy + vpb * by + dkxy_dx + dkyy_dy;
                const double divv = dvx_dx + dvy_dy;
                const double bb_gradv = (bx * (bx * dvx_dx + by * dvx_dy) + by * (bx * dvy_dx + by * dvy_dy) +
                                         bz * (bx * dvz_dx + by * dvz_dy)) * ib2;
                const double bv_gradv = (bx * (vx * dvx_dx + vy * dvx_dy) + by * (vx * dvy_dx + vy * dvy_dy) +
                                         bz * (vx * dvz_dx + vy * dvz_dy)) * ib;
                dp_dt = p * -(muf1 * divv + muf2 * bb_gradv + pitch_mu * bv_gradv / speed);
                const double div_bnorm = -(bx * db_dx + by * db_dy) * ib2;
                dmu_dt = speed * div_bnorm + pitch_mu * divv - 3.0 * pitch_mu * bb_gradv -
                         2.0 * bv_gradv / speed;
                dmu_dt *= 0.5 * (1.0 - mu2);
                double duu_du = 0.0;
                if constexpr (QLT) {
                    f_speed = speed;
                    f_divv = divv;
                    f_bb = bb_gradv;
                    f_bv = bv_gradv;
                    f_divb = div_bnorm;
                    const double omega = fabs(s.charge) * b / gamma_p;
                    q_xi = omega * s.qlt_correlation_length / speed;
                    q_xig = pow(q_xi, s.qlt_spectral_index);
                    q_pref = 0.25 * M_PI * omega * s.qlt_sigma2 * a.qlt_norm;
                    qlt_scattering(pitch_mu, q_xi, q_xig, q_pref, s.qlt_spectral_index, s.qlt_h0, duu, duu_du);
                } else {
                    const double abs_mu = fabs(pitch_mu);
                    const double dtmp = pow(abs_mu, s.gamma_turb - 1.0) + s.pitch_angle_scattering_h0;
                    duu = s.duu0 * (1.0 - mu2) * dtmp;
                    if (pitch_mu != 0.0) {
                        const double g2 = (1.0 - mu2) * pow(abs_mu, s.gamma_turb - 2.0);
                        duu_du = s.duu0 * (-2.0 * pitch_mu * dtmp + (pitch_mu > 0.0 ? g2 : -g2));
                    }
                    double duu_norm = 1.0;
                    if (s.magnetic_dependency == 1) duu_norm *= pow(b, 2.0 - s.gamma_turb);
                    if (s.momentum_dependency == 1) duu_norm *= pow(p / s.p0, s.gamma_turb - 1.0);
                    duu *= duu_norm;
                    duu_du *= duu_norm;
                }
                dmu_dt += duu_du;
                if (!isfinite(dmu_dt) || !isfinite(duu) || duu < 0.0) {
                    status = 2;
                    break;
                }
            } else {
                // Parker branch (src/main.cpp:1225-1241)
                const double vdp = 1.0 / (3.0 * s.charge) / drift_den;
                const double vdx = vdp * (dbz_dy * ib2 - 2.0 * bz * db_dy * ib3);
                const double vdy = vdp * (-dbz_dx * ib2 + 2.0 * bz * db_dx * ib3);
                dx_dt = vx + vdx + dkxx_dx + dkxy_dy;
                dy_dt = vy + vdy + dkxy_dx + dkyy_dy;
                dp_dt = -p * (dvx_dx + dvy_dy) / 3.0;
            }
            (void)dvz_dx;

            // adaptive step (src/main.cpp:1243-1277, :1603-1665)
            if (FOCUSED) {
                const double ds = skperp > 0.0 ? skperp : skpara;
                if (dx_dt != 0.0 && dy_dt != 0.0 && dp_dt != 0.0 && dmu_dt != 0.0 && ds > 0.0) {
                    dt = fmin(sqr(0.5 * s.dx / ds), sqr(0.5 * s.dy / ds));
                    dt = fmin(dt, sqr(ds / dx_dt));
                    dt = fmin(dt, sqr(ds / dy_dt));
                    dt = fmin(dt, 0.1 * p / fabs(dp_dt));
                    dt = fmin(dt, 0.1 / fabs(dmu_dt));
                    dt = fmin(dt, 2.0 * duu / sqr(dmu_dt));
                } else {
                    dt = s.dt_min;
                }
            } else if (dx_dt != 0.0 && dy_dt != 0.0 && dp_dt != 0.0) {
                const double as = skperp > 0.0 ? skperp : skpara;
                dt = fmin(sqr(0.5 * s.dx / skpara), sqr(0.5 * s.dy / skpara));
                dt = fmin(dt, sqr(as / dx_dt));
                dt = fmin(dt, sqr(as / dy_dt));
                dt = fmin(dt, 0.1 * p / fabs(dp_dt));
            } else {
                dt = s.dt_min;
            }
            if (dt < s.dt_min) dt = s.dt_min;
            if (dt > s.dt_max) dt = s.dt_max;
            diagnostic_dt = dt;
            const double remaining = interval_end - t;
            if (dt > remaining) dt = remaining;
            if (!isfinite(dt) || dt <= 0.0) {
                status = 2;
                break;
            }

            double r0, r1, r2, r3;
            const uint32_t k = static_cast<uint32_t>(step_guard - 1);
            if (a.noise != nullptr) {
                if (k >= a.noise_steps) {
                    status = 2;
                    break;
                }
                const double* row = a.noise + (static_cast<uint64_t>(idx) * a.noise_steps + k) * 4;
                r0 = row[0];
                r1 = row[1];
                r2 = row[2];
                r3 = row[3];
            } else {
                uint32_t rb[4];
                philox4x32_10(static_cast<uint32_t>(id), static_cast<uint32_t>(id >> 32), a.frame, k,
                              a.key[0], a.key[1], rb);
                r0 = -kSqrt3 + (kSqrt3 - -kSqrt3) * u01(rb[0]);
                r1 = -kSqrt3 + (kSqrt3 - -kSqrt3) * u01(rb[1]);
                r2 = -kSqrt3 + (kSqrt3 - -kSqrt3) * u01(rb[2]);
                r3 = -kSqrt3 + (kSqrt3 - -kSqrt3) * u01(rb[3]);
            }
            const double sdt = sqrt(dt);
            double delta_x, delta_y;
            if (FOCUSED) {
                // src/main.cpp:1671-1702 (ran1, ran2, ignored, pitch)
                const double bxn = bx * ib, byn = by * ib, bzn = bz * ib;
                const double bxy = sqrt(bxn * bxn + byn * byn);
                double sx = 0.0, sy = 0.0;
                if (bxy > 0.0) {
                    const double ibxy = 1.0 / bxy;
                    sx = skperp * ibxy * sdt * (-bxn * bzn * r0 - by * r1);
                    sy = skperp * ibxy * sdt * (-byn * bzn * r0 + bx * r1);
                }
                delta_x = dx_dt * dt + sx;
                delta_y = dy_dt * dt + sy;
                if constexpr (QLT) {
                    // FocusTransportSolver::picard_pitch_angle_update (FocusSolver.hpp:700-719):
                    // three passes, drift and diffusion at the trial mu, one noise value.
                    const double mu0 = pitch_mu;
                    double rate = dmu_dt, D = duu, trial = mu0;
#pragma unroll 1
                    for (int it = 0; it < 3; ++it) {
                        if (it > 0) {
                            const double m = reflect_mu(trial, s.mu_max);
                            double dD;
                            qlt_scattering(m, q_xi, q_xig, q_pref, s.qlt_spectral_index, s.qlt_h0, D, dD);
                            rate = 0.5 * (1.0 - m * m) *
                                       (f_speed * f_divb + m * f_divv - 3.0 * m * f_bb - 2.0 * f_bv / f_speed) +
                                   dD;
                        }
                        const double scale = D > 0.0 ? sqrt(2.0 * D * dt) : 0.0;
                        trial = reflect_mu(mu0 + dt * rate + scale * r3, s.mu_max);
                    }
                    pitch_mu = trial;
                } else {
                    const double m = pitch_mu + dmu_dt * dt + r3 * sqrt(2.0 * duu) * sdt;
                    pitch_mu = m > s.mu_max ? s.mu_max : (m < -s.mu_max ? -s.mu_max : m);
                }
            } else {
                // src/main.cpp:1279-1292 (ran1, ran2, ran3, ignored)
                delta_x = dx_dt * dt + r0 * skperp * sdt + r2 * skpara_perp * sdt * bx * ib;
                delta_y = dy_dt * dt + r1 * skperp * sdt + r2 * skpara_perp * sdt * by * ib;
            }
            (void)r3;
            x += delta_x;
            y += delta_y;
            p += dp_dt * dt;
            t += dt;
            ++my_steps;
            if (p < 0.25 * s.p0) p = 0.25 * s.p0;
            if (!isfinite(x) || !isfinite(y) || !isfinite(p) || !isfinite(t)) {
                status = 2;
                break;
            }
        }
        if (step_guard >= s.max_particle_steps_per_interval && t < interval_end) status = 2;

        a.P.x[idx] = wrap_periodic(x, s.x_lower, s.x_upper, s.x_width);
        a.P.y[idx] = wrap_periodic(y, s.y_lower, s.y_upper, s.y_width);
        a.P.p[idx] = p;
        if (MODEL == 21) a.P.mu[idx] = pitch_mu;
        a.P.t[idx] = t;
        a.P.step_dt[idx] = diagnostic_dt;
        a.P.status[idx] = status;
        a.P.substep[idx] = static_cast<uint32_t>(step_guard);
        my_fail = status != 0 ? 1u : 0u;
    }

    const unsigned long long steps = wave_sum(static_cast<unsigned long long>(my_steps));
    const unsigned int fails = wave_sum(my_fail);
    const unsigned int act = wave_sum(my_active);
    if ((threadIdx.x & 63) == 0 && act != 0) {
        atomicAdd(&a.counters[0], steps);
        atomicAdd(&a.counters[1], static_cast<unsigned long long>(fails));
        atomicAdd(&a.counters[2], static_cast<unsigned long long>(act));
    }
}

// ---------------------------------------------------------------------------
// Injection (src/main.cpp:1021-1044): x, y, t, mu from one Philox block with
// counter (root id, 0xFFFFFFFF, frame, root id >> 32).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) inject_kernel(ParticleArrays P, uint64_t first, uint64_t n,
                                                        uint64_t first_root, kpt_device_settings s,
                                                        kpt_inject_opts o, double t0, uint32_t frame,
                                                        uint32_t k0, uint32_t k1) {
    const uint64_t k = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x;
    if (k >= n) return;
    const uint64_t i = first + k;
    const uint64_t root = first_root + k;
    uint32_t r[4];
    philox4x32_10(static_cast<uint32_t>(root), 0xFFFFFFFFu, frame, static_cast<uint32_t>(root >> 32),
                  k0, k1, r);
    const double xl = o.use_region ? o.region_x_lo : s.x_lower;
    const double xu = o.use_region ? o.region_x_hi : s.x_upper;
    const double yl = o.use_region ? o.region_y_lo : s.y_lower;
    const double yu = o.use_region ? o.region_y_hi : s.y_upper;
    P.x[i] = xl + (xu - xl) * u01(r[0]);
    P.y[i] = yl + (yu - yl) * u01(r[1]);
    P.p[i] = o.p_init > 0.0 ? o.p_init : s.p0;
    P.t[i] = o.fixed_time ? t0 : t0 + (s.mhd_dt - 0.0) * u01(r[2]);
    P.step_dt[i] = s.dt_min;
    P.weight[i] = 1.0;
    P.mu[i] = s.mu_max * (2.0 * u01(r[3]) - 1.0);
    P.status[i] = 0;
    P.split_level[i] = 0;
    P.id[i] = root;
}

// ---------------------------------------------------------------------------
// Deterministic split (src/main.cpp:1791-1832). Pass 1 counts candidate parents
// per 256-slot tile, pass 2 scans the tile counts, pass 3 ranks candidates
// inside the tile with wavefront ballots and appends the children in parent
// slot order.
// ---------------------------------------------------------------------------
// The threshold pmin_split * ratio^level comes from a host-computed table so
// the predicate is the same double the host library (and the oracle) computes.
__device__ __forceinline__ bool split_candidate(const ParticleArrays& P, uint64_t i,
                                                const kpt_device_settings& s, const SplitTable& tab) {
    if (P.status[i] != 0) return false;
    const int level = P.split_level[i];
    const double thr = (level >= 0 && level < SplitTable::kLevels)
                           ? tab.thr[level]
                           : s.pmin_split * pow(s.split_ratio, static_cast<double>(level));
    const double p = P.p[i];
    return p > thr && p <= s.pmax;
}

__global__ void __launch_bounds__(kBlock) split_count_kernel(ParticleArrays P, uint64_t count,
                                                             kpt_device_settings s, SplitTable tab,
                                                             uint32_t* __restrict__ tile_counts) {
    __shared__ uint32_t wave_counts[kBlock / 64];
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x;
    const bool f = i < count && split_candidate(P, i, s, tab);
    const uint64_t m = __ballot(f);
    if ((threadIdx.x & 63) == 0) wave_counts[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += wave_counts[w];
        tile_counts[blockIdx.x] = t;
    }
}

// Exclusive scan of `n` tile counts in one workgroup of 1024 lanes; the total
// lands in tile_counts[n].
__global__ void __launch_bounds__(1024) split_scan_kernel(uint32_t* __restrict__ tile_counts,
                                                          uint32_t n) {
    __shared__ uint32_t buf[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < n ? tile_counts[i] : 0u;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const uint32_t add = threadIdx.x >= static_cast<unsigned>(off) ? buf[threadIdx.x - off] : 0u;
            __syncthreads();
            buf[threadIdx.x] += add;
            __syncthreads();
        }
        const uint32_t incl = buf[threadIdx.x];
        if (i < n) tile_counts[i] = carry + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) tile_counts[n] = carry;
}

__global__ void __launch_bounds__(kBlock) split_apply_kernel(ParticleArrays P, uint64_t count,
                                                             uint64_t capacity, kpt_device_settings s,
                                                             SplitTable tab,
                                                             const uint32_t* __restrict__ tile_offsets,
                                                             unsigned long long* counters) {
    __shared__ uint32_t wave_counts[kBlock / 64];
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x;
    const bool f = i < count && split_candidate(P, i, s, tab);
    const uint64_t m = __ballot(f);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (lane == 0) wave_counts[wave] = __popcll(m);
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += wave_counts[w];
    const uint64_t lt = lane == 0 ? 0ull : (m & ((1ull << lane) - 1ull));
    const uint64_t rank = static_cast<uint64_t>(tile_offsets[blockIdx.x]) + before + __popcll(lt);
    uint32_t skipped = 0, done = 0;
    if (f) {
        const uint64_t child = count + rank;
        if (child >= capacity) {
            skipped = 1;
        } else {
            const int old_level = P.split_level[i];
            const int new_level = old_level + 1;
            const double w = ldexp(1.0, -new_level);  // 0.5^level, exact
            P.x[child] = P.x[i];
            P.y[child] = P.y[i];
            P.p[child] = P.p[i];
            P.t[child] = P.t[i];
            P.step_dt[child] = P.step_dt[i];
            P.weight[i] = w;
            P.weight[child] = w;
            P.mu[child] = P.mu[i];
            P.status[child] = 0;
            P.split_level[i] = new_level;
            P.split_level[child] = new_level;
            const uint64_t pid = P.id[i];
            const uint64_t path = (pid >> 32) | (old_level < 32 ? (1ull << old_level) : 0ull);
            P.id[child] = (pid & 0xFFFFFFFFull) | (path << 32);
            done = 1;
        }
    }
    const unsigned int ws = wave_sum(skipped);
    const unsigned int wd = wave_sum(done);
    if (lane == 0 && (ws | wd) != 0) {
        atomicAdd(&counters[0], static_cast<unsigned long long>(wd));
        atomicAdd(&counters[1], static_cast<unsigned long long>(ws));
    }
}

// ---------------------------------------------------------------------------
// Histograms with LDS-privatised bins: each workgroup accumulates a private
// copy in LDS, writes it to its row of `partials`, and a second kernel sums
// the rows in fixed order.
// ---------------------------------------------------------------------------
constexpr int kHistBlocks = 512;
constexpr int kHistMaxBins = 4096;

__global__ void __launch_bounds__(kBlock) hist_logp_kernel(ParticleArrays P, uint64_t count, int bins,
                                                           double log_pmin, double inv_dlogp, double pmin,
                                                           double pmax, double* __restrict__ partials) {
    __shared__ double h[kHistMaxBins];
    for (int b = threadIdx.x; b < bins; b += kBlock) h[b] = 0.0;
    __syncthreads();
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; i < count; i += stride) {
        if (P.status[i] != 0) continue;
        const double p = P.p[i];
        if (p < pmin || p > pmax) continue;
        int bin = static_cast<int>((log(p) - log_pmin) * inv_dlogp);
        bin = bin < 0 ? 0 : (bin >= bins ? bins - 1 : bin);
        atomicAdd(&h[bin], P.weight[i]);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < bins; b += kBlock) partials[static_cast<size_t>(blockIdx.x) * bins + b] = h[b];
}

__global__ void __launch_bounds__(kBlock) hist_xy_kernel(ParticleArrays P, uint64_t count, int bx, int by,
                                                         double x_lo, double x_hi, double y_lo,
                                                         double y_hi, double sx, double sy,
                                                         double* __restrict__ partials) {
    __shared__ double h[kHistMaxBins];
    const int bins = bx * by;
    for (int b = threadIdx.x; b < bins; b += kBlock) h[b] = 0.0;
    __syncthreads();
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; i < count; i += stride) {
        if (P.status[i] != 0) continue;
        const double x = P.x[i], y = P.y[i];
        if (!(x >= x_lo && x < x_hi && y >= y_lo && y < y_hi)) continue;
        int ix = static_cast<int>((x - x_lo) * sx);
        int iy = static_cast<int>((y - y_lo) * sy);
        ix = ix >= bx ? bx - 1 : ix;
        iy = iy >= by ? by - 1 : iy;
        atomicAdd(&h[iy * bx + ix], P.weight[i]);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < bins; b += kBlock) partials[static_cast<size_t>(blockIdx.x) * bins + b] = h[b];
}

__global__ void __launch_bounds__(kBlock) hist_reduce_kernel(const double* __restrict__ partials, int rows,
                                                             int bins, double* __restrict__ out) {
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= bins) return;
    double acc = 0.0;
    for (int r = 0; r < rows; ++r) acc += partials[static_cast<size_t>(r) * bins + b];
    out[b] = acc;
}

// ---------------------------------------------------------------------------
// Summary reductions (summarize_particles, src/main.cpp:1844-1901).
// ---------------------------------------------------------------------------
struct SummaryPartial {
    unsigned long long active, inactive;
    double wsum, psum, dtsum, pmin, pmax, dtmin, dtmax;
    int32_t maxlevel, pad;
};
constexpr int kSummaryBlocks = 512;

__global__ void __launch_bounds__(kBlock) summary_kernel(ParticleArrays P, uint64_t count,
                                                         SummaryPartial* __restrict__ partials) {
    __shared__ SummaryPartial sh[kBlock / 64];
    SummaryPartial v{0, 0, 0.0, 0.0, 0.0, INFINITY, 0.0, INFINITY, 0.0, 0, 0};
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; i < count; i += stride) {
        if (P.status[i] == 0) {
            const double p = P.p[i], dt = P.step_dt[i];
            ++v.active;
            v.wsum += P.weight[i];
            v.psum += p;
            v.dtsum += dt;
            v.pmin = fmin(v.pmin, p);
            v.pmax = fmax(v.pmax, p);
            v.dtmin = fmin(v.dtmin, dt);
            v.dtmax = fmax(v.dtmax, dt);
            v.maxlevel = max(v.maxlevel, P.split_level[i]);
        } else {
            ++v.inactive;
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        v.active += __shfl_xor(v.active, off, 64);
        v.inactive += __shfl_xor(v.inactive, off, 64);
        v.wsum += __shfl_xor(v.wsum, off, 64);
        v.psum += __shfl_xor(v.psum, off, 64);
        v.dtsum += __shfl_xor(v.dtsum, off, 64);
        v.pmin = fmin(v.pmin, __shfl_xor(v.pmin, off, 64));
        v.pmax = fmax(v.pmax, __shfl_xor(v.pmax, off, 64));
        v.dtmin = fmin(v.dtmin, __shfl_xor(v.dtmin, off, 64));
        v.dtmax = fmax(v.dtmax, __shfl_xor(v.dtmax, off, 64));
        v.maxlevel = max(v.maxlevel, __shfl_xor(v.maxlevel, off, 64));
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        SummaryPartial r = sh[0];
        for (int w = 1; w < kBlock / 64; ++w) {
            const SummaryPartial& o = sh[w];
            r.active += o.active;
            r.inactive += o.inactive;
            r.wsum += o.wsum;
            r.psum += o.psum;
            r.dtsum += o.dtsum;
            r.pmin = fmin(r.pmin, o.pmin);
            r.pmax = fmax(r.pmax, o.pmax);
            r.dtmin = fmin(r.dtmin, o.dtmin);
            r.dtmax = fmax(r.dtmax, o.dtmax);
            r.maxlevel = max(r.maxlevel, o.maxlevel);
        }
        partials[blockIdx.x] = r;
    }
}

__global__ void summary_final_kernel(const SummaryPartial* __restrict__ partials, int rows, uint64_t count,
                                     kpt_summary* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    SummaryPartial r = partials[0];
    for (int k = 1; k < rows; ++k) {
        const SummaryPartial& o = partials[k];
        r.active += o.active;
        r.inactive += o.inactive;
        r.wsum += o.wsum;
        r.psum += o.psum;
        r.dtsum += o.dtsum;
        r.pmin = fmin(r.pmin, o.pmin);
        r.pmax = fmax(r.pmax, o.pmax);
        r.dtmin = fmin(r.dtmin, o.dtmin);
        r.dtmax = fmax(r.dtmax, o.dtmax);
        r.maxlevel = max(r.maxlevel, o.maxlevel);
    }
    kpt_summary s;
    s.stored = count;
    s.active = r.active;
    s.inactive = r.inactive;
    s.weight_sum = r.wsum;
    s.max_split_level = r.maxlevel;
    s.reserved = 0;
    if (r.active > 0) {
        s.p_min = r.pmin;
        s.p_max = r.pmax;
        s.p_avg = r.psum / static_cast<double>(r.active);
        s.dt_min = r.dtmin;
        s.dt_max = r.dtmax;
        s.dt_avg = r.dtsum / static_cast<double>(r.active);
    } else {
        s.p_min = 0.0;
        s.p_max = 0.0;
        s.p_avg = 0.0;
        s.dt_min = 0.0;
        s.dt_max = 0.0;
        s.dt_avg = 0.0;
    }
    *out = s;
}

inline unsigned grid_for(uint64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_pack(const double* raw, const GridDesc& g, int model, double* out64, float* out32,
                       hipStream_t stream) {
    const uint64_t nodes = static_cast<uint64_t>(g.nx + 2) * (g.ny + 2);
    hipLaunchKernelGGL(pack_kernel, dim3(grid_for(nodes)), dim3(kBlock), 0, stream, raw, g, model,
                       coefficient_plane(g.nx, g.ny), out64, out32);
    return hipGetLastError();
}

hipError_t launch_push(const PushArgs& a, int model, bool focused, bool fp32, int closure, hipStream_t stream) {
    if (a.count == 0) return hipSuccess;
    const dim3 grid(grid_for(a.count)), block(kBlock);
    constexpr int R = KPT_CLOSURE_REDUCED, Q = KPT_CLOSURE_QLT;
    if (model == KPT_MODEL_PARKER16) {
        if (fp32)
            hipLaunchKernelGGL((push_kernel<16, false, float, R>), grid, block, 0, stream, a);
        else
            hipLaunchKernelGGL((push_kernel<16, false, double, R>), grid, block, 0, stream, a);
    } else if (focused && closure == KPT_CLOSURE_QLT) {
        if (fp32)
            hipLaunchKernelGGL((push_kernel<21, true, float, Q>), grid, block, 0, stream, a);
        else
            hipLaunchKernelGGL((push_kernel<21, true, double, Q>), grid, block, 0, stream, a);
    } else if (focused) {
        if (closure != R) return hipErrorInvalidValue;
        if (fp32)
            hipLaunchKernelGGL((push_kernel<21, true, float, R>), grid, block, 0, stream, a);
        else
            hipLaunchKernelGGL((push_kernel<21, true, double, R>), grid, block, 0, stream, a);
    } else {
        if (fp32)
            hipLaunchKernelGGL((push_kernel<21, false, float, R>), grid, block, 0, stream, a);
        else
            hipLaunchKernelGGL((push_kernel<21, false, double, R>), grid, block, 0, stream, a);
    }
    return hipGetLastError();
}

hipError_t launch_inject(const ParticleArrays& P, uint64_t first, uint64_t n, uint64_t first_root,
                         const kpt_device_settings& s, const kpt_inject_opts& o, double t0,
                         uint32_t frame, const uint32_t key[2], hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(inject_kernel, dim3(grid_for(n)), dim3(kBlock), 0, stream, P, first, n, first_root,
                       s, o, t0, frame, key[0], key[1]);
    return hipGetLastError();
}

size_t split_scratch_elems(uint64_t count) { return static_cast<size_t>(grid_for(count)) + 1; }

hipError_t launch_split(const ParticleArrays& P, uint64_t count, uint64_t capacity,
                        const kpt_device_settings& s, const SplitTable& tab, uint32_t* scratch,
                        unsigned long long* counters, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const unsigned tiles = grid_for(count);
    hipLaunchKernelGGL(split_count_kernel, dim3(tiles), dim3(kBlock), 0, stream, P, count, s, tab, scratch);
    hipLaunchKernelGGL(split_scan_kernel, dim3(1), dim3(1024), 0, stream, scratch, tiles);
    hipLaunchKernelGGL(split_apply_kernel, dim3(tiles), dim3(kBlock), 0, stream, P, count, capacity, s, tab,
                       scratch, counters);
    return hipGetLastError();
}

size_t hist_partial_elems(int bins) { return static_cast<size_t>(kHistBlocks) * bins; }

hipError_t launch_hist_logp(const ParticleArrays& P, uint64_t count, int bins, double log_pmin,
                            double inv_dlogp, double pmin, double pmax, double* partials, double* out,
                            hipStream_t stream) {
    if (bins <= 0 || bins > kHistMaxBins) return hipErrorInvalidValue;
    hipLaunchKernelGGL(hist_logp_kernel, dim3(kHistBlocks), dim3(kBlock), 0, stream, P, count, bins,
                       log_pmin, inv_dlogp, pmin, pmax, partials);
    hipLaunchKernelGGL(hist_reduce_kernel, dim3((bins + kBlock - 1) / kBlock), dim3(kBlock), 0, stream,
                       partials, kHistBlocks, bins, out);
    return hipGetLastError();
}

hipError_t launch_hist_xy(const ParticleArrays& P, uint64_t count, int bx, int by, double x_lo,
                          double x_hi, double y_lo, double y_hi, double* partials, double* out,
                          hipStream_t stream) {
    const int bins = bx * by;
    if (bx <= 0 || by <= 0 || bins > kHistMaxBins) return hipErrorInvalidValue;
    const double sx = static_cast<double>(bx) / (x_hi - x_lo);
    const double sy = static_cast<double>(by) / (y_hi - y_lo);
    hipLaunchKernelGGL(hist_xy_kernel, dim3(kHistBlocks), dim3(kBlock), 0, stream, P, count, bx, by, x_lo,
                       x_hi, y_lo, y_hi, sx, sy, partials);
    hipLaunchKernelGGL(hist_reduce_kernel, dim3((bins + kBlock - 1) / kBlock), dim3(kBlock), 0, stream,
                       partials, kHistBlocks, bins, out);
    return hipGetLastError();
}

size_t summary_partial_bytes() { return sizeof(SummaryPartial) * kSummaryBlocks; }

hipError_t launch_summary(const ParticleArrays& P, uint64_t count, void* partials, kpt_summary* out_device,
                          hipStream_t stream) {
    hipLaunchKernelGGL(summary_kernel, dim3(kSummaryBlocks), dim3(kBlock), 0, stream, P, count,
                       static_cast<SummaryPartial*>(partials));
    hipLaunchKernelGGL(summary_final_kernel, dim3(1), dim3(64), 0, stream,
                       static_cast<const SummaryPartial*>(partials), kSummaryBlocks, count, out_device);
    return hipGetLastError();
}

}  // namespace kpt
