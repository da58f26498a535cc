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

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>

// Multiply-adds are fused only inside one source expression, never across
// statements: the push's rounding then does not depend on the kernel that
// inlines it (a run pushed alone and inside the batched ensemble agree bit for
// bit).
#pragma clang fp contract(on)

#include "kpt_device_common.h"
#include "kpt_internal.h"
#include "kpt_mlp_device.h"
#include "kpt_grid_device.h"

namespace kpt {
namespace {

constexpr int kBlock = 256;

constexpr double kSqrt3 = 1.7320508075688772935;

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
    // Derivative along an axis from the values at offsets -1, 0, +1: the
    // driver's central difference on a uniform axis, the reference's three-point
    // formula at the node coordinates on a stretched one.
    auto along = [&](int d, double fl, double fc, double fu) -> double {
        const int n = d == 0 ? g.nx : g.ny;
        if (n <= 1) return 0.0;
        if (g.face[d] == nullptr) return (fu - fl) / (d == 0 ? hx : hy);
        const int c = d == 0 ? i : j;
        return three_point(stretched_node(g, d, n, c - 1), stretched_node(g, d, n, c), stretched_node(g, d, n, c + 1),
                           fl, fc, fu);
    };
    // comp: 0..2 B, 3..5 V
    auto ddx = [&](int comp) -> double {
        return along(0, val(comp, i - 1, j), val(comp, i, j), val(comp, i + 1, j));
    };
    auto ddy = [&](int comp) -> double {
        return along(1, val(comp, i, j - 1), val(comp, i, j), val(comp, i, j + 1));
    };
    const double bx = val(0, i, j), by = val(1, i, j), bz = val(2, i, j);
    double c[24];
#pragma unroll
    for (int k = 0; k < 24; ++k) c[k] = 0.0;
    const double dbm_dx = along(0, bmag(i - 1, j), bmag(i, j), bmag(i + 1, j));
    const double dbm_dy = along(1, bmag(i, j - 1), bmag(i, j), bmag(i, j + 1));
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


// ---------------------------------------------------------------------------
// Arithmetic of one sub-step, templated on its type RT. The double overloads
// are the plain library operations; the float ones use the native v_rcp_f32,
// v_sqrt_f32, v_log_f32 and v_exp_f32 instructions (about 1 ulp each).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T sqr(T v) { return v * v; }
__device__ __forceinline__ double rabs(double v) { return fabs(v); }
__device__ __forceinline__ float rabs(float v) { return fabsf(v); }
__device__ __forceinline__ double rmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float rmin(float a, float b) { return fminf(a, b); }
// rt_sqrt is the square root (not its reciprocal) in the step's type.
__device__ __forceinline__ double rt_sqrt(double v) { return sqrt(v); }
__device__ __forceinline__ float rt_sqrt(float v) { return __builtin_amdgcn_sqrtf(v); }
__device__ __forceinline__ double rrcp(double v) { return 1.0 / v; }
__device__ __forceinline__ float rrcp(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ double rdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float rdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double rpow(double a, double e) { return pow(a, e); }
__device__ __forceinline__ float rpow(float a, float e) {
    // a^e = 2^(e log2 a) for normal positive a; anything else takes powf.
    return a >= 1.17549435e-38f ? __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(a)) : powf(a, e);
}
template <typename RT>
__device__ __forceinline__ bool rfinite(RT v) { return __builtin_isfinite(v); }

// One U[-sqrt3, sqrt3] draw from 32 Philox bits. Double: (r + 1/2) 2^-32 as the
// injected-noise oracle. Float: the top 23 bits, (2k + 1) 2^-24, exact in float.
template <typename RT>
__device__ __forceinline__ RT draw_sqrt3(uint32_t r);
template <>
__device__ __forceinline__ double draw_sqrt3<double>(uint32_t r) {
    return -kSqrt3 + (kSqrt3 - -kSqrt3) * u01(r);
}
template <>
__device__ __forceinline__ float draw_sqrt3<float>(uint32_t r) {
    const float u = static_cast<float>(r >> 9) * 0x1.0p-23f + 0x1.0p-24f;
    return fmaf(u, static_cast<float>(2.0 * kSqrt3), static_cast<float>(-kSqrt3));
}

// Slab-QLT D_mumu and dD_mumu/dmu (include/LegencyModel.hpp:557-643) for one
// trial mu, with omega, xi = omega Lc / v and xi^gamma_k fixed for the sub-step.
// `pref` = (pi / 4) omega sigma2 gamma_k sin(pi / gamma_k) / pi.
template <typename RT>
__device__ __forceinline__ void qlt_scattering(RT mu, RT xi, RT xi_g, RT pref, RT gk, RT h0, RT& D, RT& dD) {
    const RT one_minus_mu2 = RT(1) - mu * mu;
    D = RT(0);
    dD = RT(0);
    if (one_minus_mu2 <= RT(0)) return;
    const RT mu_h = rabs(mu) + h0;
    const RT pm1 = rpow(mu_h, gk - RT(1));  // mu_h^(gamma_k - 1)
    const RT den = pm1 * mu_h + xi_g;       // mu_h^gamma_k + xi^gamma_k
    D = rdiv(pref * one_minus_mu2 * xi * pm1, den);
    if (D == RT(0) || mu == RT(0)) return;
    const RT sign_mu = mu > RT(0) ? RT(1) : RT(-1);
    const RT shape = sign_mu * (rdiv(gk - RT(1), mu_h) - rdiv(gk * pm1, den));
    dD = D * (rdiv(RT(-2) * mu, one_minus_mu2) + shape);
}

// p / p0: double division in the FP64 path, a multiply by 1 / p0 (formed on
// the host) in the FP32 path.
template <typename RT>
__device__ __forceinline__ RT p_ratio(double p, double p0, double inv_p0) {
    if constexpr (sizeof(RT) == 4)
        return static_cast<RT>(p * inv_p0);
    else
        return p / p0;
}

// Terms of one sub-step at the particle's sampled coefficients. With the QLT
// and learned closures, dmu_dt holds only the focusing rate and duu is zero on
// return from eval_terms; the caller adds the closure's D and dD/dmu.
template <typename RT>
struct StepTerms {
    RT dx_dt, dy_dt, dmu_dt, duu;
    double dp_dt;  // adiabatic rate, formed in double in every precision
    RT skpara, skperp, skpara_perp;
    RT bx, by, bz, ib;
    // focusing pieces and QLT constants reused by the Picard passes
    RT f_speed, f_divv, f_bb, f_bv, f_divb, q_xi, q_xig, q_pref;
};

template <typename D, typename S>
__device__ __forceinline__ void narrow_terms(const StepTerms<S>& a, StepTerms<D>& o) {
    o.dx_dt = D(a.dx_dt); o.dy_dt = D(a.dy_dt); o.dmu_dt = D(a.dmu_dt); o.duu = D(a.duu);
    o.dp_dt = a.dp_dt;
    o.skpara = D(a.skpara); o.skperp = D(a.skperp); o.skpara_perp = D(a.skpara_perp);
    o.bx = D(a.bx); o.by = D(a.by); o.bz = D(a.bz); o.ib = D(a.ib);
    o.f_speed = D(a.f_speed); o.f_divv = D(a.f_divv); o.f_bb = D(a.f_bb); o.f_bv = D(a.f_bv);
    o.f_divb = D(a.f_divb); o.q_xi = D(a.q_xi); o.q_xig = D(a.q_xig); o.q_pref = D(a.q_pref);
}

// Rates of the mover at one sample: kappa model (src/main.cpp:1177-1199),
// divergence of the diffusion tensor (:1210-1223), Parker drift and rates
// (:1225-1241) or the focused branch (:1499-1589) with the reduced or the
// slab-QLT pitch-angle closure. Returns false where the reference marks the
// particle failed (status 2).
template <int MODEL, bool FOCUSED, bool QLT, typename RT>
__device__ __forceinline__ bool eval_terms(const RT* c, double p_d, double mu_d, const kpt_device_settings& s,
                                           const StepConst<RT>& k, StepTerms<RT>& o) {
    RT vx, vy, bx, by, bz, b, dvx_dx, dvx_dy = RT(0), dvy_dx = RT(0), dvy_dy, dvz_dx = RT(0), dvz_dy = RT(0),
                                  dbx_dx, dbx_dy, dby_dx, dby_dy, dbz_dx, dbz_dy, db_dx, db_dy;
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
    (void)dvz_dx;
    if (!rfinite(p_d) || p_d <= 0.0 || !rfinite(b) || b <= RT(0)) return false;
    const RT p = static_cast<RT>(p_d);
    const RT ib = rrcp(b);
    const RT ib2 = ib * ib;
    const RT ib3 = ib2 * ib;

    RT knorm = RT(1);
    if (s.magnetic_dependency == 1) knorm *= rpow(b, k.exp_b_kappa);
    if (s.momentum_dependency == 1) knorm *= rpow(p_ratio<RT>(p_d, s.p0, k.inv_p0), k.pindex);
    const RT kpara = k.kpara0 * knorm;
    const RT kperp = kpara * k.kperp_ratio;
    if (!rfinite(kpara) || kpara <= RT(0) || !rfinite(kperp) || kperp < RT(0)) return false;
    o.skpara = rt_sqrt(RT(2) * kpara);
    o.skperp = rt_sqrt(RT(2) * kperp);
    o.skpara_perp = MODEL == 16 ? rt_sqrt(RT(2) * (kpara - kperp))
                                : (kpara > kperp ? rt_sqrt(RT(2) * (kpara - kperp)) : RT(0));
    const RT kpp = FOCUSED ? -kperp : kpara - kperp;

    RT dkdx = RT(0), dkdy = RT(0);
    if (s.magnetic_dependency == 1) {
        const RT me = k.exp_b_kappa;
        dkdx = db_dx * ib * me;
        dkdy = db_dy * ib * me;
    }
    const RT dkxx_dx = kperp * dkdx + kpp * dkdx * bx * bx * ib2 +
                       RT(2) * kpp * bx * (dbx_dx * b - bx * db_dx) * ib3;
    const RT dkyy_dy = kperp * dkdy + kpp * dkdy * by * by * ib2 +
                       RT(2) * kpp * by * (dby_dy * b - by * db_dy) * ib3;
    const RT dkxy_dx = kpp * dkdx * bx * by * ib2 +
                       kpp * ((dbx_dx * by + bx * dby_dx) * ib2 - RT(2) * bx * by * db_dx * ib3);
    const RT dkxy_dy = kpp * dkdy * bx * by * ib2 +
                       kpp * ((dbx_dy * by + bx * dby_dy) * ib2 - RT(2) * bx * by * db_dy * ib3);
    const RT drift_den = rt_sqrt(sqr(rdiv(k.drift_c1, p)) + sqr(rdiv(k.drift_c2, p * p)));

    o.bx = bx;
    o.by = by;
    o.bz = bz;
    o.ib = ib;
    o.dmu_dt = RT(0);
    o.duu = RT(0);
    o.f_speed = o.f_divv = o.f_bb = o.f_bv = o.f_divb = o.q_xi = o.q_xig = o.q_pref = RT(0);
    if constexpr (FOCUSED) {
        const RT mu = static_cast<RT>(mu_d);
        const RT mu2 = mu * mu;
        const RT muf1 = RT(0.5) * (RT(1) - mu2);
        const RT muf2 = RT(0.5) * (RT(3) * mu2 - RT(1));
        const RT kx = bx * dbx_dx + by * dbx_dy;
        const RT ky = bx * dby_dx + by * dby_dy;
        const RT kz = bx * dbz_dx + by * dbz_dy;
        const RT bdot_curvb = bx * dbz_dy - by * dbz_dx + bz * (dby_dx - dbx_dy);
        const RT vdp = rdiv(k.inv_q, drift_den);
        const RT vdx = vdp * (muf1 * (-bz * db_dy) * ib2 + mu2 * (by * kz - bz * ky) * ib3 +
                              muf1 * bx * bdot_curvb * ib3);
        const RT vdy = vdp * (muf1 * (bz * db_dx) * ib2 + mu2 * (bz * kx - bx * kz) * ib3 +
                              muf1 * by * bdot_curvb * ib3);
        // QLT: relativistic speed p / gamma (m = c = 1)
        const RT gamma_p = QLT ? rt_sqrt(RT(1) + p * p) : RT(1);
        const RT speed = QLT ? rdiv(p, gamma_p) : rdiv(k.v0 * p, k.p0);
        if (!rfinite(speed) || speed <= RT(0)) return false;
        const RT vpb = speed * mu * ib;
        o.dx_dt = vx + vdx + vpb * bx + dkxx_dx + dkxy_dy;
        o.dy_dt = vy + vdy + vpb * by + dkxy_dx + dkyy_dy;
        const RT divv = dvx_dx + dvy_dy;
        const RT bb_gradv = (bx * (bx * dvx_dx + by * dvx_dy) + by * (bx * dvy_dx + by * dvy_dy) +
                             bz * (bx * dvz_dx + by * dvz_dy)) * ib2;
        const RT bv_gradv = (bx * (vx * dvx_dx + vy * dvx_dy) + by * (vx * dvy_dx + vy * dvy_dy) +
                             bz * (vx * dvz_dx + vy * dvz_dy)) * ib;
        const RT bv_over_v = rdiv(bv_gradv, speed);
        o.dp_dt = p_d * -static_cast<double>(muf1 * divv + muf2 * bb_gradv + mu * bv_over_v);
        const RT div_bnorm = -(bx * db_dx + by * db_dy) * ib2;
        RT dmu_dt = speed * div_bnorm + mu * divv - RT(3) * mu * bb_gradv - RT(2) * bv_over_v;
        dmu_dt *= RT(0.5) * (RT(1) - mu2);
        RT duu, duu_du = RT(0);
        if constexpr (QLT) {
            o.f_speed = speed;
            o.f_divv = divv;
            o.f_bb = bb_gradv;
            o.f_bv = bv_gradv;
            o.f_divb = div_bnorm;
            const RT omega = rdiv(k.abs_q * b, gamma_p);
            o.q_xi = rdiv(omega * k.qlt_lc, speed);
            o.q_xig = rpow(o.q_xi, k.qlt_gk);
            o.q_pref = omega * k.qlt_pref;
            // the closure (exact or learned) is evaluated by the caller
            duu = RT(0);
        } else {
            const RT abs_mu = rabs(mu);
            const RT dtmp = rpow(abs_mu, k.exp_mu1) + k.duu_h0;
            duu = k.duu0 * (RT(1) - mu2) * dtmp;
            if (mu != RT(0)) {
                const RT g2 = (RT(1) - mu2) * rpow(abs_mu, k.exp_mu2);
                duu_du = k.duu0 * (RT(-2) * mu * dtmp + (mu > RT(0) ? g2 : -g2));
            }
            RT duu_norm = RT(1);
            if (s.magnetic_dependency == 1) duu_norm *= rpow(b, k.exp_b_duu);
            if (s.momentum_dependency == 1) duu_norm *= rpow(p_ratio<RT>(p_d, s.p0, k.inv_p0), k.exp_mu1);
            duu *= duu_norm;
            duu_du *= duu_norm;
        }
        dmu_dt += duu_du;
        if (!rfinite(dmu_dt) || !rfinite(duu) || duu < RT(0)) return false;
        o.dmu_dt = dmu_dt;
        o.duu = duu;
    } else {
        const RT vdp = rdiv(k.inv_3q, drift_den);
        const RT vdx = vdp * (dbz_dy * ib2 - RT(2) * bz * db_dy * ib3);
        const RT vdy = vdp * (-dbz_dx * ib2 + RT(2) * bz * db_dx * ib3);
        o.dx_dt = vx + vdx + dkxx_dx + dkxy_dy;
        o.dy_dt = vy + vdy + dkxy_dx + dkyy_dy;
        if constexpr (sizeof(RT) == 4)
            o.dp_dt = -p_d * static_cast<double>((dvx_dx + dvy_dy) * RT(1.0 / 3.0));
        else
            o.dp_dt = -p_d * static_cast<double>(dvx_dx + dvy_dy) / 3.0;
    }
    return true;
}

template <typename RT>
__device__ __forceinline__ const StepConst<RT>& step_const(const PushArgs& a);
template <>
__device__ __forceinline__ const StepConst<double>& step_const<double>(const PushArgs& a) { return a.k64; }
template <>
__device__ __forceinline__ const StepConst<float>& step_const<float>(const PushArgs& a) { return a.k32; }

// Adaptive step candidates (src/main.cpp:1243-1277, :1603-1665) before the
// [dt_min, dt_max] clamp. The momentum candidate is formed in double.
template <bool FOCUSED, typename RT>
__device__ __forceinline__ double candidate_dt(const StepTerms<RT>& T, double p, const kpt_device_settings& s,
                                               const StepConst<RT>& k, RT half_dx, RT half_dy) {
    if (FOCUSED) {
        const RT ds = T.skperp > RT(0) ? T.skperp : T.skpara;
        if (T.dx_dt != RT(0) && T.dy_dt != RT(0) && T.dp_dt != 0.0 && T.dmu_dt != RT(0) && ds > RT(0)) {
            RT d = rmin(sqr(rdiv(half_dx, ds)), sqr(rdiv(half_dy, ds)));
            d = rmin(d, sqr(rdiv(ds, T.dx_dt)));
            d = rmin(d, sqr(rdiv(ds, T.dy_dt)));
            d = rmin(d, rdiv(RT(0.1), rabs(T.dmu_dt)));
            d = rmin(d, rdiv(RT(2) * T.duu, sqr(T.dmu_dt)));
            return fmin(static_cast<double>(d), 0.1 * p / fabs(T.dp_dt));
        }
        return s.dt_min;
    }
    if (T.dx_dt != RT(0) && T.dy_dt != RT(0) && T.dp_dt != 0.0) {
        const RT as = T.skperp > RT(0) ? T.skperp : T.skpara;
        RT d = rmin(sqr(rdiv(half_dx, T.skpara)), sqr(rdiv(half_dy, T.skpara)));
        d = rmin(d, sqr(rdiv(as, T.dx_dt)));
        d = rmin(d, sqr(rdiv(as, T.dy_dt)));
        return fmin(static_cast<double>(d), 0.1 * p / fabs(T.dp_dt));
    }
    return s.dt_min;
}

// Bilinear weights, both frames and the time blend of one sample.
template <int C, typename T>
__device__ __forceinline__ void sample_blend(const T* __restrict__ flo, const T* __restrict__ fhi, bool blend,
                                             double uw, int64_t n00, int nxp, uint64_t plane, double wx,
                                             double wy, T* c) {
    const T w00 = static_cast<T>((1.0 - wx) * (1.0 - wy));
    const T w10 = static_cast<T>(wx * (1.0 - wy));
    const T w01 = static_cast<T>((1.0 - wx) * wy);
    const T w11 = static_cast<T>(wx * wy);
    sample_corners<C, T>(flo, n00, nxp, plane, w00, w10, w01, w11, c);
    if (blend) {
        T hi[C];
        sample_corners<C, T>(fhi, n00, nxp, plane, w00, w10, w01, w11, hi);
        const double lw = 1.0 - uw;
        const T uwt = static_cast<T>(uw), lwt = static_cast<T>(lw);
#pragma unroll
        for (int k = 0; k < C; ++k) c[k] = lwt * c[k] + uwt * hi[k];
    }
}

// Records of the node-major pair whose loads are issued together (0: all): the
// exact-closure and Parker pushes, and the learned-closure push. Measured on the
// 1024 x 512 stretched island field, 8 M focused particles, per launch: FP32
// exact 68.6 ms unphased at 4 waves, 44.7 ms at one record per phase (4 waves),
// 42.3 ms at 3 waves; learned closure (SiLU / hard-swish) 75.1 / 57.9 ms
// unphased, 74.6 / 57.3 ms at 3 records, 74.4 / 56.1 ms at 2.
#ifndef KPT_PAIR_RECORDS_EXACT
#define KPT_PAIR_RECORDS_EXACT 1
#endif
#ifndef KPT_PAIR_RECORDS_MLP
#define KPT_PAIR_RECORDS_MLP 2
#endif
// FP32 time-blended sample from the node-major pair buffer: node n's record of
// pair_stride(C) floats starts at pair + n pair_stride(C) (16-byte aligned) and
// holds (lower, upper) of stored slot s at floats 2s, 2s + 1 (slot s is packed
// component pair_component(C, s); components the pair does not store are left
// unset). A stencil row is two adjacent records, read as 16-byte loads at
// immediate offsets from one address; the bilinear and time weights are folded
// into eight products per component.
template <int C, int RPP = 0>
__device__ __forceinline__ void sample_pair32(const float* __restrict__ pair, int64_t n00, int nxp, double uw,
                                              double wx, double wy, float* c) {
    constexpr int V = pair_stride(C) / 4;  // float4 per record
    constexpr int S = pair_components(C);
    // bilinear and time weights in float (the weights lie in [0, 1])
    const float fx = static_cast<float>(wx), fy = static_cast<float>(wy), fu = static_cast<float>(uw);
    const float gx = 1.0f - fx, gy = 1.0f - fy, gu = 1.0f - fu;
    const float w00 = gx * gy, w10 = fx * gy, w01 = gx * fy, w11 = fx * fy;
    const float a00 = w00 * gu, b00 = w00 * fu, a10 = w10 * gu, b10 = w10 * fu;
    const float a01 = w01 * gu, b01 = w01 * fu, a11 = w11 * gu, b11 = w11 * fu;
    const float4* __restrict__ q = reinterpret_cast<const float4*>(pair) + n00 * V;
    const float4* __restrict__ r = q + static_cast<int64_t>(nxp) * V;
#pragma unroll
    for (int v = 0; v < V; ++v) {
        // RPP > 0: the loads of RPP records at a time. Without the barrier the
        // scheduler clusters all 4 V loads of the stencil (same base, immediate
        // offsets) ahead of their products, which spilled to scratch in the
        // 3- and 4-wave FP32 exact push.
        if (RPP > 0 && v > 0 && v % RPP == 0) __builtin_amdgcn_sched_barrier(0);
        const float4 u0 = q[v], u1 = q[V + v], w0 = r[v], w1 = r[V + v];
        const int s0 = 2 * v;
        if (s0 < S)
            c[pair_component(C, s0)] = a00 * u0.x + b00 * u0.y + a10 * u1.x + b10 * u1.y + a01 * w0.x + b01 * w0.y +
                                       a11 * w1.x + b11 * w1.y;
        if (s0 + 1 < S)
            c[pair_component(C, s0 + 1)] = a00 * u0.z + b00 * u0.w + a10 * u1.z + b10 * u1.w + a01 * w0.z +
                                           b01 * w0.w + a11 * w1.z + b11 * w1.w;
    }
    if constexpr (C == 21) c[2] = 0.0f;  // Vz: not stored
}

// Bilinear sample of one FP32 frame from node-major records (rec_stride(C)
// floats of stored slots): the weights and the sum of sample_corners over the
// component-planar frame, so the result is the same bit for bit. q points at the
// lower-left corner's record; records are SV 16-byte units apart and the upper
// row's records start `row` records on.
template <int C, int SV>
__device__ __forceinline__ void sample_rec32(const float4* q, int64_t row, double wx, double wy, float* c) {
    constexpr int S = pair_components(C);
    const float w00 = static_cast<float>((1.0 - wx) * (1.0 - wy));
    const float w10 = static_cast<float>(wx * (1.0 - wy));
    const float w01 = static_cast<float>((1.0 - wx) * wy);
    const float w11 = static_cast<float>(wx * wy);
    const float4* r = q + row * SV;
#pragma unroll
    for (int v = 0; v < rec_stride(C) / 4; ++v) {
        const float4 u0 = q[v], u1 = q[SV + v], w0 = r[v], w1 = r[SV + v];
        const float a0[4] = {u0.x, u0.y, u0.z, u0.w}, a1[4] = {u1.x, u1.y, u1.z, u1.w};
        const float b0[4] = {w0.x, w0.y, w0.z, w0.w}, b1[4] = {w1.x, w1.y, w1.z, w1.w};
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (4 * v + i < S) c[pair_component(C, 4 * v + i)] = w00 * a0[i] + w10 * a1[i] + w01 * b0[i] + w11 * b1[i];
    }
    if constexpr (C == 21) c[2] = 0.0f;  // Vz: not stored
}

// Staged records are stage_stride(V) 16-byte units apart in LDS (V units of
// data). The FP64 Parker record (16 units, 256 B) starts every node on the same
// LDS banks; one unit of padding spreads them (FP64 Parker push 62.3 -> 52.9 ms
// per launch). Padding the other records measured no gain (FP32 Parker, 8
// units: 21.8 -> 23.0 ms from the smaller box; FP32 and FP64 focused, 10 and 21
// units: within 1%).
__host__ __device__ constexpr int stage_stride(int v) { return v % 16 == 0 ? v + 1 : v; }

// The same time-blended sample from records staged in LDS: q points at the
// lower-left corner's record, the upper row's records start `row` records on.
template <int C>
__device__ __forceinline__ void sample_staged(const float4* q, int row, double uw, double wx, double wy, float* c) {
    constexpr int V = stage_stride(pair_stride(C) / 4);  // record stride in LDS
    constexpr int S = pair_components(C);
    const float fx = static_cast<float>(wx), fy = static_cast<float>(wy), fu = static_cast<float>(uw);
    const float gx = 1.0f - fx, gy = 1.0f - fy, gu = 1.0f - fu;
    const float w00 = gx * gy, w10 = fx * gy, w01 = gx * fy, w11 = fx * fy;
    const float a00 = w00 * gu, b00 = w00 * fu, a10 = w10 * gu, b10 = w10 * fu;
    const float a01 = w01 * gu, b01 = w01 * fu, a11 = w11 * gu, b11 = w11 * fu;
    const float4* r = q + row * V;
#pragma unroll
    for (int v = 0; v < pair_stride(C) / 4; ++v) {
        const float4 u0 = q[v], u1 = q[V + v], w0 = r[v], w1 = r[V + v];
        const int s0 = 2 * v;
        if (s0 < S)
            c[pair_component(C, s0)] = a00 * u0.x + b00 * u0.y + a10 * u1.x + b10 * u1.y + a01 * w0.x + b01 * w0.y +
                                       a11 * w1.x + b11 * w1.y;
        if (s0 + 1 < S)
            c[pair_component(C, s0 + 1)] = a00 * u0.z + b00 * u0.w + a10 * u1.z + b10 * u1.w + a01 * w0.z +
                                           b01 * w0.w + a11 * w1.z + b11 * w1.w;
    }
    if constexpr (C == 21) c[2] = 0.0f;  // Vz: not stored
}

// FP64 time-blended sample from node-major (lower, upper) records: the same
// weights and the same sum per frame as sample_blend over the component-planar
// frames, so the result is the same bit for bit. q points at the lower-left
// corner's record; records are SV double2 apart (C in the pair in global
// memory, stage_stride(C) in a staged box) and the upper row's records start
// `row` records on (nx + 2 in global memory, the box width in a staged box).
template <int C, int SV = C>
__device__ __forceinline__ void sample_pair64(const double2* q, int64_t row, double uw, double wx, double wy,
                                              double* c) {
    const double w00 = (1.0 - wx) * (1.0 - wy);
    const double w10 = wx * (1.0 - wy);
    const double w01 = (1.0 - wx) * wy;
    const double w11 = wx * wy;
    const double lw = 1.0 - uw;
    const double2* r = q + row * SV;
#pragma unroll
    for (int k = 0; k < C; ++k) {
        const double2 u0 = q[k], u1 = q[SV + k], v0 = r[k], v1 = r[SV + k];
        const double lo = w00 * u0.x + w10 * u1.x + w01 * v0.x + w11 * v1.x;
        const double hi = w00 * u0.y + w10 * u1.y + w01 * v0.y + w11 * v1.y;
        c[k] = lw * lo + uw * hi;
    }
}

// ---------------------------------------------------------------------------
// Wavefront staging of the frame pair in LDS. The vector memory path (the
// texture address and data units) bounds the push when every lane loads
// its four stencil records from the L1/L2 caches on every sub-step, although
// the cell-sorted lanes of a wavefront share most of those nodes. Instead each
// wavefront copies the records of a node box around its lanes' stencils into
// its own LDS region (contiguous rows of the node-major pair, so the copy is
// coalesced); lanes whose stencil lies in the box sample from LDS, the others
// from the pair in global memory. Particles move a fraction of a cell per
// sub-step, so a box serves many sub-steps:
//  - when the bounding box of the lanes' stencils, widened by a margin where
//    it fits, holds at most CAPN records, the box covers every lane and is
//    rebuilt as soon as a lane leaves it;
//  - otherwise (a dispersed wavefront) the box is the CAPN-record window with
//    the bounding box's aspect, centred on the lanes' mean node; it is rebuilt
//    only when KPT_STAGE_OUTLIERS more lanes are outside than right after the
//    last rebuild.
// ---------------------------------------------------------------------------
// LDS for staged records per CU (0: no staging), shared by the wavefronts the
// kernel's occupancy puts on a CU: at 4 waves per SIMD, 16 wavefronts of
// 9 600 B (60 FP32 focused records of 160 B); at 2 (FP64), 8 of 19 200 B (57
// focused records of 336 B, 70 Parker records at their LDS stride of 272 B).
#ifndef KPT_STAGE_CU_BYTES
#define KPT_STAGE_CU_BYTES 153600
#endif
#ifndef KPT_STAGE_MARGIN
#define KPT_STAGE_MARGIN 2
#endif
#ifndef KPT_STAGE_OUTLIERS
#define KPT_STAGE_OUTLIERS 16
#endif
struct StageBox {
    int i_lo, j_lo, w, h;  // staged nodes (i_lo + u, j_lo + v), u < w, v < h; w = 0: none
    int out0;              // lanes outside the box after its build (-1: none; 0: the box covers all)
    __device__ __forceinline__ bool holds(int i0, int j0) const {
        return i0 >= i_lo && i0 + 1 < i_lo + w && j0 >= j_lo && j0 + 1 < j_lo + h;
    }
};

__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int wave_sum_i32(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int wave_count(bool b) { return __builtin_popcountll(__ballot(b)); }

// Keeps `box` around the stencils (i0 .. i0 + 1, j0 .. j0 + 1) of the lanes with
// `run` set (rules above). Every lane of the wavefront calls it together; `lds`
// is the wavefront's region of CAPN records of V 16-byte units (the FP32 or
// FP64 pair's records). Node indices lie in [-1, nx] x [-1, ny] (record
// (j + 1) (nx + 2) + i + 1).
template <int V, int CAPN>
__device__ __forceinline__ void stage_update(StageBox& box, float4* lds, const void* __restrict__ pair, int nx,
                                             int ny, bool run, int i0, int j0) {
    static_assert(CAPN >= 4, "a staged box holds at least one cell's four nodes");
    const int n_out = wave_count(run && !box.holds(i0, j0));
    if (n_out == 0 || (box.out0 > 0 && n_out <= box.out0 + KPT_STAGE_OUTLIERS)) return;
    const int imin = wave_min_i32(run ? i0 : INT_MAX), imax = wave_max_i32(run ? i0 + 1 : INT_MIN);
    const int jmin = wave_min_i32(run ? j0 : INT_MAX), jmax = wave_max_i32(run ? j0 + 1 : INT_MIN);
    int i_lo = 0, j_lo = 0, w = 0, h = 0;
    for (int m = KPT_STAGE_MARGIN; m >= 0; --m) {
        i_lo = max(imin - m, -1);
        j_lo = max(jmin - m, -1);
        w = min(imax + m, nx) - i_lo + 1;
        h = min(jmax + m, ny) - j_lo + 1;
        if (w * h <= CAPN) break;
    }
    bool full = w * h <= CAPN;
    if (!full) {
        // the CAPN-record window with the bounding box's aspect around the mean node
        const int W = imax - imin + 1, H = jmax - jmin + 1;  // both >= 2
        int ww = static_cast<int>(sqrtf(static_cast<float>(CAPN) * static_cast<float>(W) / static_cast<float>(H)));
        ww = max(2, min(ww, min(W, CAPN / 2)));
        const int hh = max(2, min(H, CAPN / ww));
        const int nrun = wave_count(run);
        const int ci = wave_sum_i32(run ? i0 : 0) / nrun, cj = wave_sum_i32(run ? j0 : 0) / nrun;
        i_lo = max(imin, min(ci - (ww - 1) / 2, imax - ww + 1));
        j_lo = max(jmin, min(cj - (hh - 1) / 2, jmax - hh + 1));
        w = ww;
        h = hh;
    }
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int nxp = nx + 2;
    const float4* src = static_cast<const float4*>(pair);
    for (int r = 0; r < h; ++r) {
        const float4* row = src + (static_cast<int64_t>(j_lo + r + 1) * nxp + (i_lo + 1)) * V;
        float4* dst = lds + r * w * stage_stride(V);
        for (int e = lane; e < w * V; e += 64) {
            const int node = e / V;
            dst[node * stage_stride(V) + (e - node * V)] = row[e];
        }
    }
    box.i_lo = i_lo;
    box.j_lo = j_lo;
    box.w = w;
    box.h = h;
    box.out0 = full ? 0 : wave_count(run && !box.holds(i0, j0));
}

// The same sample from the FP64 frames for the rare near-null lanes of the FP32
// push: four components at a time, so this branch does not set the register
// budget of the whole kernel.
template <int C>
__device__ __forceinline__ void sample_blend_fp64_grouped(const double* __restrict__ flo,
                                                          const double* __restrict__ fhi, bool blend,
                                                          double uw, int64_t n00, int nxp, uint64_t plane,
                                                          double wx, double wy, double* c) {
    const double w00 = (1.0 - wx) * (1.0 - wy), w10 = wx * (1.0 - wy);
    const double w01 = (1.0 - wx) * wy, w11 = wx * wy;
    const double lw = 1.0 - uw;
#pragma unroll
    for (int k = 0; k < C; ++k) {
        const double* q = flo + k * plane + n00;
        double v = w00 * q[0] + w10 * q[1] + w01 * q[nxp] + w11 * q[nxp + 1];
        if (blend) {
            const double* h = fhi + k * plane + n00;
            v = lw * v + uw * (w00 * h[0] + w10 * h[1] + w01 * h[nxp] + w11 * h[nxp + 1]);
        }
        c[k] = v;
        if ((k & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
}

// The sub-step of a near-null lane of the FP32 push in FP64: the FP64 frames
// sampled four components at a time, the rates in FP64 arithmetic, narrowed to
// float. Not inlined: its 21-double sample would otherwise set the register
// budget of the whole FP32 sub-step loop, which spilled to scratch on every
// sub-step for a branch a few lanes take; a call keeps those registers and the
// spills around it on the rare path. CALLER: one instance per calling kernel, so
// that each is compiled for its caller's register budget (a shared instance
// takes the smallest, which cost the other caller 10% per launch).
template <int MODEL, bool FOCUSED, bool QLT, int CALLER = 0>
__device__ __attribute__((noinline)) bool near_null_terms(const PushArgs* a, bool blend, double uw, int64_t n00,
                                                         int nxp, double wx, double wy, double p, double mu,
                                                         StepTerms<float>* out) {
    double cd[MODEL];
    sample_blend_fp64_grouped<MODEL>(a->field_lo64, a->field_hi64, blend, uw, n00, nxp, a->plane, wx, wy, cd);
    StepTerms<double> Td;
    const bool ok = eval_terms<MODEL, FOCUSED, QLT, double>(cd, p, mu, a->s, a->k64, Td);
    narrow_terms(Td, *out);
    return ok;
}

// Push arguments in the constant address space: the kernel's own argument block
// (kernel_args) or a run of the batched ensemble (as_args), which no kernel
// writes while a push reads it. Fields read through such a pointer load with
// scalar loads. args_reload hides the pointer's value from the compiler, so
// the fields a sub-step uses are loaded inside that sub-step rather than kept
// live across the loop: the push uses more uniform values than a wave has
// SGPRs, and values kept live were spilled to VGPR lanes and read back with
// v_readlane (VALU instructions) on every sub-step.
typedef const __attribute__((address_space(4))) PushArgs* ArgsK;

// The push arguments are the first explicit kernel argument (offset 0 of the
// kernel argument segment).
__device__ __forceinline__ ArgsK kernel_args() {
    return (ArgsK)(__builtin_amdgcn_kernarg_segment_ptr());
}
__device__ __forceinline__ ArgsK as_args(const PushArgs& a) { return (ArgsK)(&a); }
__device__ __forceinline__ const PushArgs& args_reload(ArgsK p) {
    asm volatile("" : "+s"(p));
    return *(const PushArgs*)(p);
}

// ---------------------------------------------------------------------------
// The push: one lane advances one particle through the whole interval
// (move_parker_particles_one_interval, src/main.cpp:1113-1325, and the shared
// mover move_particles_one_interval, src/main.cpp:1364-1753).
// MODEL: 16 or 21 packed components. FOCUSED: focused branch. FT: storage,
// interpolation and coefficient arithmetic type (x, y, t, p and mu are carried in
// double in both). CLOSURE: pitch-angle closure of the focused branch.
// ---------------------------------------------------------------------------
// Waves per SIMD of the push. The unified VGPR/AGPR file holds 512 registers
// per lane: 2 waves get 256 each and keep every corner load of a sub-step in
// flight without spills; 3 and 4 waves spill registers to scratch but hide more
// gather latency. Measured per launch on the 1024 x 512 stretched island field
// (8 M focused / 4 M Parker particles), against 2 waves:
//   FP64 focused: 3 waves +63%, 4 waves +106%;   FP64 Parker: 3 waves -1%, 4 +47%;
//   FP32 focused: 3 waves -20%, 4 waves -29%, 5 waves +1%;
//   FP32 Parker: 3 waves -20%, 4 waves -16%.
// With the node-major FP32 pair (one record per load phase), FP32 focused: 3
// waves 42.3 ms, 4 waves 44.7 ms per launch. With records staged in LDS (8 M
// focused particles, exact closure): 3 waves 28.3 ms, 4 waves 25.3 ms, 5 waves
// 28.1 ms per launch; FP32 Parker (4 M particles, adaptive dt): 3 waves 23.5 ms
// (44.3 ms unstaged), 4 waves 21.7 ms, 5 waves 26.9 ms. FP64 staged: focused 2
// waves 79.6 ms, 3 waves 149.6 ms; Parker 2 waves 52.8 ms, 3 waves 84.5 ms (62.0
// ms with its sample's loads issued two components at a time, 29 -> 12 VGPRs
// spilled).
#ifndef KPT_PUSH_WAVES_FP64_PARKER
#define KPT_PUSH_WAVES_FP64_PARKER 2
#endif
#ifndef KPT_PUSH_WAVES_FP64
#define KPT_PUSH_WAVES_FP64 2
#endif
#ifndef KPT_PUSH_WAVES_FP32_PARKER
#define KPT_PUSH_WAVES_FP32_PARKER 4
#endif
#ifndef KPT_PUSH_WAVES_FP32_FOCUSED
#define KPT_PUSH_WAVES_FP32_FOCUSED 4
#endif
// One lane of the push: particle `idx` of the set in `a` (lanes past a.count
// or with status != 0 only join the wave reduction of the counters). Every lane
// of a wavefront calls it together.
// STR: the frame has a stretched axis (node_stencil); the uniform kernels keep
// the plain stencil.
template <int MODEL, bool FOCUSED, typename FT, int CLOSURE, bool STR = false, int STAGE = 0>
__device__ __forceinline__ void push_lane(ArgsK ap, const int64_t idx, float4* stage_lds = nullptr) {
    const PushArgs& a = *(const PushArgs*)(ap);
    constexpr bool QLT = FOCUSED && CLOSURE != KPT_CLOSURE_REDUCED;
    constexpr bool FP32 = sizeof(FT) == 4;
    constexpr int CP = MODEL;  // components read per corner
    uint64_t my_steps = 0;
    uint32_t my_fail = 0, my_active = 0;

    const bool live = idx < static_cast<int64_t>(a.count) && a.P.status[idx] == 0;
    const kpt_device_settings& s = a.s;
    const FT* __restrict__ flo = static_cast<const FT*>(a.field_lo);
    const FT* __restrict__ fhi = static_cast<const FT*>(a.field_hi);
    const int nxp = a.g.nx + 2;
    const double interval_start = a.t0;
    const double interval_end = a.t_end;
    const bool blend = s.time_interpolation != 0;
    double x = 0.0, y = 0.0, p = 1.0, t = interval_end, dt = 0.0, diagnostic_dt = 0.0, pitch_mu = 0.0;
    uint64_t id = 0;
    // Sub-steps already taken in this interval (earlier chunks): the step guard
    // and the Philox sub-step counter continue from there.
    int step_guard = 0, status = 0;
    if (live) {
        my_active = 1;
        x = a.P.x[idx];
        y = a.P.y[idx];
        p = a.P.p[idx];
        t = a.P.t[idx];
        dt = a.P.step_dt[idx];
        diagnostic_dt = dt;
        if (MODEL == 21) {
            const double m = a.P.mu[idx];
            pitch_mu = QLT ? reflect_mu(m, s.mu_max) : (m > s.mu_max ? s.mu_max : (m < -s.mu_max ? -s.mu_max : m));
        }
        id = a.P.id[idx];
        step_guard = a.first_chunk ? 0 : static_cast<int>(a.P.substep[idx]);
    }
    const double t_stop = a.t_stop;
    // STAGE > 0: the LDS bytes of the wavefront's region for its staged node box;
    // sub-steps then run wave-uniformly (all lanes take part in restaging).
    // STAGE = 0: no staging, and each lane leaves the loop when its particle is done.
    StageBox box{0, 0, 0, 0, -1};
    // The node-major records the push samples: the frame pair of its precision, or
    // (FP32, one frame) the frame's own records.
    const bool recs = FP32 && a.field_pair == nullptr && a.field_rec != nullptr;
    const void* pair = FP32 ? (recs ? static_cast<const void*>(a.field_rec) : static_cast<const void*>(a.field_pair))
                            : static_cast<const void*>(a.field_pair64);
    constexpr int PV = FP32 ? pair_stride(CP) / 4 : CP;  // 16-byte units per pair record
    constexpr int RV = rec_stride(CP) / 4;               // and per single-frame record
    constexpr int CAP_P = STAGE / (stage_stride(PV) * 16), CAP_R = STAGE / (stage_stride(RV) * 16);
    const bool stage = STAGE != 0 && pair != nullptr && a.stage != 0;

    for (;;) {
        bool run = live && status == 0 && t < t_stop && step_guard < s.max_particle_steps_per_interval;
        if constexpr (STAGE != 0) {
            if (!__any(run)) break;
        } else {
            if (!run) break;
        }
        // The uniform arguments of this sub-step, read again from constant memory.
        const PushArgs& a = args_reload(ap);
        const kpt_device_settings& s = a.s;

        // phase 1: wrap and stencil
        int i0 = 0, j0 = 0;
        double wx = 0.0, wy = 0.0, uw = 0.0;
        FT half_dx = step_const<FT>(a).half_dx, half_dy = step_const<FT>(a).half_dy;
        do {
            if (!run) break;
            ++step_guard;
            x = wrap_periodic(x, s.x_lower, s.x_upper, s.x_width);
            y = wrap_periodic(y, s.y_lower, s.y_upper, s.y_width);
            if constexpr (STR) {
                // half cell widths of the dt limiter: the local face interval on a
                // stretched axis
                double hx, hy;
                if (!node_stencil(x + s.interpolation_shift_x, a.g, 0, a.g.nx, i0, wx, hx) ||
                    !node_stencil(y + s.interpolation_shift_y, a.g, 1, a.g.ny, j0, wy, hy)) {
                    status = 2;
                    run = false;
                    break;
                }
                if (a.g.face[0]) half_dx = static_cast<FT>(0.5 * hx);
                if (a.g.face[1]) half_dy = static_cast<FT>(0.5 * hy);
            } else if (!axis_stencil(x + s.interpolation_shift_x, a.g, 0, a.g.nx, i0, wx) ||
                       !axis_stencil(y + s.interpolation_shift_y, a.g, 1, a.g.ny, j0, wy)) {
                status = 2;
                run = false;
                break;
            }
            if (blend) {
                const double rt = FP32 ? (t - interval_start) * a.k32.inv_mhd_dt : (t - interval_start) / s.mhd_dt;
                uw = rt < 0.0 ? 0.0 : (rt > 1.0 ? 1.0 : rt);
            }
        } while (false);
        if constexpr (STAGE != 0) {
            if (stage) {
                if (recs)
                    stage_update<RV, CAP_R>(box, stage_lds, pair, a.g.nx, a.g.ny, run, i0, j0);
                else
                    stage_update<PV, CAP_P>(box, stage_lds, pair, a.g.nx, a.g.ny, run, i0, j0);
            }
        }

        // phase 2: sample, rates, closure, step (a failure ends the particle's
        // interval with status 2)
        do {
            if (!run) break;
            const int64_t n00 = static_cast<int64_t>(j0 + 1) * nxp + (i0 + 1);
            StepTerms<FT> T;
            bool ok;
            {
                FT c[CP];
                if constexpr (FP32) {
                    if (recs) {
                        if (STAGE != 0 && stage && box.holds(i0, j0)) {
                            const int local = (j0 - box.j_lo) * box.w + (i0 - box.i_lo);
                            sample_rec32<CP, stage_stride(RV)>(stage_lds + local * stage_stride(RV), box.w, wx, wy,
                                                               c);
                        } else {
                            sample_rec32<CP, RV>(reinterpret_cast<const float4*>(a.field_rec) + n00 * RV, nxp, wx,
                                                 wy, c);
                        }
                    } else if constexpr (STAGE != 0) {
                        if (stage && box.holds(i0, j0)) {
                            const int local = (j0 - box.j_lo) * box.w + (i0 - box.i_lo);
                            sample_staged<CP>(stage_lds + local * stage_stride(PV), box.w, uw, wx, wy, c);
                        } else if (a.field_pair != nullptr) {
                            sample_pair32<CP, KPT_PAIR_RECORDS_EXACT>(a.field_pair, n00, nxp, uw, wx, wy, c);
                        } else {
                            sample_blend<CP, FT>(flo, fhi, blend, uw, n00, nxp, a.plane, wx, wy, c);
                        }
                    } else if (a.field_pair != nullptr) {
                        sample_pair32<CP, KPT_PAIR_RECORDS_EXACT>(a.field_pair, n00, nxp, uw, wx, wy, c);
                    } else {
                        sample_blend<CP, FT>(flo, fhi, blend, uw, n00, nxp, a.plane, wx, wy, c);
                    }
                } else {
                    if (STAGE != 0 && stage && box.holds(i0, j0)) {
                        const int local = (j0 - box.j_lo) * box.w + (i0 - box.i_lo);
                        sample_pair64<CP, stage_stride(CP)>(
                            reinterpret_cast<const double2*>(stage_lds) + local * stage_stride(CP), box.w, uw, wx,
                            wy, c);
                    } else if (a.field_pair64 != nullptr) {
                        sample_pair64<CP>(reinterpret_cast<const double2*>(a.field_pair64) + n00 * CP, nxp, uw, wx,
                                          wy, c);
                    } else {
                        sample_blend<CP, FT>(flo, fhi, blend, uw, n00, nxp, a.plane, wx, wy, c);
                    }
                }
                if constexpr (FP32) {
                    if (!(c[MODEL == 16 ? 5 : 6] >= a.k32.null_b)) {
                        // near a null: the FP64 frames and FP64 arithmetic for this sub-step
                        StepTerms<float> Tn;
                        ok = near_null_terms<MODEL, FOCUSED, QLT>(a.self, blend, uw, n00, nxp, wx, wy, p, pitch_mu,
                                                                  &Tn);
                        T = Tn;
                    } else {
                        ok = eval_terms<MODEL, FOCUSED, QLT, float>(c, p, pitch_mu, s, a.k32, T);
                    }
                } else {
                    ok = eval_terms<MODEL, FOCUSED, QLT, double>(c, p, pitch_mu, s, a.k64, T);
                }
            }
            if (!ok) {
                status = 2;
                break;
            }
            if constexpr (QLT) {
                FT D, dD;
                qlt_scattering<FT>(static_cast<FT>(pitch_mu), T.q_xi, T.q_xig, T.q_pref, step_const<FT>(a).qlt_gk,
                                   step_const<FT>(a).qlt_h0, D, dD);
                T.dmu_dt += dD;
                T.duu = D;
                if (!rfinite(T.dmu_dt) || !rfinite(D) || D < FT(0)) {
                    status = 2;
                    break;
                }
            }

            if (s.dt_min < s.dt_max) {
                dt = candidate_dt<FOCUSED>(T, p, s, step_const<FT>(a), half_dx, half_dy);
                if (dt < s.dt_min) dt = s.dt_min;
                if (dt > s.dt_max) dt = s.dt_max;
            } else {
                // dt_min == dt_max: the clamp fixes the step whatever the
                // candidates are (they are never NaN: the first is finite and
                // fmin drops NaN operands).
                dt = s.dt_min;
            }
            diagnostic_dt = dt;
            const double remaining = interval_end - t;
            if (dt > remaining) dt = remaining;
            if (!isfinite(dt) || dt <= 0.0) {
                status = 2;
                break;
            }

            FT r0, r1, r2, r3;
            const uint32_t k = static_cast<uint32_t>(step_guard - 1);
            if (a.noise != nullptr) {
                if (k >= a.noise_steps) {
                    status = 2;
                    break;
                }
                const double* row = a.noise + (static_cast<uint64_t>(idx) * a.noise_steps + k) * 4;
                r0 = static_cast<FT>(row[0]);
                r1 = static_cast<FT>(row[1]);
                r2 = static_cast<FT>(row[2]);
                r3 = static_cast<FT>(row[3]);
            } else {
                uint32_t rb[4];
                philox4x32_10(static_cast<uint32_t>(id), static_cast<uint32_t>(id >> 32), a.frame, k,
                              a.key[0], a.key[1], rb);
                r0 = draw_sqrt3<FT>(rb[0]);
                r1 = draw_sqrt3<FT>(rb[1]);
                r2 = draw_sqrt3<FT>(rb[2]);
                r3 = draw_sqrt3<FT>(rb[3]);
            }
            const FT dtr = static_cast<FT>(dt);
            const FT sdt = rt_sqrt(dtr);
            FT delta_x, delta_y;
            if (FOCUSED) {
                // src/main.cpp:1671-1702 (ran1, ran2, ignored, pitch)
                const FT bxn = T.bx * T.ib, byn = T.by * T.ib, bzn = T.bz * T.ib;
                const FT bxy = rt_sqrt(bxn * bxn + byn * byn);
                FT sx = FT(0), sy = FT(0);
                if (bxy > FT(0)) {
                    const FT ibxy = rrcp(bxy);
                    sx = T.skperp * ibxy * sdt * (-bxn * bzn * r0 - T.by * r1);
                    sy = T.skperp * ibxy * sdt * (-byn * bzn * r0 + T.bx * r1);
                }
                delta_x = T.dx_dt * dtr + sx;
                delta_y = T.dy_dt * dtr + sy;
                if constexpr (QLT) {
                    // FocusTransportSolver::picard_pitch_angle_update (FocusSolver.hpp:700-719):
                    // three passes, drift and diffusion at the trial mu, one noise value.
                    const double mu0 = pitch_mu;
                    FT rate = T.dmu_dt, D = T.duu;
                    double trial = mu0;
#pragma unroll 1
                    for (int it = 0; it < 3; ++it) {
                        if (it > 0) {
                            const FT m = static_cast<FT>(trial);
                            FT dD;
                            qlt_scattering<FT>(m, T.q_xi, T.q_xig, T.q_pref, step_const<FT>(a).qlt_gk,
                                               step_const<FT>(a).qlt_h0, D, dD);
                            rate = FT(0.5) * (FT(1) - m * m) *
                                       (T.f_speed * T.f_divb + m * T.f_divv - FT(3) * m * T.f_bb -
                                        FT(2) * rdiv(T.f_bv, T.f_speed)) +
                                   dD;
                        }
                        const FT scale = D > FT(0) ? rt_sqrt(FT(2) * D * dtr) : FT(0);
                        trial = reflect_mu(mu0 + static_cast<double>(dtr * rate) + static_cast<double>(scale * r3),
                                           s.mu_max);
                    }
                    pitch_mu = trial;
                } else {
                    const double m = pitch_mu + static_cast<double>(T.dmu_dt * dtr) +
                                     static_cast<double>(r3 * rt_sqrt(FT(2) * T.duu) * sdt);
                    pitch_mu = m > s.mu_max ? s.mu_max : (m < -s.mu_max ? -s.mu_max : m);
                }
            } else {
                // src/main.cpp:1279-1292 (ran1, ran2, ran3, ignored)
                delta_x = T.dx_dt * dtr + r0 * T.skperp * sdt + r2 * T.skpara_perp * sdt * T.bx * T.ib;
                delta_y = T.dy_dt * dtr + r1 * T.skperp * sdt + r2 * T.skpara_perp * sdt * T.by * T.ib;
            }
            (void)r3;
            x += static_cast<double>(delta_x);
            y += static_cast<double>(delta_y);
            p += T.dp_dt * dt;
            t += dt;
            ++my_steps;
            if (p < 0.25 * s.p0) p = 0.25 * s.p0;
            if (!isfinite(x) || !isfinite(y) || !isfinite(p) || !isfinite(t)) {
                status = 2;
                break;
            }
        } while (false);
    }

    if (live) {
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

// Particle tile of a push workgroup. Workgroups are dispatched to the 8 XCDs
// round-robin (workgroup b to XCD b % 8). Runs of `run` consecutive tiles go to
// the XCDs in turn, so each XCD's L2 sees `run` spatially adjacent tiles of the
// cell-sorted particles at a time; run = 1 is the plain order. The grid is a
// multiple of 8 run and the map a bijection; ids past the last tile find no
// particles.
__device__ __forceinline__ int64_t push_first_particle(uint32_t run) {
    const uint32_t b = blockIdx.x, i = b / 8, x = b % 8;
    const uint64_t t = static_cast<uint64_t>(i / run) * (8ull * run) + static_cast<uint64_t>(x) * run + i % run;
    return static_cast<int64_t>(t) * kBlock + threadIdx.x;
}

template <int MODEL, bool FOCUSED, typename FT, int CLOSURE, bool STR>
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(sizeof(FT) == 8 ? (FOCUSED ? KPT_PUSH_WAVES_FP64 : KPT_PUSH_WAVES_FP64_PARKER)
                                                      : (FOCUSED ? KPT_PUSH_WAVES_FP32_FOCUSED
                                                                 : KPT_PUSH_WAVES_FP32_PARKER))))
    push_kernel(PushArgs a, uint32_t run) {
    (void)a;
    if constexpr (KPT_STAGE_CU_BYTES > 0) {
        constexpr int waves = sizeof(FT) == 8 ? (FOCUSED ? KPT_PUSH_WAVES_FP64 : KPT_PUSH_WAVES_FP64_PARKER)
                                              : (FOCUSED ? KPT_PUSH_WAVES_FP32_FOCUSED : KPT_PUSH_WAVES_FP32_PARKER);
        constexpr int BYTES = KPT_STAGE_CU_BYTES / (4 * waves) / 16 * 16;  // per wavefront
        __shared__ float4 stage_mem[kBlock / 64 * (BYTES / 16)];
        const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        push_lane<MODEL, FOCUSED, FT, CLOSURE, STR, BYTES>(kernel_args(), push_first_particle(run),
                                                           stage_mem + wave * (BYTES / 16));
    } else {
        push_lane<MODEL, FOCUSED, FT, CLOSURE, STR>(kernel_args(), push_first_particle(run));
    }
}

// Closure from the network outputs at pitch cosine mu (|mu| in the hull), with
// log2 R = l2r = (out_mean + out_scale y0) log2(e) and y1 = d ln R / d ln|mu|:
// D = P (1 - mu^2) R, dD/dmu = P R ((1 - mu^2) sgn(mu) y1 / |mu| - 2 mu).
__device__ __forceinline__ float mlp_log2_r(float y0, const MlpScale& ms) {
    return (ms.out_mean + ms.out_scale * y0) * 1.4426950408889634f;
}
__device__ __forceinline__ void mlp_closure_l2(float l2r, float y1, float mu, float pref, float& D, float& dD) {
    const float omm = 1.0f - mu * mu;
    const float r = __builtin_amdgcn_exp2f(l2r);
    const float sgn = mu > 0.0f ? 1.0f : -1.0f;
    D = pref * omm * r;
    dD = pref * r * (omm * sgn * y1 * __builtin_amdgcn_rcpf(fabsf(mu)) - 2.0f * mu);
}
__device__ __forceinline__ void mlp_closure(float y0, float y1, float mu, float pref, const MlpScale& ms, float& D,
                                            float& dD) {
    mlp_closure_l2(mlp_log2_r(y0, ms), y1, mu, pref, D, dD);
}

// The closure at a Picard trial value m from the evaluation at the sub-step's
// initial mu0: log2 R to first order in log2|mu| about mu0 with the network's
// slope output there, log2 R(m) = log2 R(mu0) + y1 (log2|m| - log2|mu0|), and
// the slope held at y1. R is close to a power of |mu| over one sub-step's pitch
// change (|d y1 / d ln|mu|| <= gamma_k^2 / 4 for slab QLT), so the expansion
// error is below the network's own.
__device__ __forceinline__ void mlp_closure_extend(float l2r0, float y1, float l2mu0, float m, float pref, float& D,
                                                   float& dD) {
    const float l2m = __builtin_amdgcn_logf(fabsf(m));
    mlp_closure_l2(l2r0 + y1 * (l2m - l2mu0), y1, m, pref, D, dD);
}

// One closure evaluation at mu: the network's outputs where the sample lies in
// the hull, the exact slab-QLT formula elsewhere.
__device__ __forceinline__ void closure_at(bool in_hull, float y0, float y1, float mu, const StepTerms<float>& T,
                                           const StepConst<float>& K, const MlpScale& ms, float& D, float& dD) {
    if (in_hull)
        mlp_closure(y0, y1, mu, T.q_pref, ms, D, dD);
    else
        qlt_scattering<float>(mu, T.q_xi, T.q_xig, T.q_pref, K.qlt_gk, K.qlt_h0, D, dD);
}

// ---------------------------------------------------------------------------
// Focused FP32 push with the learned closure. Same mover as push_kernel<21,
// true, float, QLT>, with the sub-step loop run wave-uniformly: every lane of
// the wavefront takes part in the network evaluations of a sub-step, lanes that
// are finished or failed with dummy features.
// PASSES = 1: one evaluation per sub-step, at the initial mu; the two Picard
// passes extend it to their trial values (mlp_closure_extend).
// PASSES = 3: one evaluation at each Picard pass (the initial mu and the two
// trial values).
// ---------------------------------------------------------------------------
// Waves per SIMD: 3 (256 VGPRs at 2 waves, 168 with spills to scratch at 3);
// measured per launch against 2 waves: 3 waves -23%, 4 waves -16%.
#ifndef KPT_PUSH_WAVES_MLP
#define KPT_PUSH_WAVES_MLP 3
#endif
// One lane of the learned-closure push (the network image already staged in
// LDS as `im`, output biases b40, b41). Every lane of a wavefront calls it
// together: the sub-step loop and the network evaluations are wave-uniform.
template <bool STR = false, int ACT = kActSilu, int PASSES = 1>
__device__ __forceinline__ void push_lane_mlp(ArgsK ap, const MlpLds& im, const float b40, const float b41,
                                              const int64_t idx) {
    static_assert(PASSES == 1 || PASSES == 3, "one or three network evaluations per sub-step");
    const PushArgs& a = *(const PushArgs*)(ap);
    const kpt_device_settings& s = a.s;
    constexpr int CP = 21;
    const bool use_net = a.mlp_on != 0;

    const bool live = idx < static_cast<int64_t>(a.count) && a.P.status[idx] == 0;
    uint64_t my_steps = 0;
    uint32_t my_fail = 0, my_active = live ? 1u : 0u;
    const int nxp = a.g.nx + 2;
    const double interval_start = a.t0, interval_end = a.t_end, t_stop = a.t_stop;
    const bool blend = s.time_interpolation != 0;
    double x = 0.0, y = 0.0, p = 1.0, t = interval_end, dt = 0.0, diagnostic_dt = 0.0, pitch_mu = 0.0;
    uint64_t id = 0;
    int step_guard = 0, status = 0;
    if (live) {
        x = a.P.x[idx];
        y = a.P.y[idx];
        p = a.P.p[idx];
        t = a.P.t[idx];
        dt = a.P.step_dt[idx];
        diagnostic_dt = dt;
        pitch_mu = reflect_mu(a.P.mu[idx], s.mu_max);
        id = a.P.id[idx];
        step_guard = a.first_chunk ? 0 : static_cast<int>(a.P.substep[idx]);
    }

    for (;;) {
        bool run = live && status == 0 && t < t_stop && step_guard < s.max_particle_steps_per_interval;
        if (!__any(run)) break;
        // The uniform arguments of this sub-step, read again from constant memory.
        const PushArgs& a = args_reload(ap);
        const kpt_device_settings& s = a.s;
        const MlpScale& ms = a.mlp_scale;
        const StepConst<float>& K = a.k32;

        // phase 1: sample, rates, features at the initial mu
        StepTerms<float> T{};
        float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f, lxi = 0.0f, l2mu0 = 0.0f;
        bool h0 = false;
        float half_dx = K.half_dx, half_dy = K.half_dy;
        if (run) {
            ++step_guard;
            x = wrap_periodic(x, s.x_lower, s.x_upper, s.x_width);
            y = wrap_periodic(y, s.y_lower, s.y_upper, s.y_width);
            int i0, j0;
            double wx, wy, hx = 0.0, hy = 0.0;
            bool ok_stencil;
            if constexpr (STR)
                ok_stencil = node_stencil(x + s.interpolation_shift_x, a.g, 0, a.g.nx, i0, wx, hx) &&
                             node_stencil(y + s.interpolation_shift_y, a.g, 1, a.g.ny, j0, wy, hy);
            else
                ok_stencil = axis_stencil(x + s.interpolation_shift_x, a.g, 0, a.g.nx, i0, wx) &&
                             axis_stencil(y + s.interpolation_shift_y, a.g, 1, a.g.ny, j0, wy);
            if (!ok_stencil) {
                status = 2;
                run = false;
            } else {
                if constexpr (STR) {
                    if (a.g.face[0]) half_dx = static_cast<float>(0.5 * hx);
                    if (a.g.face[1]) half_dy = static_cast<float>(0.5 * hy);
                }
                const int64_t n00 = static_cast<int64_t>(j0 + 1) * nxp + (i0 + 1);
                double uw = 0.0;
                if (blend) {
                    const double rt = (t - interval_start) * K.inv_mhd_dt;
                    uw = rt < 0.0 ? 0.0 : (rt > 1.0 ? 1.0 : rt);
                }
                float c[CP];
                if (a.field_pair != nullptr)
                    sample_pair32<CP, KPT_PAIR_RECORDS_MLP>(a.field_pair, n00, nxp, uw, wx, wy, c);
                else if (a.field_rec != nullptr)
                    sample_rec32<CP, rec_stride(CP) / 4>(reinterpret_cast<const float4*>(a.field_rec) +
                                                             n00 * (rec_stride(CP) / 4),
                                                         nxp, wx, wy, c);
                else
                    sample_blend<CP, float>(static_cast<const float*>(a.field_lo), static_cast<const float*>(a.field_hi),
                                            blend, uw, n00, nxp, a.plane, wx, wy, c);
                bool ok;
                if (!(c[6] >= K.null_b)) {
                    StepTerms<float> Tn;
                    ok = near_null_terms<21, true, true, 1>(a.self, blend, uw, n00, nxp, wx, wy, p, pitch_mu, &Tn);
                    T = Tn;
                } else {
                    ok = eval_terms<21, true, true, float>(c, p, pitch_mu, s, K, T);
                }
                if (!ok) {
                    status = 2;
                    run = false;
                } else {
                    lxi = __builtin_amdgcn_logf(T.q_xi) * 0.6931471805599453f;
                    h0 = use_net &&
                         mlp_features(ms, fabsf(static_cast<float>(pitch_mu)), lxi, K.qlt_gk, f0, f1, f2, l2mu0);
                }
            }
        }
        float y0 = 0.0f, y1 = 0.0f;
        if (use_net) mlp_eval_wave<ACT>(im, f0, f1, f2, b40, b41, y0, y1);  // use_net is uniform over the launch

        // phase 2: closure at mu0, step size, noise, transverse increments, first Picard pass
        float r3 = 0.0f, dtr = 0.0f, delta_x = 0.0f, delta_y = 0.0f;
        const double mu0 = pitch_mu;
        double trial = mu0;
        bool h1 = false;
        f0 = f1 = f2 = 0.0f;
        if (run) {
            float D, dD;
            closure_at(h0, y0, y1, static_cast<float>(mu0), T, K, ms, D, dD);
            T.dmu_dt += dD;
            T.duu = D;
            if (!rfinite(T.dmu_dt) || !rfinite(D) || D < 0.0f) {
                status = 2;
                run = false;
            }
            if (run) {
                if (s.dt_min < s.dt_max) {
                    dt = candidate_dt<true>(T, p, s, K, half_dx, half_dy);
                    if (dt < s.dt_min) dt = s.dt_min;
                    if (dt > s.dt_max) dt = s.dt_max;
                } else {
                    dt = s.dt_min;
                }
                diagnostic_dt = dt;
                const double remaining = interval_end - t;
                if (dt > remaining) dt = remaining;
                if (!isfinite(dt) || dt <= 0.0) {
                    status = 2;
                    run = false;
                }
            }
            if (run) {
                float r0, r1;
                const uint32_t k = static_cast<uint32_t>(step_guard - 1);
                if (a.noise != nullptr) {
                    if (k >= a.noise_steps) {
                        status = 2;
                        run = false;
                    } else {
                        const double* row = a.noise + (static_cast<uint64_t>(idx) * a.noise_steps + k) * 4;
                        r0 = static_cast<float>(row[0]);
                        r1 = static_cast<float>(row[1]);
                        r3 = static_cast<float>(row[3]);
                    }
                } else {
                    uint32_t rb[4];
                    philox4x32_10(static_cast<uint32_t>(id), static_cast<uint32_t>(id >> 32), a.frame, k, a.key[0],
                                  a.key[1], rb);
                    r0 = draw_sqrt3<float>(rb[0]);
                    r1 = draw_sqrt3<float>(rb[1]);
                    r3 = draw_sqrt3<float>(rb[3]);
                }
                if (run) {
                    dtr = static_cast<float>(dt);
                    const float sdt = rt_sqrt(dtr);
                    const float bxn = T.bx * T.ib, byn = T.by * T.ib, bzn = T.bz * T.ib;
                    const float bxy = rt_sqrt(bxn * bxn + byn * byn);
                    float sx = 0.0f, sy = 0.0f;
                    if (bxy > 0.0f) {
                        const float ibxy = rrcp(bxy);
                        sx = T.skperp * ibxy * sdt * (-bxn * bzn * r0 - T.by * r1);
                        sy = T.skperp * ibxy * sdt * (-byn * bzn * r0 + T.bx * r1);
                    }
                    delta_x = T.dx_dt * dtr + sx;
                    delta_y = T.dy_dt * dtr + sy;
                    const float scale = D > 0.0f ? rt_sqrt(2.0f * D * dtr) : 0.0f;
                    trial = reflect_mu(mu0 + static_cast<double>(dtr * T.dmu_dt) + static_cast<double>(scale * r3),
                                       s.mu_max);
                    if constexpr (PASSES == 3) {
                        const float m1 = static_cast<float>(trial);
                        float l2m;
                        h1 = use_net && mlp_features(ms, fabsf(m1), lxi, K.qlt_gk, f0, f1, f2, l2m);
                    }
                }
            }
        }

        if constexpr (PASSES == 1) {
            // phases 3 and 4: the two Picard passes, with the closure extended from mu0
            if (run) {
                const float l2r0 = mlp_log2_r(y0, ms);
#pragma unroll 1
                for (int it = 1; it < 3; ++it) {
                    const float m = static_cast<float>(trial);
                    float D, dD;
                    if (h0 && mlp_mu_in_hull(ms, fabsf(m)))
                        mlp_closure_extend(l2r0, y1, l2mu0, m, T.q_pref, D, dD);
                    else
                        qlt_scattering<float>(m, T.q_xi, T.q_xig, T.q_pref, K.qlt_gk, K.qlt_h0, D, dD);
                    const float rate = 0.5f * (1.0f - m * m) *
                                           (T.f_speed * T.f_divb + m * T.f_divv - 3.0f * m * T.f_bb -
                                            2.0f * rdiv(T.f_bv, T.f_speed)) +
                                       dD;
                    const float scale = D > 0.0f ? rt_sqrt(2.0f * D * dtr) : 0.0f;
                    trial = reflect_mu(mu0 + static_cast<double>(dtr * rate) + static_cast<double>(scale * r3),
                                       s.mu_max);
                }
            }
        } else {
            // phases 3 and 4: the two Picard passes, each with its own network evaluation
#pragma unroll 1
            for (int it = 1; it < 3; ++it) {
                if (use_net) mlp_eval_wave<ACT>(im, f0, f1, f2, b40, b41, y0, y1);
                bool hn = false;
                f0 = f1 = f2 = 0.0f;
                if (run) {
                    const float m = static_cast<float>(trial);
                    float D, dD;
                    closure_at(h1, y0, y1, m, T, K, ms, D, dD);
                    const float rate = 0.5f * (1.0f - m * m) *
                                           (T.f_speed * T.f_divb + m * T.f_divv - 3.0f * m * T.f_bb -
                                            2.0f * rdiv(T.f_bv, T.f_speed)) +
                                       dD;
                    const float scale = D > 0.0f ? rt_sqrt(2.0f * D * dtr) : 0.0f;
                    trial = reflect_mu(mu0 + static_cast<double>(dtr * rate) + static_cast<double>(scale * r3),
                                       s.mu_max);
                    if (it == 1) {
                        const float m2 = static_cast<float>(trial);
                        float l2m;
                        hn = use_net && mlp_features(ms, fabsf(m2), lxi, K.qlt_gk, f0, f1, f2, l2m);
                    }
                }
                h1 = hn;
            }
        }

        if (run) {
            pitch_mu = trial;
            x += static_cast<double>(delta_x);
            y += static_cast<double>(delta_y);
            p += T.dp_dt * dt;
            t += dt;
            ++my_steps;
            if (p < 0.25 * s.p0) p = 0.25 * s.p0;
            if (!isfinite(x) || !isfinite(y) || !isfinite(p) || !isfinite(t)) status = 2;
        }
    }

    if (live) {
        if (step_guard >= s.max_particle_steps_per_interval && t < interval_end) status = 2;
        a.P.x[idx] = wrap_periodic(x, s.x_lower, s.x_upper, s.x_width);
        a.P.y[idx] = wrap_periodic(y, s.y_lower, s.y_upper, s.y_width);
        a.P.p[idx] = p;
        a.P.mu[idx] = pitch_mu;
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

// The network image of one launch goes to LDS once per workgroup.
__device__ __forceinline__ MlpLds mlp_stage(const void* image, uint4* smem, float& b40, float& b41) {
    mlp_load_lds(image, smem);
    __syncthreads();
    const float* b4 = reinterpret_cast<const float*>(static_cast<const char*>(image) + kMlpLdsBytes);
    b40 = b4[0];
    b41 = b4[1];
    return mlp_lds_view(smem);
}

template <bool STR, int ACT, int PASSES>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(KPT_PUSH_WAVES_MLP)))
    push_kernel_mlp(PushArgs a, uint32_t run) {
    __shared__ uint4 mlp_smem[kMlpLdsBytes / 16];
    float b40, b41;
    const MlpLds im = mlp_stage(a.mlp_image, mlp_smem, b40, b41);
    push_lane_mlp<STR, ACT, PASSES>(kernel_args(), im, b40, b41, push_first_particle(run));
}

#if !KPT_NO_BUDGET_KERNEL
// Never launched. It calls the learned-closure push's near-null instance from a
// 4-wave kernel, so the compiler builds that instance within 128 VGPRs: the
// registers a call clobbers are the ones its caller spills around it, and a
// callee built for the learned-closure kernel's own 168 cost that kernel 11% per
// launch in spills.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 4), used))
    near_null_budget_kernel(const PushArgs* a, StepTerms<float>* out) {
    (void)near_null_terms<21, true, true, 1>(a, true, 0.5, 0, a->g.nx + 2, 0.5, 0.5, 1.0, 0.5, out);
}
#endif

// ---------------------------------------------------------------------------
// Batched ensemble: one persistent launch advances many independent runs.
// Workgroups take 256-particle tiles from work queues; with 8 queues, queue g
// holds the tiles of the runs assigned to XCD g (workgroups are dispatched to
// the XCDs round-robin, so workgroup b starts at queue b % 8 and keeps its
// run's field in that XCD's L2), and a workgroup whose queue is empty takes
// tiles from the next queues. Every workgroup leaves once all queues are
// exhausted.
// ---------------------------------------------------------------------------
#ifndef KPT_ENSEMBLE_WAVES
#define KPT_ENSEMBLE_WAVES KPT_PUSH_WAVES_MLP
#endif
// LDS per wavefront for staged records of the Parker and exact-closure runs (0:
// none). 3 840 B is what 3 workgroups per CU leave beside their network images
// (60 Parker or 48 focused single-frame records); on the 64-run ensemble of 512 x 256 fields it
// measured 1.86e10 particle-steps/s against 2.23e10 without staging (the runs
// read their node-major single-frame records from L1/L2, and each lane leaves
// the sub-step loop when its particle is done).
#ifndef KPT_ENSEMBLE_STAGE_BYTES
#define KPT_ENSEMBLE_STAGE_BYTES 0
#endif
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(KPT_ENSEMBLE_WAVES)))
    ensemble_kernel(const EnsembleRun* __restrict__ runs, const uint2* __restrict__ tiles,
                    const uint32_t* __restrict__ group_start, int groups, unsigned int* __restrict__ queues,
                    const void* __restrict__ mlp_image, int has_mlp) {
    __shared__ uint4 mlp_smem[kMlpLdsBytes / 16];
    __shared__ uint32_t s_tile;
    constexpr int SB = KPT_ENSEMBLE_STAGE_BYTES;
    __shared__ float4 ens_stage[kBlock / 64 * (SB > 0 ? SB / 16 : 1)];
    float4* const my_stage = ens_stage + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * (SB / 16);
    float b40 = 0.0f, b41 = 0.0f;
    MlpLds im{};
    if (has_mlp) im = mlp_stage(mlp_image, mlp_smem, b40, b41);
    int g = static_cast<int>(blockIdx.x % static_cast<unsigned>(groups));
    int visited = 0;
    while (visited < groups) {
        const uint32_t size = group_start[g + 1] - group_start[g];
        if (threadIdx.x == 0) s_tile = atomicAdd(&queues[g], 1u);
        __syncthreads();
        const uint32_t k = __builtin_amdgcn_readfirstlane(s_tile);
        __syncthreads();
        if (k >= size) {  // this queue is exhausted: move to the next one
            g = g + 1 == groups ? 0 : g + 1;
            ++visited;
            continue;
        }
        const uint2 t = tiles[group_start[g] + k];
        const uint32_t r = __builtin_amdgcn_readfirstlane(t.x);
        const int64_t idx = static_cast<int64_t>(__builtin_amdgcn_readfirstlane(t.y)) + threadIdx.x;
        const EnsembleRun& R = runs[r];
        if (R.kind == kEnsParker32)
            push_lane<16, false, float, KPT_CLOSURE_REDUCED, false, SB>(as_args(R.a), idx, my_stage);
        else if (R.kind == kEnsFocusedQlt32)
            push_lane<21, true, float, KPT_CLOSURE_QLT, false, SB>(as_args(R.a), idx, my_stage);
        else if (R.a.mlp_act == kActHardSwish)
            push_lane_mlp<false, kActHardSwish, 1>(as_args(R.a), im, b40, b41, idx);
        else
            push_lane_mlp<false, kActSilu, 1>(as_args(R.a), im, b40, b41, idx);
    }
}

// Sets every run's interval for the next launch of the frame loop and clears
// the queue counters (a kernel rather than a memset, so the loop can be
// captured in a graph).
__global__ void __launch_bounds__(kBlock) ensemble_advance_kernel(EnsembleRun* __restrict__ runs, int n_runs,
                                                                  EnsembleLoop* __restrict__ loop,
                                                                  unsigned int* __restrict__ queues, int groups) {
    const uint32_t frame = loop->first_frame + loop->done;
    const double window = loop->window;
    for (int r = threadIdx.x; r < n_runs; r += kBlock) {
        PushArgs& a = runs[r].a;
        a.frame = frame;
        a.t0 = static_cast<double>(frame) * a.s.mhd_dt;
        a.t_end = a.t0 + a.s.mhd_dt;
        a.t_stop = window >= 1.0 ? a.t_end : a.t0 + a.s.mhd_dt * window;
        a.first_chunk = 1;
    }
    for (int q = threadIdx.x; q < groups; q += kBlock) queues[q] = 0u;
    __syncthreads();
    if (threadIdx.x == 0) loop->done = loop->done + 1u;
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
    // split baseline in the driver's units (rest mass = c = 1), the E0 its
    // snapshots record (src/main.cpp:2160-2170)
    P.e0[i] = sqrt(1.0 + P.p[i] * P.p[i]) - 1.0;
    // ParticleSystem state of a new particle (set_particle,
    // include/ParticleSystem.hpp:516-522)
    P.prev_x[i] = P.prev_step_x[i] = P.x[i];
    P.prev_y[i] = P.prev_step_y[i] = P.y[i];
    P.sort_key[i] = 0;
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
            P.e0[child] = P.e0[i];
            P.prev_x[child] = P.prev_x[i];
            P.prev_y[child] = P.prev_y[i];
            P.prev_step_x[child] = P.prev_step_x[i];
            P.prev_step_y[child] = P.prev_step_y[i];
            P.sort_key[child] = P.sort_key[i];
            P.status[child] = 0;
            P.split_level[i] = new_level;
            P.split_level[child] = new_level;
            P.id[child] = split_child_id(P.id[i], old_level);
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

// Frame-interleaved FP32 pair of the two coefficient frames of an interval.
// Node-major record: float 2s + f of node n holds stored slot s (packed
// component pair_component(model, s)) of the lower (f = 0) or upper (f = 1)
// frame; the floats up to the record stride are 0.
__global__ void __launch_bounds__(kBlock) pair_kernel(const float* __restrict__ lo, const float* __restrict__ hi,
                                                      uint64_t plane, int model, float* __restrict__ out) {
    const int stride = pair_stride(model);
    const uint64_t e = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x;
    if (e >= plane * stride) return;
    const uint64_t n = e / stride;
    const int j = static_cast<int>(e - n * stride);
    float v = 0.0f;
    if (j < 2 * pair_components(model))
        v = ((j & 1) ? hi : lo)[static_cast<uint64_t>(pair_component(model, j >> 1)) * plane + n];
    out[e] = v;
}

uint64_t pair_elems(uint64_t plane, int model) { return plane * static_cast<uint64_t>(pair_stride(model)); }

// Single-frame FP32 records: float s of node n's record holds stored slot s
// (packed component pair_component(model, s)); the floats up to the stride are 0.
__global__ void __launch_bounds__(kBlock) rec_kernel(const float* __restrict__ c32, uint64_t plane, int model,
                                                     float* __restrict__ out) {
    const int stride = rec_stride(model);
    const uint64_t e = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x;
    if (e >= plane * stride) return;
    const uint64_t n = e / stride;
    const int j = static_cast<int>(e - n * stride);
    out[e] = j < pair_components(model) ? c32[static_cast<uint64_t>(pair_component(model, j)) * plane + n] : 0.0f;
}

uint64_t rec_elems(uint64_t plane, int model) { return plane * static_cast<uint64_t>(rec_stride(model)); }

hipError_t launch_rec(const float* c32, uint64_t plane, int model, float* out, hipStream_t stream) {
    const uint64_t n = rec_elems(plane, model);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(rec_kernel, dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream,
                       c32, plane, model, out);
    return hipGetLastError();
}

// FP64 pair: double 2k + f of node n's record holds component k of the lower
// (f = 0) or upper (f = 1) frame.
__global__ void __launch_bounds__(kBlock) pair64_kernel(const double* __restrict__ lo, const double* __restrict__ hi,
                                                        uint64_t plane, int model, double* __restrict__ out) {
    const int stride = pair64_stride(model);
    const uint64_t e = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x;
    if (e >= plane * stride) return;
    const uint64_t n = e / stride;
    const int j = static_cast<int>(e - n * stride);
    out[e] = ((j & 1) ? hi : lo)[static_cast<uint64_t>(j >> 1) * plane + n];
}

uint64_t pair64_elems(uint64_t plane, int model) { return plane * static_cast<uint64_t>(pair64_stride(model)); }

hipError_t launch_pair64(const double* lo, const double* hi, uint64_t plane, int model, double* out,
                         hipStream_t stream) {
    const uint64_t n = pair64_elems(plane, model);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pair64_kernel, dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream,
                       lo, hi, plane, model, out);
    return hipGetLastError();
}

hipError_t launch_pair(const float* lo, const float* hi, uint64_t plane, int model, float* out, hipStream_t stream) {
    const uint64_t n = pair_elems(plane, model);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pair_kernel, dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream,
                       lo, hi, plane, model, out);
    return hipGetLastError();
}

template <bool STR>
hipError_t launch_push_t(const PushArgs& a, int model, bool focused, bool fp32, int closure, hipStream_t stream) {
    // 256 runs, 32 per XCD: each XCD gets 32 bands spread over the sorted set,
    // which balances work that varies with position (adaptive dt), and each band
    // is long enough for vertical neighbours to share its L2. On 4 M (Parker) and
    // 8 M (focused) particles of the 1024 x 512 field this measured +26% (Parker
    // FP64) and +6% (focused FP64) on the stretched grid against the plain order,
    // and within 1% on the uniform grid.
    const uint32_t tiles = grid_for(a.count);
    const uint32_t run = std::min(128u, std::max(1u, tiles / 256));
    const dim3 grid((tiles + 8 * run - 1) / (8 * run) * (8 * run)), block(kBlock);
    constexpr int R = KPT_CLOSURE_REDUCED, Q = KPT_CLOSURE_QLT;
    if (model == KPT_MODEL_PARKER16) {
        if (fp32)
            hipLaunchKernelGGL((push_kernel<16, false, float, R, STR>), grid, block, 0, stream, a, run);
        else
            hipLaunchKernelGGL((push_kernel<16, false, double, R, STR>), grid, block, 0, stream, a, run);
    } else if (focused && closure == KPT_CLOSURE_MLP) {
        if (!fp32 || a.mlp_image == nullptr || !mlp_activation_ok(a.mlp_act)) return hipErrorInvalidValue;
        const bool three = a.s.mlp_passes == 3;
        if (a.mlp_act == kActHardSwish) {
            if (three)
                hipLaunchKernelGGL((push_kernel_mlp<STR, kActHardSwish, 3>), grid, block, 0, stream, a, run);
            else
                hipLaunchKernelGGL((push_kernel_mlp<STR, kActHardSwish, 1>), grid, block, 0, stream, a, run);
        } else {
            if (three)
                hipLaunchKernelGGL((push_kernel_mlp<STR, kActSilu, 3>), grid, block, 0, stream, a, run);
            else
                hipLaunchKernelGGL((push_kernel_mlp<STR, kActSilu, 1>), grid, block, 0, stream, a, run);
        }
    } else if (focused && closure == KPT_CLOSURE_QLT) {
        if (fp32)
            hipLaunchKernelGGL((push_kernel<21, true, float, Q, STR>), grid, block, 0, stream, a, run);
        else
            hipLaunchKernelGGL((push_kernel<21, true, double, Q, STR>), grid, block, 0, stream, a, run);
    } else if (focused) {
        if (closure != R) return hipErrorInvalidValue;
        if (fp32)
            hipLaunchKernelGGL((push_kernel<21, true, float, R, STR>), grid, block, 0, stream, a, run);
        else
            hipLaunchKernelGGL((push_kernel<21, true, double, R, STR>), grid, block, 0, stream, a, run);
    } else {
        if (fp32)
            hipLaunchKernelGGL((push_kernel<21, false, float, R, STR>), grid, block, 0, stream, a, run);
        else
            hipLaunchKernelGGL((push_kernel<21, false, double, R, STR>), grid, block, 0, stream, a, run);
    }
    return hipGetLastError();
}

hipError_t launch_push(const PushArgs& a, int model, bool focused, bool fp32, int closure, hipStream_t stream) {
    if (a.count == 0) return hipSuccess;
    if (a.g.face[0] != nullptr || a.g.face[1] != nullptr)
        return launch_push_t<true>(a, model, focused, fp32, closure, stream);
    return launch_push_t<false>(a, model, focused, fp32, closure, stream);
}

hipError_t launch_ensemble_advance(EnsembleRun* runs, int n_runs, EnsembleLoop* loop, unsigned int* queues,
                                   int groups, hipStream_t stream) {
    hipLaunchKernelGGL(ensemble_advance_kernel, dim3(1), dim3(kBlock), 0, stream, runs, n_runs, loop, queues, groups);
    return hipGetLastError();
}

hipError_t launch_ensemble(const EnsembleRun* runs, const uint2* tiles, const uint32_t* group_start, int groups,
                           unsigned int* queues, int grid, const void* mlp_image, int has_mlp, hipStream_t stream) {
    if (groups < 1 || grid < 1 || (has_mlp && mlp_image == nullptr)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ensemble_kernel, dim3(grid), dim3(kBlock), 0, stream, runs, tiles, group_start, groups, queues,
                       mlp_image, has_mlp);
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

hipError_t launch_scan_tiles(uint32_t* tile_counts, uint32_t n, hipStream_t stream) {
    hipLaunchKernelGGL(split_scan_kernel, dim3(1), dim3(1024), 0, stream, tile_counts, n);
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
