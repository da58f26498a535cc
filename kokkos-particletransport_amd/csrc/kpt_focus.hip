// kpt_focus.hip — the generic focused-transport solver on the GPU (FP64): its
// packed 27-component coefficient basis and its global-time-step SDE step with
// the Picard slab-QLT pitch-angle update.
//
// Basis (FocusTransportSolver::build_packed_coefficient_basis,
// include/FocusSolver.hpp:106-136, :497-590, over
// FocusCoefficient::make_legacy_transport_coefficient_fields_with_time_derivative,
// include/FocusCoefficient.hpp:858-990), cell-centred and component-planar:
//   0-2 B, 3-5 b = B / |B|, 6-8 V, 9-11 b x grad ln|B| / |B|,
//   12-14 b x (b.grad)b / |B|, 15-17 div(kappa_perp(gamma=1) (I - bb)), 18 |B|,
//   19 kappa_par(gamma=1), 20 kappa_perp(gamma=1), 21 div V, 22 b.grad ln|B|,
//   23 bb:grad V, 24 b.dV/dt + b.(V.grad)V, 25 sigma2, 26 Lc.
// Every stored node is evaluated at its periodic image from the raw B and V
// planes; derived fields that the reference differentiates (b, ln|B|, the
// kappa tensor) are re-evaluated at the neighbours, which equals its
// ghost-filled intermediate fields on a periodic grid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "kpt_device_common.h"
#include "kpt_generic_device.h"
#include "kpt_grid_device.h"
#include "kpt_internal.h"

namespace kpt {
namespace {

constexpr int kBlk = 256;

// Waves per SIMD the time-step and step kernels are built for (0: the
// compiler's choice). Left to itself the compiler builds the time-step kernel
// with 262 VGPRs, one wave per SIMD; held to 2 waves it takes 256 and a 28-byte
// spill, and the solver step (4 M particles, 1024 x 512 island field) went
// from 18.8 to 16.7 ms. Holding either kernel to 3 or 4 waves spills 228-568
// bytes per lane and measured 1.5-1.9x slower.
#ifndef KPT_FOCUS_DT_WAVES
#define KPT_FOCUS_DT_WAVES 2
#endif
#ifndef KPT_FOCUS_STEP_WAVES
#define KPT_FOCUS_STEP_WAVES 0
#endif
#if KPT_FOCUS_DT_WAVES > 0
#define KPT_FOCUS_DT_ATTR __attribute__((amdgpu_waves_per_eu(KPT_FOCUS_DT_WAVES, KPT_FOCUS_DT_WAVES)))
#else
#define KPT_FOCUS_DT_ATTR
#endif
#if KPT_FOCUS_STEP_WAVES > 0
#define KPT_FOCUS_STEP_ATTR __attribute__((amdgpu_waves_per_eu(KPT_FOCUS_STEP_WAVES, KPT_FOCUS_STEP_WAVES)))
#else
#define KPT_FOCUS_STEP_ATTR
#endif
constexpr int kFC = KPT_GENERIC_FOCUS_COMPONENTS;

// Quantities the focused basis takes from B, V and dV/dt at one physical point.
struct FocusPoint {
    const RawView& r;
    const double* dvdt;  // three planes nx * ny, or null for dV/dt = 0
    double ff;           // floor of b, ln|B| and the drift bases
    int* bad;

    __device__ double bmag(int i, int j) const {
        double s = 0.0;
        for (int c = 0; c < 3; ++c) s += r.at(c, i, j) * r.at(c, i, j);
        return sqrt(s);
    }
    // FieldOperator::fill_direction with ZeroMagnitudePolicy::Abort
    // (include/FieldOperator.hpp:736-778)
    __device__ double dir(int c, int i, int j) const {
        const double m = bmag(i, j);
        if (m <= ff) *bad = 1;
        const double inv = 1.0 / m;
        return r.at(c, i, j) * inv;
    }
    // fill_log_magnetic_field (include/FocusCoefficient.hpp:486-520)
    __device__ double lnb(int i, int j) const {
        const double m = bmag(i, j);
        const double safe = m > ff ? m : ff;
        if (safe <= 0.0) *bad = 1;
        return log(safe);
    }
    __device__ double tdv(int c, int i, int j) const {
        if (dvdt == nullptr) return 0.0;
        return dvdt[static_cast<int64_t>(c) * r.nx * r.ny + static_cast<int64_t>(wrapi(j, r.ny)) * r.nx + wrapi(i, r.nx)];
    }
};

__device__ __forceinline__ double cross_c(const double a[3], const double b[3], int c) {
    return c == 0 ? a[1] * b[2] - a[2] * b[1] : (c == 1 ? a[2] * b[0] - a[0] * b[2] : a[0] * b[1] - a[1] * b[0]);
}

__global__ void __launch_bounds__(kBlk) focus_pack_kernel(const double* __restrict__ raw,
                                                          const double* __restrict__ dvdt, GridDesc g,
                                                          kpt_generic_focus_params p, uint64_t plane,
                                                          double* __restrict__ out, int* __restrict__ bad) {
    const int nxp = g.nx + 2, nyp = g.ny + 2;
    const int64_t node = blockIdx.x * static_cast<int64_t>(kBlk) + threadIdx.x;
    if (node >= static_cast<int64_t>(nxp) * nyp) return;
    const int i = wrapi(static_cast<int>(node % nxp) - 1, g.nx);
    const int j = wrapi(static_cast<int>(node / nxp) - 1, g.ny);
    const RawView r{raw, g.nx, g.ny};
    int flag = 0;
    const FocusPoint fp{r, dvdt, p.minimum_magnetic_field_magnitude, &flag};
    const PointEval e{r, p.model, &flag};
    double b[3], glnb[3], curv[3];
    for (int c = 0; c < 3; ++c) b[c] = fp.dir(c, i, j);
    // FieldOperator::make_gradient_field of ln|B| (Lame coefficients 1)
    for (int d = 0; d < 2; ++d) glnb[d] = deriv(g, i, j, d, [&](int a, int bb) { return fp.lnb(a, bb); });
    glnb[2] = 0.0;
    // magnetic_curvature_component (include/FocusCoefficient.hpp:430-450)
    for (int c = 0; c < 3; ++c) {
        double s = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double grad = d < 2 ? deriv(g, i, j, d, [&](int a, int bb) { return fp.dir(c, a, bb); }) : 0.0;
            s += b[d] * grad;
        }
        curv[c] = s;
    }
    double v[kFC];
    for (int c = 0; c < 3; ++c) {
        v[c] = r.at(c, i, j);
        v[3 + c] = b[c];
        v[6 + c] = r.at(3 + c, i, j);
    }
    // fill_focused_drift_basis_fields (include/FocusCoefficient.hpp:600-660)
    const double m = fp.bmag(i, j);
    const double mm = m > fp.ff ? m : fp.ff;
    if (mm <= 0.0) flag = 1;
    const double inv = 1.0 / mm;
    for (int c = 0; c < 3; ++c) {
        v[9 + c] = cross_c(b, glnb, c) * inv;
        v[12 + c] = cross_c(b, curv, c) * inv;
    }
    // div(kappa_perp (I - bb)) with the diffusion model's |B| floor
    // (include/FocusCoefficient.hpp:840-856 over ParkerCoefficient.hpp:611-640)
    for (int c = 0; c < 3; ++c) {
        double s = 0.0;
        for (int d = 0; d < 2; ++d) s += deriv(g, i, j, d, [&](int a, int bb) { return e.tensor(1, c, d, a, bb); });
        v[15 + c] = s;
    }
    v[18] = m;
    e.kappa(i, j, v[19], v[20]);
    double dv = 0.0;
    for (int d = 0; d < 2; ++d) dv += deriv(g, i, j, d, [&](int a, int bb) { return r.at(3 + d, a, bb); });
    v[21] = dv;
    // fill_focused_transport_contraction_fields (include/FocusCoefficient.hpp:700-778)
    double gv[3][3];
    for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 3; ++d)
            gv[a][d] = d < 2 ? deriv(g, i, j, d, [&](int x, int y) { return r.at(3 + a, x, y); }) : 0.0;
    double foc = 0.0, bbg = 0.0, part = 0.0, conv = 0.0;
    for (int c = 0; c < 3; ++c) foc += b[c] * glnb[c];
    for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 3; ++d) bbg += b[a] * b[d] * gv[a][d];
    if (dvdt != nullptr)
        for (int c = 0; c < 3; ++c) part += b[c] * fp.tdv(c, i, j);
    for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 3; ++d) conv += b[a] * r.at(3 + d, i, j) * gv[a][d];
    v[22] = foc;
    v[23] = bbg;
    v[24] = part + conv;
    v[25] = p.model.sigma2;
    v[26] = p.model.lc > 0.0 ? p.model.lc : p.model.lslab;
    for (int k = 0; k < kFC; ++k) out[k * plane + node] = v[k];
    if (flag) atomicOr(bad, 1);
}

struct FocTerms {
    double b[3], u[2], keff[2];
    double p, mu, gamma, speed, kperp, prate;
    double bmag, divv, bbg, bdv, foc, s2, lc;
    double D, mu_rate;
    double h[2];  // width of the particle's cell along each axis
};

__device__ __forceinline__ double floor_positive(double v, double f) { return f <= 0.0 ? v : (v < f ? f : v); }

// evaluate_pitch_angle_at_mu (include/FocusSolver.hpp:665-694) with
// evaluate_pitch_angle_scattering and the slab-QLT D_mumu, dD_mumu/dmu
// (include/LegencyModel.hpp:539-689).
__device__ __forceinline__ bool focus_pitch(const kpt_generic_focus_params& P, const FocTerms& T, double trial,
                                            double& D, double& rate) {
    const double mu = reflect_mu(trial, P.maximum_pitch_angle_cosine);
    const double omm = 1.0 - mu * mu;
    if (T.speed <= 0.0) return false;
    const double focusing = 0.5 * omm * (-T.speed * T.foc + mu * T.divv - 3.0 * mu * T.bbg - 2.0 * T.bdv / T.speed);
    const double sb = floor_positive(T.bmag, P.scattering_minimum_magnetic_field_magnitude);
    const double ss = floor_positive(T.s2, P.scattering_minimum_sigma2);
    const double sl = floor_positive(T.lc, P.scattering_minimum_correlation_length);
    const double sv = floor_positive(T.speed, P.scattering_minimum_speed);
    if (sb <= 0.0 || ss <= 0.0 || sl <= 0.0 || sv <= 0.0) return false;
    const double omega = fabs(P.model.charge) * sb / (P.model.rest_mass * P.model.speed_of_light) / T.gamma;
    double dd = 0.0;
    D = 0.0;
    if (omm > 0.0) {
        const double pi = 3.14159265358979323846;
        const double gk = P.spectral_index;
        const double mu_h = fabs(mu) + P.resonance_regularization_h0;
        const double xi = omega * sl / sv;
        const double den = pow(mu_h, gk) + pow(xi, gk);
        const double norm = gk * sin(pi / gk) / pi;
        const double pm1 = pow(mu_h, gk - 1.0);
        D = 0.25 * pi * omega * ss * norm * omm * xi * pm1 / den;
        if (D != 0.0 && mu != 0.0) {
            const double sign_mu = mu > 0.0 ? 1.0 : -1.0;
            const double shape = sign_mu * ((gk - 1.0) / mu_h - gk * pm1 / den);
            dd = D * (-2.0 * mu / omm + shape);
        }
    }
    rate = focusing + dd;
    return true;
}

// evaluate_terms_from_packed (include/FocusSolver.hpp:905-1033) at a wrapped position.
__device__ __forceinline__ bool focus_terms(const double* __restrict__ f, uint64_t plane, const GridDesc& g,
                                            const kpt_generic_focus_params& P, double x, double y, double p_in,
                                            double mu_in, FocTerms& T) {
    const kpt_generic_parker_params& M = P.model;
    int i0, j0;
    double wx, wy;
    if (!cell_stencil(x, g, 0, g.nx, i0, wx, T.h[0]) || !cell_stencil(y, g, 1, g.ny, j0, wy, T.h[1])) return false;
    const int nxp = g.nx + 2;
    const int64_t n00 = static_cast<int64_t>(j0 + 1) * nxp + (i0 + 1);
    const double w00 = (1.0 - wx) * (1.0 - wy), w10 = wx * (1.0 - wy), w01 = (1.0 - wx) * wy, w11 = wx * wy;
    double v[kFC];
#pragma unroll
    for (int k = 0; k < kFC; ++k) {
        const double* q = f + k * plane + n00;
        v[k] = w00 * q[0] + w10 * q[1] + w01 * q[nxp] + w11 * q[nxp + 1];
    }
    T.p = fabs(p_in);
    T.mu = reflect_mu(mu_in, P.maximum_pitch_angle_cosine);
    const double nm = T.p / (M.rest_mass * M.speed_of_light);
    T.gamma = sqrt(1.0 + nm * nm);
    T.speed = T.p / (T.gamma * M.rest_mass);
    T.bmag = v[18];
    if (!(isfinite(T.bmag) && T.bmag > 0.0)) return false;
    const double dm = sqrt(v[3] * v[3] + v[4] * v[4] + v[5] * v[5]);
    if (!(isfinite(dm) && dm > 0.0)) return false;
    const double idm = 1.0 / dm;
    for (int c = 0; c < 3; ++c) T.b[c] = v[3 + c] * idm;
    const double fperp = M.perpendicular_model == 1 ? pow(T.gamma, 1.0 / 3.0) : pow(T.gamma, 1.0 / 9.0);
    const double kpar = v[19] * pow(T.gamma, 1.0 / 3.0);
    T.kperp = v[20] * fperp;
    if (!(isfinite(kpar) && kpar >= 0.0 && isfinite(T.kperp) && T.kperp >= 0.0)) return false;
    const double mu2 = T.mu * T.mu;
    const double gdf = 0.5 * (1.0 - mu2), cdf = mu2;
    const double pref = T.p * T.speed * M.speed_of_light / M.charge;
    for (int d = 0; d < 2; ++d) {
        const double stream = T.speed * T.mu * T.b[d];
        const double drift = pref * (gdf * v[9 + d] + cdf * v[12 + d]);
        const double dadv = fperp * v[15 + d];
        T.u[d] = stream + v[6 + d] + drift + dadv;
        const double proj = 1.0 - T.b[d] * T.b[d];
        T.keff[d] = proj > 0.0 ? T.kperp * proj : 0.0;
    }
    T.divv = v[21];
    T.bbg = v[23];
    T.bdv = v[24];
    T.foc = v[22];
    T.s2 = v[25];
    T.lc = v[26];
    const double omm = 1.0 - mu2;
    const double accel = T.speed > 0.0 ? T.mu * T.bdv / T.speed : 0.0;
    const double mf = 0.5 * omm * (T.divv - T.bbg) + mu2 * T.bbg + accel;
    T.prate = -T.p * mf;
    return focus_pitch(P, T, T.mu, T.D, T.mu_rate);
}

// local_stable_time_step (include/FocusSolver.hpp:1149-1232): diffusion,
// advection, momentum and the three pitch-angle bounds (relaxed by 10).
__device__ __forceinline__ double focus_local_dt(const FocTerms& T) {
    const double large = 1.0e300;
    double ddt = large, adt = large;
    for (int d = 0; d < 2; ++d) {
        const double dx = T.h[d];
        double di = large, ai = large;
        if (T.keff[d] > 0.0) {
            const double lim = 0.5 * dx;
            di = lim * lim / (2.0 * T.keff[d]);
        }
        const double a = fabs(T.u[d]);
        if (a > 0.0) ai = 1.0 * dx / a;
        ddt = ddt < di ? ddt : di;
        adt = adt < ai ? adt : ai;
    }
    double mdt = large, padt = large, pddt = large, pbdt = large;
    const double pr = fabs(T.prate);
    if (T.p > 0.0 && pr > 0.0) mdt = 1.0 * T.p / pr;
    const double mr = fabs(T.mu_rate);
    if (mr > 0.0) padt = 10.0 * 0.1 / mr;
    if (T.D > 0.0) pddt = 10.0 * 0.1 * 0.1 / (2.0 * T.D);
    if (T.D > 0.0 && mr > 0.0) pbdt = 10.0 * 2.0 * T.D / (mr * mr);
    const double a = ddt < adt ? ddt : adt;
    const double b = a < mdt ? a : mdt;
    const double c = pddt < pbdt ? pddt : pbdt;
    const double e = padt < c ? padt : c;
    return b < e ? b : e;
}

__global__ void __launch_bounds__(kBlk) KPT_FOCUS_DT_ATTR focus_dt_kernel(ParticleArrays P, uint64_t count,
                                                        const double* __restrict__ f, uint64_t plane, GridDesc g,
                                                        kpt_generic_focus_params p,
                                                        unsigned long long* __restrict__ min_bits) {
    double m = 1.0e300;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x; i < count;
         i += static_cast<uint64_t>(gridDim.x) * kBlk) {
        if (P.status[i] != 0) continue;
        FocTerms T;
        if (focus_terms(f, plane, g, p, wrap_axis(g, 0, P.x[i]), wrap_axis(g, 1, P.y[i]), P.p[i], P.mu[i], T)) {
            const double l = focus_local_dt(T);
            m = l < m ? l : m;
        }
    }
    block_min_to(m, min_bits);
}

__global__ void __launch_bounds__(kBlk) KPT_FOCUS_STEP_ATTR focus_step_kernel(ParticleArrays P, uint64_t count,
                                                          const double* __restrict__ f, uint64_t plane, GridDesc g,
                                                          kpt_generic_focus_params p, double dt, uint32_t step,
                                                          uint32_t k0, uint32_t k1, const double* __restrict__ noise,
                                                          uint32_t noise_steps, int deterministic,
                                                          unsigned long long* __restrict__ failures) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    uint32_t fail = 0;
    if (i < count) {
        // apply_grid_boundary_conditions (periodic wrap of alive particles), then
        // capture_previous_positions (include/ParticleSystem.hpp:623-628) for
        // every stored slot, as the solver's advance does before its update
        const bool alive = P.status[i] == 0;
        double x = P.x[i], y = P.y[i];
        if (alive) {
            x = wrap_axis(g, 0, x);
            y = wrap_axis(g, 1, y);
            P.x[i] = x;
            P.y[i] = y;
        }
        P.prev_step_x[i] = P.prev_x[i];
        P.prev_step_y[i] = P.prev_y[i];
        P.prev_x[i] = x;
        P.prev_y[i] = y;
    }
    if (i < count && P.status[i] == 0) {
        const double x = P.x[i], y = P.y[i];
        FocTerms T;
        if (!focus_terms(f, plane, g, p, x, y, P.p[i], P.mu[i], T)) {
            P.status[i] = 2;
            fail = 1;
        } else {
            double n[4] = {0.0, 0.0, 0.0, 0.0};  // advance_deterministic: no increments
            if (!deterministic) {
                if (noise != nullptr) {
                    const double* row = noise + (i * noise_steps + step) * 4;
                    for (int d = 0; d < 4; ++d) n[d] = row[d];
                } else {
                    philox_normals4(P.id[i], 0x464F4355u, step, k0, k1, n);
                }
            }
            // sample_perpendicular_diffusion_increment (include/FocusSolver.hpp:635-659)
            double par = 0.0;
            for (int c = 0; c < 3; ++c) par += n[c] * T.b[c];
            const double scale = T.kperp > 0.0 ? sqrt(2.0 * T.kperp * dt) : 0.0;
            const double nx = x + (dt * T.u[0] + scale * (n[0] - par * T.b[0]));
            const double ny = y + (dt * T.u[1] + scale * (n[1] - par * T.b[1]));
            P.x[i] = wrap_axis(g, 0, nx);
            P.y[i] = wrap_axis(g, 1, ny);
            const double pn = T.p + dt * T.prate;
            P.p[i] = pn > 0.0 ? pn : 0.0;
            // picard_pitch_angle_update (include/FocusSolver.hpp:700-719): three
            // passes with the same normal
            const double mu0 = reflect_mu(T.mu, p.maximum_pitch_angle_cosine);
            double trial = mu0;
            for (int it = 0; it < 3; ++it) {
                double D = 0.0, rate = 0.0;
                focus_pitch(p, T, trial, D, rate);  // the inputs that can fail were checked at T.mu
                const double sc = D > 0.0 ? sqrt(2.0 * D * dt) : 0.0;
                trial = reflect_mu(mu0 + dt * rate + sc * n[3], p.maximum_pitch_angle_cosine);
            }
            P.mu[i] = trial;
        }
    }
    const unsigned int w = wave_sum(fail);
    if ((threadIdx.x & 63) == 0 && w) atomicAdd(failures, static_cast<unsigned long long>(w));
}

inline unsigned blocks(uint64_t n) { return static_cast<unsigned>((n + kBlk - 1) / kBlk); }

}  // namespace

hipError_t launch_focus_pack(const double* raw, const double* dvdt, const GridDesc& g,
                             const kpt_generic_focus_params& p, double* out, int* bad, hipStream_t stream) {
    const uint64_t plane = coefficient_plane(g.nx, g.ny);
    hipLaunchKernelGGL(focus_pack_kernel, dim3(blocks(plane)), dim3(kBlk), 0, stream, raw, dvdt, g, p, plane, out,
                       bad);
    return hipGetLastError();
}

hipError_t launch_focus_min_dt(const ParticleArrays& P, uint64_t count, const double* f, const GridDesc& g,
                               const kpt_generic_focus_params& p, unsigned long long* min_bits, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const unsigned grid = static_cast<unsigned>(std::min<uint64_t>(blocks(count), kMinDtGrid));
    hipLaunchKernelGGL(focus_dt_kernel, dim3(grid), dim3(kBlk), 0, stream, P, count, f,
                       coefficient_plane(g.nx, g.ny), g, p, min_bits);
    return hipGetLastError();
}

hipError_t launch_focus_step(const ParticleArrays& P, uint64_t count, const double* f, const GridDesc& g,
                             const kpt_generic_focus_params& p, double dt, uint32_t step, const uint32_t key[2],
                             const double* noise, uint32_t noise_steps, int deterministic, unsigned long long* failures,
                             hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(focus_step_kernel, dim3(blocks(count)), dim3(kBlk), 0, stream, P, count, f,
                       coefficient_plane(g.nx, g.ny), g, p, dt, step, key[0], key[1], noise, noise_steps, deterministic, failures);
    return hipGetLastError();
}

}  // namespace kpt
