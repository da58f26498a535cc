// kpt_generic.hip — the generic Parker solver on the GPU (FP64): its packed
// coefficient basis and its global-time-step stochastic step.
//
// Basis (ParkerSolver::build_packed_coefficient_basis, include/ParkerSolver.hpp:
// 452-640, over ParkerCoefficient::make_legacy_transport_coefficient_fields,
// include/ParkerCoefficient.hpp:1090-1140), cell-centred, component-planar like
// the driver frames (plane k, nodes i in [-1, nx], j in [-1, ny]):
//   0-2 B, 3-5 V, 6-8 curl(B / B^2), 9-11 primary and 12-14 secondary
//   div(kappa_ij) basis, 15 kappa_par(gamma = 1), 16 kappa_perp(gamma = 1), 17 div V.
// Each stored node is computed at its periodic image in the physical domain, so
// ghost nodes equal the reference's periodic ghost fill exactly. The grid is the
// driver's uniform Cartesian one (Lame coefficients 1, Jacobian 1); z is the
// invariant direction.
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
// compiler's choice: 211 and 164 VGPRs, 2 and 3 waves). Held to 3 or 4 waves
// they spill 148-300 bytes per lane and the solver step (4 M particles,
// 1024 x 512 island field) measured 1.3-2.3x slower.
#ifndef KPT_GENERIC_DT_WAVES
#define KPT_GENERIC_DT_WAVES 0
#endif
#ifndef KPT_GENERIC_STEP_WAVES
#define KPT_GENERIC_STEP_WAVES 0
#endif
#if KPT_GENERIC_DT_WAVES > 0
#define KPT_GENERIC_DT_ATTR __attribute__((amdgpu_waves_per_eu(KPT_GENERIC_DT_WAVES, KPT_GENERIC_DT_WAVES)))
#else
#define KPT_GENERIC_DT_ATTR
#endif
#if KPT_GENERIC_STEP_WAVES > 0
#define KPT_GENERIC_STEP_ATTR __attribute__((amdgpu_waves_per_eu(KPT_GENERIC_STEP_WAVES, KPT_GENERIC_STEP_WAVES)))
#else
#define KPT_GENERIC_STEP_ATTR
#endif
constexpr int kGC = KPT_GENERIC_PARKER_COMPONENTS;

__global__ void __launch_bounds__(kBlk) generic_pack_kernel(const double* __restrict__ raw, GridDesc g,
                                                            kpt_generic_parker_params p, uint64_t plane,
                                                            double* __restrict__ out, int* __restrict__ bad) {
    const int nxp = g.nx + 2, nyp = g.ny + 2;
    const int64_t node = blockIdx.x * static_cast<int64_t>(kBlk) + threadIdx.x;
    if (node >= static_cast<int64_t>(nxp) * nyp) return;
    const int i = wrapi(static_cast<int>(node % nxp) - 1, g.nx);
    const int j = wrapi(static_cast<int>(node / nxp) - 1, g.ny);
    const RawView r{raw, g.nx, g.ny};
    int flag = 0;
    const PointEval e{r, p, &flag};
    double v[kGC];
    for (int c = 0; c < 3; ++c) {
        v[c] = r.at(c, i, j);
        v[3 + c] = r.at(3 + c, i, j);
    }
    for (int c = 0; c < 3; ++c) {
        const int da = (c + 1) % 3, db = (c + 2) % 3;
        v[6 + c] = deriv(g, i, j, da, [&](int a, int b) { return e.over_b2(db, a, b); }) -
                   deriv(g, i, j, db, [&](int a, int b) { return e.over_b2(da, a, b); });
    }
    const bool ratio = p.perpendicular_model == 1;
    const int b0 = ratio ? 2 : 0;
    for (int c = 0; c < 3; ++c) {
        double prim = 0.0, sec = 0.0;
        for (int d = 0; d < 2; ++d) {
            prim += deriv(g, i, j, d, [&](int a, int b) { return e.tensor(b0, c, d, a, b); });
            if (!ratio) sec += deriv(g, i, j, d, [&](int a, int b) { return e.tensor(1, c, d, a, b); });
        }
        v[9 + c] = prim;
        v[12 + c] = ratio ? prim : sec;
    }
    e.kappa(i, j, v[15], v[16]);
    double s = 0.0;
    for (int d = 0; d < 2; ++d) s += deriv(g, i, j, d, [&](int a, int b) { return r.at(3 + d, a, b); });
    v[17] = s;
    for (int k = 0; k < kGC; ++k) out[k * plane + node] = v[k];
    if (flag) atomicOr(bad, 1);
}

struct GenTerms {
    double u[2], b[3], kpar, kperp, keff[2], divv, p, prate;
    double h[2];  // width of the particle's cell along each axis
};

// evaluate_terms_from_packed (ParkerSolver.hpp:771-875) at a wrapped position.
__device__ __forceinline__ bool generic_terms(const double* __restrict__ f, uint64_t plane, const GridDesc& g,
                                              const kpt_generic_parker_params& P, double x, double y, double p_in,
                                              GenTerms& T) {
    int i0, j0;
    double wx, wy;
    if (!cell_stencil(x, g, 0, g.nx, i0, wx, T.h[0]) || !cell_stencil(y, g, 1, g.ny, j0, wy, T.h[1])) return false;
    const int nxp = g.nx + 2;
    const int64_t n00 = static_cast<int64_t>(j0 + 1) * nxp + (i0 + 1);
    const double w00 = (1.0 - wx) * (1.0 - wy), w10 = wx * (1.0 - wy), w01 = (1.0 - wx) * wy, w11 = wx * wy;
    double v[kGC];
#pragma unroll
    for (int k = 0; k < kGC; ++k) {
        const double* q = f + k * plane + n00;
        v[k] = w00 * q[0] + w10 * q[1] + w01 * q[nxp] + w11 * q[nxp + 1];
    }
    const double pm = fabs(p_in);
    const double nm = pm / (P.rest_mass * P.speed_of_light);
    const double gamma = sqrt(1.0 + nm * nm);
    const double speed = pm / (gamma * P.rest_mass);
    const double bm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (!(isfinite(bm) && bm > 0.0)) return false;
    const double ib = 1.0 / bm;
    for (int c = 0; c < 3; ++c) T.b[c] = v[c] * ib;
    const double fpar = pow(gamma, 1.0 / 3.0);
    const bool ratio = P.perpendicular_model == 1;
    const double fperp = ratio ? fpar : pow(gamma, 1.0 / 9.0);
    T.kpar = v[15] * fpar;
    T.kperp = v[16] * fperp;
    if (!(isfinite(T.kpar) && T.kpar >= 0.0 && isfinite(T.kperp) && T.kperp >= 0.0)) return false;
    for (int d = 0; d < 2; ++d) {
        T.keff[d] = T.kperp + (T.kpar - T.kperp) * T.b[d] * T.b[d];
        if (!(isfinite(T.keff[d]) && T.keff[d] >= 0.0)) return false;
    }
    T.divv = v[17];
    const double drift = pm * speed * P.speed_of_light / (3.0 * P.charge);
    for (int d = 0; d < 2; ++d) {
        const double adv = ratio ? fpar * v[9 + d] : fpar * v[9 + d] + fperp * v[12 + d];
        T.u[d] = v[3 + d] + drift * v[6 + d] + adv;
    }
    T.p = pm;
    T.prate = -(pm / 3.0) * T.divv;
    return true;
}

// local_stable_time_step (ParkerSolver.hpp:956-1007) with the width of the
// particle's cell.
__device__ __forceinline__ double generic_local_dt(const GenTerms& T) {
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
    double mdt = large;
    const double dvm = fabs(T.divv);
    if (dvm > 0.0) mdt = 3.0 * 1.0 / dvm;
    const double m = ddt < adt ? ddt : adt;
    return m < mdt ? m : mdt;
}

__global__ void __launch_bounds__(kBlk) KPT_GENERIC_DT_ATTR generic_dt_kernel(ParticleArrays P, uint64_t count,
                                                          const double* __restrict__ f, uint64_t plane, GridDesc g,
                                                          kpt_generic_parker_params p,
                                                          unsigned long long* __restrict__ min_bits) {
    double m = 1.0e300;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x; i < count;
         i += static_cast<uint64_t>(gridDim.x) * kBlk) {
        if (P.status[i] != 0) continue;
        GenTerms T;
        if (generic_terms(f, plane, g, p, wrap_axis(g, 0, P.x[i]), wrap_axis(g, 1, P.y[i]), P.p[i], T)) {
            const double l = generic_local_dt(T);
            m = l < m ? l : m;
        }
    }
    block_min_to(m, min_bits);
}

__global__ void __launch_bounds__(kBlk) KPT_GENERIC_STEP_ATTR generic_step_kernel(ParticleArrays P, uint64_t count,
                                                            const double* __restrict__ f, uint64_t plane, GridDesc g,
                                                            kpt_generic_parker_params p, double dt, uint32_t step,
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
        GenTerms T;
        if (!generic_terms(f, plane, g, p, x, y, P.p[i], T)) {
            P.status[i] = 2;
            fail = 1;
        } else {
            double n[4] = {0.0, 0.0, 0.0, 0.0};  // advance_deterministic: no increments
            if (!deterministic) {
                if (noise != nullptr) {
                    const double* row = noise + (i * noise_steps + step) * 4;
                    for (int d = 0; d < 4; ++d) n[d] = row[d];
                } else {
                    philox_normals4(P.id[i], 0x47454E50u, step, k0, k1, n);
                }
            }
            // sample_magnetic_diffusion_increment (ParkerSolver.hpp:443-477)
            double iso = 0.0;
            for (int c = 0; c < 3; ++c) iso += n[c] * T.b[c];
            const double sperp = T.kperp > 0.0 ? sqrt(2.0 * T.kperp * dt) : 0.0;
            const double spar = T.kpar > 0.0 ? sqrt(2.0 * T.kpar * dt) : 0.0;
            const double repl = spar * n[3] - sperp * iso;
            double pos[2] = {x, y};
            for (int d = 0; d < 2; ++d) {
                const double det = dt * T.u[d];
                const double sto = sperp * n[d] + repl * T.b[d];
                pos[d] += det + sto;
            }
            P.x[i] = wrap_axis(g, 0, pos[0]);
            P.y[i] = wrap_axis(g, 1, pos[1]);
            const double pn = T.p + dt * T.prate;
            P.p[i] = pn > 0.0 ? pn : 0.0;
        }
    }
    const unsigned int w = wave_sum(fail);
    if ((threadIdx.x & 63) == 0 && w) atomicAdd(failures, static_cast<unsigned long long>(w));
}

inline unsigned blocks(uint64_t n) { return static_cast<unsigned>((n + kBlk - 1) / kBlk); }

}  // namespace

hipError_t launch_generic_pack(const double* raw, const GridDesc& g, const kpt_generic_parker_params& p,
                               double* out, int* bad, hipStream_t stream) {
    const uint64_t plane = coefficient_plane(g.nx, g.ny);
    hipLaunchKernelGGL(generic_pack_kernel, dim3(blocks(plane)), dim3(kBlk), 0, stream, raw, g, p, plane, out, bad);
    return hipGetLastError();
}

hipError_t launch_generic_min_dt(const ParticleArrays& P, uint64_t count, const double* f, const GridDesc& g,
                                 const kpt_generic_parker_params& p, unsigned long long* min_bits,
                                 hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const unsigned grid = static_cast<unsigned>(std::min<uint64_t>(blocks(count), kMinDtGrid));
    hipLaunchKernelGGL(generic_dt_kernel, dim3(grid), dim3(kBlk), 0, stream, P, count, f,
                       coefficient_plane(g.nx, g.ny), g, p, min_bits);
    return hipGetLastError();
}

hipError_t launch_generic_step(const ParticleArrays& P, uint64_t count, const double* f, const GridDesc& g,
                               const kpt_generic_parker_params& p, double dt, uint32_t step, const uint32_t key[2],
                               const double* noise, uint32_t noise_steps, int deterministic, unsigned long long* failures,
                               hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(generic_step_kernel, dim3(blocks(count)), dim3(kBlk), 0, stream, P, count, f,
                       coefficient_plane(g.nx, g.ny), g, p, dt, step, key[0], key[1], noise, noise_steps, deterministic, failures);
    return hipGetLastError();
}

}  // namespace kpt
