// kpt_generic_device.h — device helpers shared by the generic Parker and
// focused-transport solvers: periodic index wrap, cell-centre coordinates, the
// reference's three-point derivative, point evaluations of B-derived quantities
// on the raw B and V planes, and the Box-Muller normals of one Philox block.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "kpt_device_common.h"
#include "kpt_grid_device.h"
#include "kpt_internal.h"

namespace kpt {

__device__ __forceinline__ int wrapi(int i, int n) {
    const int w = i % n;
    return w < 0 ? w + n : w;
}

// Cell-centre coordinate (AnalyticCoordinateGrid, include/AnalyticCoordinateGrid.hpp:159-176;
// on a stretched axis StoredCoordinateGrid's (e_k + e_{k+1}) / 2 of the periodic faces,
// include/StoredCoordinateGrid.hpp:411-425), k in [-2, n].
__device__ __forceinline__ double cell_coord(const GridDesc& g, int d, int k) {
    if (g.face[d] != nullptr) {
        const int n = d == 0 ? g.nx : g.ny;
        return 0.5 * (stretched_node(g, d, n, k) + stretched_node(g, d, n, k + 1));
    }
    return 0.5 * ((g.start[d] + k * g.step[d]) + (g.start[d] + (k + 1.0) * g.step[d]));
}


struct RawView {
    const double* raw;  // 6 planes of nx * ny (Bx, By, Bz, Vx, Vy, Vz), x fastest
    int nx, ny;
    __device__ double at(int c, int i, int j) const {
        return raw[static_cast<int64_t>(c) * nx * ny + static_cast<int64_t>(wrapi(j, ny)) * nx + wrapi(i, nx)];
    }
};

struct PointEval {
    const RawView& r;
    const kpt_generic_parker_params& p;
    int* bad;

    __device__ double bmag(int i, int j) const {
        double s = 0.0;
        for (int c = 0; c < 3; ++c) s += r.at(c, i, j) * r.at(c, i, j);
        return sqrt(s);
    }
    // magnetic_direction_component (ParkerCoefficient.hpp:255-270)
    __device__ double bdir(int c, int i, int j) const {
        const double m = bmag(i, j);
        const double f = p.minimum_magnetic_field_magnitude;
        if (m <= 0.0 && f <= 0.0) *bad = 1;
        return r.at(c, i, j) / (m > f ? m : f);
    }
    // magnetic_field_over_magnitude_squared_component (ParkerCoefficient.hpp:279-306)
    __device__ double over_b2(int c, int i, int j) const {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += r.at(k, i, j) * r.at(k, i, j);
        const double f = p.minimum_magnetic_field_squared_for_drift;
        if (s <= 0.0 && f <= 0.0) *bad = 1;
        return r.at(c, i, j) / (s > f ? s : f);
    }
    // evaluate_gamma_independent_coefficients (LegencyModel.hpp:430-470) with the
    // constant turbulence state of TurbulencePropertyModel::evaluate_constant.
    __device__ void kappa(int i, int j, double& kpar, double& kperp) const {
        auto floor_pos = [](double v, double f) { return f <= 0.0 ? v : (v < f ? f : v); };
        double vs = 0.0;
        for (int c = 3; c < 6; ++c) vs += r.at(c, i, j) * r.at(c, i, j);
        const double b = floor_pos(bmag(i, j), p.minimum_magnetic_field_magnitude);
        const double v = floor_pos(sqrt(vs), p.minimum_solar_wind_speed);
        const double s2 = floor_pos(p.sigma2, p.minimum_sigma2);
        const double lc = p.lc > 0.0 ? p.lc : p.lslab;
        if (b <= 0.0) *bad = 1;
        const double omega = fabs(p.charge) * b / (p.rest_mass * p.speed_of_light);
        if (v <= 0.0 || lc <= 0.0 || omega <= 0.0 || s2 <= 0.0) *bad = 1;
        kpar = p.parallel_prefactor * pow(v, 4.0 / 3.0) * pow(lc, 2.0 / 3.0) / (pow(omega, 1.0 / 3.0) * s2);
        if (p.perpendicular_model == 1) {
            kperp = p.perpendicular_parallel_ratio * kpar;
        } else {
            if (p.l2d <= 0.0 || kpar <= 0.0) *bad = 1;
            kperp = (v / 3.0) * pow(3.0 * kpar / v, 1.0 / 3.0) * pow(p.perpendicular_prefactor * s2 * p.l2d, 2.0 / 3.0);
        }
    }
    // diffusion_tensor_basis_component (ParkerCoefficient.hpp:611-640):
    // basis 0 parallel, 1 perpendicular, 2 constant ratio
    __device__ double tensor(int basis, int a, int b, int i, int j) const {
        const double ba = bdir(a, i, j), bb = bdir(b, i, j);
        const double proj = ba * bb, id = a == b ? 1.0 : 0.0;
        double kpar, kperp;
        kappa(i, j, kpar, kperp);
        if (basis == 0) return kpar * proj;
        if (basis == 1) return kperp * (id - proj);
        const double rr = p.perpendicular_parallel_ratio;
        return kpar * (rr * id + (1.0 - rr) * proj);
    }
};

// Central derivative along dim of f(i, j) at physical (i, j); z -> 0.
template <typename F>
__device__ __forceinline__ double deriv(const GridDesc& g, int i, int j, int dim, F f) {
    if (dim >= 2 || (dim == 0 ? g.nx : g.ny) <= 1) return 0.0;
    const int c = dim == 0 ? i : j;
    const double fl = dim == 0 ? f(i - 1, j) : f(i, j - 1);
    const double fu = dim == 0 ? f(i + 1, j) : f(i, j + 1);
    return three_point(cell_coord(g, dim, c - 1), cell_coord(g, dim, c), cell_coord(g, dim, c + 1), fl, f(i, j), fu);
}

// MappedCoordinateGrid::wrap_coordinate for the physical domain of axis d.
__device__ __forceinline__ double wrap_axis(const GridDesc& g, int d, double x) {
    double w = x - floor((x - g.phys_lo[d]) * g.inv_width[d]) * g.width[d];
    if (w >= g.phys_hi[d]) w -= g.width[d];
    if (w < g.phys_lo[d]) w += g.width[d];
    return w;
}


// Four standard normals from one Philox4x32-10 block with counter
// (id_lo, id_hi, tag, step): Box-Muller on (r0, r1) and (r2, r3).
__device__ __forceinline__ void philox_normals4(uint64_t id, uint32_t tag, uint32_t step, uint32_t k0, uint32_t k1,
                                                double n[4]) {
    uint32_t r[4];
    philox4x32_10(static_cast<uint32_t>(id), static_cast<uint32_t>(id >> 32), tag, step, k0, k1, r);
    for (int k = 0; k < 2; ++k) {
        const double rad = sqrt(-2.0 * log(u01(r[2 * k])));
        const double ang = 6.283185307179586476925286766559 * u01(r[2 * k + 1]);
        n[2 * k] = rad * cos(ang);
        n[2 * k + 1] = rad * sin(ang);
    }
}

// Workgroup-wide minimum of a positive double, then one atomicMin on its bit
// pattern per workgroup: positive doubles order like their bits, so the result
// is exact and independent of the order of the atomics. (One atomic per
// wavefront on the single counter serialises in the L2.) Every thread of the
// workgroup must call it.
__device__ __forceinline__ void block_min_to(double m, unsigned long long* min_bits) {
    __shared__ double part[16];  // up to 1024 threads
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = o < m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (unsigned w = 1; w < (blockDim.x + 63) / 64; ++w) m = part[w] < m ? part[w] : m;
        atomicMin(min_bits, static_cast<unsigned long long>(__double_as_longlong(m)));
    }
}

// Workgroups of the dt reductions: a grid-stride loop over at most this many.
constexpr uint64_t kMinDtGrid = 2048;

}  // namespace kpt
