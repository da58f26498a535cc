// kpt_grid_device.h — periodic wrap and bilinear stencil on the uniform grid,
// shared by the driver push and the generic solver.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "kpt_internal.h"

namespace kpt {

// MappedCoordinateGrid::wrap_coordinate (include/MappedCoordinateGrid.hpp:295-315).
__device__ __forceinline__ double grid_wrap(double x, double lo, double hi, double width,
                                            double inv_width) {
    double w = x - floor((x - lo) * inv_width) * width;
    if (w >= hi) w -= width;
    if (w < lo) w += width;
    return w;
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

// centered_three_point_derivative (include/ParkerCoefficient.hpp:215-233).
__device__ __forceinline__ double three_point(double xl, double xc, double xu, double fl, double fc, double fu) {
    const double lw = xc - xl, uw = xu - xc, tw = xu - xl;
    return (-uw / (lw * tw)) * fl + ((uw - lw) / (lw * uw)) * fc + (lw / (uw * tw)) * fu;
}

// Stencil of the driver's node convention. Uniform axis: value k sits at
// start + k step, reached through the +step/2 interpolation shift and the
// cell-centred inverse map (axis_stencil). Stretched axis (no counterpart in the
// driver, whose reader rejects non-uniform faces): value k sits at face k and q
// (the shift is 0 there) falls in the face interval [face k, face k+1), read
// from one bin record (interval k or k + 1 of the bin). `h` receives the width
// of the interval (the uniform step on a uniform axis).
__device__ __forceinline__ bool node_stencil(double q, const GridDesc& g, int d, int n, int& i0, double& w,
                                             double& h) {
    if (g.face[d] == nullptr) {
        h = g.step[d];
        return axis_stencil(q, g, d, n, i0, w);
    }
    const double qw = grid_wrap(q, g.phys_lo[d], g.phys_hi[d], g.width[d], g.inv_width[d]);
    if (!(qw == qw)) return false;
    const double bf = (qw - g.start[d]) * g.bins_scale[d];
    const int b = bf < 0.0 ? 0 : (bf >= static_cast<double>(g.bins_n[d] - 1) ? g.bins_n[d] - 1 : static_cast<int>(bf));
    const FaceBin fb = g.bins[d][b];
    int k = fb.k;
    double el = fb.lo, eu = fb.mid;
    if (qw >= fb.mid && k + 1 < n) {
        ++k;
        el = fb.mid;
        eu = fb.hi;
    }
    h = eu - el;
    double ww = (qw - el) / h;
    ww = ww < 0.0 ? 0.0 : (ww > 1.0 ? 1.0 : ww);
    i0 = k;
    w = ww;
    return true;
}

// Coordinate of stored node k (k in [-2, n + 1]) along a stretched axis: the
// face of its periodic image shifted by whole periods.
__device__ __forceinline__ double stretched_node(const GridDesc& g, int d, int n, int k) {
    const int m = ((k % n) + n) % n;
    const int period = (k - m) / n;
    return g.face[d][m] + static_cast<double>(period) * g.width[d];
}

// The cell-centred stencil of the generic solvers (no interpolation shift), with
// `h` the width of the cell containing q (the physical_cell_width of their dt
// bound). Uniform axis: the analytic inverse map. Stretched axis: cell centres
// c_k = (e_k + e_{k+1}) / 2 of StoredCoordinateGrid (include/
// StoredCoordinateGrid.hpp:411-425); q lies in face interval m (one bin record),
// and the centre interval is [c_m, c_{m+1}) or [c_{m-1}, c_m) by the side of c_m,
// as the reference's bisection over cell coordinates finds it (:128-173).
__device__ __forceinline__ bool cell_stencil(double q, const GridDesc& g, int d, int n, int& i0, double& w,
                                             double& h) {
    if (g.face[d] == nullptr) {
        h = fabs(g.step[d]);
        return axis_stencil(q, g, d, n, i0, w);
    }
    int m;
    double wf;
    if (!node_stencil(q, g, d, n, m, wf, h)) return false;
    const double qw = grid_wrap(q, g.phys_lo[d], g.phys_hi[d], g.width[d], g.inv_width[d]);
    const double el = stretched_node(g, d, n, m), eu = stretched_node(g, d, n, m + 1);
    const double cm = 0.5 * (el + eu);
    double clo, chi;
    if (qw >= cm) {
        i0 = m;
        clo = cm;
        chi = 0.5 * (eu + stretched_node(g, d, n, m + 2));
    } else {
        i0 = m - 1;
        clo = 0.5 * (stretched_node(g, d, n, m - 1) + el);
        chi = cm;
    }
    double ww = (qw - clo) / (chi - clo);
    w = ww < 0.0 ? 0.0 : (ww > 1.0 ? 1.0 : ww);
    return true;
}

}  // namespace kpt
