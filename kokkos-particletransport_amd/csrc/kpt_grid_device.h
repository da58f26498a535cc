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

// The cell-centred stencil of the generic solvers (no interpolation shift).
__device__ __forceinline__ bool axis_stencil_cc(double q, const GridDesc& g, int d, int n, int& i0, double& w) {
    return axis_stencil(q, g, d, n, i0, w);
}

}  // namespace kpt
