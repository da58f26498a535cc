// kpt_oracle.cpp — CPU FP64 restatement of the reference particle push.
//
// TEST INFRASTRUCTURE ONLY. This file is the oracle that the HIP library is
// checked against, and the timed CPU baseline ("restatement of the Kokkos OpenMP
// path"). Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it, and only as the checker or the CPU baseline: the product path
// never calls it.
//
// It follows the reference code literally rather than efficiently: fields are
// stored on the ghosted domain (2 ghost layers, include/Grid.hpp:65) with the
// x-fastest flattening of include/Field.hpp:184-192, ghosts are filled by
// periodic copy (include/Field.hpp:274-316), the interpolator is the generic
// multilinear corner loop of include/FieldInterpolator.hpp:346-380 with the
// stencil of :509-560, and the movers restate src/main.cpp:1082-1755 line by line.
//
// Parity status per function (how each is pinned; the tests name the checks):
//   philox4x32_10      Random123 known-answer vectors (Salmon et al. 2011).
//   field packing      closed forms on linear / quadratic / trigonometric fields,
//                      periodic-ghost identities.
//   interpolation      exact on bilinear fields, node values, periodicity.
//   parker push        closed forms: pure diffusion variance, V=0 -> p const,
//                      constant div V -> exponential p, dt limiter cases.
//   focused push       closed forms: streaming along uniform B, D_mumu formula,
//                      mu clamp; otherwise shares the Parker branch structure.
//   split / histogram / summary / inject   invariants (weight conservation,
//                      totals) and the reference's tracked histogram CSV edges.
#include <algorithm>
#include <cmath>
#include <omp.h>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "../include/kpt.h"

namespace {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): 10 rounds of the 4x32 S-box with Weyl key schedule.
// ---------------------------------------------------------------------------
inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c[0];
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c[2];
    const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
    const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
}

inline void philox10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k[0] += 0x9E3779B9u;
            k[1] += 0xBB67AE85u;
        }
        philox_round(c, k);
    }
    for (int i = 0; i < 4; ++i) out[i] = c[i];
}

// U(0,1) from 32 random bits, open at both ends: (r + 1/2) / 2^32.
inline double u01(uint32_t r) {
    return (static_cast<double>(r) + 0.5) * (1.0 / 4294967296.0);
}

// StochasticSampler::uniform_sqrt3 (include/RandomManager.hpp:110-114):
// drand(-sqrt3, sqrt3) = lower + (upper - lower) * U.
inline double uniform_sqrt3(uint32_t r) {
    constexpr double sqrt3 = 1.7320508075688772935;
    return -sqrt3 + (sqrt3 - -sqrt3) * u01(r);
}

// ---------------------------------------------------------------------------
// Ghosted, periodic, uniform 2-D grid and field (include/Grid.hpp,
// include/AnalyticCoordinateGrid.hpp, include/MappedCoordinateGrid.hpp).
// ---------------------------------------------------------------------------
constexpr int kGhost = 2;

struct OGrid {
    int n[2];
    double start[2];  // x0 = first face coordinate
    double step[2];   // uniform face spacing (AthenaFieldReader.hpp:167-193)
    double phys_lo[2], phys_hi[2], width[2], inv_width[2];
    // stretched axis (non-uniform faces; the driver's reader rejects these):
    // value k sits at face k, extended periodically
    bool stretched[2] = {false, false};
    std::vector<double> faces[2];
    // the generic solvers' convention: value k sits at cell centre k, the mean of
    // faces k and k + 1 (StoredCoordinateGrid, include/StoredCoordinateGrid.hpp:411-425)
    bool cell_centred = false;

    double node(int d, int k) const {
        const int m = ((k % n[d]) + n[d]) % n[d];
        return faces[d][m] + static_cast<double>((k - m) / n[d]) * width[d];
    }

    // AnalyticCoordinateGrid face / cell coordinate (:159-176).
    double face(int d, double i) const { return start[d] + i * step[d]; }
    double cell(int d, double i) const { return 0.5 * (face(d, i) + face(d, i + 1.0)); }
    // Cell-centre coordinate of stored cell k (any k; ghost cells by periodic faces).
    double cell_at(int d, int k) const {
        return stretched[d] ? 0.5 * (node(d, k) + node(d, k + 1)) : cell(d, static_cast<double>(k));
    }
    // Cell-centred inverse map, uniform spacing (:205-210).
    double cell_grid_index(int d, double x) const { return (x - start[d]) / step[d] - 0.5; }
    int lower_ghosted(int) const { return -kGhost; }
    int upper_ghosted(int d) const { return n[d] + kGhost - 1; }
    // MappedCoordinateGrid::wrap_coordinate (:295-315).
    double wrap(int d, double x) const {
        double w = x - std::floor((x - phys_lo[d]) * inv_width[d]) * width[d];
        if (w >= phys_hi[d]) w -= width[d];
        if (w < phys_lo[d]) w += width[d];
        return w;
    }
};

struct OField {
    OGrid g;
    int comps = 0;
    int ext[2];                  // ghosted extents n + 4
    std::vector<double> data;    // [point][comp], point = (j+2)*ext0 + (i+2)

    size_t flat(int i, int j) const {
        return static_cast<size_t>(j + kGhost) * ext[0] + static_cast<size_t>(i + kGhost);
    }
    double at(int i, int j, int c) const { return data[flat(i, j) * comps + c]; }
    double& at(int i, int j, int c) { return data[flat(i, j) * comps + c]; }
};

// validate_uniform_face_edges' spacing test (AthenaFieldReader.hpp:167-193).
bool faces_uniform(const double* e, int count) {
    const double span = e[count - 1] - e[0];
    const double step = span / static_cast<double>(count - 1);
    for (int i = 1; i < count; ++i) {
        const double cur = e[i] - e[i - 1];
        const double mag = std::max({std::fabs(span), std::fabs(step), std::fabs(cur), 1.0});
        if (std::fabs(cur - step) > 1024.0 * std::numeric_limits<double>::epsilon() * mag) return false;
    }
    return true;
}

void make_grid(OGrid& g, int nx, int ny, const double* xe, const double* ye) {
    g.n[0] = nx;
    g.n[1] = ny;
    const double* e[2] = {xe, ye};
    for (int d = 0; d < 2; ++d) {
        g.start[d] = e[d][0];
        g.step[d] = (e[d][g.n[d]] - e[d][0]) / static_cast<double>(g.n[d]);
        g.stretched[d] = g.n[d] > 1 && !faces_uniform(e[d], g.n[d] + 1);
        if (g.stretched[d]) g.faces[d].assign(e[d], e[d] + g.n[d] + 1);
        // refresh_bounds_cache_on_host (MappedCoordinateGrid.hpp:117-145); a
        // stretched axis spans exactly [e_0, e_n]
        g.phys_lo[d] = g.stretched[d] ? e[d][0] : g.face(d, 0.0);
        g.phys_hi[d] = g.stretched[d] ? e[d][g.n[d]] : g.face(d, static_cast<double>(g.n[d]));
        g.width[d] = g.phys_hi[d] - g.phys_lo[d];
        g.inv_width[d] = g.width[d] != 0.0 ? 1.0 / g.width[d] : 0.0;
    }
}

// Face interval k (e_k <= q < e_{k+1}) of a wrapped coordinate on a stretched
// axis, by bisection like StoredCoordinateGrid::grid_index
// (include/StoredCoordinateGrid.hpp:128-173), with its weight and width.
void stretched_interval(const OGrid& g, int d, double q, int& k, double& w, double& h) {
    const std::vector<double>& e = g.faces[d];
    int lo = 0, hi = g.n[d] - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (e[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    k = lo;
    h = e[k + 1] - e[k];
    w = (q - e[k]) / h;
    w = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
}

// physical_cell_width at a coordinate: the uniform spacing, or on a stretched axis
// the width of the face interval holding the wrapped coordinate.
double cell_width(const OGrid& g, int d, double q) {
    if (!g.stretched[d]) return std::fabs(g.step[d]);
    int k;
    double w, h;
    stretched_interval(g, d, g.wrap(d, q), k, w, h);
    return h;
}

// Field::fill_ghost_cells, periodic in both axes (Field.hpp:274-316).
void fill_ghosts_periodic(OField& f) {
    const int nx = f.g.n[0], ny = f.g.n[1];
    for (int j = -kGhost; j < ny + kGhost; ++j) {
        for (int i = -kGhost; i < nx + kGhost; ++i) {
            if (i >= 0 && i < nx && j >= 0 && j < ny) continue;
            int si = i % nx;
            if (si < 0) si += nx;
            int sj = j % ny;
            if (sj < 0) sj += ny;
            for (int c = 0; c < f.comps; ++c) f.at(i, j, c) = f.at(si, sj, c);
        }
    }
}

double centered_three_point(double xl, double xc, double xu, double fl, double fc, double fu);

// Derivative at node k of a stretched axis: the three-point formula at the node
// coordinates (FieldOperator.hpp:85-105); at the outermost ghost layer the
// one-sided second-order formula over the two inner neighbours.
template <typename F>
double stretched_derivative(const OGrid& g, int dim, int k, F val) {
    const int lower = g.lower_ghosted(dim), upper = g.upper_ghosted(dim);
    if (k <= lower) return (-3.0 * val(lower) + 4.0 * val(lower + 1) - val(lower + 2)) / (g.node(dim, lower + 2) - g.node(dim, lower));
    if (k >= upper) return (3.0 * val(upper) - 4.0 * val(upper - 1) + val(upper - 2)) / (g.node(dim, upper) - g.node(dim, upper - 2));
    return centered_three_point(g.node(dim, k - 1), g.node(dim, k), g.node(dim, k + 1), val(k - 1), val(k), val(k + 1));
}

// derivative_component (src/main.cpp:714-770): central difference on the ghosted
// domain, one-sided second order at the outermost ghost layer.
double derivative(const OField& f, int i, int j, int dim, int comp) {
    if (f.g.n[dim] <= 1) return 0.0;
    const int lower = f.g.lower_ghosted(dim);
    const int upper = f.g.upper_ghosted(dim);
    int idx[2] = {i, j};
    auto val = [&](int k) {
        int a[2] = {i, j};
        a[dim] = k;
        return f.at(a[0], a[1], comp);
    };
    if (f.g.stretched[dim]) return stretched_derivative(f.g, dim, idx[dim], val);
    if (idx[dim] <= lower) {
        const double x0 = f.g.cell(dim, lower);
        const double x2 = f.g.cell(dim, lower + 2);
        return (-3.0 * val(lower) + 4.0 * val(lower + 1) - val(lower + 2)) / (x2 - x0);
    }
    if (idx[dim] >= upper) {
        const double x0 = f.g.cell(dim, upper - 2);
        const double x2 = f.g.cell(dim, upper);
        return (3.0 * val(upper) - 4.0 * val(upper - 1) + val(upper - 2)) / (x2 - x0);
    }
    const double xl = f.g.cell(dim, idx[dim] - 1);
    const double xr = f.g.cell(dim, idx[dim] + 1);
    return (val(idx[dim] + 1) - val(idx[dim] - 1)) / (xr - xl);
}

double bmag_at(const OField& b, int i, int j) {
    const double bx = b.at(i, j, 0), by = b.at(i, j, 1), bz = b.at(i, j, 2);
    return std::sqrt(bx * bx + by * by + bz * bz);
}

// derivative_magnetic_magnitude (src/main.cpp:775-834).
double derivative_bmag(const OField& b, int i, int j, int dim) {
    if (b.g.n[dim] <= 1) return 0.0;
    const int lower = b.g.lower_ghosted(dim);
    const int upper = b.g.upper_ghosted(dim);
    int idx[2] = {i, j};
    auto val = [&](int k) {
        int a[2] = {i, j};
        a[dim] = k;
        return bmag_at(b, a[0], a[1]);
    };
    if (b.g.stretched[dim]) return stretched_derivative(b.g, dim, idx[dim], val);
    if (idx[dim] <= lower) {
        const double x0 = b.g.cell(dim, lower);
        const double x2 = b.g.cell(dim, lower + 2);
        return (-3.0 * val(lower) + 4.0 * val(lower + 1) - val(lower + 2)) / (x2 - x0);
    }
    if (idx[dim] >= upper) {
        const double x0 = b.g.cell(dim, upper - 2);
        const double x2 = b.g.cell(dim, upper);
        return (3.0 * val(upper) - 4.0 * val(upper - 1) + val(upper - 2)) / (x2 - x0);
    }
    const double xl = b.g.cell(dim, idx[dim] - 1);
    const double xr = b.g.cell(dim, idx[dim] + 1);
    return (val(idx[dim] + 1) - val(idx[dim] - 1)) / (xr - xl);
}

// make_parker_reconnection_coefficients (src/main.cpp:839-892) and
// make_reconnection_coefficients (src/main.cpp:897-957), over every ghosted point.
void pack_coefficients(const OField& B, const OField& V, OField& out, int model) {
    out.g = B.g;
    out.comps = model;
    out.ext[0] = B.ext[0];
    out.ext[1] = B.ext[1];
    out.data.assign(static_cast<size_t>(out.ext[0]) * out.ext[1] * model, 0.0);
    const int nx = B.g.n[0], ny = B.g.n[1];
    for (int j = -kGhost; j < ny + kGhost; ++j) {
        for (int i = -kGhost; i < nx + kGhost; ++i) {
            const double bx = B.at(i, j, 0), by = B.at(i, j, 1), bz = B.at(i, j, 2);
            const double bm = std::sqrt(bx * bx + by * by + bz * bz);
            if (model == KPT_MODEL_PARKER16) {
                out.at(i, j, 0) = V.at(i, j, 0);
                out.at(i, j, 1) = V.at(i, j, 1);
                out.at(i, j, 2) = bx;
                out.at(i, j, 3) = by;
                out.at(i, j, 4) = bz;
                out.at(i, j, 5) = bm;
                out.at(i, j, 6) = derivative(V, i, j, 0, 0);
                out.at(i, j, 7) = derivative(V, i, j, 1, 1);
                out.at(i, j, 8) = derivative(B, i, j, 0, 0);
                out.at(i, j, 9) = derivative(B, i, j, 1, 0);
                out.at(i, j, 10) = derivative(B, i, j, 0, 1);
                out.at(i, j, 11) = derivative(B, i, j, 1, 1);
                out.at(i, j, 12) = derivative(B, i, j, 0, 2);
                out.at(i, j, 13) = derivative(B, i, j, 1, 2);
                out.at(i, j, 14) = derivative_bmag(B, i, j, 0);
                out.at(i, j, 15) = derivative_bmag(B, i, j, 1);
            } else {
                out.at(i, j, 0) = V.at(i, j, 0);
                out.at(i, j, 1) = V.at(i, j, 1);
                out.at(i, j, 2) = V.at(i, j, 2);
                out.at(i, j, 3) = bx;
                out.at(i, j, 4) = by;
                out.at(i, j, 5) = bz;
                out.at(i, j, 6) = bm;
                out.at(i, j, 7) = derivative(V, i, j, 0, 0);
                out.at(i, j, 8) = derivative(V, i, j, 1, 0);
                out.at(i, j, 9) = derivative(V, i, j, 0, 1);
                out.at(i, j, 10) = derivative(V, i, j, 1, 1);
                out.at(i, j, 11) = derivative(V, i, j, 0, 2);
                out.at(i, j, 12) = derivative(V, i, j, 1, 2);
                out.at(i, j, 13) = derivative(B, i, j, 0, 0);
                out.at(i, j, 14) = derivative(B, i, j, 1, 0);
                out.at(i, j, 15) = derivative(B, i, j, 0, 1);
                out.at(i, j, 16) = derivative(B, i, j, 1, 1);
                out.at(i, j, 17) = derivative(B, i, j, 0, 2);
                out.at(i, j, 18) = derivative(B, i, j, 1, 2);
                out.at(i, j, 19) = derivative_bmag(B, i, j, 0);
                out.at(i, j, 20) = derivative_bmag(B, i, j, 1);
            }
        }
    }
}

// LinearPositionInterpolator::sample with WrapPeriodic on the ghosted domain
// (include/FieldInterpolator.hpp:346-380, :465-480, :509-560).
void sample(const OField& f, const double q_in[2], double* result) {
    double q[2];
    for (int d = 0; d < 2; ++d) q[d] = f.g.wrap(d, q_in[d]);
    int lo[2], hi[2];
    double w[2];
    for (int d = 0; d < 2; ++d) {
        if (f.g.stretched[d] && f.g.cell_centred) {
            // bisection over the cell coordinates (StoredCoordinateGrid::grid_index,
            // include/StoredCoordinateGrid.hpp:128-173): q lies in face interval m,
            // hence in [c_m, c_{m+1}) or [c_{m-1}, c_m)
            int m;
            double wf, h;
            stretched_interval(f.g, d, q[d], m, wf, h);
            const int k = q[d] >= f.g.cell_at(d, m) ? m : m - 1;
            const double clo = f.g.cell_at(d, k), chi = f.g.cell_at(d, k + 1);
            lo[d] = k;
            hi[d] = k + 1;
            w[d] = std::min(1.0, std::max(0.0, (q[d] - clo) / (chi - clo)));
            continue;
        }
        if (f.g.stretched[d]) {
            double h;
            stretched_interval(f.g, d, q[d], lo[d], w[d], h);
            hi[d] = lo[d] + 1;
            continue;
        }
        const int lb = f.g.lower_ghosted(d);
        const int ub = f.g.upper_ghosted(d);
        const double logical = f.g.cell_grid_index(d, q[d]);
        if (logical <= static_cast<double>(lb)) {
            lo[d] = hi[d] = lb;
            w[d] = 0.0;
            continue;
        }
        if (logical >= static_cast<double>(ub)) {
            lo[d] = hi[d] = ub;
            w[d] = 0.0;
            continue;
        }
        const int cand = static_cast<int>(std::floor(logical));
        lo[d] = cand;
        hi[d] = cand + 1;
        w[d] = logical - static_cast<double>(cand);
        if (lo[d] < lb) {
            lo[d] = hi[d] = lb;
            w[d] = 0.0;
        } else if (hi[d] > ub) {
            lo[d] = hi[d] = ub;
            w[d] = 0.0;
        }
    }
    for (int c = 0; c < f.comps; ++c) result[c] = 0.0;
    for (int corner = 0; corner < 4; ++corner) {
        int idx[2];
        double weight = 1.0;
        for (int d = 0; d < 2; ++d) {
            const bool up = ((corner >> d) & 1) != 0;
            idx[d] = up ? hi[d] : lo[d];
            weight *= up ? w[d] : (1.0 - w[d]);
        }
        for (int c = 0; c < f.comps; ++c) result[c] += weight * f.at(idx[0], idx[1], c);
    }
}

// sample_time_interpolated_coefficients (src/main.cpp:1053-1077).
void sample_time(const OField& lo, const OField& hi, const double q[2], double rt, int use_time,
                 double* out) {
    double a[32], b[32];
    sample(lo, q, a);
    if (use_time == 0) {
        for (int c = 0; c < lo.comps; ++c) out[c] = a[c];
        return;
    }
    sample(hi, q, b);
    const double uw = rt < 0.0 ? 0.0 : (rt > 1.0 ? 1.0 : rt);
    const double lw = 1.0 - uw;
    for (int c = 0; c < lo.comps; ++c) out[c] = lw * a[c] + uw * b[c];
}

// wrap_periodic_coordinate (src/main.cpp:641-665).
double wrap_periodic_coordinate(double value, double lower, double upper, double width) {
    double wrapped = value;
    if (wrapped >= upper) {
        wrapped -= width;
        if (wrapped >= upper) wrapped -= std::floor((wrapped - lower) / width) * width;
    } else if (wrapped < lower) {
        wrapped += width;
        if (wrapped < lower) wrapped -= std::floor((wrapped - lower) / width) * width;
    }
    if (wrapped >= upper) wrapped -= width;
    if (wrapped < lower) wrapped += width;
    return wrapped;
}

inline double sq(double v) { return v * v; }
inline double min_positive(double a, double b) { return a < b ? a : b; }
inline double clamp_mu(double v, double m) { return v > m ? m : (v < -m ? -m : v); }

// FocusTransportSolver::reflected_pitch_angle_cosine (include/FocusSolver.hpp:
// 454-470): mirror at +-upper until inside. The mirror sequence is bounded
// here; past it the same map is applied in closed form (a triangle wave of
// period 4 upper), which only differs for |mu| in the thousands.
double reflect_mu(double mu, double upper) {
    const double lower = -upper;
    for (int k = 0; k < 16 && (mu > upper || mu < lower); ++k) {
        if (mu > upper) mu = 2.0 * upper - mu;
        if (mu < lower) mu = 2.0 * lower - mu;
    }
    if (mu > upper || mu < lower) {
        const double period = 4.0 * upper;
        double m = mu + upper;
        m -= period * std::floor(m / period);
        if (m > 2.0 * upper) m = period - m;
        mu = m - upper;
    }
    return mu;
}

// Slab-QLT pitch-angle scattering: LegencyModel::pitch_angle_spectrum_normalization,
// pitch_angle_diffusion_coefficient and pitch_angle_diffusion_derivative
// (include/LegencyModel.hpp:539-643), with omega = |q| |B| / gamma (m = c = 1).
struct QltD {
    double D, dD;
};
QltD qlt_scattering(double mu, double omega, double speed, double sigma2, double lc, double gk, double h0) {
    QltD r{0.0, 0.0};
    const double one_minus_mu2 = 1.0 - mu * mu;
    if (one_minus_mu2 <= 0.0) return r;
    const double pi = 3.14159265358979323846;
    const double mu_h = std::fabs(mu) + h0;
    const double xi = omega * lc / speed;
    const double abs_mu_gamma = std::pow(mu_h, gk);
    const double xi_gamma = std::pow(xi, gk);
    const double den = abs_mu_gamma + xi_gamma;
    const double norm = gk * std::sin(pi / gk) / pi;
    const double prefactor = 0.25 * pi * omega * sigma2 * norm;
    r.D = prefactor * one_minus_mu2 * xi * std::pow(mu_h, gk - 1.0) / den;
    if (r.D == 0.0) return r;
    const double sign_mu = mu > 0.0 ? 1.0 : (mu < 0.0 ? -1.0 : 0.0);
    if (sign_mu == 0.0) return r;
    const double shape = sign_mu * ((gk - 1.0) / mu_h - gk * std::pow(mu_h, gk - 1.0) / den);
    const double endpoint = -2.0 * mu / one_minus_mu2;
    r.dD = r.D * (endpoint + shape);
    return r;
}

struct NoiseSource {
    const double* table;  // injected draws or nullptr
    uint32_t steps;
    uint32_t key[2];
    uint32_t frame;

    // The four U[-sqrt3, sqrt3] values of sub-step k of the particle in slot i
    // with id `id`.
    void draws(uint64_t i, uint64_t id, uint32_t k, double r[4]) const {
        if (table) {
            const double* row = table + (static_cast<size_t>(i) * steps + k) * 4;
            for (int d = 0; d < 4; ++d) r[d] = row[d];
            return;
        }
        const uint32_t ctr[4] = {static_cast<uint32_t>(id), static_cast<uint32_t>(id >> 32), frame,
                                 k};
        uint32_t out[4];
        philox10(ctr, key, out);
        for (int d = 0; d < 4; ++d) r[d] = uniform_sqrt3(out[d]);
    }
};

// One particle through one interval: move_parker_particles_one_interval
// (src/main.cpp:1117-1325) when model == 16, move_particles_one_interval
// (src/main.cpp:1368-1753) when model == 21. Returns the executed sub-steps.
uint64_t move_one(const OField& lower, const OField& upper, const kpt_device_settings& S,
                  double interval_start, const NoiseSource& noise, uint64_t index,
                  kpt_particles_host& P, bool* failed) {
    *failed = false;
    if (P.status[index] != 0) return 0;
    const int model = lower.comps;
    const bool focused_transport = model == KPT_MODEL_FOCUSED21 && S.focused_transport != 0;
    const bool qlt = focused_transport && S.closure != KPT_CLOSURE_REDUCED;
    // the pitch-angle cosine: clamped by the driver, reflected by the generic solver
    auto bound_mu = [&](double m) { return qlt ? reflect_mu(m, S.mu_max) : clamp_mu(m, S.mu_max); };
    const double interval_end = interval_start + S.mhd_dt;
    double x = P.x[index], y = P.y[index], p = P.p[index], t = P.t[index];
    double pitch_mu = model == KPT_MODEL_FOCUSED21 ? bound_mu(P.mu[index]) : 0.0;
    double dt = P.step_dt[index];
    double diagnostic_dt = dt;
    int step_guard = 0;
    int status = 0;
    uint64_t steps = 0;
    const uint64_t id = P.id ? P.id[index] : index;
    double c[32];

    while (t < interval_end && step_guard < S.max_particle_steps_per_interval && status == 0) {
        ++step_guard;
        x = wrap_periodic_coordinate(x, S.x_lower, S.x_upper, S.x_width);
        y = wrap_periodic_coordinate(y, S.y_lower, S.y_upper, S.y_width);
        const double q[2] = {x + S.interpolation_shift_x, y + S.interpolation_shift_y};
        const double rt = (t - interval_start) / S.mhd_dt;
        sample_time(lower, upper, q, rt, S.time_interpolation, c);
        // half cell widths of the dt limiter: the settings' uniform ones, the
        // particle's face interval on a stretched axis
        double half_dx = 0.5 * S.dx, half_dy = 0.5 * S.dy;
        for (int d = 0; d < 2; ++d) {
            if (!lower.g.stretched[d]) continue;
            int k;
            double w, h;
            stretched_interval(lower.g, d, lower.g.wrap(d, q[d]), k, w, h);
            (d == 0 ? half_dx : half_dy) = 0.5 * h;
        }

        double vx, vy, bx, by, bz, b, dvx_dx, dvx_dy = 0, dvy_dx = 0, dvy_dy, dvz_dx = 0,
                                          dvz_dy = 0, dbx_dx, dbx_dy, dby_dx, dby_dy, dbz_dx,
                                          dbz_dy, db_dx, db_dy;
        if (model == KPT_MODEL_PARKER16) {
            vx = c[0]; vy = c[1]; bx = c[2]; by = c[3]; bz = c[4]; b = c[5];
            dvx_dx = c[6]; dvy_dy = c[7]; dbx_dx = c[8]; dbx_dy = c[9]; dby_dx = c[10];
            dby_dy = c[11]; dbz_dx = c[12]; dbz_dy = c[13]; db_dx = c[14]; db_dy = c[15];
        } else {
            vx = c[0]; vy = c[1]; bx = c[3]; by = c[4]; bz = c[5]; b = c[6];
            dvx_dx = c[7]; dvx_dy = c[8]; dvy_dx = c[9]; dvy_dy = c[10]; dvz_dx = c[11];
            dvz_dy = c[12]; dbx_dx = c[13]; dbx_dy = c[14]; dby_dx = c[15]; dby_dy = c[16];
            dbz_dx = c[17]; dbz_dy = c[18]; db_dx = c[19]; db_dy = c[20];
        }

        if (!std::isfinite(p) || p <= 0.0 || !std::isfinite(b) || b <= 0.0) {
            status = 2;
            break;
        }
        const double ib = 1.0 / b;
        const double ib2 = ib * ib;
        const double ib3 = ib2 * ib;

        double knorm_para = 1.0;
        if (S.magnetic_dependency == 1) knorm_para *= std::pow(b, S.gamma_turb - 2.0);
        double knorm = knorm_para;
        if (S.momentum_dependency == 1) knorm *= std::pow(p / S.p0, S.pindex);
        const double kpara = S.kpara0 * knorm;
        const double kperp = kpara * S.kperp_over_kpara;
        if (!std::isfinite(kpara) || kpara <= 0.0 || !std::isfinite(kperp) || kperp < 0.0) {
            status = 2;
            break;
        }
        const double skpara = std::sqrt(2.0 * kpara);
        const double skperp = std::sqrt(2.0 * kperp);
        // Parker-only mover: unguarded sqrt (src/main.cpp:1197-1198); shared mover
        // guards it (src/main.cpp:1456-1457).
        const double skpara_perp = model == KPT_MODEL_PARKER16
                                       ? std::sqrt(2.0 * (kpara - kperp))
                                       : (kpara > kperp ? std::sqrt(2.0 * (kpara - kperp)) : 0.0);
        const double kpp = focused_transport ? -kperp : kpara - kperp;

        double dkdx = 0.0, dkdy = 0.0;
        if (S.magnetic_dependency == 1) {
            const double me = S.gamma_turb - 2.0;
            dkdx = db_dx * ib * me;
            dkdy = db_dy * ib * me;
        }
        const double dkxx_dx =
            kperp * dkdx + kpp * dkdx * bx * bx * ib2 + 2.0 * kpp * bx * (dbx_dx * b - bx * db_dx) * ib3;
        const double dkyy_dy =
            kperp * dkdy + kpp * dkdy * by * by * ib2 + 2.0 * kpp * by * (dby_dy * b - by * db_dy) * ib3;
        const double dkxy_dx = kpp * dkdx * bx * by * ib2 +
                               kpp * ((dbx_dx * by + bx * dby_dx) * ib2 - 2.0 * bx * by * db_dx * ib3);
        const double dkxy_dy = kpp * dkdy * bx * by * ib2 +
                               kpp * ((dbx_dy * by + bx * dby_dy) * ib2 - 2.0 * bx * by * db_dy * ib3);
        const double drift_denominator =
            std::sqrt(sq(S.drift_param1 * S.p0 / p) + sq(S.drift_param2 * S.p0 * S.p0 / (p * p)));

        double dx_dt = 0.0, dy_dt = 0.0, dp_dt = 0.0, dmu_dt = 0.0, duu = 0.0;
        // pieces of the focusing rate, kept for the Picard iterations of the QLT closure
        double f_speed = 0.0, f_divv = 0.0, f_bb = 0.0, f_bv = 0.0, f_divb = 0.0, f_omega = 0.0;
        if (focused_transport) {
            // src/main.cpp:1499-1589
            pitch_mu = bound_mu(pitch_mu);
            const double mu2 = pitch_mu * pitch_mu;
            const double muf1 = 0.5 * (1.0 - mu2);
            const double muf2 = 0.5 * (3.0 * mu2 - 1.0);
            const double kx = bx * dbx_dx + by * dbx_dy;
            const double ky = bx * dby_dx + by * dby_dy;
            const double kz = bx * dbz_dx + by * dbz_dy;
            const double bdot_curvb = bx * dbz_dy - by * dbz_dx + bz * (dby_dx - dbx_dy);
            const double vdp = 1.0 / S.charge / drift_denominator;
            const double vdx = vdp * (muf1 * (-bz * db_dy) * ib2 + mu2 * (by * kz - bz * ky) * ib3 +
                                      muf1 * bx * bdot_curvb * ib3);
            const double vdy = vdp * (muf1 * (bz * db_dx) * ib2 + mu2 * (bz * kx - bx * kz) * ib3 +
                                      muf1 * by * bdot_curvb * ib3);
            // QLT closure: relativistic speed p / gamma with m = c = 1
            // (ParticleProperties::speed_from_momentum_magnitude, ParticleSystem.hpp:170-184)
            const double gamma_p = qlt ? std::sqrt(1.0 + p * p) : 1.0;
            const double speed = qlt ? p / gamma_p : S.particle_v0 * p / S.p0;
            if (!std::isfinite(speed) || speed <= 0.0) {
                status = 2;
                break;
            }
            const double vpb = speed * pitch_mu * ib;
            const double vbx = vpb * bx;
            const double vby = vpb * by;
            dx_dt = vx + vdx + vbx + dkxx_dx + dkxy_dy;
            dy_dt = vy + vdy + vby + dkxy_dx + dkyy_dy;
            const double divv = dvx_dx + dvy_dy;
            const double bb_gradv = (bx * (bx * dvx_dx + by * dvx_dy) + by * (bx * dvy_dx + by * dvy_dy) +
                                     bz * (bx * dvz_dx + by * dvz_dy)) * ib2;
            const double bv_gradv = (bx * (vx * dvx_dx + vy * dvx_dy) + by * (vx * dvy_dx + vy * dvy_dy) +
                                     bz * (vx * dvz_dx + vy * dvz_dy)) * ib;
            const double acc_rate = -(muf1 * divv + muf2 * bb_gradv + pitch_mu * bv_gradv / speed);
            dp_dt = p * acc_rate;
            const double div_bnorm = -(bx * db_dx + by * db_dy) * ib2;
            dmu_dt = speed * div_bnorm + pitch_mu * divv - 3.0 * pitch_mu * bb_gradv -
                     2.0 * bv_gradv / speed;
            dmu_dt *= 0.5 * (1.0 - mu2);
            double duu_du = 0.0;
            if (qlt) {
                // LegencyModel::evaluate_pitch_angle_scattering (LegencyModel.hpp:650-689)
                f_speed = speed;
                f_divv = divv;
                f_bb = bb_gradv;
                f_bv = bv_gradv;
                f_divb = div_bnorm;
                f_omega = std::fabs(S.charge) * b / gamma_p;
                const QltD q = qlt_scattering(pitch_mu, f_omega, speed, S.qlt_sigma2, S.qlt_correlation_length,
                                              S.qlt_spectral_index, S.qlt_h0);
                duu = q.D;
                duu_du = q.dD;
            } else {
                const double abs_mu = std::fabs(pitch_mu);
                const double dtmp = std::pow(abs_mu, S.gamma_turb - 1.0) + S.pitch_angle_scattering_h0;
                duu = S.duu0 * (1.0 - mu2) * dtmp;
                if (pitch_mu > 0.0) {
                    duu_du = S.duu0 * (-2.0 * pitch_mu * dtmp + (1.0 - mu2) * std::pow(abs_mu, S.gamma_turb - 2.0));
                } else if (pitch_mu < 0.0) {
                    duu_du = S.duu0 * (-2.0 * pitch_mu * dtmp - (1.0 - mu2) * std::pow(abs_mu, S.gamma_turb - 2.0));
                }
                double duu_norm = 1.0;
                if (S.magnetic_dependency == 1) duu_norm *= std::pow(b, 2.0 - S.gamma_turb);
                if (S.momentum_dependency == 1) duu_norm *= std::pow(p / S.p0, S.gamma_turb - 1.0);
                duu *= duu_norm;
                duu_du *= duu_norm;
            }
            dmu_dt += duu_du;
            if (!std::isfinite(dmu_dt) || !std::isfinite(duu) || duu < 0.0) {
                status = 2;
                break;
            }
        } else {
            // src/main.cpp:1225-1241 (Parker mover), :1590-1601 (shared mover)
            const double vdp = 1.0 / (3.0 * S.charge) / drift_denominator;
            const double vdx = vdp * (dbz_dy * ib2 - 2.0 * bz * db_dy * ib3);
            const double vdy = vdp * (-dbz_dx * ib2 + 2.0 * bz * db_dx * ib3);
            dx_dt = vx + vdx + dkxx_dx + dkxy_dy;
            dy_dt = vy + vdy + dkxy_dx + dkyy_dy;
            const double divv = dvx_dx + dvy_dy;
            dp_dt = -p * divv / 3.0;
        }
        (void)dvz_dx;

        // dt limiter (src/main.cpp:1243-1277, :1603-1665)
        if (focused_transport) {
            const double ds = skperp > 0.0 ? skperp : skpara;
            if (dx_dt != 0.0 && dy_dt != 0.0 && dp_dt != 0.0 && dmu_dt != 0.0 && ds > 0.0) {
                dt = min_positive(sq(half_dx / ds), sq(half_dy / ds));
                dt = min_positive(dt, sq(ds / dx_dt));
                dt = min_positive(dt, sq(ds / dy_dt));
                dt = min_positive(dt, 0.1 * p / std::fabs(dp_dt));
                dt = min_positive(dt, 0.1 / std::fabs(dmu_dt));
                dt = min_positive(dt, 2.0 * duu / sq(dmu_dt));
            } else {
                dt = S.dt_min;
            }
        } else if (dx_dt != 0.0 && dy_dt != 0.0 && dp_dt != 0.0) {
            const double as = skperp > 0.0 ? skperp : skpara;
            dt = min_positive(sq(half_dx / skpara), sq(half_dy / skpara));
            dt = min_positive(dt, sq(as / dx_dt));
            dt = min_positive(dt, sq(as / dy_dt));
            dt = min_positive(dt, 0.1 * p / std::fabs(dp_dt));
        } else {
            dt = S.dt_min;
        }
        if (dt < S.dt_min) dt = S.dt_min;
        if (dt > S.dt_max) dt = S.dt_max;
        diagnostic_dt = dt;
        const double remaining = interval_end - t;
        if (dt > remaining) dt = remaining;
        if (!std::isfinite(dt) || dt <= 0.0) {
            status = 2;
            break;
        }

        const double sdt = std::sqrt(dt);
        double r[4];
        noise.draws(index, id, static_cast<uint32_t>(step_guard - 1), r);
        double delta_x, delta_y;
        const double delta_p = dp_dt * dt;
        if (focused_transport) {
            // src/main.cpp:1671-1702
            const double ran1 = r[0], ran2 = r[1], pitch_random = r[3];
            const double bxn = bx * ib, byn = by * ib, bzn = bz * ib;
            const double bxy_norm = std::sqrt(bxn * bxn + byn * byn);
            double sx = 0.0, sy = 0.0;
            if (bxy_norm > 0.0) {
                const double ibxy = 1.0 / bxy_norm;
                sx = skperp * ibxy * sdt * (-bxn * bzn * ran1 - by * ran2);
                sy = skperp * ibxy * sdt * (-byn * bzn * ran1 + bx * ran2);
            }
            delta_x = dx_dt * dt + sx;
            delta_y = dy_dt * dt + sy;
            if (qlt) {
                // FocusTransportSolver::picard_pitch_angle_update (FocusSolver.hpp:700-719),
                // the drift and diffusion re-evaluated at the trial mu
                // (evaluate_pitch_angle_at_mu, :665-694), the same noise in every pass.
                const double mu0 = pitch_mu;
                double rate = dmu_dt, D = duu, trial = mu0;
                for (int it = 0; it < 3; ++it) {
                    if (it > 0) {
                        const double m = bound_mu(trial);
                        const double m2 = m * m;
                        const QltD q = qlt_scattering(m, f_omega, f_speed, S.qlt_sigma2, S.qlt_correlation_length,
                                                      S.qlt_spectral_index, S.qlt_h0);
                        rate = 0.5 * (1.0 - m2) *
                                   (f_speed * f_divb + m * f_divv - 3.0 * m * f_bb - 2.0 * f_bv / f_speed) +
                               q.dD;
                        D = q.D;
                    }
                    const double scale = D > 0.0 ? std::sqrt(2.0 * D * dt) : 0.0;
                    trial = bound_mu(mu0 + dt * rate + scale * pitch_random);
                }
                pitch_mu = trial;
            } else {
                pitch_mu = clamp_mu(pitch_mu + dmu_dt * dt + pitch_random * std::sqrt(2.0 * duu) * sdt, S.mu_max);
            }
        } else {
            // src/main.cpp:1279-1292
            const double ran1 = r[0], ran2 = r[1], ran3 = r[2];
            delta_x = dx_dt * dt + ran1 * skperp * sdt + ran3 * skpara_perp * sdt * bx * ib;
            delta_y = dy_dt * dt + ran2 * skperp * sdt + ran3 * skpara_perp * sdt * by * ib;
        }
        x += delta_x;
        y += delta_y;
        p += delta_p;
        t += dt;
        ++steps;
        if (p < 0.25 * S.p0) p = 0.25 * S.p0;
        if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(p) || !std::isfinite(t)) {
            status = 2;
            break;
        }
    }
    if (step_guard >= S.max_particle_steps_per_interval && t < interval_end) status = 2;

    P.x[index] = wrap_periodic_coordinate(x, S.x_lower, S.x_upper, S.x_width);
    P.y[index] = wrap_periodic_coordinate(y, S.y_lower, S.y_upper, S.y_width);
    P.p[index] = p;
    if (model == KPT_MODEL_FOCUSED21) P.mu[index] = pitch_mu;
    P.t[index] = t;
    P.step_dt[index] = diagnostic_dt;
    P.status[index] = status;
    *failed = status != 0;
    return steps;
}

}  // namespace

// ===========================================================================
// C interface of the oracle.
// ===========================================================================
extern "C" {

void kpto_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox10(ctr, key, out);
}

double kpto_uniform_sqrt3(uint32_t r) { return uniform_sqrt3(r); }

// Slab-QLT D_mumu and dD_mumu/dmu at one state (out[0], out[1]).
void kpto_qlt_scattering(double mu, double omega, double speed, double sigma2, double lc, double gk,
                         double h0, double* out) {
    const QltD q = qlt_scattering(mu, omega, speed, sigma2, lc, gk, h0);
    out[0] = q.D;
    out[1] = q.dD;
}

double kpto_reflect_mu(double mu, double upper) { return reflect_mu(mu, upper); }

// Build one packed coefficient frame from compact-field arrays.
void* kpto_field_create(int32_t nx, int32_t ny, const double* xe, const double* ye, const double* bx,
                        const double* by, const double* bz, const double* vx, const double* vy,
                        const double* vz, int32_t model) {
    if (nx <= 0 || ny <= 0 || (model != KPT_MODEL_PARKER16 && model != KPT_MODEL_FOCUSED21)) {
        return nullptr;
    }
    OField B, V;
    for (OField* f : {&B, &V}) {
        make_grid(f->g, nx, ny, xe, ye);
        f->comps = 3;
        f->ext[0] = nx + 2 * kGhost;
        f->ext[1] = ny + 2 * kGhost;
        f->data.assign(static_cast<size_t>(f->ext[0]) * f->ext[1] * 3, 0.0);
    }
    const double* bs[3] = {bx, by, bz};
    const double* vs[3] = {vx, vy, vz};
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            const size_t s = static_cast<size_t>(j) * nx + i;
            for (int c = 0; c < 3; ++c) {
                B.at(i, j, c) = bs[c][s];
                V.at(i, j, c) = vs[c][s];
            }
        }
    }
    fill_ghosts_periodic(B);
    fill_ghosts_periodic(V);
    OField* out = new OField();
    pack_coefficients(B, V, *out, model);
    return out;
}

void kpto_field_destroy(void* f) { delete static_cast<OField*>(f); }

// Packed coefficients of nodes i in [-1, nx], j in [-1, ny], as
// [(ny+2)][(nx+2)][C]. Returns the element count written, or -1.
int64_t kpto_field_coefficients(const void* handle, double* out, size_t len) {
    const OField& f = *static_cast<const OField*>(handle);
    const int nx = f.g.n[0], ny = f.g.n[1];
    const size_t need = static_cast<size_t>(nx + 2) * (ny + 2) * f.comps;
    if (len < need) return -1;
    size_t k = 0;
    for (int j = -1; j <= ny; ++j)
        for (int i = -1; i <= nx; ++i)
            for (int c = 0; c < f.comps; ++c) out[k++] = f.at(i, j, c);
    return static_cast<int64_t>(need);
}

// Interpolate all components of `handle` at one position (q is the sampling
// coordinate, already shifted). For interpolator tests.
void kpto_field_sample(const void* handle, double qx, double qy, double* out) {
    const double q[2] = {qx, qy};
    sample(*static_cast<const OField*>(handle), q, out);
}

// move_*_particles_one_interval over `count` particles with OpenMP threads.
// noise == nullptr selects Philox keyed by (seed, id, frame, sub-step).
void kpto_push_interval(const void* lower, const void* upper, const kpt_device_settings* s,
                        double t0, uint32_t frame, uint64_t seed, kpt_particles_host* P,
                        uint64_t count, const double* noise, uint32_t noise_steps, int threads,
                        uint64_t* steps_out, uint64_t* failures_out) {
    const OField& lo = *static_cast<const OField*>(lower);
    const OField& hi = *static_cast<const OField*>(upper ? upper : lower);
    NoiseSource ns{noise, noise_steps,
                   {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)}, frame};
    uint64_t steps = 0, failures = 0;
    const int64_t n = static_cast<int64_t>(count);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : steps, failures) num_threads(threads > 0 ? threads : 1)
    for (int64_t i = 0; i < n; ++i) {
        bool failed = false;
        const bool was_active = P->status[i] == 0;
        steps += move_one(lo, hi, *s, t0, ns, static_cast<uint64_t>(i), *P, &failed);
        if (was_active && failed) ++failures;
    }
    if (steps_out) *steps_out = steps;
    if (failures_out) *failures_out = failures;
}

// inject_uniform_particles (src/main.cpp:995-1048) with the axis-aligned region
// option of include/ParticleGenerator.hpp:17-62. Draws: Philox with counter
// (root id, 0xFFFFFFFF, frame, 0) -> x, y, t, mu. Returns the inserted count.
uint64_t kpto_inject(const kpt_device_settings* s, const kpt_inject_opts* o, uint64_t requested,
                     double t0, uint32_t frame, uint64_t seed, uint64_t first_root,
                     kpt_particles_host* P, uint64_t count, uint64_t capacity) {
    if (count >= capacity) return 0;
    const uint64_t n = requested < capacity - count ? requested : capacity - count;
    const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
    const double xl = o && o->use_region ? o->region_x_lo : s->x_lower;
    const double xu = o && o->use_region ? o->region_x_hi : s->x_upper;
    const double yl = o && o->use_region ? o->region_y_lo : s->y_lower;
    const double yu = o && o->use_region ? o->region_y_hi : s->y_upper;
    const double pinit = o && o->p_init > 0.0 ? o->p_init : s->p0;
    for (uint64_t k = 0; k < n; ++k) {
        const uint64_t i = count + k;
        const uint64_t root = first_root + k;
        const uint32_t ctr[4] = {static_cast<uint32_t>(root), 0xFFFFFFFFu, frame,
                                 static_cast<uint32_t>(root >> 32)};
        uint32_t r[4];
        philox10(ctr, key, r);
        P->x[i] = xl + (xu - xl) * u01(r[0]);
        P->y[i] = yl + (yu - yl) * u01(r[1]);
        P->p[i] = pinit;
        P->t[i] = (o && o->fixed_time) ? t0 : t0 + (s->mhd_dt - 0.0) * u01(r[2]);
        P->step_dt[i] = s->dt_min;
        P->weight[i] = 1.0;
        P->mu[i] = s->mu_max * (2.0 * u01(r[3]) - 1.0);
        P->status[i] = 0;
        P->split_level[i] = 0;
        if (P->id) P->id[i] = root;
        if (P->e0) P->e0[i] = std::sqrt(1.0 + pinit * pinit) - 1.0;
    }
    return n;
}

// Id of a split child whose parent (id pid) was at level L: (root | path << 32)
// gets path | (1 << L) for 0 <= L < 32; other levels take the splitmix64 finalizer
// (Steele, Lea and Flood 2014) of pid + (L + 1) * 0x9E3779B97F4A7C15.
uint64_t kpto_split_child_id(uint64_t pid, int32_t level) {
    if (level >= 0 && level < 32) {
        const uint64_t path = (pid >> 32) | (static_cast<uint64_t>(1) << level);
        return (pid & 0xFFFFFFFFull) | (path << 32);
    }
    uint64_t z = pid + static_cast<uint64_t>(level + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// split_particles (src/main.cpp:1760-1839), children appended in parent slot
// order; child ids from kpto_split_child_id. Returns the new count; split /
// skipped through pointers.
uint64_t kpto_split(const kpt_device_settings* s, kpt_particles_host* P, uint64_t count,
                    uint64_t capacity, uint64_t* split_out, uint64_t* skipped_out) {
    uint64_t split = 0, skipped = 0;
    if (count >= capacity || s->split_particles == 0) {
        if (split_out) *split_out = 0;
        if (skipped_out) *skipped_out = 0;
        return count;
    }
    uint64_t counter = count;
    for (uint64_t parent = 0; parent < count; ++parent) {
        if (P->status[parent] != 0) continue;
        const int old_level = P->split_level[parent];
        const double threshold = s->pmin_split * std::pow(s->split_ratio, static_cast<double>(old_level));
        const double p = P->p[parent];
        if (!(p > threshold && p <= s->pmax)) continue;
        const uint64_t child = counter++;
        if (child >= capacity) {
            ++skipped;
            continue;
        }
        const int new_level = old_level + 1;
        const double w = std::pow(0.5, static_cast<double>(new_level));
        P->x[child] = P->x[parent];
        P->y[child] = P->y[parent];
        P->p[child] = P->p[parent];
        P->t[child] = P->t[parent];
        P->step_dt[child] = P->step_dt[parent];
        P->weight[parent] = w;
        P->weight[child] = w;
        P->mu[child] = P->mu[parent];
        if (P->e0) P->e0[child] = P->e0[parent];
        P->status[child] = 0;
        P->split_level[parent] = new_level;
        P->split_level[child] = new_level;
        if (P->id) P->id[child] = kpto_split_child_id(P->id[parent], old_level);
        ++split;
    }
    if (split_out) *split_out = split;
    if (skipped_out) *skipped_out = skipped;
    return counter < capacity ? counter : capacity;
}

// write_momentum_histogram binning (src/main.cpp:1910-1941) with `bins` bins.
void kpto_histogram_logp(const kpt_particles_host* P, uint64_t count, int32_t bins, double pmin,
                         double pmax, double* out) {
    for (int b = 0; b < bins; ++b) out[b] = 0.0;
    const double log_pmin = std::log(pmin);
    const double inv_dlogp = static_cast<double>(bins) / (std::log(pmax) - log_pmin);
    for (uint64_t i = 0; i < count; ++i) {
        if (P->status[i] != 0) continue;
        const double p = P->p[i];
        if (p < pmin || p > pmax) continue;
        int bin = static_cast<int>((std::log(p) - log_pmin) * inv_dlogp);
        if (bin < 0) bin = 0;
        if (bin >= bins) bin = bins - 1;
        out[bin] += P->weight[i];
    }
}

// Spatial density: weight per cell of a uniform bins_x x bins_y grid.
void kpto_histogram_xy(const kpt_particles_host* P, uint64_t count, int32_t bx, int32_t by,
                       double x_lo, double x_hi, double y_lo, double y_hi, double* out) {
    for (int64_t k = 0; k < static_cast<int64_t>(bx) * by; ++k) out[k] = 0.0;
    const double sx = static_cast<double>(bx) / (x_hi - x_lo);
    const double sy = static_cast<double>(by) / (y_hi - y_lo);
    for (uint64_t i = 0; i < count; ++i) {
        if (P->status[i] != 0) continue;
        const double x = P->x[i], y = P->y[i];
        if (!(x >= x_lo && x < x_hi && y >= y_lo && y < y_hi)) continue;
        int ix = static_cast<int>((x - x_lo) * sx);
        int iy = static_cast<int>((y - y_lo) * sy);
        if (ix >= bx) ix = bx - 1;
        if (iy >= by) iy = by - 1;
        out[static_cast<int64_t>(iy) * bx + ix] += P->weight[i];
    }
}

// summarize_particles (src/main.cpp:1844-1901).
void kpto_summary(const kpt_particles_host* P, uint64_t count, kpt_summary* s) {
    std::memset(s, 0, sizeof(*s));
    s->stored = count;
    if (count == 0) return;
    double p_sum = 0.0, dt_sum = 0.0;
    s->p_min = std::numeric_limits<double>::infinity();
    s->dt_min = std::numeric_limits<double>::infinity();
    s->p_max = 0.0;
    s->dt_max = 0.0;
    for (uint64_t i = 0; i < count; ++i) {
        if (P->status[i] == 0) {
            ++s->active;
            const double p = P->p[i], dt = P->step_dt[i];
            s->weight_sum += P->weight[i];
            p_sum += p;
            dt_sum += dt;
            s->p_min = p < s->p_min ? p : s->p_min;
            s->p_max = p > s->p_max ? p : s->p_max;
            s->dt_min = dt < s->dt_min ? dt : s->dt_min;
            s->dt_max = dt > s->dt_max ? dt : s->dt_max;
            if (P->split_level[i] > s->max_split_level) s->max_split_level = P->split_level[i];
        } else {
            ++s->inactive;
        }
    }
    if (s->active > 0) {
        s->p_avg = p_sum / static_cast<double>(s->active);
        s->dt_avg = dt_sum / static_cast<double>(s->active);
    } else {
        s->p_min = 0.0;
        s->dt_min = 0.0;
    }
}

}  // extern "C"

// ===========================================================================
// Generic ParkerSolver path (include/ParkerSolver.hpp, ParkerCoefficient.hpp,
// LegencyModel.hpp, FieldOperator.hpp) on the uniform periodic 2-D grid, three
// vector components, cell-centred fields with two ghost layers.
// ===========================================================================
namespace {

// centered_three_point_derivative (include/ParkerCoefficient.hpp:215-233,
// include/FieldOperator.hpp:86-103).
double centered_three_point(double xl, double xc, double xu, double fl, double fc, double fu) {
    const double lw = xc - xl, uw = xu - xc, tw = xu - xl;
    return (-uw / (lw * tw)) * fl + ((uw - lw) / (lw * uw)) * fc + (lw / (uw * tw)) * fu;
}

// Central derivative along `dim` (0, 1; 0 for the invariant z) of a per-point
// quantity f(i, j) at physical point (i, j).
template <typename F>
double gen_derivative(const OGrid& g, int i, int j, int dim, F f) {
    if (dim >= 2 || g.n[dim] <= 1) return 0.0;
    int lo[2] = {i, j}, up[2] = {i, j};
    lo[dim] -= 1;
    up[dim] += 1;
    const int c = dim == 0 ? i : j;
    return centered_three_point(g.cell_at(dim, c - 1), g.cell_at(dim, c), g.cell_at(dim, c + 1), f(lo[0], lo[1]),
                                f(i, j), f(up[0], up[1]));
}

double gen_bmag(const OField& B, int i, int j) {
    double s = 0.0;
    for (int c = 0; c < 3; ++c) s += B.at(i, j, c) * B.at(i, j, c);
    return std::sqrt(s);
}

// magnetic_direction_component (ParkerCoefficient.hpp:255-270).
double gen_bdir(const OField& B, int i, int j, int c, double bmin, bool& bad) {
    const double m = gen_bmag(B, i, j);
    if (m <= 0.0 && bmin <= 0.0) bad = true;
    return B.at(i, j, c) / (m > bmin ? m : bmin);
}

// LegencyModel::parallel_diffusion_gamma_one / perpendicular_diffusion_gamma_one
// (LegencyModel.hpp:359-424) with the floors of evaluate_gamma_independent_coefficients.
void gen_kappa(const kpt_generic_parker_params& P, double bmag, double vsw, double& kpar, double& kperp,
               bool& bad) {
    auto floor_pos = [](double v, double f) { return f <= 0.0 ? v : (v < f ? f : v); };
    const double b = floor_pos(bmag, P.minimum_magnetic_field_magnitude);
    const double v = floor_pos(vsw, P.minimum_solar_wind_speed);
    const double s2 = floor_pos(P.sigma2, P.minimum_sigma2);
    const double lc = P.lc > 0.0 ? P.lc : P.lslab;
    if (b <= 0.0) bad = true;
    const double omega = std::fabs(P.charge) * b / (P.rest_mass * P.speed_of_light);
    if (v <= 0.0 || lc <= 0.0 || omega <= 0.0 || s2 <= 0.0) bad = true;
    kpar = P.parallel_prefactor * std::pow(v, 4.0 / 3.0) * std::pow(lc, 2.0 / 3.0) /
           (std::pow(omega, 1.0 / 3.0) * s2);
    if (P.perpendicular_model == 1) {
        kperp = P.perpendicular_parallel_ratio * kpar;
    } else {
        if (P.l2d <= 0.0 || kpar <= 0.0) bad = true;
        kperp = (v / 3.0) * std::pow(3.0 * kpar / v, 1.0 / 3.0) *
                std::pow(P.perpendicular_prefactor * s2 * P.l2d, 2.0 / 3.0);
    }
}

OField gen_alloc(const OGrid& g, int comps) {
    OField f;
    f.g = g;
    f.comps = comps;
    f.ext[0] = g.n[0] + 2 * kGhost;
    f.ext[1] = g.n[1] + 2 * kGhost;
    f.data.assign(static_cast<size_t>(f.ext[0]) * f.ext[1] * comps, 0.0);
    return f;
}

// make_legacy_transport_coefficient_fields (ParkerCoefficient.hpp:1090-1140) packed
// as ParkerSolver::build_packed_coefficient_basis (ParkerSolver.hpp:452-640):
// B(0-2) V(3-5) curl(B/B^2)(6-8) primary(9-11) secondary(12-14) kpar1 kperp1 divV.
bool gen_build(const OField& B, const OField& V, const kpt_generic_parker_params& P, OField& out) {
    const OGrid& g = B.g;
    const int nx = g.n[0], ny = g.n[1];
    bool bad = false;
    OField curl = gen_alloc(g, 3), kap = gen_alloc(g, 2), div = gen_alloc(g, 6), dv = gen_alloc(g, 1);
    // curl(B / B^2) (ParkerCoefficient.hpp:390-419, :476-511)
    auto over_b2 = [&](int c) {
        return [&, c](int i, int j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += B.at(i, j, k) * B.at(i, j, k);
            const double fl = P.minimum_magnetic_field_squared_for_drift;
            if (s <= 0.0 && fl <= 0.0) bad = true;
            return B.at(i, j, c) / (s > fl ? s : fl);
        };
    };
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            for (int c = 0; c < 3; ++c) {
                const int da = (c + 1) % 3, db = (c + 2) % 3;
                curl.at(i, j, c) = gen_derivative(g, i, j, da, over_b2(db)) - gen_derivative(g, i, j, db, over_b2(da));
            }
    fill_ghosts_periodic(curl);
    // gamma-independent kappa fields (LegencyModel.hpp:801-871)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double vs = 0.0;
            for (int k = 0; k < 3; ++k) vs += V.at(i, j, k) * V.at(i, j, k);
            gen_kappa(P, gen_bmag(B, i, j), std::sqrt(vs), kap.at(i, j, 0), kap.at(i, j, 1), bad);
        }
    fill_ghosts_periodic(kap);
    // div(kappa_ij) bases (ParkerCoefficient.hpp:611-640, :800-869): Cartesian,
    // so div_i = sum_j d_j T_ij. basis 0 parallel, 1 perpendicular, 2 constant ratio.
    auto tensor = [&](int basis, int a, int b) {
        return [&, basis, a, b](int i, int j) {
            const double bi = gen_bdir(B, i, j, a, P.minimum_magnetic_field_magnitude, bad);
            const double bj = gen_bdir(B, i, j, b, P.minimum_magnetic_field_magnitude, bad);
            const double proj = bi * bj, id = a == b ? 1.0 : 0.0;
            if (basis == 0) return kap.at(i, j, 0) * proj;
            if (basis == 1) return kap.at(i, j, 1) * (id - proj);
            const double r = P.perpendicular_parallel_ratio;
            return kap.at(i, j, 0) * (r * id + (1.0 - r) * proj);
        };
    };
    const int b0 = P.perpendicular_model == 1 ? 2 : 0;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            for (int c = 0; c < 3; ++c) {
                double prim = 0.0, sec = 0.0;
                for (int d = 0; d < 2; ++d) {
                    prim += gen_derivative(g, i, j, d, tensor(b0, c, d));
                    if (P.perpendicular_model != 1) sec += gen_derivative(g, i, j, d, tensor(1, c, d));
                }
                div.at(i, j, c) = prim;
                div.at(i, j, 3 + c) = P.perpendicular_model == 1 ? prim : sec;
            }
    fill_ghosts_periodic(div);
    // FieldOperator::fill_divergence (FieldOperator.hpp:259-290, :487-513)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double s = 0.0;
            for (int d = 0; d < 2; ++d)
                s += gen_derivative(g, i, j, d, [&, d](int a, int b) { return V.at(a, b, d); });
            dv.at(i, j, 0) = s;
        }
    fill_ghosts_periodic(dv);
    out = gen_alloc(g, KPT_GENERIC_PARKER_COMPONENTS);
    for (int j = -kGhost; j < ny + kGhost; ++j)
        for (int i = -kGhost; i < nx + kGhost; ++i) {
            for (int c = 0; c < 3; ++c) {
                out.at(i, j, c) = B.at(i, j, c);
                out.at(i, j, 3 + c) = V.at(i, j, c);
                out.at(i, j, 6 + c) = curl.at(i, j, c);
                out.at(i, j, 9 + c) = div.at(i, j, c);
                out.at(i, j, 12 + c) = div.at(i, j, 3 + c);
            }
            out.at(i, j, 15) = kap.at(i, j, 0);
            out.at(i, j, 16) = kap.at(i, j, 1);
            out.at(i, j, 17) = dv.at(i, j, 0);
        }
    return !bad;
}

struct GenTerms {
    bool ok = false;
    double u[3] = {0, 0, 0}, b[3] = {0, 0, 0}, kpar = 0, kperp = 0, keff[2] = {0, 0}, divv = 0, p = 0,
           prate = 0;
    double h[2] = {0, 0};  // width of the particle's cell along each axis
};

// evaluate_terms_from_packed (ParkerSolver.hpp:771-875).
GenTerms gen_terms(const OField& F, const kpt_generic_parker_params& P, const double pos[2], double p_in) {
    GenTerms T;
    double v[KPT_GENERIC_PARKER_COMPONENTS];
    sample(F, pos, v);
    for (int d = 0; d < 2; ++d) T.h[d] = cell_width(F.g, d, pos[d]);
    const double pm = std::fabs(p_in);
    const double nm = pm / (P.rest_mass * P.speed_of_light);
    const double gamma = std::sqrt(1.0 + nm * nm);
    const double speed = pm / (gamma * P.rest_mass);
    const double bm = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (!(std::isfinite(bm) && bm > 0.0)) return T;
    const double ib = 1.0 / bm;
    for (int c = 0; c < 3; ++c) T.b[c] = v[c] * ib;
    double adv[3];
    const double fpar = std::pow(gamma, 1.0 / 3.0);
    if (P.perpendicular_model == 1) {
        for (int c = 0; c < 3; ++c) adv[c] = fpar * v[9 + c];
    } else {
        const double fperp = std::pow(gamma, 1.0 / 9.0);
        for (int c = 0; c < 3; ++c) adv[c] = fpar * v[9 + c] + fperp * v[12 + c];
    }
    T.kpar = v[15] * fpar;
    T.kperp = v[16] * (P.perpendicular_model == 1 ? fpar : std::pow(gamma, 1.0 / 9.0));
    if (!(std::isfinite(T.kpar) && T.kpar >= 0.0 && std::isfinite(T.kperp) && T.kperp >= 0.0)) return T;
    for (int d = 0; d < 2; ++d) {
        T.keff[d] = T.kperp + (T.kpar - T.kperp) * T.b[d] * T.b[d];
        if (!(std::isfinite(T.keff[d]) && T.keff[d] >= 0.0)) return T;
    }
    T.divv = v[17];
    const double drift = pm * speed * P.speed_of_light / (3.0 * P.charge);
    for (int c = 0; c < 3; ++c) T.u[c] = v[3 + c] + drift * v[6 + c] + adv[c];
    T.p = pm;
    T.prate = -(pm / 3.0) * T.divv;
    T.ok = true;
    return T;
}

// local_stable_time_step (ParkerSolver.hpp:956-1007) with the width of the
// particle's cell.
double gen_local_dt(const GenTerms& T) {
    const double large = 1.0e300;
    double ddt = large, adt = large;
    for (int d = 0; d < 2; ++d) {
        const double dx = T.h[d];
        double di = large, ai = large;
        if (T.keff[d] > 0.0) {
            const double lim = 0.5 * dx;
            di = lim * lim / (2.0 * T.keff[d]);
        }
        const double a = std::fabs(T.u[d]);
        if (a > 0.0) ai = 1.0 * dx / a;
        ddt = ddt < di ? ddt : di;
        adt = adt < ai ? adt : ai;
    }
    double mdt = large;
    const double dvm = std::fabs(T.divv);
    if (dvm > 0.0) mdt = 3.0 * 1.0 / dvm;
    const double m = ddt < adt ? ddt : adt;
    return m < mdt ? m : mdt;
}

// Four standard normals: Box-Muller on the four Philox uniforms (0, 1), or the
// injected row (g_x, g_y, g_z, g_parallel).
void gen_normals(const double* table, uint32_t steps, uint64_t i, uint64_t id, uint32_t step, const uint32_t key[2],
                 double n[4]) {
    if (table) {
        const double* row = table + (static_cast<size_t>(i) * steps + step) * 4;
        for (int d = 0; d < 4; ++d) n[d] = row[d];
        return;
    }
    const uint32_t ctr[4] = {static_cast<uint32_t>(id), static_cast<uint32_t>(id >> 32), 0x47454E50u, step};
    uint32_t r[4];
    philox10(ctr, key, r);
    for (int k = 0; k < 2; ++k) {
        const double rad = std::sqrt(-2.0 * std::log(u01(r[2 * k])));
        const double ang = 6.283185307179586476925286766559 * u01(r[2 * k + 1]);
        n[2 * k] = rad * std::cos(ang);
        n[2 * k + 1] = rad * std::sin(ang);
    }
}

}  // namespace

namespace {

// apply_grid_boundary_conditions (periodic wrap of alive particles) followed by
// capture_previous_positions (include/ParticleSystem.hpp:623-628) for every
// stored slot, as ParkerSolver / FocusTransportSolver::advance_* do before their
// update (include/ParkerSolver.hpp:265-282, include/FocusSolver.hpp:276-293);
// the capture is skipped when the caller passes no previous-position arrays.
void wrap_and_capture(const OGrid& g, kpt_particles_host* P, uint64_t count) {
    const bool capture = P->prev_x && P->prev_y && P->prev_step_x && P->prev_step_y;
    for (uint64_t i = 0; i < count; ++i) {
        if (P->status[i] == 0) {
            P->x[i] = g.wrap(0, P->x[i]);
            P->y[i] = g.wrap(1, P->y[i]);
        }
        if (!capture) continue;
        P->prev_step_x[i] = P->prev_x[i];
        P->prev_step_y[i] = P->prev_y[i];
        P->prev_x[i] = P->x[i];
        P->prev_y[i] = P->y[i];
    }
}

}  // namespace

extern "C" {

void kpto_generic_default_params(kpt_generic_parker_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->perpendicular_model = 0;
    p->parallel_prefactor = 1.622;
    p->perpendicular_prefactor = 0.198;
    p->sigma2 = 0.5;
    p->l2d = 1.0;
    p->lslab = 1.0;
    p->lc = -1.0;
    p->rest_mass = 1.0;
    p->charge = 1.0;
    p->speed_of_light = 1.0;
    p->courant = 0.3;
}

// Packed generic basis from B and V planes [ny][nx]; nullptr where the reference
// aborts (|B| = 0 without floor, non-positive kappa inputs).
void* kpto_generic_create(int32_t nx, int32_t ny, const double* xe, const double* ye, const double* const planes[6],
                          const kpt_generic_parker_params* p) {
    OField B, V;
    for (OField* f : {&B, &V}) {
        make_grid(f->g, nx, ny, xe, ye);
        f->g.cell_centred = true;
        *f = gen_alloc(f->g, 3);
    }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            for (int c = 0; c < 3; ++c) {
                B.at(i, j, c) = planes[c][static_cast<size_t>(j) * nx + i];
                V.at(i, j, c) = planes[3 + c][static_cast<size_t>(j) * nx + i];
            }
    fill_ghosts_periodic(B);
    fill_ghosts_periodic(V);
    OField* out = new OField();
    if (!gen_build(B, V, *p, *out)) {
        delete out;
        return nullptr;
    }
    return out;
}

// compute_adaptive_time_step (ParkerSolver.hpp:1084-1130): courant * min, 0 when
// no alive particle gives a finite bound.
// OpenMP threads the generic solver loops use (OMP_NUM_THREADS or every core).
int32_t kpto_max_threads() { return omp_get_max_threads(); }

double kpto_generic_time_step(const void* field, const kpt_generic_parker_params* p, const kpt_particles_host* P,
                              uint64_t count) {
    const OField& F = *static_cast<const OField*>(field);
    double m = 1.0e300;
    // particles are independent and the minimum is exact in any order
#pragma omp parallel for schedule(static, 1024) reduction(min : m)
    for (uint64_t i = 0; i < count; ++i) {
        if (P->status[i] != 0) continue;
        const double pos[2] = {F.g.wrap(0, P->x[i]), F.g.wrap(1, P->y[i])};
        const GenTerms T = gen_terms(F, *p, pos, P->p[i]);
        if (!T.ok) continue;
        const double d = gen_local_dt(T);
        m = m < d ? m : d;
    }
    if (!std::isfinite(m) || m >= 0.5e300) return 0.0;
    return p->courant * m;
}

// advance_stochastic (ParkerSolver.hpp:265-299, :1462-1512) with the increment of
// sample_magnetic_diffusion_increment (:443-477). Returns dt (0 = no step).
double kpto_generic_step(const void* field, const kpt_generic_parker_params* p, kpt_particles_host* P, uint64_t count,
                         uint32_t step, uint64_t seed, const double* noise, uint32_t noise_steps,
                         uint64_t* failures) {
    const OField& F = *static_cast<const OField*>(field);
    const double dt = kpto_generic_time_step(field, p, P, count);
    if (failures) *failures = 0;
    if (!(dt > 0.0)) return 0.0;
    wrap_and_capture(F.g, P, count);
    const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
    uint64_t fails = 0;
#pragma omp parallel for schedule(static, 1024) reduction(+ : fails)
    for (uint64_t i = 0; i < count; ++i) {
        if (P->status[i] != 0) continue;
        const double pos[2] = {F.g.wrap(0, P->x[i]), F.g.wrap(1, P->y[i])};
        const GenTerms T = gen_terms(F, *p, pos, P->p[i]);
        if (!T.ok) {
            P->status[i] = 2;
            ++fails;
            continue;
        }
        double n[4];
        gen_normals(noise, noise_steps, i, P->id ? P->id[i] : i, step, key, n);
        double iso = 0.0;
        for (int c = 0; c < 3; ++c) iso += n[c] * T.b[c];
        const double sperp = T.kperp > 0.0 ? std::sqrt(2.0 * T.kperp * dt) : 0.0;
        const double spar = T.kpar > 0.0 ? std::sqrt(2.0 * T.kpar * dt) : 0.0;
        const double repl = spar * n[3] - sperp * iso;
        double x[2] = {pos[0], pos[1]};
        for (int d = 0; d < 2; ++d) {
            const double det = dt * T.u[d];
            const double sto = sperp * n[d] + repl * T.b[d];
            x[d] += det + sto;
        }
        P->x[i] = F.g.wrap(0, x[0]);
        P->y[i] = F.g.wrap(1, x[1]);
        const double pn = T.p + dt * T.prate;
        P->p[i] = pn > 0.0 ? pn : 0.0;
    }
    if (failures) *failures = fails;
    return dt;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// ParticleSystem lifecycle (include/ParticleSystem.hpp): energy-growth split,
// baseline reset and grid boundary conditions, serial over slots.
// ---------------------------------------------------------------------------
namespace {

// ParticleProperties::kinetic_energy_from_momentum_magnitude (:190-194).
double kinetic_energy(double pm, const kpt_particle_properties& pr) {
    const double nm = pm / (pr.rest_mass * pr.speed_of_light);
    const double gamma = std::sqrt(1.0 + nm * nm);
    return (gamma - 1.0) * pr.rest_mass * pr.speed_of_light * pr.speed_of_light;
}

}  // namespace

extern "C" {

// split_particles_on_energy_growth (include/ParticleSystem.hpp:802-917) with the
// parents taken in slot order; children take ids first_root, first_root + 1, ...
// Returns the new count; stats[5] = alive, split, skipped capacity, invalid
// baseline, light child.
uint64_t kpto_split_energy_growth(const kpt_particle_properties* pr, double ratio_min, double min_child_weight,
                                  kpt_particles_host* P, uint64_t count, uint64_t capacity, uint64_t first_root,
                                  uint64_t stats[5]) {
    for (int k = 0; k < 5; ++k) stats[k] = 0;
    if (count == 0 || count >= capacity) return count;
    const double log_ratio = std::log(ratio_min);
    uint64_t appended = 0;
    for (uint64_t i = 0; i < count; ++i) {
        if (P->status[i] != 0) continue;
        ++stats[0];
        const double pm = P->p[i] >= 0.0 ? P->p[i] : -P->p[i];
        const double energy = kinetic_energy(pm, *pr);
        const double base = P->e0[i];
        if (!std::isfinite(base) || base <= 0.0) {
            if (std::isfinite(energy) && energy > 0.0) P->e0[i] = energy;
            ++stats[3];
            continue;
        }
        const double ratio = energy / base;
        if (!std::isfinite(ratio) || ratio < ratio_min) continue;
        const int target = static_cast<int>(std::floor(std::log(ratio) / log_ratio));
        if (target <= P->split_level[i]) continue;
        const double cw = 0.5 * P->weight[i];
        if (!std::isfinite(cw) || cw <= 0.0 || cw < min_child_weight) {
            ++stats[4];
            continue;
        }
        const uint64_t rank = appended++;
        const uint64_t child = count + rank;
        if (child >= capacity) {
            ++stats[2];
            continue;
        }
        P->x[child] = P->x[i];
        P->y[child] = P->y[i];
        P->p[child] = P->p[i];
        P->t[child] = P->t[i];
        P->step_dt[child] = P->step_dt[i];
        P->mu[child] = P->mu[i];
        P->e0[child] = base;
        P->weight[i] = cw;
        P->weight[child] = cw;
        P->status[child] = 0;
        P->split_level[i] += 1;
        P->split_level[child] = P->split_level[i];
        if (P->id) P->id[child] = first_root + rank;
        // both previous positions and the sort key follow the parent (:883-896)
        if (P->prev_x) P->prev_x[child] = P->prev_x[i];
        if (P->prev_y) P->prev_y[child] = P->prev_y[i];
        if (P->prev_step_x) P->prev_step_x[child] = P->prev_step_x[i];
        if (P->prev_step_y) P->prev_step_y[child] = P->prev_step_y[i];
        if (P->sort_key) P->sort_key[child] = P->sort_key[i];
        ++stats[1];
    }
    return count + stats[1];
}

// reset_split_baselines (include/ParticleSystem.hpp:633-653).
void kpto_reset_split_baselines(const kpt_particle_properties* pr, kpt_particles_host* P, uint64_t count) {
    for (uint64_t i = 0; i < count; ++i) {
        P->e0[i] = kinetic_energy(P->p[i] >= 0.0 ? P->p[i] : -P->p[i], *pr);
        P->split_level[i] = 0;
    }
}

// apply_grid_boundary_conditions (include/ParticleSystem.hpp:661-706) on the
// grid of edges xe[0..nx], ye[0..ny]. Returns the number of particles escaped.
uint64_t kpto_apply_boundaries(int32_t nx, int32_t ny, const double* xe, const double* ye, int32_t mask,
                               kpt_particles_host* P, uint64_t count) {
    OGrid g;
    make_grid(g, nx, ny, xe, ye);
    uint64_t escaped = 0;
    double* pos[2] = {P->x, P->y};
    for (uint64_t i = 0; i < count; ++i) {
        if (P->status[i] != 0) continue;
        bool out = false;
        for (int d = 0; d < 2; ++d) {
            const double x = pos[d][i];
            if (mask & (1 << d))
                pos[d][i] = g.wrap(d, x);
            else if (!(x >= g.phys_lo[d] && x <= g.phys_hi[d]))
                out = true;
        }
        if (out) {
            P->status[i] = 1;
            ++escaped;
        }
    }
    return escaped;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Generic focused-transport solver: the packed 27-component basis of
// FocusCoefficient::make_legacy_transport_coefficient_fields_with_time_derivative
// (include/FocusCoefficient.hpp:858-990) as FocusTransportSolver::
// build_packed_coefficient_basis (include/FocusSolver.hpp:106-136, :497-590)
// packs it, and the solver's global-dt step (include/FocusSolver.hpp:276-299,
// :635-719, :905-1033, :1149-1232, :1560-1620). Intermediate fields keep the
// reference's order: |B| and b and ln|B| on every stored point, derivatives on
// the physical points followed by the periodic ghost fill.
// ---------------------------------------------------------------------------
namespace {

constexpr int kFC = KPT_GENERIC_FOCUS_COMPONENTS;

double floor_positive(double v, double f) { return f <= 0.0 ? v : (v < f ? f : v); }

// physical_vector_gradient_component, Cartesian (FocusCoefficient.hpp:184-216):
// d V_out / d x_der, zero for the invariant z.
double vec_grad(const OField& f, int i, int j, int out, int der) {
    if (der >= 2) return 0.0;
    return gen_derivative(f.g, i, j, der, [&](int a, int b) { return f.at(a, b, out); });
}

bool foc_build(const OField& B, const OField& V, const OField* dvdt, const kpt_generic_focus_params& P, OField& out) {
    const kpt_generic_parker_params& M = P.model;
    const OGrid& g = B.g;
    const int nx = g.n[0], ny = g.n[1];
    const double ff = P.minimum_magnetic_field_magnitude;
    bool bad = false;
    OField mag = gen_alloc(g, 1), dir = gen_alloc(g, 3), lnb = gen_alloc(g, 1), glnb = gen_alloc(g, 3),
           curv = gen_alloc(g, 3), drift = gen_alloc(g, 6), kap = gen_alloc(g, 2), dperp = gen_alloc(g, 3),
           dv = gen_alloc(g, 1), con = gen_alloc(g, 3);
    // make_magnitude_field, make_direction_field (Abort at |B| <= floor),
    // make_log_magnetic_field (FieldOperator.hpp:616-640, :700-800; FocusCoefficient.hpp:486-520)
    for (int j = -kGhost; j < ny + kGhost; ++j)
        for (int i = -kGhost; i < nx + kGhost; ++i) {
            const double m = gen_bmag(B, i, j);
            mag.at(i, j, 0) = m;
            if (m <= ff) {
                bad = true;
                continue;
            }
            const double inv = 1.0 / m;
            for (int c = 0; c < 3; ++c) dir.at(i, j, c) = B.at(i, j, c) * inv;
            const double safe = m > ff ? m : ff;
            if (safe <= 0.0) bad = true;
            lnb.at(i, j, 0) = std::log(safe);
        }
    if (bad) return false;
    // FieldOperator::make_gradient_field of ln|B|; magnetic curvature (b.grad)b
    // (FocusCoefficient.hpp:430-450, :545-580)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            for (int d = 0; d < 2; ++d)
                glnb.at(i, j, d) = gen_derivative(g, i, j, d, [&](int a, int b) { return lnb.at(a, b, 0); });
            glnb.at(i, j, 2) = 0.0;
            for (int c = 0; c < 3; ++c) {
                double r = 0.0;
                for (int d = 0; d < 3; ++d) r += dir.at(i, j, d) * vec_grad(dir, i, j, c, d);
                curv.at(i, j, c) = r;
            }
        }
    fill_ghosts_periodic(glnb);
    fill_ghosts_periodic(curv);
    // fill_focused_drift_basis_fields (FocusCoefficient.hpp:600-660), every point
    auto cross = [](const double a[3], const double b[3], int c) {
        return c == 0 ? a[1] * b[2] - a[2] * b[1] : (c == 1 ? a[2] * b[0] - a[0] * b[2] : a[0] * b[1] - a[1] * b[0]);
    };
    for (int j = -kGhost; j < ny + kGhost; ++j)
        for (int i = -kGhost; i < nx + kGhost; ++i) {
            const double mm = mag.at(i, j, 0) > ff ? mag.at(i, j, 0) : ff;
            if (mm <= 0.0) bad = true;
            const double inv = 1.0 / mm;
            const double bb[3] = {dir.at(i, j, 0), dir.at(i, j, 1), dir.at(i, j, 2)};
            const double gg[3] = {glnb.at(i, j, 0), glnb.at(i, j, 1), glnb.at(i, j, 2)};
            const double cc[3] = {curv.at(i, j, 0), curv.at(i, j, 1), curv.at(i, j, 2)};
            for (int c = 0; c < 3; ++c) {
                drift.at(i, j, c) = cross(bb, gg, c) * inv;
                drift.at(i, j, 3 + c) = cross(bb, cc, c) * inv;
            }
        }
    // gamma-independent kappa (LegencyModel.hpp:801-871) and
    // div(kappa_perp (I - bb)) with the diffusion floor (FocusCoefficient.hpp:840-856)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double vs = 0.0;
            for (int k = 0; k < 3; ++k) vs += V.at(i, j, k) * V.at(i, j, k);
            gen_kappa(M, gen_bmag(B, i, j), std::sqrt(vs), kap.at(i, j, 0), kap.at(i, j, 1), bad);
        }
    fill_ghosts_periodic(kap);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            for (int c = 0; c < 3; ++c) {
                double s = 0.0;
                for (int d = 0; d < 2; ++d)
                    s += gen_derivative(g, i, j, d, [&](int a, int b) {
                        const double bi = gen_bdir(B, a, b, c, M.minimum_magnetic_field_magnitude, bad);
                        const double bj = gen_bdir(B, a, b, d, M.minimum_magnetic_field_magnitude, bad);
                        return kap.at(a, b, 1) * ((c == d ? 1.0 : 0.0) - bi * bj);
                    });
                dperp.at(i, j, c) = s;
            }
    fill_ghosts_periodic(dperp);
    // div V; focusing b.grad ln|B|, bb:grad V, b.dV/dt + b.(V.grad)V
    // (FocusCoefficient.hpp:218-330, :700-778)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double s = 0.0;
            for (int d = 0; d < 2; ++d)
                s += gen_derivative(g, i, j, d, [&](int a, int b) { return V.at(a, b, d); });
            dv.at(i, j, 0) = s;
            double foc = 0.0, bbg = 0.0, conv = 0.0, part = 0.0;
            for (int c = 0; c < 3; ++c) foc += dir.at(i, j, c) * glnb.at(i, j, c);
            for (int a = 0; a < 3; ++a)
                for (int d = 0; d < 3; ++d) bbg += dir.at(i, j, a) * dir.at(i, j, d) * vec_grad(V, i, j, a, d);
            if (dvdt)
                for (int c = 0; c < 3; ++c) part += dir.at(i, j, c) * dvdt->at(i, j, c);
            for (int a = 0; a < 3; ++a)
                for (int d = 0; d < 3; ++d) conv += dir.at(i, j, a) * V.at(i, j, d) * vec_grad(V, i, j, a, d);
            con.at(i, j, 0) = foc;
            con.at(i, j, 1) = bbg;
            con.at(i, j, 2) = part + conv;
        }
    fill_ghosts_periodic(dv);
    fill_ghosts_periodic(con);
    if (bad) return false;
    out = gen_alloc(g, kFC);
    const double lc = M.lc > 0.0 ? M.lc : M.lslab;
    for (int j = -kGhost; j < ny + kGhost; ++j)
        for (int i = -kGhost; i < nx + kGhost; ++i) {
            for (int c = 0; c < 3; ++c) {
                out.at(i, j, c) = B.at(i, j, c);
                out.at(i, j, 3 + c) = dir.at(i, j, c);
                out.at(i, j, 6 + c) = V.at(i, j, c);
                out.at(i, j, 9 + c) = drift.at(i, j, c);
                out.at(i, j, 12 + c) = drift.at(i, j, 3 + c);
                out.at(i, j, 15 + c) = dperp.at(i, j, c);
            }
            out.at(i, j, 18) = mag.at(i, j, 0);
            out.at(i, j, 19) = kap.at(i, j, 0);
            out.at(i, j, 20) = kap.at(i, j, 1);
            out.at(i, j, 21) = dv.at(i, j, 0);
            out.at(i, j, 22) = con.at(i, j, 0);
            out.at(i, j, 23) = con.at(i, j, 1);
            out.at(i, j, 24) = con.at(i, j, 2);
            out.at(i, j, 25) = M.sigma2;
            out.at(i, j, 26) = lc;
        }
    return true;
}

struct FocTerms {
    bool ok = false;
    double b[3] = {0, 0, 0}, u[2] = {0, 0}, keff[2] = {0, 0};
    double p = 0, mu = 0, gamma = 0, speed = 0, kpar = 0, kperp = 0, prate = 0;
    double bmag = 0, divv = 0, bbg = 0, bdv = 0, foc = 0, s2 = 0, lc = 0;
    double D = 0, dD = 0, mu_rate = 0;
    double h[2] = {0, 0};  // width of the particle's cell along each axis
};

// evaluate_pitch_angle_at_mu (FocusSolver.hpp:665-694) with
// LegencyModel::evaluate_pitch_angle_scattering (LegencyModel.hpp:650-689).
bool foc_pitch(const kpt_generic_focus_params& P, const FocTerms& T, double trial, double& D, double& dD,
               double& rate) {
    const kpt_generic_parker_params& M = P.model;
    const double mu = reflect_mu(trial, P.maximum_pitch_angle_cosine);
    const double omm = 1.0 - mu * mu;
    if (T.speed <= 0.0) return false;
    const double focusing = 0.5 * omm * (-T.speed * T.foc + mu * T.divv - 3.0 * mu * T.bbg - 2.0 * T.bdv / T.speed);
    const double sb = floor_positive(T.bmag, P.scattering_minimum_magnetic_field_magnitude);
    const double ss = floor_positive(T.s2, P.scattering_minimum_sigma2);
    const double sl = floor_positive(T.lc, P.scattering_minimum_correlation_length);
    const double sv = floor_positive(T.speed, P.scattering_minimum_speed);
    if (sb <= 0.0 || ss <= 0.0 || sl <= 0.0 || sv <= 0.0) return false;
    const double omega = std::fabs(M.charge) * sb / (M.rest_mass * M.speed_of_light) / T.gamma;
    const QltD q = qlt_scattering(mu, omega, sv, ss, sl, P.spectral_index, P.resonance_regularization_h0);
    D = q.D;
    dD = q.dD;
    rate = focusing + q.dD;
    return true;
}

// evaluate_terms_from_packed (FocusSolver.hpp:905-1033) at a wrapped position.
FocTerms foc_terms(const OField& F, const kpt_generic_focus_params& P, const double pos[2], double p_in,
                   double mu_in) {
    const kpt_generic_parker_params& M = P.model;
    FocTerms T;
    double v[kFC];
    sample(F, pos, v);
    for (int d = 0; d < 2; ++d) T.h[d] = cell_width(F.g, d, pos[d]);
    T.p = std::fabs(p_in);
    T.mu = reflect_mu(mu_in, P.maximum_pitch_angle_cosine);
    const double nm = T.p / (M.rest_mass * M.speed_of_light);
    T.gamma = std::sqrt(1.0 + nm * nm);
    T.speed = T.p / (T.gamma * M.rest_mass);
    T.bmag = v[18];
    if (!(std::isfinite(T.bmag) && T.bmag > 0.0)) return T;
    const double dm = std::sqrt(v[3] * v[3] + v[4] * v[4] + v[5] * v[5]);
    if (!(std::isfinite(dm) && dm > 0.0)) return T;
    const double idm = 1.0 / dm;
    for (int c = 0; c < 3; ++c) T.b[c] = v[3 + c] * idm;
    const double fperp = M.perpendicular_model == 1 ? std::pow(T.gamma, 1.0 / 3.0) : std::pow(T.gamma, 1.0 / 9.0);
    T.kpar = v[19] * std::pow(T.gamma, 1.0 / 3.0);
    T.kperp = v[20] * fperp;
    if (!(std::isfinite(T.kpar) && T.kpar >= 0.0 && std::isfinite(T.kperp) && T.kperp >= 0.0)) return T;
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
    if (!foc_pitch(P, T, T.mu, T.D, T.dD, T.mu_rate)) return T;
    T.ok = true;
    return T;
}

// local_stable_time_step (FocusSolver.hpp:1149-1232).
double foc_local_dt(const FocTerms& T) {
    const double large = 1.0e300;
    double ddt = large, adt = large;
    for (int d = 0; d < 2; ++d) {
        const double dx = T.h[d];
        double di = large, ai = large;
        if (T.keff[d] > 0.0) {
            const double lim = 0.5 * dx;
            di = lim * lim / (2.0 * T.keff[d]);
        }
        const double a = std::fabs(T.u[d]);
        if (a > 0.0) ai = 1.0 * dx / a;
        ddt = ddt < di ? ddt : di;
        adt = adt < ai ? adt : ai;
    }
    double mdt = large, padt = large, pddt = large, pbdt = large;
    const double pr = std::fabs(T.prate);
    if (T.p > 0.0 && pr > 0.0) mdt = 1.0 * T.p / pr;
    const double mr = std::fabs(T.mu_rate);
    if (mr > 0.0) padt = 10.0 * 0.1 / mr;
    if (T.D > 0.0) pddt = 10.0 * 0.1 * 0.1 / (2.0 * T.D);
    if (T.D > 0.0 && mr > 0.0) pbdt = 10.0 * 2.0 * T.D / (mr * mr);
    const double a = ddt < adt ? ddt : adt;
    const double b = a < mdt ? a : mdt;
    const double c = pddt < pbdt ? pddt : pbdt;
    const double e = padt < c ? padt : c;
    return b < e ? b : e;
}

// Four standard normals of step `step` (Box-Muller on one Philox block with
// counter (id_lo, id_hi, 0x464F4355, step)) or a noise-table row.
void foc_normals(const double* table, uint32_t steps, uint64_t i, uint64_t id, uint32_t step, const uint32_t key[2],
                 double n[4]) {
    if (table) {
        const double* row = table + (static_cast<size_t>(i) * steps + step) * 4;
        for (int d = 0; d < 4; ++d) n[d] = row[d];
        return;
    }
    const uint32_t ctr[4] = {static_cast<uint32_t>(id), static_cast<uint32_t>(id >> 32), 0x464F4355u, step};
    uint32_t r[4];
    philox10(ctr, key, r);
    for (int k = 0; k < 2; ++k) {
        const double rad = std::sqrt(-2.0 * std::log(u01(r[2 * k])));
        const double ang = 6.283185307179586476925286766559 * u01(r[2 * k + 1]);
        n[2 * k] = rad * std::cos(ang);
        n[2 * k + 1] = rad * std::sin(ang);
    }
}

}  // namespace

extern "C" {

void kpto_generic_focus_default_params(kpt_generic_focus_params* p) {
    std::memset(p, 0, sizeof(*p));
    kpto_generic_default_params(&p->model);
    p->spectral_index = 5.0 / 3.0;
    p->resonance_regularization_h0 = 1.0e-6;
    p->maximum_pitch_angle_cosine = 1.0;
}

// dvdt: nullptr or three planes [3][ny][nx]. nullptr where the reference aborts.
void* kpto_generic_focus_create(int32_t nx, int32_t ny, const double* xe, const double* ye,
                                const double* const planes[6], const double* dvdt,
                                const kpt_generic_focus_params* p) {
    OField B, V, T;
    for (OField* f : {&B, &V, &T}) {
        make_grid(f->g, nx, ny, xe, ye);
        f->g.cell_centred = true;
        *f = gen_alloc(f->g, 3);
    }
    const size_t plane = static_cast<size_t>(nx) * ny;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            for (int c = 0; c < 3; ++c) {
                const size_t k = static_cast<size_t>(j) * nx + i;
                B.at(i, j, c) = planes[c][k];
                V.at(i, j, c) = planes[3 + c][k];
                if (dvdt) T.at(i, j, c) = dvdt[c * plane + k];
            }
    fill_ghosts_periodic(B);
    fill_ghosts_periodic(V);
    fill_ghosts_periodic(T);
    OField* out = new OField();
    if (!foc_build(B, V, dvdt ? &T : nullptr, *p, *out)) {
        delete out;
        return nullptr;
    }
    return out;
}

double kpto_generic_focus_time_step(const void* field, const kpt_generic_focus_params* p, const kpt_particles_host* P,
                                    uint64_t count) {
    const OField& F = *static_cast<const OField*>(field);
    double m = 1.0e300;
#pragma omp parallel for schedule(static, 1024) reduction(min : m)
    for (uint64_t i = 0; i < count; ++i) {
        if (P->status[i] != 0) continue;
        const double pos[2] = {F.g.wrap(0, P->x[i]), F.g.wrap(1, P->y[i])};
        const FocTerms T = foc_terms(F, *p, pos, P->p[i], P->mu[i]);
        if (!T.ok) continue;
        const double d = foc_local_dt(T);
        m = m < d ? m : d;
    }
    if (!std::isfinite(m) || m >= 0.5e300) return 0.0;
    return p->model.courant * m;
}

// Terms of one particle for tests: out = {dt_local, u_x, u_y, p_rate, mu_rate, D,
// dD, kpar, kperp, keff_x, keff_y, b_x, b_y, b_z}; returns 0 where the reference aborts.
int kpto_generic_focus_terms(const void* field, const kpt_generic_focus_params* p, double x, double y, double pm,
                             double mu, double* out) {
    const OField& F = *static_cast<const OField*>(field);
    const double pos[2] = {F.g.wrap(0, x), F.g.wrap(1, y)};
    const FocTerms T = foc_terms(F, *p, pos, pm, mu);
    if (!T.ok) return 0;
    const double v[14] = {foc_local_dt(T), T.u[0], T.u[1], T.prate, T.mu_rate, T.D, T.dD,
                          T.kpar, T.kperp, T.keff[0], T.keff[1], T.b[0], T.b[1], T.b[2]};
    for (int k = 0; k < 14; ++k) out[k] = v[k];
    return 1;
}

double kpto_generic_focus_step(const void* field, const kpt_generic_focus_params* p, kpt_particles_host* P,
                               uint64_t count, uint32_t step, uint64_t seed, const double* noise,
                               uint32_t noise_steps, uint64_t* failures) {
    const OField& F = *static_cast<const OField*>(field);
    const double dt = kpto_generic_focus_time_step(field, p, P, count);
    if (failures) *failures = 0;
    if (!(dt > 0.0)) return 0.0;
    wrap_and_capture(F.g, P, count);
    const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
    uint64_t fails = 0;
#pragma omp parallel for schedule(static, 1024) reduction(+ : fails)
    for (uint64_t i = 0; i < count; ++i) {
        if (P->status[i] != 0) continue;
        const double pos[2] = {F.g.wrap(0, P->x[i]), F.g.wrap(1, P->y[i])};
        const FocTerms T = foc_terms(F, *p, pos, P->p[i], P->mu[i]);
        if (!T.ok) {
            P->status[i] = 2;
            ++fails;
            continue;
        }
        double n[4];
        foc_normals(noise, noise_steps, i, P->id ? P->id[i] : i, step, key, n);
        // sample_perpendicular_diffusion_increment (FocusSolver.hpp:635-659)
        double par = 0.0;
        for (int c = 0; c < 3; ++c) par += n[c] * T.b[c];
        const double scale = T.kperp > 0.0 ? std::sqrt(2.0 * T.kperp * dt) : 0.0;
        double x[2] = {pos[0], pos[1]};
        for (int d = 0; d < 2; ++d) x[d] += dt * T.u[d] + scale * (n[d] - par * T.b[d]);
        P->x[i] = F.g.wrap(0, x[0]);
        P->y[i] = F.g.wrap(1, x[1]);
        const double pn = T.p + dt * T.prate;
        P->p[i] = pn > 0.0 ? pn : 0.0;
        // picard_pitch_angle_update (FocusSolver.hpp:700-719)
        const double mu0 = reflect_mu(T.mu, p->maximum_pitch_angle_cosine);
        double trial = mu0;
        for (int it = 0; it < 3; ++it) {
            double D = 0.0, dD = 0.0, rate = 0.0;
            foc_pitch(*p, T, trial, D, dD, rate);  // the inputs that can fail were checked at T.mu
            const double sc = D > 0.0 ? std::sqrt(2.0 * D * dt) : 0.0;
            trial = reflect_mu(mu0 + dt * rate + sc * n[3], p->maximum_pitch_angle_cosine);
        }
        P->mu[i] = trial;
    }
    if (failures) *failures = fails;
    return dt;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Generic particle generator (include/ParticleGenerator.hpp): candidate sets over
// the cells of a field (field threshold in a region, :463-784; every cell,
// :545-845), injection from them (:957-1117), uniform injection with rejection
// against a field-threshold predicate (:875-955) and recycling (:1176-1470).
// Raw planes are cell-centred values; cell areas are products of cell widths.
// Draws: Philox with counter (id, tag | attempt, call, id >> 32) and the seed
// as key. Coordinates follow the library's arithmetic: explicit fma for the
// analytic faces and draws, an unfused product for the analytic upper bound.
// ---------------------------------------------------------------------------
namespace {

constexpr uint32_t kGenTagInject = 0x494E0000u;
constexpr uint32_t kGenTagRecycle = 0x52430000u;

struct GenGrid {
    int n[2];
    double start[2], step[2], inv_step[2], lo[2], hi[2];
    bool stretched[2];
    const double* e[2];
};

double unfused_bound(double start, int n, double step) {
    double t = static_cast<double>(n) * step;
    asm volatile("" : "+x"(t));  // keeps the product apart from the sum
    return start + t;
}

GenGrid gen_grid(int nx, int ny, const double* xe, const double* ye) {
    GenGrid g{};
    g.n[0] = nx;
    g.n[1] = ny;
    g.e[0] = xe;
    g.e[1] = ye;
    for (int d = 0; d < 2; ++d) {
        const double* e = g.e[d];
        g.start[d] = e[0];
        g.step[d] = (e[g.n[d]] - e[0]) / static_cast<double>(g.n[d]);
        g.inv_step[d] = 1.0 / g.step[d];
        g.stretched[d] = g.n[d] > 1 && !faces_uniform(e, g.n[d] + 1);
        g.lo[d] = g.stretched[d] ? e[0] : g.start[d];
        g.hi[d] = g.stretched[d] ? e[g.n[d]] : unfused_bound(g.start[d], g.n[d], g.step[d]);
    }
    return g;
}

double gen_face(const GenGrid& g, int d, int k) {
    return g.stretched[d] ? g.e[d][k] : std::fma(static_cast<double>(k), g.step[d], g.start[d]);
}

double gen_centre(const GenGrid& g, int d, int k) {
    return g.stretched[d] ? 0.5 * (g.e[d][k] + g.e[d][k + 1])
                          : std::fma(static_cast<double>(k) + 0.5, g.step[d], g.start[d]);
}

// grid.cell_index on the physical domain: the interval e_k <= q < e_(k+1),
// clamped to the cells (q = e_n falls in the last one).
int gen_cell_of(const GenGrid& g, int d, double q) {
    const int n = g.n[d];
    if (!g.stretched[d]) {
        const int k = static_cast<int>(std::floor((q - g.start[d]) * g.inv_step[d]));
        return std::min(std::max(k, 0), n - 1);
    }
    const double* e = g.e[d];
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (e[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

double gen_area(const GenGrid& g, int i, int j) {
    const double a = (gen_face(g, 0, i + 1) - gen_face(g, 0, i)) * (gen_face(g, 1, j + 1) - gen_face(g, 1, j));
    return std::isfinite(a) && a > 0.0 ? a : 0.0;
}

bool gen_compare(double v, double thr, int cmp) {
    switch (cmp) {
        case KPT_CMP_GREATER: return v > thr;
        case KPT_CMP_GREATER_EQUAL: return v >= thr;
        case KPT_CMP_LESS: return v < thr;
        case KPT_CMP_LESS_EQUAL: return v <= thr;
        case KPT_CMP_ABS_GREATER: return std::fabs(v) > thr;
        case KPT_CMP_ABS_LESS: return std::fabs(v) < thr;
    }
    return false;
}

struct GenSource {
    GenGrid g;
    const double* raw;
    double region[4];
    int use_pred, comp, cmp, max_attempts;
    double thr;
};

GenSource gen_source(const double* raw, int nx, int ny, const double* xe, const double* ye, const double* region,
                     int use_pred, int comp, int cmp, double thr, int max_attempts) {
    GenSource s{};
    s.g = gen_grid(nx, ny, xe, ye);
    s.raw = raw;
    for (int q = 0; q < 4; ++q) s.region[q] = region ? region[q] : (q < 2 ? (q == 0 ? s.g.lo[0] : s.g.hi[0])
                                                                         : (q == 2 ? s.g.lo[1] : s.g.hi[1]));
    s.use_pred = use_pred;
    s.comp = comp;
    s.cmp = cmp;
    s.thr = thr;
    s.max_attempts = max_attempts;
    return s;
}

bool gen_predicate(const GenSource& s, double x, double y) {
    const GenGrid& g = s.g;
    if (!(x >= g.lo[0] && x <= g.hi[0] && y >= g.lo[1] && y <= g.hi[1])) return false;
    const int i = gen_cell_of(g, 0, x), j = gen_cell_of(g, 1, y);
    const size_t cells = static_cast<size_t>(g.n[0]) * g.n[1];
    return gen_compare(s.raw[static_cast<size_t>(s.comp) * cells + static_cast<size_t>(j) * g.n[0] + i], s.thr,
                       s.cmp);
}

void gen_draw(uint64_t id, uint32_t tag, uint32_t attempt, uint32_t call, uint64_t seed, uint32_t r[4]) {
    const uint32_t ctr[4] = {static_cast<uint32_t>(id), tag | attempt, call, static_cast<uint32_t>(id >> 32)};
    const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
    philox10(ctr, key, r);
}

bool gen_uniform_if(const GenSource& s, uint64_t id, uint32_t tag, uint32_t call, uint64_t seed, double& x,
                    double& y) {
    for (int a = 0; a < s.max_attempts; ++a) {
        uint32_t r[4];
        gen_draw(id, tag, static_cast<uint32_t>(a), call, seed, r);
        x = std::fma(s.region[1] - s.region[0], u01(r[0]), s.region[0]);
        y = std::fma(s.region[3] - s.region[2], u01(r[1]), s.region[2]);
        if (!s.use_pred || gen_predicate(s, x, y)) return true;
    }
    return false;
}

// sample_cumulative_weight_index (include/ParticleGenerator.hpp:436-460).
uint64_t gen_sample_index(const double* cum, uint64_t count, double total, double u) {
    const double target = total * u;
    uint64_t left = 0, right = count - 1;
    while (left < right) {
        const uint64_t mid = left + (right - left) / 2;
        if (target < cum[mid]) right = mid;
        else left = mid + 1;
    }
    return left;
}

void gen_candidate_position(int kind, const uint32_t* cells, const double* cum, uint64_t count, double total,
                            const GenGrid& g, const uint32_t r[4], double& x, double& y) {
    const uint64_t idx = gen_sample_index(cum, count, total, u01(r[0]));
    const uint64_t cell = cells ? cells[idx] : idx;
    const int i = static_cast<int>(cell % g.n[0]), j = static_cast<int>(cell / g.n[0]);
    if (kind == 0) {
        x = gen_centre(g, 0, i);
        y = gen_centre(g, 1, j);
    } else {
        const double x0 = gen_face(g, 0, i), x1 = gen_face(g, 0, i + 1);
        const double y0 = gen_face(g, 1, j), y1 = gen_face(g, 1, j + 1);
        x = std::fma(x1 - x0, u01(r[1]), x0);
        y = std::fma(y1 - y0, u01(r[2]), y0);
    }
}

void gen_place(kpt_particles_host* P, uint64_t i, double x, double y, const kpt_injection_kinematics* k, double e0,
               int status) {
    P->x[i] = x;
    P->y[i] = y;
    P->p[i] = k->momentum;
    P->mu[i] = k->mu;
    P->weight[i] = k->weight;
    P->t[i] = k->time;
    P->step_dt[i] = k->step_dt;
    P->status[i] = status;
    P->split_level[i] = 0;
    if (P->e0) P->e0[i] = e0;
}

bool gen_selected(int status, int selection) {
    switch (selection) {
        case KPT_RECYCLE_ESCAPED: return status == 1;
        case KPT_RECYCLE_INACTIVE: return status == 2;
        case KPT_RECYCLE_ANY: return status == 1 || status == 2;
    }
    return false;
}

}  // namespace

extern "C" {

// make_field_particle_candidate_set_in_region (include/ParticleGenerator.hpp:
// 658-769): accepted cells in x-fastest order with their cumulative areas (a
// sequential sum). region may be NULL (the physical domain). Returns the count.
uint64_t kpto_candidates_field(const double* raw, int32_t nx, int32_t ny, const double* xe, const double* ye,
                               const double* region, int32_t comp, int32_t cmp, double thr, uint32_t* cells,
                               double* cumulative, double* total) {
    const GenSource s = gen_source(raw, nx, ny, xe, ye, region, 1, comp, cmp, thr, 1);
    const size_t n = static_cast<size_t>(nx) * ny;
    uint64_t count = 0;
    double sum = 0.0;
    for (size_t c = 0; c < n; ++c) {
        const int i = static_cast<int>(c % nx), j = static_cast<int>(c / nx);
        const double cx = gen_centre(s.g, 0, i), cy = gen_centre(s.g, 1, j);
        if (!(cx >= s.region[0] && cx <= s.region[1] && cy >= s.region[2] && cy <= s.region[3])) continue;
        if (!gen_compare(raw[static_cast<size_t>(comp) * n + c], thr, cmp)) continue;
        sum += gen_area(s.g, i, j);
        cells[count] = static_cast<uint32_t>(c);
        cumulative[count] = sum;
        ++count;
    }
    *total = sum;
    return count;
}

// make_grid_particle_candidate_set (include/ParticleGenerator.hpp:786-845).
void kpto_candidates_grid(int32_t nx, int32_t ny, const double* xe, const double* ye, double* cumulative,
                          double* total) {
    const GenGrid g = gen_grid(nx, ny, xe, ye);
    double sum = 0.0;
    for (int64_t c = 0; c < static_cast<int64_t>(nx) * ny; ++c) {
        sum += gen_area(g, static_cast<int>(c % nx), static_cast<int>(c / nx));
        cumulative[c] = sum;
    }
    *total = sum;
}

// inject_particles_from_field_candidates / _from_grid_candidates
// (include/ParticleGenerator.hpp:957-1117) into slots first .. first + n - 1;
// kind 0 places at cell centres, 1 uniformly in the cell. cells may be NULL
// (every cell).
void kpto_generator_inject_candidates(int32_t kind, const uint32_t* cells, const double* cumulative, uint64_t count,
                                      double total, int32_t nx, int32_t ny, const double* xe, const double* ye,
                                      uint64_t n, const kpt_injection_kinematics* k, double e0, uint64_t seed,
                                      uint32_t call, uint64_t first_root, kpt_particles_host* P, uint64_t first) {
    const GenGrid g = gen_grid(nx, ny, xe, ye);
    for (uint64_t m = 0; m < n; ++m) {
        const uint64_t root = first_root + m;
        uint32_t r[4];
        gen_draw(root, kGenTagInject, 0, call, seed, r);
        double x, y;
        gen_candidate_position(kind, cells, cumulative, count, total, g, r, x, y);
        gen_place(P, first + m, x, y, k, e0, k->status);
        if (P->id) P->id[first + m] = root;
    }
}

// inject_particles_uniform_if (include/ParticleGenerator.hpp:875-955). region
// may be NULL (the physical domain). Returns the rejected count.
uint64_t kpto_generator_inject_uniform(const double* raw, int32_t nx, int32_t ny, const double* xe, const double* ye,
                                       const double* region, int32_t use_pred, int32_t comp, int32_t cmp, double thr,
                                       int32_t max_attempts, uint64_t n, const kpt_injection_kinematics* k, double e0,
                                       uint64_t seed, uint32_t call, uint64_t first_root, kpt_particles_host* P,
                                       uint64_t first) {
    const GenSource s = gen_source(raw, nx, ny, xe, ye, region, use_pred, comp, cmp, thr, max_attempts);
    uint64_t rejected = 0;
    for (uint64_t m = 0; m < n; ++m) {
        const uint64_t root = first_root + m;
        double x = 0.0, y = 0.0;
        const bool ok = gen_uniform_if(s, root, kGenTagInject, call, seed, x, y);
        gen_place(P, first + m, x, y, k, e0, ok ? k->status : 2);
        if (P->id) P->id[first + m] = root;
        if (!ok) ++rejected;
    }
    return rejected;
}

// recycle_particles_uniform_if / _from_*_candidates (include/ParticleGenerator.hpp:
// 1176-1470): mode 0 from the candidate set, 1 uniform with the predicate. ids
// are kept. Returns the recycled count; *rejected those left inactive.
uint64_t kpto_generator_recycle(int32_t mode, int32_t kind, const uint32_t* cells, const double* cumulative,
                                uint64_t count, double total, const double* raw, int32_t nx, int32_t ny,
                                const double* xe, const double* ye, const double* region, int32_t use_pred,
                                int32_t comp, int32_t cmp, double thr, int32_t max_attempts, int32_t selection,
                                const kpt_injection_kinematics* k, double e0, uint64_t seed, uint32_t call,
                                kpt_particles_host* P, uint64_t stored, uint64_t* rejected) {
    const GenSource s = gen_source(raw, nx, ny, xe, ye, region, use_pred, comp, cmp, thr, max_attempts);
    uint64_t recycled = 0, rej = 0;
    for (uint64_t i = 0; i < stored; ++i) {
        if (!gen_selected(P->status[i], selection)) continue;
        ++recycled;
        const uint64_t id = P->id[i];
        double x = 0.0, y = 0.0;
        if (mode == 0) {
            uint32_t r[4];
            gen_draw(id, kGenTagRecycle, 0, call, seed, r);
            gen_candidate_position(kind, cells, cumulative, count, total, s.g, r, x, y);
            gen_place(P, i, x, y, k, e0, k->status);
        } else {
            const bool ok = gen_uniform_if(s, id, kGenTagRecycle, call, seed, x, y);
            gen_place(P, i, x, y, k, e0, ok ? k->status : 2);
            if (!ok) ++rej;
        }
    }
    if (rejected) *rejected = rej;
    return recycled;
}

}  // extern "C"
