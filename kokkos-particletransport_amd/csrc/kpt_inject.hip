// kpt_inject.hip — the generic particle generator (include/ParticleGenerator.hpp):
// volume-weighted candidate sets over the cells of a field slot, injection from
// them, uniform injection with a field-threshold predicate and the recycling of
// escaped or inactive slots.
//
// Candidate sets are built with device scans: a flag and a cell area per cell,
// an exclusive scan of the flags that compacts the accepted cells (x-fastest
// order, as the reference's parallel_scan), and an inclusive scan of their areas
// into the cumulative weights a particle's draw is binary-searched against.
// Every particle draws from its own Philox stream keyed by (id, stream tag,
// attempt, generator call), so results do not depend on the launch shape.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdint>

#include "kpt_device_common.h"
#include "kpt_grid_device.h"
#include "kpt_internal.h"

namespace kpt {
namespace {

constexpr int kBlk = 256;

inline unsigned blocks_for(uint64_t n) { return static_cast<unsigned>((n + kBlk - 1) / kBlk); }

// Face k of axis d (k in [0, n]): the stored face on a stretched axis, start +
// k step on a uniform one (the analytic grid's face coordinate). Multiply-adds
// are explicit fma here, so the oracle rounds them the same way.
__device__ __forceinline__ double face_at(const GridDesc& g, int d, int k) {
    return g.face[d] ? g.face[d][k] : fma(static_cast<double>(k), g.step[d], g.start[d]);
}

__device__ __forceinline__ double cell_centre(const GridDesc& g, int d, int k) {
    return g.face[d] ? 0.5 * (g.face[d][k] + g.face[d][k + 1])
                     : fma(static_cast<double>(k) + 0.5, g.step[d], g.start[d]);
}

// approximate_grid_cell_physical_volume (include/ParticleGenerator.hpp:230-245)
// on a Cartesian grid: the product of the cell widths (the Jacobian is 1).
__device__ __forceinline__ double cell_area(const GridDesc& g, int i, int j) {
    const double a = (face_at(g, 0, i + 1) - face_at(g, 0, i)) * (face_at(g, 1, j + 1) - face_at(g, 1, j));
    return isfinite(a) && a > 0.0 ? a : 0.0;
}

// Cell of coordinate q inside [e_0, e_n] (grid.cell_index of the physical
// domain): the face interval containing q, the last cell for q = e_n.
__device__ __forceinline__ int cell_of(const GridDesc& g, int d, int n, double q) {
    if (g.face[d] == nullptr) {
        const int k = static_cast<int>(floor((q - g.start[d]) * g.inv_step[d]));
        return k < 0 ? 0 : (k > n - 1 ? n - 1 : k);
    }
    const double bf = (q - g.start[d]) * g.bins_scale[d];
    const int b = bf < 0.0 ? 0 : (bf >= static_cast<double>(g.bins_n[d] - 1) ? g.bins_n[d] - 1 : static_cast<int>(bf));
    const FaceBin fb = g.bins[d][b];
    int k = fb.k;
    if (q >= fb.mid && k + 1 < n) ++k;
    return k;
}

// compare_field_predicate_value (include/ParticleGenerator.hpp:230-251).
__device__ __forceinline__ bool compare_value(double v, double thr, int cmp) {
    switch (cmp) {
        case KPT_CMP_GREATER: return v > thr;
        case KPT_CMP_GREATER_EQUAL: return v >= thr;
        case KPT_CMP_LESS: return v < thr;
        case KPT_CMP_LESS_EQUAL: return v <= thr;
        case KPT_CMP_ABS_GREATER: return fabs(v) > thr;
        case KPT_CMP_ABS_LESS: return fabs(v) < thr;
    }
    return false;
}

// FieldThresholdParticlePredicate::operator() (include/ParticleGenerator.hpp:
// 297-315) for cell-centred raw values: inside the physical domain, then the
// value of the containing cell.
__device__ __forceinline__ bool field_predicate(const GeneratorSource& src, double x, double y) {
    const GridDesc& g = src.g;
    if (!(x >= g.phys_lo[0] && x <= g.phys_hi[0] && y >= g.phys_lo[1] && y <= g.phys_hi[1])) return false;
    const int i = cell_of(g, 0, g.nx, x), j = cell_of(g, 1, g.ny, y);
    const double v = src.raw[static_cast<uint64_t>(src.component) * g.nx * g.ny + static_cast<uint64_t>(j) * g.nx + i];
    return compare_value(v, src.threshold, src.comparison);
}

// particle_recycle_selection_matches (include/ParticleGenerator.hpp:138-152).
__device__ __forceinline__ bool recycle_selected(int status, int selection) {
    switch (selection) {
        case KPT_RECYCLE_ESCAPED: return status == 1;
        case KPT_RECYCLE_INACTIVE: return status == 2;
        case KPT_RECYCLE_ANY: return status == 1 || status == 2;
    }
    return false;
}

__global__ void __launch_bounds__(kBlk) candidate_flag_kernel(GeneratorSource src, int use_predicate,
                                                              uint32_t* __restrict__ flags,
                                                              double* __restrict__ area) {
    const GridDesc& g = src.g;
    const uint64_t cells = static_cast<uint64_t>(g.nx) * g.ny;
    const uint64_t c = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    if (c >= cells) return;
    const int i = static_cast<int>(c % g.nx), j = static_cast<int>(c / g.nx);
    bool ok = true;
    if (use_predicate) {
        const double cx = cell_centre(g, 0, i), cy = cell_centre(g, 1, j);
        ok = cx >= src.region[0] && cx <= src.region[1] && cy >= src.region[2] && cy <= src.region[3];
        if (ok) ok = compare_value(src.raw[static_cast<uint64_t>(src.component) * cells + c], src.threshold,
                                   src.comparison);
    }
    flags[c] = ok ? 1u : 0u;
    area[c] = cell_area(g, i, j);
}

__global__ void __launch_bounds__(kBlk) candidate_scatter_kernel(uint64_t cells, const uint32_t* __restrict__ flags,
                                                                 const uint32_t* __restrict__ offsets,
                                                                 const double* __restrict__ area,
                                                                 uint32_t* __restrict__ out_cells,
                                                                 double* __restrict__ out_area) {
    const uint64_t c = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    if (c >= cells || flags[c] == 0) return;
    out_cells[offsets[c]] = static_cast<uint32_t>(c);
    out_area[offsets[c]] = area[c];
}

// sample_cumulative_weight_index (include/ParticleGenerator.hpp:436-460): the
// first index whose cumulative weight exceeds the draw.
__device__ __forceinline__ uint64_t sample_index(const CandidateView& c, double u) {
    const double target = c.total * u;
    uint64_t left = 0, right = c.count - 1;
    while (left < right) {
        const uint64_t mid = left + (right - left) / 2;
        if (target < c.cumulative[mid]) right = mid;
        else left = mid + 1;
    }
    return left;
}

// The state every generator path writes: ParticleInjectionKinematics plus the
// driver fields the push reads; fresh particles also get their id.
__device__ __forceinline__ void place(const ParticleArrays& P, uint64_t i, double x, double y,
                                      const kpt_injection_kinematics& k, double e0, int status) {
    P.x[i] = x;
    P.y[i] = y;
    P.p[i] = k.momentum;
    P.mu[i] = k.mu;
    P.weight[i] = k.weight;
    P.t[i] = k.time;
    P.step_dt[i] = k.step_dt;
    P.status[i] = status;
    P.split_level[i] = 0;
    P.sort_key[i] = 0;
    P.e0[i] = e0;
    // both previous positions start at the new position, the sort key at 0
    // (include/ParticleGenerator.hpp:939-953)
    P.prev_x[i] = P.prev_step_x[i] = x;
    P.prev_y[i] = P.prev_step_y[i] = y;
}

// One particle's draw from a candidate set: the candidate by weight, then its
// cell centre (field candidates) or a uniform point in its cell (grid).
__device__ __forceinline__ void candidate_position(const CandidateView& c, const uint32_t r[4], double& x,
                                                   double& y) {
    const uint64_t idx = sample_index(c, u01(r[0]));
    const uint64_t cell = c.cells ? c.cells[idx] : idx;
    const int i = static_cast<int>(cell % c.g.nx), j = static_cast<int>(cell / c.g.nx);
    if (c.kind == 0) {
        x = cell_centre(c.g, 0, i);
        y = cell_centre(c.g, 1, j);
    } else {
        const double x0 = face_at(c.g, 0, i), x1 = face_at(c.g, 0, i + 1);
        const double y0 = face_at(c.g, 1, j), y1 = face_at(c.g, 1, j + 1);
        x = fma(x1 - x0, u01(r[1]), x0);
        y = fma(y1 - y0, u01(r[2]), y0);
    }
}

__device__ __forceinline__ void draw(uint64_t id, uint32_t tag, uint32_t attempt, uint32_t call, uint32_t k0,
                                     uint32_t k1, uint32_t r[4]) {
    philox4x32_10(static_cast<uint32_t>(id), tag | attempt, call, static_cast<uint32_t>(id >> 32), k0, k1, r);
}

// Uniform draws in the region, up to max_attempts while the predicate fails
// (inject_particles_uniform_if, include/ParticleGenerator.hpp:875-955): the last
// draw is kept either way.
__device__ __forceinline__ bool uniform_if(const GeneratorSource& src, uint64_t id, uint32_t tag, uint32_t call,
                                           uint32_t k0, uint32_t k1, double& x, double& y) {
    for (int a = 0; a < src.max_attempts; ++a) {
        uint32_t r[4];
        draw(id, tag, static_cast<uint32_t>(a), call, k0, k1, r);
        x = fma(src.region[1] - src.region[0], u01(r[0]), src.region[0]);
        y = fma(src.region[3] - src.region[2], u01(r[1]), src.region[2]);
        if (!src.use_predicate || field_predicate(src, x, y)) return true;
    }
    return false;
}

__global__ void __launch_bounds__(kBlk) inject_candidates_kernel(ParticleArrays P, uint64_t first, uint64_t n,
                                                                 uint64_t first_root, CandidateView c,
                                                                 kpt_injection_kinematics k, double e0,
                                                                 uint32_t call, uint32_t k0, uint32_t k1) {
    const uint64_t m = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    if (m >= n) return;
    const uint64_t i = first + m, root = first_root + m;
    uint32_t r[4];
    draw(root, kTagInject, 0, call, k0, k1, r);
    double x, y;
    candidate_position(c, r, x, y);
    place(P, i, x, y, k, e0, k.status);
    P.id[i] = root;
}

__global__ void __launch_bounds__(kBlk) inject_uniform_kernel(ParticleArrays P, uint64_t first, uint64_t n,
                                                              uint64_t first_root, GeneratorSource src,
                                                              kpt_injection_kinematics k, double e0, uint32_t call,
                                                              uint32_t k0, uint32_t k1,
                                                              unsigned long long* __restrict__ rejected) {
    const uint64_t m = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    bool rej = false;
    if (m < n) {
        const uint64_t i = first + m, root = first_root + m;
        double x, y;
        const bool ok = uniform_if(src, root, kTagInject, call, k0, k1, x, y);
        place(P, i, x, y, k, e0, ok ? k.status : 2);
        P.id[i] = root;
        rej = !ok;
    }
    const uint64_t b = __ballot(rej);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(rejected, static_cast<unsigned long long>(__popcll(b)));
}

// Recycling keeps the slot's id (recycle_particles_*, include/ParticleGenerator.hpp:
// 1176-1470); `mode` 0 draws from the candidate set, 1 uniformly with the
// predicate. counters[0] counts recycled slots, counters[1] those left inactive.
__global__ void __launch_bounds__(kBlk) recycle_kernel(ParticleArrays P, uint64_t count, int selection, int mode,
                                                       CandidateView c, GeneratorSource src,
                                                       kpt_injection_kinematics k, double e0, uint32_t call,
                                                       uint32_t k0, uint32_t k1,
                                                       unsigned long long* __restrict__ counters) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    bool hit = false, rej = false;
    if (i < count && recycle_selected(P.status[i], selection)) {
        hit = true;
        const uint64_t id = P.id[i];
        double x, y;
        if (mode == 0) {
            uint32_t r[4];
            draw(id, kTagRecycle, 0, call, k0, k1, r);
            candidate_position(c, r, x, y);
            place(P, i, x, y, k, e0, k.status);
        } else {
            const bool ok = uniform_if(src, id, kTagRecycle, call, k0, k1, x, y);
            place(P, i, x, y, k, e0, ok ? k.status : 2);
            rej = !ok;
        }
    }
    const uint64_t bh = __ballot(hit), br = __ballot(rej);
    if ((threadIdx.x & 63) == 0) {
        if (bh) atomicAdd(&counters[0], static_cast<unsigned long long>(__popcll(bh)));
        if (br) atomicAdd(&counters[1], static_cast<unsigned long long>(__popcll(br)));
    }
}

}  // namespace

hipError_t candidate_scratch_bytes(uint64_t cells, size_t* bytes) {
    size_t a = 0, b = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, a, static_cast<const uint32_t*>(nullptr),
                                                    static_cast<uint32_t*>(nullptr), static_cast<int>(cells));
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::InclusiveSum(nullptr, b, static_cast<const double*>(nullptr),
                                         static_cast<double*>(nullptr), static_cast<int>(cells));
    *bytes = a > b ? a : b;
    return e;
}

hipError_t launch_candidates(const GeneratorSource& src, int use_predicate, CandidateBuild& w, hipStream_t stream) {
    const uint64_t cells = static_cast<uint64_t>(src.g.nx) * src.g.ny;
    hipLaunchKernelGGL(candidate_flag_kernel, dim3(blocks_for(cells)), dim3(kBlk), 0, stream, src, use_predicate,
                       w.flags, w.area);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = w.temp_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(w.temp, bytes, w.flags, w.offsets, static_cast<int>(cells), stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(candidate_scatter_kernel, dim3(blocks_for(cells)), dim3(kBlk), 0, stream, cells, w.flags,
                       w.offsets, w.area, w.out_cells, w.out_area);
    return hipGetLastError();
}

hipError_t launch_cumulative(const double* area, double* cumulative, uint64_t count, void* temp, size_t temp_bytes,
                             hipStream_t stream) {
    if (count == 0) return hipSuccess;
    size_t bytes = temp_bytes;
    return hipcub::DeviceScan::InclusiveSum(temp, bytes, area, cumulative, static_cast<int>(count), stream);
}

hipError_t launch_inject_candidates(const ParticleArrays& P, uint64_t first, uint64_t n, uint64_t first_root,
                                    const CandidateView& c, const kpt_injection_kinematics& k, double e0,
                                    uint32_t call, const uint32_t key[2], hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(inject_candidates_kernel, dim3(blocks_for(n)), dim3(kBlk), 0, stream, P, first, n,
                       first_root, c, k, e0, call, key[0], key[1]);
    return hipGetLastError();
}

hipError_t launch_inject_uniform(const ParticleArrays& P, uint64_t first, uint64_t n, uint64_t first_root,
                                 const GeneratorSource& src, const kpt_injection_kinematics& k, double e0,
                                 uint32_t call, const uint32_t key[2], unsigned long long* rejected,
                                 hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(inject_uniform_kernel, dim3(blocks_for(n)), dim3(kBlk), 0, stream, P, first, n, first_root,
                       src, k, e0, call, key[0], key[1], rejected);
    return hipGetLastError();
}

hipError_t launch_recycle(const ParticleArrays& P, uint64_t count, int selection, int mode, const CandidateView& c,
                          const GeneratorSource& src, const kpt_injection_kinematics& k, double e0, uint32_t call,
                          const uint32_t key[2], unsigned long long* counters, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(recycle_kernel, dim3(blocks_for(count)), dim3(kBlk), 0, stream, P, count, selection, mode, c,
                       src, k, e0, call, key[0], key[1], counters);
    return hipGetLastError();
}

}  // namespace kpt
