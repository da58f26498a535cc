// kpt_particles.hip — ParticleSystem lifecycle passes (include/ParticleSystem.hpp):
// the energy-growth split, split-baseline reset, grid boundary conditions and
// the relativistic kinematics of ParticleProperties. One lane per particle over
// the SoA arrays; the split appends children in parent slot order through a
// ballot rank inside each 256-slot tile plus a scan of the tile counts.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "kpt_device_common.h"
#include "kpt_grid_device.h"
#include "kpt_internal.h"

namespace kpt {
namespace {

constexpr int kBlk = 256;

inline unsigned tiles_for(uint64_t n) { return static_cast<unsigned>((n + kBlk - 1) / kBlk); }

// ParticleProperties::gamma_from_momentum_magnitude / speed_from_momentum_magnitude /
// kinetic_energy_from_momentum_magnitude (include/ParticleSystem.hpp:170-195).
__device__ __forceinline__ double rel_gamma(double pm, const kpt_particle_properties& pr) {
    const double nm = pm / (pr.rest_mass * pr.speed_of_light);
    return sqrt(1.0 + nm * nm);
}

__device__ __forceinline__ double kinetic_energy(double pm, const kpt_particle_properties& pr) {
    return (rel_gamma(pm, pr) - 1.0) * pr.rest_mass * pr.speed_of_light * pr.speed_of_light;
}

enum SplitOutcome : int { kNotAlive = 0, kNoSplit = 1, kBadBaseline = 2, kLightChild = 3, kSplits = 4 };

// The decision of split_particles_on_energy_growth for one parent
// (include/ParticleSystem.hpp:836-878). It reads only the parent's own slot, so
// the counting and the appending pass reach the same outcome.
__device__ __forceinline__ int split_outcome(const ParticleArrays& P, uint64_t i, const kpt_particle_properties& pr,
                                             double ratio_min, double log_ratio, double min_child_weight,
                                             double& energy) {
    if (P.status[i] != 0) return kNotAlive;
    const double pv = P.p[i];
    energy = kinetic_energy(pv >= 0.0 ? pv : -pv, pr);
    const double base = P.e0[i];
    if (!isfinite(base) || base <= 0.0) return kBadBaseline;
    const double ratio = energy / base;
    if (!isfinite(ratio) || ratio < ratio_min) return kNoSplit;
    const int target = static_cast<int>(floor(log(ratio) / log_ratio));
    if (target <= P.split_level[i]) return kNoSplit;
    const double cw = 0.5 * P.weight[i];
    if (!isfinite(cw) || cw <= 0.0 || cw < min_child_weight) return kLightChild;
    return kSplits;
}

__global__ void __launch_bounds__(kBlk) esplit_count_kernel(ParticleArrays P, uint64_t count,
                                                            kpt_particle_properties pr, double ratio_min,
                                                            double log_ratio, double min_child_weight,
                                                            uint32_t* __restrict__ tile_counts) {
    __shared__ uint32_t wave_counts[kBlk / 64];
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    double e = 0.0;
    const bool f = i < count && split_outcome(P, i, pr, ratio_min, log_ratio, min_child_weight, e) == kSplits;
    const uint64_t m = __ballot(f);
    if ((threadIdx.x & 63) == 0) wave_counts[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kBlk / 64; ++w) t += wave_counts[w];
        tile_counts[blockIdx.x] = t;
    }
}

// counters: [0] alive parents, [1] children written, [2] skipped for capacity,
// [3] invalid baselines, [4] children below the minimum weight.
__global__ void __launch_bounds__(kBlk) esplit_apply_kernel(ParticleArrays P, uint64_t count, uint64_t capacity,
                                                            uint64_t first_root, kpt_particle_properties pr,
                                                            double ratio_min, double log_ratio,
                                                            double min_child_weight,
                                                            const uint32_t* __restrict__ tile_offsets,
                                                            unsigned long long* counters) {
    __shared__ uint32_t wave_counts[kBlk / 64];
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    double energy = 0.0;
    const int o = i < count ? split_outcome(P, i, pr, ratio_min, log_ratio, min_child_weight, energy) : kNotAlive;
    const bool f = o == kSplits;
    const uint64_t m = __ballot(f);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (lane == 0) wave_counts[wave] = __popcll(m);
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += wave_counts[w];
    const uint64_t lt = lane == 0 ? 0ull : (m & ((1ull << lane) - 1ull));
    const uint64_t rank = static_cast<uint64_t>(tile_offsets[blockIdx.x]) + before + __popcll(lt);
    uint32_t written = 0, skipped = 0;
    if (o == kBadBaseline && isfinite(energy) && energy > 0.0) P.e0[i] = energy;
    if (f) {
        const uint64_t child = count + rank;
        if (child >= capacity) {
            skipped = 1;
        } else {
            const double cw = 0.5 * P.weight[i];
            const int level = P.split_level[i] + 1;
            P.x[child] = P.x[i];
            P.y[child] = P.y[i];
            P.p[child] = P.p[i];
            P.t[child] = P.t[i];
            P.step_dt[child] = P.step_dt[i];
            P.mu[child] = P.mu[i];
            P.e0[child] = P.e0[i];
            // split_particles_on_energy_growth copies both previous positions
            // and the sort key (include/ParticleSystem.hpp:883-896)
            P.prev_x[child] = P.prev_x[i];
            P.prev_y[child] = P.prev_y[i];
            P.prev_step_x[child] = P.prev_step_x[i];
            P.prev_step_y[child] = P.prev_step_y[i];
            P.sort_key[child] = P.sort_key[i];
            P.weight[i] = cw;
            P.weight[child] = cw;
            P.status[child] = 0;
            P.split_level[i] = level;
            P.split_level[child] = level;
            P.id[child] = first_root + rank;
            written = 1;
        }
    }
    const unsigned int a = wave_sum(o != kNotAlive ? 1u : 0u);
    const unsigned int w = wave_sum(written);
    const unsigned int s = wave_sum(skipped);
    const unsigned int b = wave_sum(o == kBadBaseline ? 1u : 0u);
    const unsigned int l = wave_sum(o == kLightChild ? 1u : 0u);
    if (lane == 0 && a != 0) {
        atomicAdd(&counters[0], static_cast<unsigned long long>(a));
        if (w) atomicAdd(&counters[1], static_cast<unsigned long long>(w));
        if (s) atomicAdd(&counters[2], static_cast<unsigned long long>(s));
        if (b) atomicAdd(&counters[3], static_cast<unsigned long long>(b));
        if (l) atomicAdd(&counters[4], static_cast<unsigned long long>(l));
    }
}

// reset_split_baselines (include/ParticleSystem.hpp:633-653).
__global__ void __launch_bounds__(kBlk) reset_baselines_kernel(ParticleArrays P, uint64_t count,
                                                               kpt_particle_properties pr) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    if (i >= count) return;
    const double pv = P.p[i];
    P.e0[i] = kinetic_energy(pv >= 0.0 ? pv : -pv, pr);
    P.split_level[i] = 0;
}

// apply_grid_boundary_conditions (include/ParticleSystem.hpp:661-706): wrap
// periodic axes, mark alive particles outside a non-periodic physical extent as
// Escaped (status 1).
__global__ void __launch_bounds__(kBlk) boundary_kernel(ParticleArrays P, uint64_t count, GridDesc g, int mask,
                                                        unsigned long long* escaped) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    uint32_t esc = 0;
    if (i < count && P.status[i] == 0) {
        double* pos[2] = {P.x, P.y};
        bool out = false;
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const double x = pos[d][i];
            if (mask & (1 << d)) {
                pos[d][i] = grid_wrap(x, g.phys_lo[d], g.phys_hi[d], g.width[d], g.inv_width[d]);
            } else if (!(x >= g.phys_lo[d] && x <= g.phys_hi[d])) {
                out = true;
            }
        }
        if (out) {
            P.status[i] = 1;
            esc = 1;
        }
    }
    const unsigned int e = wave_sum(esc);
    if ((threadIdx.x & 63) == 0 && e != 0) atomicAdd(escaped, static_cast<unsigned long long>(e));
}

__global__ void __launch_bounds__(kBlk) kinematics_kernel(const double* __restrict__ p, uint64_t count,
                                                          kpt_particle_properties pr, double* __restrict__ gamma,
                                                          double* __restrict__ speed, double* __restrict__ ekin) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    if (i >= count) return;
    const double pm = p[i];
    const double g = rel_gamma(pm, pr);
    if (gamma) gamma[i] = g;
    if (speed) speed[i] = pm / (g * pr.rest_mass);
    if (ekin) ekin[i] = (g - 1.0) * pr.rest_mass * pr.speed_of_light * pr.speed_of_light;
}

// capture_previous_positions (include/ParticleSystem.hpp:623-628).
__global__ void __launch_bounds__(kBlk) capture_kernel(ParticleArrays P, uint64_t count) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlk) + threadIdx.x;
    if (i >= count) return;
    P.prev_step_x[i] = P.prev_x[i];
    P.prev_step_y[i] = P.prev_y[i];
    P.prev_x[i] = P.x[i];
    P.prev_y[i] = P.y[i];
}

}  // namespace

hipError_t launch_capture_previous(const ParticleArrays& P, uint64_t count, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(capture_kernel, dim3(static_cast<unsigned>((count + kBlk - 1) / kBlk)), dim3(kBlk), 0, stream,
                       P, count);
    return hipGetLastError();
}

size_t esplit_scratch_elems(uint64_t count) { return static_cast<size_t>(tiles_for(count)) + 1; }

hipError_t launch_energy_split(const ParticleArrays& P, uint64_t count, uint64_t capacity, uint64_t first_root,
                               const kpt_particle_properties& pr, double ratio_min, double log_ratio,
                               double min_child_weight, uint32_t* scratch, unsigned long long* counters,
                               hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const unsigned tiles = tiles_for(count);
    hipLaunchKernelGGL(esplit_count_kernel, dim3(tiles), dim3(kBlk), 0, stream, P, count, pr, ratio_min, log_ratio,
                       min_child_weight, scratch);
    hipError_t e = launch_scan_tiles(scratch, tiles, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(esplit_apply_kernel, dim3(tiles), dim3(kBlk), 0, stream, P, count, capacity, first_root, pr,
                       ratio_min, log_ratio, min_child_weight, scratch, counters);
    return hipGetLastError();
}

hipError_t launch_reset_baselines(const ParticleArrays& P, uint64_t count, const kpt_particle_properties& pr,
                                  hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(reset_baselines_kernel, dim3(tiles_for(count)), dim3(kBlk), 0, stream, P, count, pr);
    return hipGetLastError();
}

hipError_t launch_boundaries(const ParticleArrays& P, uint64_t count, const GridDesc& g, int mask,
                             unsigned long long* escaped, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(boundary_kernel, dim3(tiles_for(count)), dim3(kBlk), 0, stream, P, count, g, mask, escaped);
    return hipGetLastError();
}

hipError_t launch_kinematics(const double* p, uint64_t count, const kpt_particle_properties& pr, double* gamma,
                             double* speed, double* ekin, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(kinematics_kernel, dim3(tiles_for(count)), dim3(kBlk), 0, stream, p, count, pr, gamma, speed,
                       ekin);
    return hipGetLastError();
}

}  // namespace kpt
