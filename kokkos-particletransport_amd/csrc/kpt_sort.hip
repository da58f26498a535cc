// kpt_sort.hip — cell sort and compaction of the particle set
// (ParticleSystem::assign_sort_keys / sort_by_key / compact_active_particles,
// include/ParticleSystem.hpp:715-793, which use Kokkos::BinSort).
//
// Key = the lower-left stencil node on the lower frame's grid (the node the push
// gathers first), numbered in blocks of 16 x 16 nodes: blocks in row-major order,
// nodes row-major inside a block. After the sort the 64 lanes of a wavefront
// gather overlapping 128 B lines, and the workgroups that one XCD runs together
// cover a compact patch, so the field rows a patch shares stay in that XCD's L2.
// A plain row-major order puts those workgroups on a long strip (1-3 rows of
// 1024 cells); the blocks measured 15-21% faster for the FP64 pushes on the
// 1024 x 512 fields. Inactive particles get the largest key and land at the
// end. The stable LSD radix sort (rocPRIM through hipCUB) keeps equal keys in
// slot order, so the result is deterministic.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdint>

#include "kpt_device_common.h"
#include "kpt_internal.h"
#include "kpt_grid_device.h"

namespace kpt {
namespace {

constexpr int kBlock = 256;
constexpr uint32_t kSortBlock = 16;  // nodes per side of a key block

// Blocked node key of an active particle at (x, y): the stencil's lower-left node
// (I, J) over the stored nodes, in 16 x 16 blocks.
__device__ __forceinline__ uint32_t node_key(const GridDesc& g, double x, double y, double shift_x, double shift_y) {
    const double q[2] = {x + shift_x, y + shift_y};
    const int n[2] = {g.nx, g.ny};
    int idx[2];
    for (int d = 0; d < 2; ++d) {
        int i0;
        double w, h;
        // NaN positions sort first; the push marks them
        idx[d] = node_stencil(q[d], g, d, n[d], i0, w, h) ? i0 : -1;
    }
    const uint32_t I = static_cast<uint32_t>(idx[0] + 1), J = static_cast<uint32_t>(idx[1] + 1);
    const uint32_t blocks_x = (static_cast<uint32_t>(g.nx) + 2 + kSortBlock - 1) / kSortBlock;
    return (((J / kSortBlock) * blocks_x + I / kSortBlock) * kSortBlock + J % kSortBlock) * kSortBlock +
           I % kSortBlock;
}

// Adds the workgroup's count of active slots (one per lane, summed over the
// grid-stride loop) to *active with one atomic: a single counter hit once per
// wavefront serialises in the L2 (about 12 ns per atomic, 1.3 ms per 8 M slots).
__device__ __forceinline__ void add_block_count(uint32_t c, unsigned long long* active) {
    __shared__ uint32_t part[kBlock / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += part[w];
        if (t != 0) atomicAdd(active, static_cast<unsigned long long>(t));
    }
}

// Workgroups of the key kernels: a grid-stride loop over at most this many.
constexpr uint64_t kKeyGrid = 2048;

__global__ void __launch_bounds__(kBlock) sort_key_kernel(ParticleArrays P, uint64_t count, GridDesc g,
                                                          double shift_x, double shift_y, double t_done,
                                                          uint32_t inactive_key,
                                                          uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                          unsigned long long* __restrict__ active) {
    uint32_t n_act = 0;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; i < count;
         i += static_cast<uint64_t>(gridDim.x) * kBlock) {
        uint32_t key = inactive_key;
        if (P.status[i] == 0 && P.t[i] >= t_done) {
            ++n_act;
            key = inactive_key - 1;  // finished this interval: after the moving ones
        } else if (P.status[i] == 0) {
            ++n_act;
            key = node_key(g, P.x[i], P.y[i], shift_x, shift_y);
        }
        keys[i] = key;
        vals[i] = static_cast<uint32_t>(i);
    }
    add_block_count(n_act, active);
}

__global__ void __launch_bounds__(kBlock) gather_kernel(ParticleArrays src, ParticleArrays dst, uint64_t count,
                                                        const uint32_t* __restrict__ perm) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x;
    if (i >= count) return;
    const uint32_t s = perm[i];
    dst.x[i] = src.x[s];
    dst.y[i] = src.y[s];
    dst.p[i] = src.p[s];
    dst.t[i] = src.t[s];
    dst.step_dt[i] = src.step_dt[s];
    dst.weight[i] = src.weight[s];
    dst.mu[i] = src.mu[s];
    dst.status[i] = src.status[s];
    dst.split_level[i] = src.split_level[s];
    dst.e0[i] = src.e0[s];
    dst.id[i] = src.id[s];
    dst.substep[i] = src.substep[s];
    dst.prev_x[i] = src.prev_x[s];
    dst.prev_y[i] = src.prev_y[s];
    dst.prev_step_x[i] = src.prev_step_x[s];
    dst.prev_step_y[i] = src.prev_step_y[s];
    dst.sort_key[i] = src.sort_key[s];
}

int key_bits(uint32_t max_key) {
    int b = 1;
    while (b < 32 && (static_cast<uint64_t>(1) << b) <= max_key) ++b;
    return b;
}

// Keys of all slots of a batched ensemble: run r's slots [off, off + cap) get
// (r << run_shift) | node key for active particles, | (2^run_shift - 2) for
// inactive ones and | (2^run_shift - 1) for slots past the run's count. A stable
// sort of all slots then keeps every run inside its own slot range and orders
// it exactly as the run's own cell sort would.
__global__ void __launch_bounds__(kBlock) ensemble_key_kernel(const SortRun* __restrict__ runs, int n_runs,
                                                              uint64_t total, ParticleArrays P, uint32_t run_shift,
                                                              uint32_t* __restrict__ keys,
                                                              uint32_t* __restrict__ vals,
                                                              unsigned long long* __restrict__ active) {
    const uint32_t top = (1u << run_shift) - 1u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t n_act = 0;
    for (uint64_t base = blockIdx.x * static_cast<uint64_t>(kBlock); base < total;
         base += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint64_t i = base + threadIdx.x;
        // A wavefront's 64 slots are contiguous: the run of its first slot is found
        // once, with scalar loads, and lanes past the next run's start step forward
        // (only a wavefront that straddles a run boundary has such lanes).
        const uint64_t i0 = base + 64u * wave;
        int lo = 0, hi = n_runs - 1;  // last run with off <= i0
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (runs[mid].off <= i0) lo = mid;
            else hi = mid - 1;
        }
        int rl = lo;
        while (rl + 1 < n_runs && runs[rl + 1].off <= i) ++rl;
        if (i < total) {
            auto slot_key = [&](const SortRun& r) {
                uint32_t key = top;
                if (i - r.off < r.count) {
                    if (P.status[i] == 0) {
                        ++n_act;
                        key = node_key(r.g, P.x[i], P.y[i], r.shift_x, r.shift_y);
                    } else {
                        key = top - 1u;
                    }
                }
                return key;
            };
            // one run per wavefront (the common case): its record by scalar loads
            const uint32_t key = __ballot(rl != lo) == 0 ? slot_key(runs[lo]) : slot_key(runs[rl]);
            keys[i] = (static_cast<uint32_t>(rl) << run_shift) | key;
            vals[i] = static_cast<uint32_t>(i);
        }
    }
    add_block_count(n_act, active);
}

}  // namespace

// Keys of ParticleSystem::sort_by_key (include/ParticleSystem.hpp:727-760):
// the stored key less min_key, clamped to [0, max_key - min_key]. With
// `compact`, the key first becomes 0 for alive particles and 1 otherwise and is
// stored (compact_active_particles, :769-793); `active` counts alive slots.
__global__ void __launch_bounds__(kBlock) stored_key_kernel(ParticleArrays P, uint64_t count, int32_t min_key,
                                                            int32_t max_key, int compact, uint32_t* __restrict__ keys,
                                                            uint32_t* __restrict__ vals,
                                                            unsigned long long* __restrict__ active) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x;
    uint32_t alive = 0;
    if (i < count) {
        alive = P.status[i] == 0 ? 1u : 0u;
        int32_t k;
        if (compact) {
            k = alive ? 0 : 1;
            P.sort_key[i] = k;
        } else {
            k = P.sort_key[i];
        }
        k = k < min_key ? min_key : (k > max_key ? max_key : k);
        keys[i] = static_cast<uint32_t>(static_cast<int64_t>(k) - min_key);
        vals[i] = static_cast<uint32_t>(i);
    }
    const unsigned int a = wave_sum(alive);
    if ((threadIdx.x & 63) == 0 && a) atomicAdd(active, static_cast<unsigned long long>(a));
}

uint32_t sort_key_limit(const GridDesc& g) {
    const uint32_t bx = (static_cast<uint32_t>(g.nx) + 2 + kSortBlock - 1) / kSortBlock;
    const uint32_t by = (static_cast<uint32_t>(g.ny) + 2 + kSortBlock - 1) / kSortBlock;
    return bx * by * kSortBlock * kSortBlock;
}

hipError_t launch_ensemble_sort(const SortRun* runs, int n_runs, uint64_t total, uint32_t run_shift,
                                const ParticleArrays& src, const ParticleArrays& dst, SortScratch& w,
                                unsigned long long* active, hipStream_t stream) {
    if (total == 0) return hipSuccess;
    const unsigned blocks = static_cast<unsigned>((total + kBlock - 1) / kBlock);
    const unsigned key_blocks = static_cast<unsigned>(std::min<uint64_t>(blocks, kKeyGrid));
    hipLaunchKernelGGL(ensemble_key_kernel, dim3(key_blocks), dim3(kBlock), 0, stream, runs, n_runs, total, src,
                       run_shift, w.keys[0], w.vals[0], active);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipcub::DoubleBuffer<uint32_t> k(w.keys[0], w.keys[1]), v(w.vals[0], w.vals[1]);
    size_t bytes = w.temp_bytes;
    const uint32_t top = static_cast<uint32_t>(n_runs - 1) << run_shift | ((1u << run_shift) - 1u);
    e = hipcub::DeviceRadixSort::SortPairs(w.temp, bytes, k, v, static_cast<int>(total), 0, key_bits(top), stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(gather_kernel, dim3(blocks), dim3(kBlock), 0, stream, src, dst, total, v.Current());
    return hipGetLastError();
}

size_t sort_temp_bytes(uint64_t capacity) {
    size_t bytes = 0;
    hipcub::DoubleBuffer<uint32_t> k(nullptr, nullptr), v(nullptr, nullptr);
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k, v, static_cast<int>(capacity), 0, 32);
    return bytes;
}

hipError_t launch_sort_stored(const ParticleArrays& src, const ParticleArrays& dst, uint64_t count, int32_t min_key,
                              int32_t max_key, int compact, SortScratch& w, unsigned long long* active,
                              hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const unsigned blocks = static_cast<unsigned>((count + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(stored_key_kernel, dim3(blocks), dim3(kBlock), 0, stream, src, count, min_key, max_key, compact,
                       w.keys[0], w.vals[0], active);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipcub::DoubleBuffer<uint32_t> k(w.keys[0], w.keys[1]), v(w.vals[0], w.vals[1]);
    size_t bytes = w.temp_bytes;
    const uint32_t top = static_cast<uint32_t>(static_cast<int64_t>(max_key) - min_key);
    e = hipcub::DeviceRadixSort::SortPairs(w.temp, bytes, k, v, static_cast<int>(count), 0, key_bits(top), stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(gather_kernel, dim3(blocks), dim3(kBlock), 0, stream, src, dst, count, v.Current());
    return hipGetLastError();
}

hipError_t launch_sort(const ParticleArrays& src, const ParticleArrays& dst, uint64_t count, const GridDesc& g,
                       double shift_x, double shift_y, double t_done, SortScratch& w,
                       unsigned long long* active, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    // node keys are below the padded block count times 16 x 16; the two classes
    // after them are "finished this interval" and "inactive"
    const uint32_t blocks_x = (static_cast<uint32_t>(g.nx) + 2 + kSortBlock - 1) / kSortBlock;
    const uint32_t blocks_y = (static_cast<uint32_t>(g.ny) + 2 + kSortBlock - 1) / kSortBlock;
    const uint32_t inactive_key = blocks_x * blocks_y * kSortBlock * kSortBlock + 1;
    const unsigned blocks = static_cast<unsigned>((count + kBlock - 1) / kBlock);
    const unsigned key_blocks = static_cast<unsigned>(std::min<uint64_t>(blocks, kKeyGrid));
    hipLaunchKernelGGL(sort_key_kernel, dim3(key_blocks), dim3(kBlock), 0, stream, src, count, g, shift_x, shift_y,
                       t_done, inactive_key, w.keys[0], w.vals[0], active);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipcub::DoubleBuffer<uint32_t> k(w.keys[0], w.keys[1]), v(w.vals[0], w.vals[1]);
    size_t bytes = w.temp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(w.temp, bytes, k, v, static_cast<int>(count), 0, key_bits(inactive_key),
                                           stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(gather_kernel, dim3(blocks), dim3(kBlock), 0, stream, src, dst, count, v.Current());
    return hipGetLastError();
}

}  // namespace kpt
