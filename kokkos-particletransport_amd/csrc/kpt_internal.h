// kpt_internal.h — launch interface between the C-ABI host code and the gfx950
// kernels. Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/kpt.h"
#include "kpt_mlp.h"

namespace kpt {

// Coefficient frames are component-planar: component k of node (i, j) sits at
// k * plane + (j + 1) * (nx + 2) + (i + 1), with nodes i in [-1, nx], j in
// [-1, ny] (the stencil of the mover never leaves that range) and
// plane = (nx + 2) * (ny + 2). Cell-sorted lanes of a wavefront then read
// neighbouring 8-byte (or 4-byte) values of one plane, so a wavefront-wide load
// touches a few cache lines instead of one line per lane.
inline uint64_t coefficient_plane(int nx, int ny) {
    return static_cast<uint64_t>(nx + 2) * static_cast<uint64_t>(ny + 2);
}

// Components the FP32 frame pair stores per node: every packed component but
// the focused model's Vz (component 2), which no rate of the FP32 push reads
// (the focused terms take only the gradients of Vz).
__host__ __device__ constexpr int pair_components(int model) { return model == 21 ? 20 : model; }
// Packed component held in stored slot s of the pair.
__host__ __device__ constexpr int pair_component(int model, int s) { return model == 21 && s >= 2 ? s + 1 : s; }
// Floats per node record of the FP32 frame pair: both frames' stored
// components, rounded up to whole 16-byte loads.
__host__ __device__ constexpr int pair_stride(int model) { return (2 * pair_components(model) + 3) / 4 * 4; }
// Floats per node record of a single FP32 frame (static-field pushes and the
// batched ensemble): the stored components of the pair, one frame, rounded up
// to whole 16-byte loads.
__host__ __device__ constexpr int rec_stride(int model) { return (pair_components(model) + 3) / 4 * 4; }
// Doubles per node record of the FP64 frame pair: (lower, upper) of every
// packed component, k at doubles 2k, 2k + 1 (one 16-byte load per component).
__host__ __device__ constexpr int pair64_stride(int model) { return 2 * model; }

// One bin of a stretched axis' lookup table, read with one 32-byte load: the
// face interval k that contains the bin's lower end (less a guard band) and the
// faces e_k, e_{k+1}, e_{k+2} (the last two repeat e_n past the end). The host
// sizes the bins so that every query that rounds into the bin lies in [e_k,
// e_{k+2}): its interval is k, or k + 1 when it reaches e_{k+1}.
struct alignas(16) FaceBin {
    double lo, mid, hi;
    int32_t k, pad;
};

// Grid of a loaded frame: uniform cell-centred spacing (AnalyticCoordinateGrid
// with UniformSpacing, include/AnalyticCoordinateGrid.hpp:159-210, and the
// bounds cache of include/MappedCoordinateGrid.hpp:117-145). An axis with
// non-uniform faces is "stretched": face[d] points to its n + 1 face coordinates
// on the device and bins[d] to bins_n[d] equal bins over [e_0, e_n), with
// bins_scale[d] bins per unit length. Null pointers mark a uniform axis.
struct GridDesc {
    int32_t nx, ny;
    double start[2];
    double step[2];
    double inv_step[2];
    double phys_lo[2], phys_hi[2], width[2], inv_width[2];
    const double* face[2];
    const FaceBin* bins[2];
    int32_t bins_n[2];
    double bins_scale[2];
};

struct ParticleArrays {
    double* x;
    double* y;
    double* p;
    double* t;
    double* step_dt;
    double* weight;
    double* mu;
    int32_t* status;
    int32_t* split_level;
    uint64_t* id;
    double* e0;         // split baseline: initial kinetic energy
    uint32_t* substep;  // sub-steps taken in the current interval (push chunks only)
    // ParticleSystem::previous_position and previous_step_position
    // (include/ParticleSystem.hpp:342-343): the positions before the last and the
    // last-but-one generic solver step (capture_previous_positions, :623-628)
    double* prev_x;
    double* prev_y;
    double* prev_step_x;
    double* prev_step_y;
    int32_t* sort_key;  // ParticleSystem::sort_key (:349), the key of sort_by_key
};

// Every array of a particle set with its element size, for the host code that
// allocates, frees, copies or offsets all of them.
template <typename F>
inline void for_each_array(ParticleArrays& P, F&& f) {
    f(reinterpret_cast<void**>(&P.x), 8);
    f(reinterpret_cast<void**>(&P.y), 8);
    f(reinterpret_cast<void**>(&P.p), 8);
    f(reinterpret_cast<void**>(&P.t), 8);
    f(reinterpret_cast<void**>(&P.step_dt), 8);
    f(reinterpret_cast<void**>(&P.weight), 8);
    f(reinterpret_cast<void**>(&P.mu), 8);
    f(reinterpret_cast<void**>(&P.status), 4);
    f(reinterpret_cast<void**>(&P.split_level), 4);
    f(reinterpret_cast<void**>(&P.id), 8);
    f(reinterpret_cast<void**>(&P.e0), 8);
    f(reinterpret_cast<void**>(&P.substep), 4);
    f(reinterpret_cast<void**>(&P.prev_x), 8);
    f(reinterpret_cast<void**>(&P.prev_y), 8);
    f(reinterpret_cast<void**>(&P.prev_step_x), 8);
    f(reinterpret_cast<void**>(&P.prev_step_y), 8);
    f(reinterpret_cast<void**>(&P.sort_key), 4);
}

// The particle set starting at slot `o` of P.
inline ParticleArrays arrays_at(const ParticleArrays& P, uint64_t o) {
    ParticleArrays v = P;
    for_each_array(v, [&](void** a, size_t es) { *a = static_cast<char*>(*a) + o * es; });
    return v;
}

inline hipError_t alloc_arrays(ParticleArrays& P, uint64_t n) {
    hipError_t e = hipSuccess;
    for_each_array(P, [&](void** a, size_t es) {
        *a = nullptr;
        if (e == hipSuccess && n) e = hipMalloc(a, n * es);
    });
    return e;
}

inline void free_arrays(ParticleArrays& P) {
    for_each_array(P, [](void** a, size_t) {
        if (*a) (void)hipFree(*a);
        *a = nullptr;
    });
}

// Stream-ordered copy of n slots of every array.
inline hipError_t copy_arrays(const ParticleArrays& dst, const ParticleArrays& src, uint64_t n, hipStream_t st) {
    ParticleArrays d = dst, s = src;
    void* sp[17];
    int k = 0;
    for_each_array(s, [&](void** a, size_t) { sp[k++] = *a; });
    hipError_t e = hipSuccess;
    k = 0;
    for_each_array(d, [&](void** a, size_t es) {
        if (e == hipSuccess && n) e = hipMemcpyAsync(*a, sp[k], n * es, hipMemcpyDeviceToDevice, st);
        ++k;
    });
    return e;
}

// Constants of the sub-step arithmetic in its own type, formed once on the host
// (kpt_push_interval) so that the float push reads 32-bit scalars directly.
template <typename RT>
struct StepConst {
    RT kpara0, kperp_ratio;
    RT exp_b_kappa;        // gamma_turb - 2: kappa ~ B^(gamma-2)
    RT pindex;             // kappa ~ (p/p0)^pindex
    RT drift_c1, drift_c2; // drift_param1 p0, drift_param2 p0^2
    RT inv_q, inv_3q, abs_q;
    RT v0, p0;
    RT half_dx, half_dy;
    RT duu0, duu_h0;       // reduced closure
    RT exp_mu1, exp_mu2;   // gamma_turb - 1, gamma_turb - 2
    RT exp_b_duu;          // 2 - gamma_turb
    RT qlt_lc, qlt_gk, qlt_h0;
    RT qlt_pref;           // (pi / 4) sigma2 gamma_k sin(pi / gamma_k) / pi
    RT null_b;             // fp64_null_b
    double inv_mhd_dt;     // 1 / mhd_dt and 1 / p0, correctly rounded (the float
    double inv_p0;         // push multiplies by them instead of dividing)
};

template <typename RT>
inline StepConst<RT> make_step_const(const kpt_device_settings& s, double qlt_norm) {
    StepConst<RT> k;
    k.kpara0 = static_cast<RT>(s.kpara0);
    k.kperp_ratio = static_cast<RT>(s.kperp_over_kpara);
    k.exp_b_kappa = static_cast<RT>(s.gamma_turb - 2.0);
    k.pindex = static_cast<RT>(s.pindex);
    k.drift_c1 = static_cast<RT>(s.drift_param1 * s.p0);
    k.drift_c2 = static_cast<RT>(s.drift_param2 * s.p0 * s.p0);
    k.inv_q = static_cast<RT>(1.0 / s.charge);
    k.inv_3q = static_cast<RT>(1.0 / (3.0 * s.charge));
    k.abs_q = static_cast<RT>(s.charge < 0.0 ? -s.charge : s.charge);
    k.v0 = static_cast<RT>(s.particle_v0);
    k.p0 = static_cast<RT>(s.p0);
    k.half_dx = static_cast<RT>(0.5 * s.dx);
    k.half_dy = static_cast<RT>(0.5 * s.dy);
    k.duu0 = static_cast<RT>(s.duu0);
    k.duu_h0 = static_cast<RT>(s.pitch_angle_scattering_h0);
    k.exp_mu1 = static_cast<RT>(s.gamma_turb - 1.0);
    k.exp_mu2 = static_cast<RT>(s.gamma_turb - 2.0);
    k.exp_b_duu = static_cast<RT>(2.0 - s.gamma_turb);
    k.qlt_lc = static_cast<RT>(s.qlt_correlation_length);
    k.qlt_gk = static_cast<RT>(s.qlt_spectral_index);
    k.qlt_h0 = static_cast<RT>(s.qlt_h0);
    k.qlt_pref = static_cast<RT>(0.25 * 3.14159265358979323846 * s.qlt_sigma2 * qlt_norm);
    k.null_b = static_cast<RT>(s.fp64_null_b);
    k.inv_mhd_dt = 1.0 / s.mhd_dt;
    k.inv_p0 = 1.0 / s.p0;
    return k;
}

struct PushArgs {
    ParticleArrays P;
    uint64_t count;
    const void* field_lo;  // double or float, component-planar (see coefficient_plane)
    const void* field_hi;
    const float* field_pair;   // FP32 time-blended push: lower/upper interleaved, or nullptr
    const double* field_pair64;  // FP64 time-blended push: node-major (lower, upper) records, or nullptr
    const float* field_rec;      // FP32 push of one frame: node-major records, or nullptr
    int32_t stage;             // 1: FP32 pushes stage field_pair records in LDS
    const double* field_lo64;  // FP64 frames, read by FP32 lanes below fp64_null_b
    const double* field_hi64;
    uint64_t plane;
    GridDesc g;
    kpt_device_settings s;
    double t0;
    double t_end;
    double t_stop;       // particles stop when t >= t_stop (chunk end, <= t_end)
    int32_t first_chunk; // 1: the sub-step counter starts at 0, else it is read
    uint32_t key[2];
    uint32_t frame;
    const double* noise;  // injected draw table or nullptr
    uint32_t noise_steps;
    double qlt_norm;      // gamma_k sin(pi / gamma_k) / pi (pitch_angle_spectrum_normalization)
    StepConst<double> k64;
    StepConst<float> k32;
    // learned closure (KPT_CLOSURE_MLP): network image, feature scaling and
    // activation (KPT_MLP_*); mlp_on = 0 sends every evaluation to the exact formula
    const void* mlp_image;
    MlpScale mlp_scale;
    int32_t mlp_on;
    int32_t mlp_act;
    unsigned long long* counters;  // [0] particle steps, [1] failures, [2] active at entry
    // These arguments in device memory: the FP32 push's rare near-null sub-step
    // (a non-inlined call) reads the FP64 frames and constants through it.
    const PushArgs* self;
};

// One run of a batched ensemble: its full push arguments (particles point into
// the ensemble's arrays) and which lane function advances it.
enum EnsembleKind : int32_t { kEnsParker32 = 0, kEnsFocusedQlt32 = 1, kEnsFocusedMlp32 = 2 };
struct EnsembleRun {
    PushArgs a;
    int32_t kind;
    int32_t pad;
};
// Frame-loop state of an ensemble in device memory: intervals run so far and the
// first frame of the loop.
struct EnsembleLoop {
    uint32_t first_frame;
    uint32_t done;
    double window;  // fraction of the interval pushed (1 in the frame loop)
};

// Device-side entry points (kpt_kernels.hip). All launch on `stream`.
hipError_t launch_pack(const double* raw /* 6 planes nx*ny */, const GridDesc& g, int model,
                       double* out64, float* out32, hipStream_t stream);
hipError_t launch_push(const PushArgs& a, int model, bool focused, bool fp32, int closure, hipStream_t stream);
// Generic Parker solver (kpt_generic.hip): packed 18-component basis, global
// minimum of the local time-step bounds (as double bits), one stochastic step.
hipError_t launch_generic_pack(const double* raw, const GridDesc& g, const kpt_generic_parker_params& p,
                               double* out, int* bad, hipStream_t stream);
hipError_t launch_generic_min_dt(const ParticleArrays& P, uint64_t count, const double* f, const GridDesc& g,
                                 const kpt_generic_parker_params& p, unsigned long long* min_bits,
                                 hipStream_t stream);
hipError_t launch_generic_step(const ParticleArrays& P, uint64_t count, const double* f, const GridDesc& g,
                               const kpt_generic_parker_params& p, double dt, uint32_t step, const uint32_t key[2],
                               const double* noise, uint32_t noise_steps, int deterministic, unsigned long long* failures,
                               hipStream_t stream);
// The FP32 frame pair of a time-blended push: floats of the buffer for `plane`
// nodes of `model` components, and its build from the two component-planar
// frames (layout in kpt_kernels.hip, sample_pair32).
uint64_t pair_elems(uint64_t plane, int model);
hipError_t launch_pair(const float* lo, const float* hi, uint64_t plane, int model, float* out, hipStream_t stream);
// Node-major records of one FP32 frame (rec_stride(model) floats per node).
uint64_t rec_elems(uint64_t plane, int model);
hipError_t launch_rec(const float* c32, uint64_t plane, int model, float* out, hipStream_t stream);
// The FP64 frame pair (node-major records of pair64_stride(model) doubles).
uint64_t pair64_elems(uint64_t plane, int model);
hipError_t launch_pair64(const double* lo, const double* hi, uint64_t plane, int model, double* out,
                         hipStream_t stream);
hipError_t launch_inject(const ParticleArrays& P, uint64_t first, uint64_t n, uint64_t first_root,
                         const kpt_device_settings& s, const kpt_inject_opts& o, double t0,
                         uint32_t frame, const uint32_t key[2], hipStream_t stream);
// Id of the child of a split whose parent (id pid) was at split level `level`.
// Ids are root | path << 32: the root is the injection counter and path bit L
// is set when an ancestor split at level L, so every particle of a tree has its
// own id (and Philox stream) while the tree is at most 32 levels deep. Deeper
// children (and those of a parent with a negative level) take the splitmix64
// finalizer (Steele, Lea and Flood 2014) of pid + (level + 1) * golden gamma, a
// 64-bit mix under which distinct particles collide with probability about
// 2^-64 per pair.
__host__ __device__ constexpr uint64_t split_child_id(uint64_t pid, int level) {
    if (level >= 0 && level < 32) return (pid & 0xFFFFFFFFull) | (((pid >> 32) | (1ull << level)) << 32);
    uint64_t z = pid + static_cast<uint64_t>(level + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Split thresholds pmin_split * split_ratio^level for level < kLevels.
struct SplitTable {
    static constexpr int kLevels = 64;
    double thr[kLevels];
};

// Deterministic split. `scratch` must hold split_scratch_elems(count) uint32.
size_t split_scratch_elems(uint64_t count);
hipError_t launch_split(const ParticleArrays& P, uint64_t count, uint64_t capacity,
                        const kpt_device_settings& s, const SplitTable& tab, uint32_t* scratch,
                        unsigned long long* counters /* [0] split, [1] skipped */,
                        hipStream_t stream);
// Generic focused-transport solver (kpt_focus.hip): the 27-component basis, the
// per-particle stable-dt minimum (as double bits) and one step.
hipError_t launch_focus_pack(const double* raw, const double* dvdt, const GridDesc& g,
                             const kpt_generic_focus_params& p, double* out, int* bad, hipStream_t stream);
hipError_t launch_focus_min_dt(const ParticleArrays& P, uint64_t count, const double* f, const GridDesc& g,
                               const kpt_generic_focus_params& p, unsigned long long* min_bits, hipStream_t stream);
hipError_t launch_focus_step(const ParticleArrays& P, uint64_t count, const double* f, const GridDesc& g,
                             const kpt_generic_focus_params& p, double dt, uint32_t step, const uint32_t key[2],
                             const double* noise, uint32_t noise_steps, int deterministic, unsigned long long* failures,
                             hipStream_t stream);
// Ensemble: `tiles` lists (run, first particle) of every 256-particle tile,
// grouped per work queue: group g owns tiles [group_start[g], group_start[g+1]).
// launch_ensemble_advance sets every run's interval from `loop` (frame
// first_frame + done, t0 = frame mhd_dt) and clears the queue counters;
// launch_ensemble runs one persistent launch of `grid` workgroups.
hipError_t launch_ensemble_advance(EnsembleRun* runs, int n_runs, EnsembleLoop* loop, unsigned int* queues,
                                   int groups, hipStream_t stream);
hipError_t launch_ensemble(const EnsembleRun* runs, const uint2* tiles, const uint32_t* group_start, int groups,
                           unsigned int* queues, int grid, const void* mlp_image, int has_mlp, hipStream_t stream);
// Exclusive scan of n uint32 tile counts in place; the total lands in [n].
hipError_t launch_scan_tiles(uint32_t* tile_counts, uint32_t n, hipStream_t stream);
// ParticleSystem lifecycle passes. The energy split needs esplit_scratch_elems(count)
// uint32 of scratch; counters [0..4] follow ParticleSplitStatistics.
size_t esplit_scratch_elems(uint64_t count);
hipError_t launch_energy_split(const ParticleArrays& P, uint64_t count, uint64_t capacity, uint64_t first_root,
                               const kpt_particle_properties& pr, double ratio_min, double log_ratio,
                               double min_child_weight, uint32_t* scratch, unsigned long long* counters,
                               hipStream_t stream);
hipError_t launch_reset_baselines(const ParticleArrays& P, uint64_t count, const kpt_particle_properties& pr,
                                  hipStream_t stream);
hipError_t launch_boundaries(const ParticleArrays& P, uint64_t count, const GridDesc& g, int mask,
                             unsigned long long* escaped, hipStream_t stream);
hipError_t launch_kinematics(const double* p, uint64_t count, const kpt_particle_properties& pr, double* gamma,
                             double* speed, double* ekin, hipStream_t stream);
// Histograms: `partials` holds hist_partial_elems(bins) doubles.
size_t hist_partial_elems(int bins);
hipError_t launch_hist_logp(const ParticleArrays& P, uint64_t count, int bins, double log_pmin,
                            double inv_dlogp, double pmin, double pmax, double* partials,
                            double* out, hipStream_t stream);
hipError_t launch_hist_xy(const ParticleArrays& P, uint64_t count, int bx, int by, double x_lo,
                          double x_hi, double y_lo, double y_hi, double* partials, double* out,
                          hipStream_t stream);
// Cell sort (kpt_sort.hip): keys/values double buffers of `capacity` u32 each
// plus radix-sort temporary storage of sort_temp_bytes(capacity).
// ---- generic particle generator (kpt_inject.hip) ----
// Philox stream tags of the generator: counter word 1 is tag | attempt.
constexpr uint32_t kTagInject = 0x494E0000u;   // "IN"
constexpr uint32_t kTagRecycle = 0x52430000u;  // "RC"

// The field slot a generator call reads: its grid, raw planes (cell-centred
// Bx, By, Bz, Vx, Vy, Vz), the region (x_lo, x_hi, y_lo, y_hi) and the
// field-threshold predicate.
struct GeneratorSource {
    GridDesc g;
    const double* raw;
    double region[4];
    int32_t use_predicate;
    int32_t component;
    int32_t comparison;
    int32_t max_attempts;
    double threshold;
};

// A built candidate set on the device: accepted cells (nullptr: every cell,
// candidate index = cell index) and cumulative cell areas; kind 0 places
// particles at cell centres (field candidates), 1 uniformly in the cell (grid).
struct CandidateView {
    GridDesc g;
    const uint32_t* cells;
    const double* cumulative;
    uint64_t count;
    double total;
    int32_t kind;
};

// Scratch of one candidate build over `cells` cells.
struct CandidateBuild {
    uint32_t* flags;
    uint32_t* offsets;
    double* area;
    uint32_t* out_cells;
    double* out_area;
    void* temp;
    size_t temp_bytes;
};

hipError_t candidate_scratch_bytes(uint64_t cells, size_t* bytes);
hipError_t launch_candidates(const GeneratorSource& src, int use_predicate, CandidateBuild& w, hipStream_t stream);
hipError_t launch_cumulative(const double* area, double* cumulative, uint64_t count, void* temp, size_t temp_bytes,
                             hipStream_t stream);
hipError_t launch_inject_candidates(const ParticleArrays& P, uint64_t first, uint64_t n, uint64_t first_root,
                                    const CandidateView& c, const kpt_injection_kinematics& k, double e0,
                                    uint32_t call, const uint32_t key[2], hipStream_t stream);
hipError_t launch_inject_uniform(const ParticleArrays& P, uint64_t first, uint64_t n, uint64_t first_root,
                                 const GeneratorSource& src, const kpt_injection_kinematics& k, double e0,
                                 uint32_t call, const uint32_t key[2], unsigned long long* rejected,
                                 hipStream_t stream);
hipError_t launch_recycle(const ParticleArrays& P, uint64_t count, int selection, int mode, const CandidateView& c,
                          const GeneratorSource& src, const kpt_injection_kinematics& k, double e0, uint32_t call,
                          const uint32_t key[2], unsigned long long* counters, hipStream_t stream);

// One run of a batched ensemble sort: its grid and interpolation shifts, its
// slot range start and its stored count.
struct SortRun {
    GridDesc g;
    double shift_x, shift_y;
    uint64_t off, count;
};

struct SortScratch {
    uint32_t* keys[2];
    uint32_t* vals[2];
    void* temp;
    size_t temp_bytes;
};
size_t sort_temp_bytes(uint64_t capacity);
// Node keys of a grid lie below this (16 x 16 blocks over the stored nodes).
uint32_t sort_key_limit(const GridDesc& g);
// Cell sort of every run of a batched ensemble in one key pass, one radix sort
// of all `total` slots and one gather (src -> dst); `runs` is a device array.
hipError_t launch_ensemble_sort(const SortRun* runs, int n_runs, uint64_t total, uint32_t run_shift,
                                const ParticleArrays& src, const ParticleArrays& dst, SortScratch& w,
                                unsigned long long* active, hipStream_t stream);
// Writes the sorted particle set into `dst`; `active` (device) accumulates the
// number of status-0 particles.
// Particles with t >= t_done (finished their interval) sort after the active
// ones and before the inactive ones; pass +inf to treat every active particle
// alike.
hipError_t launch_sort(const ParticleArrays& src, const ParticleArrays& dst, uint64_t count, const GridDesc& g,
                       double shift_x, double shift_y, double t_done, SortScratch& w,
                       unsigned long long* active, hipStream_t stream);

// ParticleSystem::sort_by_key / compact_active_particles on the stored sort keys
// (see stored_key_kernel); `active` (device) accumulates the alive count.
hipError_t launch_sort_stored(const ParticleArrays& src, const ParticleArrays& dst, uint64_t count, int32_t min_key,
                              int32_t max_key, int compact, SortScratch& w, unsigned long long* active,
                              hipStream_t stream);
// ParticleSystem::capture_previous_positions of the first `count` slots.
hipError_t launch_capture_previous(const ParticleArrays& P, uint64_t count, hipStream_t stream);

// Summary: `partials` holds summary_partial_bytes() bytes.
size_t summary_partial_bytes();
hipError_t launch_summary(const ParticleArrays& P, uint64_t count, void* partials,
                          kpt_summary* out_device, hipStream_t stream);

}  // namespace kpt
