/*
 * kpt.h — C ABI of the MI355X (gfx950) Parker / focused-transport SDE particle push.
 *
 * Every entry point is a plain-C replacement for one function of the Kokkos driver
 * path in LiuYihanSDU/Kokkos-ParticleTransport (src/main.cpp, namespace
 * Reconnection2DCalibration) or of its generic solver / ParticleSystem API. The
 * reference function each one replaces is cited beside its declaration.
 *
 * Conventions
 *  - All functions return kpt_status: 0 on success, a negative KPT_ERR_* code on
 *    failure. kpt_last_error(ctx) returns the message of the last failure on that
 *    context (the reference throws std::runtime_error with the same text where one
 *    exists).
 *  - Device buffers are owned by the context. Host pointers passed in are borrowed
 *    for the duration of the call only.
 *  - One context per host thread; a context is not thread-safe. Calls are
 *    synchronous: they return after the device work they enqueue has finished
 *    (the reference fences after every kernel).
 *  - Numerical failures inside the push never abort: the particle gets status 2
 *    and is counted in kpt_push_stats.failures (src/main.cpp:1167-1171,
 *    :1189-1193, :1274-1277, :1301-1305).
 */
#ifndef KPT_H
#define KPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int kpt_status;

enum {
    KPT_OK = 0,
    KPT_ERR_ARG = -1,      /* invalid argument or settings                      */
    KPT_ERR_IO = -2,       /* file open/read/write or format failure            */
    KPT_ERR_DEVICE = -3,   /* HIP runtime failure or no gfx950 device           */
    KPT_ERR_CAPACITY = -4, /* particle capacity exceeded                        */
    KPT_ERR_STATE = -5     /* call made in the wrong state (e.g. no field)      */
};

/* Packed coefficient layouts (component counts of the reference's packed fields). */
typedef enum {
    KPT_MODEL_PARKER16 = 16,  /* src/main.cpp:28, :62-79   */
    KPT_MODEL_FOCUSED21 = 21  /* src/main.cpp:29, :38-60   */
} kpt_model;

typedef enum { KPT_SLOT_LOWER = 0, KPT_SLOT_UPPER = 1 } kpt_slot;

/* Arithmetic of the push. FP64: everything in double, as the reference.
 * FP32: coefficient frames stored and interpolated in float; transport
 * coefficients, dt candidates, noise and increments in float; x, y, t, mu and p
 * accumulated in double, and the adiabatic dp = p (rate) dt formed in double.
 * A lane whose interpolated |B| is below fp64_null_b re-samples the FP64
 * frames and takes that sub-step's coefficient arithmetic in double (the drift
 * and the diffusion-tensor divergence scale as 1/|B|^2 and 1/|B|^3 near
 * current-sheet nulls). */
typedef enum { KPT_PREC_FP64 = 0, KPT_PREC_FP32 = 1 } kpt_precision;

/* Noise source of the push. */
typedef enum {
    KPT_NOISE_PHILOX = 0,  /* counter-based Philox4x32-10 keyed by seed and particle id */
    KPT_NOISE_INJECTED = 1 /* fixed host-provided draw table, see kpt_set_noise        */
} kpt_noise_mode;

/*
 * Device parameter block. Mirrors Reconnection2DDeviceSettings field by field
 * (src/main.cpp:246-283); the trailing fields are extensions of this library.
 * Fill it with kpt_make_device_settings (src/main.cpp:595-636) or by hand.
 */
typedef struct kpt_device_settings {
    int32_t max_particle_steps_per_interval;
    int32_t momentum_dependency;
    int32_t magnetic_dependency;
    int32_t split_particles;
    int32_t time_interpolation;
    int32_t focused_transport;

    double x_lower;
    double x_upper;
    double y_lower;
    double y_upper;
    double x_width;
    double y_width;
    double dx;
    double dy;
    double interpolation_shift_x;
    double interpolation_shift_y;

    double mhd_dt;
    double p0;
    double pmax;
    double gamma_turb;
    double pindex;
    double kpara0;
    double kperp_over_kpara;
    double dt_min;
    double dt_max;
    double drift_param1;
    double drift_param2;
    double charge;
    double particle_v0;
    double duu0;
    double pitch_angle_scattering_h0;
    double split_ratio;
    double pmin_split;
    double mu_max;

    /* extensions */
    int32_t precision;   /* kpt_precision */
    /* kpt_push_interval splits the interval into this many equal time chunks and
     * launches the push once per chunk, re-sorting particles by cell before each
     * chunk (kpt_sort_particles; skipped in the injected-noise mode). Particles
     * are not stopped at chunk ends: the sub-step counter, the step guard and the
     * truncation of the last step to the interval end carry across chunks, so the
     * result equals one launch. 0 or 1: one launch, no sort. */
    int32_t push_chunks;

    /* Pitch-angle closure of the focused branch (FOCUSED21 with focused_transport):
     *  KPT_CLOSURE_REDUCED: the driver's D_mumu = duu0 (1-mu^2)(|mu|^(g-1) + h0)
     *    B^(2-g) (p/p0)^(g-1), one Euler mu step, clamp to +-mu_max
     *    (src/main.cpp:1551-1584, :1699-1702);
     *  KPT_CLOSURE_QLT: slab-QLT D_mumu and dD_mumu/dmu of
     *    LegencyModel::evaluate_pitch_angle_scattering (include/LegencyModel.hpp:
     *    539-689) with Omega = |q| |B| / gamma, gamma = sqrt(1 + p^2),
     *    speed = p / gamma (m = c = 1), and the mu update of
     *    FocusTransportSolver::picard_pitch_angle_update (3 Picard iterations,
     *    include/FocusSolver.hpp:700-719) reflecting at +-mu_max (:454-470);
     *  KPT_CLOSURE_MLP: as QLT, with D_mumu / (Omega sigma2) and its mu
     *    derivative taken from the loaded MLP inside its training hull
     *    (kpt_mlp_set) and from the exact QLT outside it. */
    int32_t closure;
    /* KPT_CLOSURE_MLP: network evaluations per sub-step.
     *  0 or 1: one, at the sub-step's initial mu; the two Picard passes take
     *    ln(D / (1 - mu^2)) at their trial mu to first order in ln|mu| from that
     *    evaluation, with the network's own slope output d ln R / d ln|mu|;
     *  3: one at each of the three Picard passes.
     * A trial mu outside the training hull takes the exact formula either way. */
    int32_t mlp_passes;
    double qlt_sigma2;          /* delta B^2 / B0^2 */
    double qlt_correlation_length;
    double qlt_spectral_index;  /* gamma_k, PitchAngleScatteringParameters default 5/3 */
    double qlt_h0;              /* resonance regularisation, default 1e-6 */
    double fp64_null_b;         /* KPT_PREC_FP32 only: |B| below which a sub-step
                                   runs in FP64 (default 0.02; 0 = never) */
} kpt_device_settings;

enum { KPT_CLOSURE_REDUCED = 0, KPT_CLOSURE_QLT = 1, KPT_CLOSURE_MLP = 2 };

/* Host-side controls that kpt_make_device_settings reads (the physics subset of
 * Reconnection2DSettings, src/main.cpp:114-154, with the same defaults). */
typedef struct kpt_host_settings {
    int32_t max_particle_steps_per_interval; /* 200000 */
    int32_t momentum_dependency;             /* 1 */
    int32_t magnetic_dependency;             /* 1 */
    int32_t split_particles;                 /* 1 */
    int32_t time_interpolation;              /* 1 */
    int32_t focused_transport;               /* 0 */
    double mhd_dt;                           /* 0.1 */
    double p0;                               /* 0.1 */
    double pmin;                             /* 1e-2 (histogram lower edge) */
    double pmax;                             /* 1e1  */
    double gamma_turb;                       /* 1.6666667 */
    double kpara0;                           /* 0.00743592 */
    double kperp_over_kpara;                 /* 0.01 */
    double dt_min_rel;                       /* 1e-7 */
    double dt_max_rel;                       /* 1e-2 */
    double drift_param1;                     /* 850964.408 */
    double drift_param2;                     /* 13575468.975 */
    double charge;                           /* -1 */
    double particle_v0;                      /* 17.20195 */
    double duu0;                             /* 5578.445 */
    double split_ratio;                      /* 2 */
    double pmin_split_over_p0;               /* 2 */
    double mu_max;                           /* 0.99 */
} kpt_host_settings;

/* Particle arrays on the host, structure of arrays (the fields of
 * ReconnectionParticleStorage, src/main.cpp:288-333, plus the particle id that
 * keys the counter-based RNG and the energy-split baseline of ParticleSystem,
 * include/ParticleSystem.hpp:347). Any pointer may be NULL on download to
 * skip it. */
typedef struct kpt_particles_host {
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
    double* e0;  /* initial_kinetic_energy: baseline of the energy-growth split;
                    NULL on upload stores 0 (an unset baseline) */
    /* ParticleSystem::previous_position and previous_step_position
     * (include/ParticleSystem.hpp:342-343): NULL on upload stores the current
     * position (as ParticleSystem::set_particle does, :516-522). */
    double* prev_x;
    double* prev_y;
    double* prev_step_x;
    double* prev_step_y;
    int32_t* sort_key; /* ParticleSystem::sort_key (:349); NULL on upload stores 0 */
} kpt_particles_host;

typedef struct kpt_push_stats {
    uint64_t particle_steps; /* executed Euler–Maruyama sub-steps (src/main.cpp:1283-1305) */
    uint64_t failures;       /* particles that went from status 0 to status 2 in this call   */
    uint64_t active_in;      /* particles with status 0 when the call started                */
    double kernel_ms;        /* device time of the push kernel(s), from HIP events            */
    uint32_t launches;       /* push kernel launches in this call                            */
    uint32_t reserved;
} kpt_push_stats;

/* Mirrors ParticleSummary (src/main.cpp:338-350). */
typedef struct kpt_summary {
    uint64_t stored;
    uint64_t active;
    uint64_t inactive;
    double weight_sum;
    double p_min;
    double p_max;
    double p_avg;
    double dt_min;
    double dt_max;
    double dt_avg;
    int32_t max_split_level;
    int32_t reserved;
} kpt_summary;

/* Injection region and timing. The driver's inject_uniform_particles
 * (src/main.cpp:995-1048) is the default: whole domain, t ~ U(t0, t0 + dt_out).
 * The axis-aligned region of inject_particles_uniform_if
 * (include/ParticleGenerator.hpp:17-62, :884-958) is selected with use_region. */
typedef struct kpt_inject_opts {
    int32_t use_region;   /* 0: whole domain; 1: [region_x_lo, region_x_hi) x [region_y_lo, region_y_hi) */
    int32_t fixed_time;   /* 0: t ~ U(t0, t0 + mhd_dt) as the driver; 1: t = t0 */
    double region_x_lo;
    double region_x_hi;
    double region_y_lo;
    double region_y_hi;
    double p_init;        /* momentum of new particles; <= 0 selects settings.p0 */
} kpt_inject_opts;

typedef struct kpt_ctx kpt_ctx;

/* ---- context ---------------------------------------------------------------- */

/* Create a context on HIP device `device` with particle capacity `capacity` and
 * RNG seed `seed` (the reference's RandomManager(seed), include/RandomManager.hpp:28;
 * ReconnectionParticleStorage(cap), src/main.cpp:318-332).
 * The environment variable KPT_PUSH_STAGE=0, read here, turns off the staging of
 * stencil records in LDS by the context's pushes (results are the same bit for
 * bit; the switch is for A/B measurements). */
kpt_status kpt_create(kpt_ctx** out, int device, uint64_t capacity, uint64_t seed);
kpt_status kpt_destroy(kpt_ctx* ctx);
const char* kpt_last_error(const kpt_ctx* ctx);
/* Message of the last failure of kpt_create (when no context exists yet). */
const char* kpt_last_create_error(void);

/* Library version string and the device architecture string of the context. */
const char* kpt_version(void);
kpt_status kpt_device_name(kpt_ctx* ctx, char* buf, size_t len);

/* Default host settings (the member initialisers of Reconnection2DSettings,
 * src/main.cpp:129-154). */
kpt_status kpt_default_host_settings(kpt_host_settings* out);

/* make_device_settings (src/main.cpp:595-636): derive the device block from host
 * controls and the uniform grid given by its face coordinates. */
kpt_status kpt_make_device_settings(const kpt_host_settings* host, int32_t nx, int32_t ny,
                                    const double* x_edges, const double* y_edges,
                                    kpt_device_settings* out);

kpt_status kpt_set_seed(kpt_ctx* ctx, uint64_t seed);

/* ---- fields ----------------------------------------------------------------- */

/* load_parker_coefficient_frame / load_coefficient_frame (src/main.cpp:962-990):
 * read one compact-field binary (include/AthenaFieldReader.hpp:274-360) and build
 * the packed coefficient field of `model` in `slot`. */
kpt_status kpt_field_load_compact(kpt_ctx* ctx, const char* path, int32_t slot, int32_t model);
/* The same reader with stretched (non-uniform, strictly ascending) faces accepted
 * when allow_stretched != 0; the reference's reader rejects them. On a stretched
 * axis value k sits at face k (the driver's node convention, whose value k sits at
 * x0 + k dx on a uniform axis), the packed derivatives use the three-point
 * formula at those coordinates (include/FieldOperator.hpp:85-105), and the dt
 * limiter's half cell width is that of the particle's face interval.
 * kpt_field_upload accepts stretched faces as well. */
kpt_status kpt_field_load_compact_ex(kpt_ctx* ctx, const char* path, int32_t slot, int32_t model,
                                     int32_t allow_stretched);

/* Background load of the next frame, for the frame loop of run_transport_loop
 * (src/main.cpp:2305-2313: lower = upper, then read frame + 1 into upper). The
 * reference reads each frame synchronously between intervals; here a host
 * thread reads and validates `path` (as kpt_field_load_compact_ex) into pinned
 * memory, then copies it into a third, staging slot on a second stream while
 * the caller's pushes run. One stage at a time. */
kpt_status kpt_field_stage_compact(kpt_ctx* ctx, const char* path, int32_t model, int32_t allow_stretched);
/* Finish the stage and rotate the slots: lower <- upper, upper <- staged, the
 * same frames as kpt_field_advance followed by kpt_field_load_compact_ex into
 * the upper slot. The staged frame is packed here, on the context's stream
 * (about 0.1 ms at 1024 x 512; beside a push its frame writes would evict the
 * push's frames from the caches), and later device work is ordered after it. `read_seconds` (may be NULL) receives the host time of the
 * file read. A read or validation error of the stage is returned here, with
 * the slots unchanged. */
kpt_status kpt_field_advance_staged(kpt_ctx* ctx, double* read_seconds);

/* Same as kpt_field_load_compact, from host arrays in the compact-field order
 * (Bx, By, Bz, Vx, Vy, Vz each [ny][nx], x fastest). */
kpt_status kpt_field_upload(kpt_ctx* ctx, int32_t slot, int32_t model, int32_t nx, int32_t ny,
                            const double* x_edges, const double* y_edges,
                            const double* bx, const double* by, const double* bz,
                            const double* vx, const double* vy, const double* vz);

/* lower = upper (src/main.cpp:2311): the upper slot becomes the lower slot. */
kpt_status kpt_field_advance(kpt_ctx* ctx);

/* Download the packed coefficients of `slot` for nodes i in [-1, nx], j in [-1, ny]
 * as double[(ny+2)][(nx+2)][C] (C = model). `out_len` is the element count of `out`. */
kpt_status kpt_field_download_coefficients(kpt_ctx* ctx, int32_t slot, double* out, size_t out_len);

/* ---- compact-field files (host) ------------------------------------------- */

/* Parse a compact-field file into caller buffers. Pass NULL buffers to query
 * nx, ny only. Validation follows include/AthenaFieldReader.hpp:86-270, except
 * that non-uniform edges are accepted when allow_nonuniform != 0. */
kpt_status kpt_compact_field_read(const char* path, int32_t* nx, int32_t* ny,
                                  double* x_edges, double* y_edges, double* bx, double* by,
                                  double* bz, double* vx, double* vy, double* vz,
                                  int32_t allow_nonuniform, char* err, size_t err_len);
kpt_status kpt_compact_field_write(const char* path, int32_t nx, int32_t ny,
                                   const double* x_edges, const double* y_edges,
                                   const double* bx, const double* by, const double* bz,
                                   const double* vx, const double* vy, const double* vz);

/* ---- particles -------------------------------------------------------------- */

kpt_status kpt_particles_count(kpt_ctx* ctx, uint64_t* count);
kpt_status kpt_particles_capacity(kpt_ctx* ctx, uint64_t* capacity);
/* Replace the particle set with `count` host particles (status, split level and
 * weight are taken from the arrays; a NULL id array assigns id = index). */
kpt_status kpt_particles_upload(kpt_ctx* ctx, const kpt_particles_host* h, uint64_t count);
/* Copy the first `count` stored particles to the host. */
kpt_status kpt_particles_download(kpt_ctx* ctx, kpt_particles_host* h, uint64_t count);

/* ---- the hot path ------------------------------------------------------------ */

/* inject_uniform_particles (src/main.cpp:995-1048): append up to n particles. */
kpt_status kpt_inject(kpt_ctx* ctx, const kpt_device_settings* s, const kpt_inject_opts* opts,
                      uint64_t n, double t0, uint32_t frame, uint64_t* injected);

/* move_parker_particles_one_interval (src/main.cpp:1082-1327) when
 * s->focused_transport == 0 and the field model is PARKER16, and
 * move_particles_one_interval (src/main.cpp:1332-1755) when the field model is
 * FOCUSED21 (Parker or reduced focused branch per s->focused_transport).
 * Advances every status-0 particle from its time to t0 + s->mhd_dt. `frame`
 * keys the Philox stream of this interval. */
kpt_status kpt_push_interval(kpt_ctx* ctx, const kpt_device_settings* s, double t0,
                             uint32_t frame, kpt_push_stats* stats);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). kpt_push_interval with its per-chunk launch arguments and
 * counters in a file-scope device array of the library, KPT_NOALLOC_SCRATCH_BYTES
 * on each GPU the library runs on, held until the process ends; the call makes
 * no device allocation (the particles, frames and noise table are the ones
 * already held by the context). Differences from kpt_push_interval:
 *  - at most KPT_NOALLOC_MAX_CHUNKS push chunks (KPT_ERR_ARG beyond);
 *  - no cell re-sort between chunks: the particle order is kept, and every
 *    particle ends where kpt_push_interval moves it;
 *  - an FP32 time-blended push reads the two frames directly instead of the
 *    interleaved copy kpt_push_interval builds, so its blend weights round in
 *    another order (FP32-level differences; FP64, and FP32 without blending,
 *    are bit-identical).
 * The call returns when the push has finished; `stats` may be NULL (its
 * kernel_ms also covers one small staging kernel per chunk). Every call
 * on one GPU shares that GPU's array and overwrites what the last call left
 * there: calls that overlap on one GPU (other contexts, other host threads)
 * race. */
#define KPT_NOALLOC_MAX_CHUNKS 16
#define KPT_NOALLOC_SCRATCH_BYTES 24576
kpt_status kpt_push_interval_noalloc(kpt_ctx* ctx, const kpt_device_settings* s, double t0,
                                     uint32_t frame, kpt_push_stats* stats);

/* split_particles (src/main.cpp:1760-1839). Children are appended in parent slot
 * order (one admissible schedule of the reference's atomic counter). */
kpt_status kpt_split(kpt_ctx* ctx, const kpt_device_settings* s, uint64_t* split,
                     uint64_t* skipped_capacity);

/* write_momentum_histogram binning (src/main.cpp:1906-1942): weight per log-p bin
 * over [pmin, pmax], active particles only, out-of-range dropped. */
kpt_status kpt_histogram_logp(kpt_ctx* ctx, int32_t bins, double pmin, double pmax, double* out);

/* Spatial density of active particles: weight per cell of a bins_x x bins_y
 * uniform grid over [x_lo, x_hi) x [y_lo, y_hi) (out is [bins_y][bins_x]). */
kpt_status kpt_histogram_xy(kpt_ctx* ctx, int32_t bins_x, int32_t bins_y, double x_lo, double x_hi,
                            double y_lo, double y_hi, double* out);

/* ParticleSystem::sort_by_key + compact_active_particles
 * (include/ParticleSystem.hpp:715-793): reorder the stored particles by the
 * lower-left node of their interpolation stencil on the lower frame's grid
 * (stable), inactive particles last. Nodes are numbered in blocks of 16 x 16
 * (blocks row-major, nodes row-major inside a block), which keeps the particles
 * that run together on one XCD in a compact patch. With compact != 0 the inactive particles
 * are dropped and the count shrinks to the active count. Slot order changes, so
 * the injected-noise mode (which is keyed by slot) rejects this call. */
kpt_status kpt_sort_particles(kpt_ctx* ctx, const kpt_device_settings* s, int32_t compact,
                              uint64_t* active);

/* summarize_particles (src/main.cpp:1844-1901), computed on the device. */
kpt_status kpt_summary_compute(kpt_ctx* ctx, kpt_summary* out);

/* ---- noise ------------------------------------------------------------------ */

/* Select the fixed-noise mode: the push of particle slot i at its k-th sub-step of
 * an interval uses draws[(i * steps_per_particle + k) * 4 + d], d = 0..3, as the
 * four U[-sqrt3, sqrt3] values of that step (src/main.cpp:1280-1291 for Parker;
 * :1672-1698 for focused: ran1, ran2, ignored, pitch). The table is copied.
 * s->max_particle_steps_per_interval must not exceed steps_per_particle. */
kpt_status kpt_set_noise(kpt_ctx* ctx, const double* draws, uint64_t n_particles,
                         uint32_t steps_per_particle);
/* Return to Philox noise. */
kpt_status kpt_clear_noise(kpt_ctx* ctx);

/* Philox4x32-10 block (Salmon et al. 2011), exposed for known-answer tests. */
void kpt_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- v4 particle snapshots -------------------------------------------------- */

/* write_reconnection_particle_snapshot (src/main.cpp:2131-2202): byte-identical
 * version-4 KPTPRT file of the stored particles. */
kpt_status kpt_snapshot_write(kpt_ctx* ctx, const char* path, double charge, double split_ratio,
                              double p0);
/* No counterpart in the reference: the same v4 file with each record's id field
 * holding the particle id (which keys the particle's random stream) and its E0
 * field holding the stored split baseline. kpt_snapshot_read of such a file
 * followed by the same frames reproduces an uninterrupted run bit for bit. */
kpt_status kpt_snapshot_write_resumable(kpt_ctx* ctx, const char* path, double charge, double split_ratio,
                                        double p0);
/* load_reconnection_particle_snapshot (src/main.cpp:2009-2126): replace the
 * particle set; time = restart_time, step_dt = dt_min. The records' ids become
 * the particle ids (slot indices in a file of the driver). */
kpt_status kpt_snapshot_read(kpt_ctx* ctx, const char* path, double charge, double split_ratio,
                             double restart_time, double dt_min);

/* Host-only halves of the two calls above (no device involved): write `count`
 * host particles as a v4 file with the given header capacity, and read a v4
 * file into host arrays (pass NULL arrays to query the count only; `count` is
 * in/out: capacity of the arrays in, stored particles out). The reader applies
 * the header checks of src/main.cpp:2018-2073; `err` receives the message. */
kpt_status kpt_snapshot_write_host(const char* path, const kpt_particles_host* h, uint64_t count,
                                   uint64_t capacity, double charge, double split_ratio, double p0);
kpt_status kpt_snapshot_read_host(const char* path, kpt_particles_host* h, uint64_t* count,
                                  double charge, double split_ratio, char* err, size_t err_len);

/* ---- generic Parker solver (global time step) ------------------------------------
 * ParkerSolver::advance_stochastic (include/ParkerSolver.hpp:265-299) on the
 * packed 18-component basis of its coefficient fields
 * (ParkerCoefficient::make_legacy_transport_coefficient_fields,
 * include/ParkerCoefficient.hpp:1090-1140): B, V, curl(B/B^2), the primary and
 * secondary div(kappa_ij) bases, kappa_par and kappa_perp at gamma = 1
 * (LegencyModel, include/LegencyModel.hpp:359-507, constant turbulence of
 * include/TurbulenceProperties.hpp:59-210) and div V, cell-centred on the slot's
 * uploaded B and V (no interpolation shift). One step: dt = courant * min over
 * alive particles of the local limits (diffusion (dx/2)^2 / (2 kappa_eff),
 * advection dx / |u|, momentum 3 / |div V|), then x += dt u + dW with
 * Cov(dW) = 2 kappa dt split into perpendicular and field-aligned Gaussians,
 * p = max(p - dt p div V / 3, 0), periodic wrap. Where the reference aborts
 * (|B| = 0, invalid kappa), the building call returns an error and the step
 * marks the particle with status 2. */
typedef struct kpt_generic_parker_params {
    int32_t perpendicular_model;           /* 0 README formula, 1 constant ratio  */
    int32_t reserved;
    double parallel_prefactor;             /* 1.622                               */
    double perpendicular_prefactor;        /* 0.198                               */
    double perpendicular_parallel_ratio;   /* used by the constant-ratio model    */
    double minimum_magnetic_field_magnitude;
    double minimum_solar_wind_speed;
    double minimum_sigma2;
    double minimum_magnetic_field_squared_for_drift;
    double sigma2;                         /* constant turbulence: 0.5            */
    double l2d;                            /* 1                                   */
    double lslab;                          /* 1                                   */
    double lc;                             /* <= 0: lc = lslab                    */
    double rest_mass;                      /* ParticleProperties: 1               */
    double charge;                         /* 1                                   */
    double speed_of_light;                 /* 1                                   */
    double courant;                        /* SolverBase courant number: 0.3      */
} kpt_generic_parker_params;

enum { KPT_GENERIC_PARKER_COMPONENTS = 18 };

kpt_status kpt_generic_parker_default_params(kpt_generic_parker_params* out);
/* Build the packed basis of `slot` from the B and V uploaded there. */
kpt_status kpt_generic_parker_build(kpt_ctx* ctx, const kpt_generic_parker_params* p, int32_t slot);
/* The packed basis of nodes i in [-1, nx], j in [-1, ny] as
 * double[(ny+2)][(nx+2)][18]. */
kpt_status kpt_generic_parker_download(kpt_ctx* ctx, int32_t slot, double* out, size_t out_len);
/* compute_adaptive_time_step: courant * min over alive particles. */
kpt_status kpt_generic_parker_time_step(kpt_ctx* ctx, const kpt_generic_parker_params* p, int32_t slot,
                                        double* dt);
/* advance_stochastic: one step of every alive particle. `step` keys the Philox
 * stream (id, step) and selects the row of an injected noise table, which then
 * holds the four Gaussians (g_x, g_y, g_z, g_parallel) of each step. */
kpt_status kpt_generic_parker_step(kpt_ctx* ctx, const kpt_generic_parker_params* p, int32_t slot,
                                   uint32_t step, double* dt_used, uint64_t* failures);
/* advance_deterministic (include/ParkerSolver.hpp:228-255): the same step with no
 * stochastic increments, x += dt u, p = max(p + dt p_rate, 0).
 * Both advance calls follow the reference's order: the adaptive dt (an error
 * with the reference's message when it is not positive and finite), the periodic
 * wrap of alive particles, capture_previous_positions of every stored particle,
 * the update. */
kpt_status kpt_generic_parker_advance_deterministic(kpt_ctx* ctx, const kpt_generic_parker_params* p,
                                                    int32_t slot, double* dt_used, uint64_t* failures);

/* ---- batched ensemble: many independent runs in one persistent launch ----------
 * No counterpart in the reference, which runs one configuration per process. An
 * ensemble holds R runs, each with its own field frame, settings, Philox seed and
 * particle range. kpt_ensemble_push advances every run through whole intervals
 * in ONE persistent launch per interval: workgroups take 256-particle tiles from
 * per-XCD work queues. A run is either the Parker FP32 push (field model
 * PARKER16) or the focused FP32 push with the slab-QLT closure (FOCUSED21,
 * closure KPT_CLOSURE_QLT, or KPT_CLOSURE_MLP with the context's network). Each
 * run holds one frame: time interpolation is off and the interval is pushed in
 * one chunk. A run's result equals a kpt_push_interval of the same run in a
 * context of its own (same settings, seed, frame, particles). The frame loop can
 * be captured once in a hipGraph and replayed. */
typedef struct kpt_ensemble kpt_ensemble;

typedef struct kpt_ensemble_tuning {
    int32_t grid_blocks;  /* persistent workgroups of one launch (0: two per CU)           */
    int32_t groups;       /* work queues: 1 (one global queue) or 8 (one per XCD, runs
                             assigned longest-first, idle XCDs steal)                     */
    double ms;            /* autotune: time of the chosen configuration on the window      */
} kpt_ensemble_tuning;

typedef struct kpt_ensemble_run_stats {
    uint64_t particle_steps; /* sub-steps executed by the run in the call */
    uint64_t failures;       /* particles that went to status 2           */
    uint64_t active_in;      /* status-0 particles summed over the frames */
} kpt_ensemble_run_stats;

/* The ensemble uses ctx's device, stream and learned closure; destroy it before ctx. */
kpt_status kpt_ensemble_create(kpt_ctx* ctx, int32_t runs, const uint64_t* capacities, kpt_ensemble** out);
kpt_status kpt_ensemble_destroy(kpt_ensemble* e);
kpt_status kpt_ensemble_field_upload(kpt_ensemble* e, int32_t run, int32_t model, int32_t nx, int32_t ny,
                                     const double* x_edges, const double* y_edges, const double* bx,
                                     const double* by, const double* bz, const double* vx, const double* vy,
                                     const double* vz);
/* Settings and Philox seed of one run (precision must be KPT_PREC_FP32). */
kpt_status kpt_ensemble_set_run(kpt_ensemble* e, int32_t run, const kpt_device_settings* s, uint64_t seed);
kpt_status kpt_ensemble_particles_upload(kpt_ensemble* e, int32_t run, const kpt_particles_host* h,
                                         uint64_t count);
kpt_status kpt_ensemble_particles_download(kpt_ensemble* e, int32_t run, kpt_particles_host* h, uint64_t count);
kpt_status kpt_ensemble_count(kpt_ensemble* e, int32_t run, uint64_t* count);
/* inject_uniform_particles into one run (its settings, seed and root counter). */
kpt_status kpt_ensemble_inject(kpt_ensemble* e, int32_t run, uint64_t n, double t0, uint32_t frame,
                               uint64_t* injected);
/* Advance every run through `frames` intervals, frame f = first_frame .. + frames - 1
 * starting at t0 = f * mhd_dt of the run. use_graph != 0 records one interval in a
 * hipGraph and replays it. per_run (runs entries, may be NULL) receives the
 * counters; kernel_ms the device time of the push launches. */
kpt_status kpt_ensemble_push(kpt_ensemble* e, uint32_t first_frame, uint32_t frames, int32_t use_graph,
                             kpt_ensemble_run_stats* per_run, double* kernel_ms);
/* Time the candidate launch shapes ({1, 2, 3, 4} workgroups per CU x {8, 1}
 * queues) on the first `window` (0 < window <= 1) of the interval starting at
 * frame `frame`, keep the fastest, and leave the particles as they were. `table`
 * (may be NULL) receives up to `table_len` candidates tried. */
kpt_status kpt_ensemble_autotune(kpt_ensemble* e, uint32_t frame, double window, kpt_ensemble_tuning* chosen,
                                 kpt_ensemble_tuning* table, int32_t table_len, int32_t* tried);
kpt_status kpt_ensemble_set_tuning(kpt_ensemble* e, const kpt_ensemble_tuning* t);
kpt_status kpt_ensemble_histogram_logp(kpt_ensemble* e, int32_t run, int32_t bins, double pmin, double pmax,
                                       double* out);
/* Reorder every run's particles by the lower-left node of their stencil (stable;
 * inactive particles last), so the gathers of one tile share cache lines.
 * Results per particle id do not change. `active` (may be NULL) receives the
 * number of status-0 particles over all runs. */
kpt_status kpt_ensemble_sort(kpt_ensemble* e, uint64_t* active);

/* ---- generic focused-transport solver (global time step) -------------------------
 * FocusTransportSolver (include/FocusSolver.hpp:74-1890) on the packed basis of
 * FocusCoefficient::make_legacy_transport_coefficient_fields_with_time_derivative
 * (include/FocusCoefficient.hpp:858-990), FP64, on the periodic grid of a field
 * slot (cell-centred, like the generic Parker solver). Packed components:
 *   0-2 B, 3-5 b, 6-8 V, 9-11 b x grad ln|B| / |B|, 12-14 b x (b.grad)b / |B|,
 *   15-17 div(kappa_perp(gamma=1) (I - bb)), 18 |B|, 19 kappa_par(gamma=1),
 *   20 kappa_perp(gamma=1), 21 div V, 22 b.grad ln|B|, 23 bb:grad V,
 *   24 b.dV/dt + b.(V.grad)V, 25 sigma2, 26 Lc. */
typedef struct kpt_generic_focus_params {
    kpt_generic_parker_params model;      /* diffusion model, turbulence, particle, courant */
    double minimum_magnetic_field_magnitude; /* floor of b, ln|B| and the drift bases: 0 */
    /* PitchAngleScatteringParameters (include/LegencyModel.hpp:196-223) */
    double spectral_index;                /* 5/3  */
    double resonance_regularization_h0;   /* 1e-6 */
    double scattering_minimum_magnetic_field_magnitude;
    double scattering_minimum_sigma2;
    double scattering_minimum_correlation_length;
    double scattering_minimum_speed;
    double maximum_pitch_angle_cosine;    /* set_maximum_pitch_angle_cosine: 1 */
} kpt_generic_focus_params;

enum { KPT_GENERIC_FOCUS_COMPONENTS = 27 };

kpt_status kpt_generic_focus_default_params(kpt_generic_focus_params* out);
/* Build the packed basis of `slot` from the B and V uploaded there; dvdt is NULL
 * (dV/dt = 0, make_legacy_transport_coefficient_fields) or three planes
 * [3][ny][nx] of dV/dt, x fastest. */
kpt_status kpt_generic_focus_build(kpt_ctx* ctx, const kpt_generic_focus_params* p, int32_t slot,
                                   const double* dvdt);
/* The packed basis of nodes i in [-1, nx], j in [-1, ny] as double[(ny+2)][(nx+2)][27]. */
kpt_status kpt_generic_focus_download(kpt_ctx* ctx, int32_t slot, double* out, size_t out_len);
/* compute_adaptive_time_step (include/FocusSolver.hpp:208): courant * min over
 * alive particles of the six local bounds; 0 when none is finite. */
kpt_status kpt_generic_focus_time_step(kpt_ctx* ctx, const kpt_generic_focus_params* p, int32_t slot,
                                       double* dt);
/* advance_stochastic (include/FocusSolver.hpp:276): one step of every alive
 * particle (x, y, p, mu). `step` keys the Philox stream and selects the row of
 * an injected noise table: (g_x, g_y, g_z) of the perpendicular increment and
 * the pitch-angle normal. Particles where the reference aborts (|B| <= 0, zero
 * speed, invalid coefficients) get status 2 and are counted in failures. */
kpt_status kpt_generic_focus_step(kpt_ctx* ctx, const kpt_generic_focus_params* p, int32_t slot,
                                  uint32_t step, double* dt_used, uint64_t* failures);
/* advance_deterministic (include/FocusSolver.hpp:251-271): the same step with the
 * perpendicular increment and the pitch-angle normal zero (the Picard passes
 * then integrate the drift only); dt, wrap and capture as the Parker calls. */
kpt_status kpt_generic_focus_advance_deterministic(kpt_ctx* ctx, const kpt_generic_focus_params* p,
                                                   int32_t slot, double* dt_used, uint64_t* failures);

/* ---- ParticleSystem lifecycle (include/ParticleSystem.hpp) ----------------------- */

/* ParticleProperties (include/ParticleSystem.hpp:145-213). */
typedef struct kpt_particle_properties {
    double rest_mass;        /* 1 */
    double charge;           /* 1 */
    double speed_of_light;   /* 1 */
    double energy_scale_erg; /* 1 */
} kpt_particle_properties;

/* ParticleSplittingPolicy (include/ParticleSystem.hpp:280-295). */
typedef struct kpt_energy_split_policy {
    double energy_split_ratio;   /* > 1; 2 */
    double minimum_child_weight; /* >= 0; 0 */
} kpt_energy_split_policy;

/* ParticleSplitStatistics (include/ParticleSystem.hpp:268-275). */
typedef struct kpt_split_statistics {
    uint64_t active_count;
    uint64_t split_count;
    uint64_t skipped_capacity_count;
    uint64_t skipped_invalid_baseline_count;
    uint64_t skipped_minimum_weight_count;
} kpt_split_statistics;

kpt_status kpt_default_particle_properties(kpt_particle_properties* out);
kpt_status kpt_default_energy_split_policy(kpt_energy_split_policy* out);

/* split_particles_on_energy_growth (include/ParticleSystem.hpp:802-917): an alive
 * parent with E_k / e0 >= ratio and floor(ln(E_k/e0)/ln ratio) > level appends one
 * child at the tail; both carry half the weight and level + 1. A non-finite or
 * non-positive e0 is replaced by E_k and counted. Children are appended in parent
 * slot order and take fresh root ids (the ids key the RNG streams). */
kpt_status kpt_split_energy_growth(kpt_ctx* ctx, const kpt_particle_properties* props,
                                   const kpt_energy_split_policy* policy, kpt_split_statistics* stats);
/* reset_split_baselines (include/ParticleSystem.hpp:630-653): e0 = E_k(|p|) and
 * split level 0 for every stored particle. */
kpt_status kpt_reset_split_baselines(kpt_ctx* ctx, const kpt_particle_properties* props);
/* apply_grid_boundary_conditions (include/ParticleSystem.hpp:661-706) on the grid of
 * field `slot`: alive particles are wrapped along periodic axes (bit 0: x, bit 1:
 * y of periodic_mask) and get status 1 (Escaped) when they leave the physical
 * domain [lo, hi] along a non-periodic axis. */
kpt_status kpt_apply_boundaries(kpt_ctx* ctx, int32_t slot, int32_t periodic_mask, uint64_t* escaped);
/* gamma_from_momentum_magnitude, speed_from_momentum_magnitude and
 * kinetic_energy_from_momentum_magnitude (include/ParticleSystem.hpp:170-195) of the
 * first `count` stored particles; any output may be NULL. */
kpt_status kpt_particle_kinematics(kpt_ctx* ctx, const kpt_particle_properties* props, double* gamma,
                                   double* speed, double* kinetic_energy, uint64_t count);
/* capture_previous_positions (include/ParticleSystem.hpp:623-628):
 * previous_step_position = previous_position, previous_position = position, for
 * every stored particle. */
kpt_status kpt_capture_previous_positions(kpt_ctx* ctx);
/* sort_by_key (include/ParticleSystem.hpp:727-760): every stored array permuted by
 * the stored sort keys (kpt_particles_host.sort_key) in [min_key, max_key]; keys
 * outside the range take its nearest end. Stable: equal keys keep their slot
 * order (one admissible order of Kokkos::BinSort). */
kpt_status kpt_sort_by_key(kpt_ctx* ctx, int32_t min_key, int32_t max_key);
/* compact_active_particles (include/ParticleSystem.hpp:769-793): sort key 0 for
 * alive particles and 1 for the others, sort_by_key(0, 1), and the stored count
 * becomes the alive count (returned in `active`). */
kpt_status kpt_compact_active_particles(kpt_ctx* ctx, uint64_t* active);
/* write_binary_snapshot (include/ParticleSystem.hpp:936-1006): the version-4 file
 * of the stored particles with their ids, sort keys, both previous positions and
 * split baselines, and `species` (ParticleSpecies: 0 proton .. 3 custom), the
 * properties and the splitting policy in the header. */
kpt_status kpt_particle_system_write_snapshot(kpt_ctx* ctx, const char* path, const kpt_particle_properties* props,
                                              int32_t species, const kpt_energy_split_policy* policy);
/* read_binary_snapshot (include/ParticleSystem.hpp:1011-1145) with its checks and
 * messages: the file's particles replace the stored set (time and step_dt 0), and
 * its header's properties, species and policy are returned (any output may be
 * NULL). A file with more particles than the capacity is refused. */
kpt_status kpt_particle_system_read_snapshot(kpt_ctx* ctx, const char* path, kpt_particle_properties* props,
                                             int32_t* species, kpt_energy_split_policy* policy);

/* ---- generic particle generator (include/ParticleGenerator.hpp) ------------------
 * Injection and recycling over the cells of a field slot. The slot's raw planes
 * (Bx, By, Bz, Vx, Vy, Vz) are taken as cell-centred values of its grid, as the
 * generic solvers do; cell areas are the products of the cell widths (the
 * Cartesian Jacobian is 1). Each particle draws from its own Philox stream keyed
 * by its id, a stream tag, the attempt and a per-context generator call counter,
 * so results do not depend on the launch shape; the reference draws from its
 * Kokkos random pool. New particles take fresh root ids. */

/* FieldPredicateComparison (include/ParticleGenerator.hpp:117-124). */
enum {
    KPT_CMP_GREATER = 0,
    KPT_CMP_GREATER_EQUAL = 1,
    KPT_CMP_LESS = 2,
    KPT_CMP_LESS_EQUAL = 3,
    KPT_CMP_ABS_GREATER = 4,
    KPT_CMP_ABS_LESS = 5
};
/* ParticleRecycleSelection (include/ParticleGenerator.hpp:129-133) over the
 * statuses of ParticleStatus (include/ParticleSystem.hpp:35-39): 0 active,
 * 1 escaped, 2 inactive. */
enum { KPT_RECYCLE_ESCAPED = 0, KPT_RECYCLE_INACTIVE = 1, KPT_RECYCLE_ANY = 2 };

/* ParticleInjectionKinematics (include/ParticleGenerator.hpp:67-74), plus the
 * time and first step the driver's push reads. */
typedef struct kpt_injection_kinematics {
    double momentum;  /* |p|, finite and >= 0 */
    double mu;        /* finite */
    double weight;    /* finite and > 0 */
    int32_t status;   /* 0, 1 or 2 */
    int32_t pad;
    double time;      /* t of the placed particles */
    double step_dt;   /* their stored step (the push's first adaptive guess) */
} kpt_injection_kinematics;

/* AxisAlignedParticleRegion (include/ParticleGenerator.hpp:17-60): finite and
 * ascending bounds; membership is inclusive, draws are uniform in it. */
typedef struct kpt_region {
    double x_lo, x_hi, y_lo, y_hi;
} kpt_region;

/* FieldThresholdParticlePredicate (include/ParticleGenerator.hpp:254-358):
 * raw component `component` (0..5) of the containing cell compared with
 * `threshold`; false outside the physical domain. */
typedef struct kpt_field_predicate {
    int32_t component;
    int32_t comparison; /* KPT_CMP_* */
    double threshold;
} kpt_field_predicate;

typedef struct kpt_candidates kpt_candidates;

/* make_field_particle_candidate_set(_in_region) (include/ParticleGenerator.hpp:
 * 658-784): the cells of `slot` whose centre lies in `region` (NULL: the
 * physical domain) and whose value passes `pred`, in x-fastest order, with
 * cumulative cell areas. An empty set is valid; injecting from it is not. */
kpt_status kpt_candidates_from_field(kpt_ctx* ctx, int32_t slot, const kpt_region* region,
                                     const kpt_field_predicate* pred, kpt_candidates** out);
/* make_grid_particle_candidate_set (include/ParticleGenerator.hpp:786-845): every
 * physical cell of the slot's grid with cumulative areas. */
kpt_status kpt_candidates_from_grid(kpt_ctx* ctx, int32_t slot, kpt_candidates** out);
/* candidate_count() and the total weight. */
kpt_status kpt_candidates_info(const kpt_candidates* c, uint64_t* count, double* total_weight);
/* The accepted cell indices (i + nx j) and cumulative weights; either output
 * may be NULL, each holds `count` values. */
kpt_status kpt_candidates_download(kpt_ctx* ctx, const kpt_candidates* c, uint32_t* cells, double* cumulative);
kpt_status kpt_candidates_destroy(kpt_candidates* c);

/* inject_particles_from_field_candidates / _from_grid_candidates
 * (include/ParticleGenerator.hpp:957-1117): append n particles, each at a
 * candidate drawn by weight (sample_cumulative_weight_index, :436-460): its cell
 * centre (field candidates) or a uniform point in its cell (grid candidates).
 * e0 is the kinetic energy of `momentum` under `props` (NULL: defaults).
 * ParticleInjectionPlan's capacity check (:846-870) applies. */
kpt_status kpt_inject_from_candidates(kpt_ctx* ctx, const kpt_candidates* c, uint64_t n,
                                      const kpt_injection_kinematics* k, const kpt_particle_properties* props,
                                      uint64_t* injected);
/* inject_particles_uniform(_if) (include/ParticleGenerator.hpp:875-955,
 * :1122-1132): append n particles uniform in `region` (NULL: the slot's physical
 * domain), each redrawn up to max_attempts times while `pred` (NULL: always
 * true) fails; those that never pass keep their last draw with status 2
 * (inactive) and are counted in `rejected`. */
kpt_status kpt_inject_uniform_if(kpt_ctx* ctx, int32_t slot, const kpt_region* region,
                                 const kpt_field_predicate* pred, int32_t max_attempts, uint64_t n,
                                 const kpt_injection_kinematics* k, const kpt_particle_properties* props,
                                 uint64_t* injected, uint64_t* rejected);
/* recycle_particles_uniform(_if) (include/ParticleGenerator.hpp:1176-1262):
 * every stored particle whose status matches `selection` (KPT_RECYCLE_*) is
 * placed anew as kpt_inject_uniform_if places one, keeping its slot and id. */
kpt_status kpt_recycle_uniform_if(kpt_ctx* ctx, int32_t slot, const kpt_region* region,
                                  const kpt_field_predicate* pred, int32_t max_attempts, int32_t selection,
                                  const kpt_injection_kinematics* k, const kpt_particle_properties* props,
                                  uint64_t* recycled, uint64_t* rejected);
/* recycle_particles_from_grid_candidates / _from_field_candidates
 * (include/ParticleGenerator.hpp:1314-1436): the same from a candidate set. */
kpt_status kpt_recycle_from_candidates(kpt_ctx* ctx, const kpt_candidates* c, int32_t selection,
                                       const kpt_injection_kinematics* k, const kpt_particle_properties* props,
                                       uint64_t* recycled);
/* The generator call counter: 0 after kpt_create, +1 per inject or recycle call
 * above; it is the third Philox counter word of the next call's draws. */
kpt_status kpt_generator_calls(const kpt_ctx* ctx, uint32_t* calls);

/* ---- learned pitch-angle closure (KPT_CLOSURE_MLP) ------------------------------
 * No counterpart in the reference. A small network stands in for the resonance
 * part of the slab-QLT D_mumu of LegencyModel::evaluate_pitch_angle_scattering
 * (include/LegencyModel.hpp:539-689):
 *   D = (pi/4) Omega sigma2 C_gamma (1 - mu^2) R(|mu|, xi, gamma_k),
 *   R = xi mu_h^(gamma_k - 1) / (mu_h^gamma_k + xi^gamma_k),
 * network inputs (ln|mu|, ln xi, gamma_k), outputs (ln R - out_mean) / out_scale
 * and |mu| d ln R / d|mu|; 3 -> 64 -> 64 -> 64 -> 2 with SiLU (default) or
 * hard-swish hidden units (kpt_mlp_hull.activation). It is trained on
 * the GPU (kpt_mlp_train) on exact-QLT samples and evaluated inside the FP32
 * focused push with BF16 MFMA and FP32 accumulation; samples outside the hull
 * (and every sample when the push's gamma_k or h0 differ from the hull's) take
 * the exact formula. */
enum { KPT_MLP_INPUTS = 3, KPT_MLP_HIDDEN = 64, KPT_MLP_OUTPUTS = 2, KPT_MLP_PARAMS = 8706 };

/* Hidden-layer activation of the network. */
enum {
    KPT_MLP_SILU = 0,       /* z / (1 + e^-z)                                        */
    KPT_MLP_HARD_SWISH = 1  /* z clamp(z / 6 + 1/2, 0, 1): no transcendental per unit */
};

/* The trained network's description: its input hull, output scaling and
 * activation. */
typedef struct kpt_mlp_hull {
    double mu_lo;                 /* |mu| >= mu_lo (and < 1)            default 1e-3  */
    double log_xi_lo, log_xi_hi;  /* ln xi range                        ln 1e-7, 0    */
    double gamma_lo, gamma_hi;    /* gamma_k range                      1.5, 2        */
    double h0;                    /* resonance regularisation trained   1e-6          */
    double out_mean, out_scale;   /* ln R = out_mean + out_scale * y0 (set by training) */
    int32_t activation;           /* KPT_MLP_SILU (default) or KPT_MLP_HARD_SWISH      */
    int32_t reserved;             /* 0                                                  */
} kpt_mlp_hull;

typedef struct kpt_mlp_train_opts {
    uint64_t samples;  /* training set size, default 2^22 */
    uint32_t steps;    /* Adam steps, default 8000 */
    uint32_t batch;    /* default 4096 */
    double lr;         /* peak learning rate, cosine decay to 0, default 3e-3 */
    uint64_t seed;
    kpt_mlp_hull hull;
} kpt_mlp_train_opts;

typedef struct kpt_mlp_report {
    double final_loss;               /* mean squared error of the last batch (scaled targets) */
    double val_rms_lnr, val_max_lnr; /* held-out samples, BF16-MFMA inference vs exact FP64  */
    double val_rms_y1, val_max_y1;
    double val_rms_rel_d;            /* rms relative error of D over the held-out set        */
    double train_ms;
    uint32_t steps;
    uint32_t validation_samples;
} kpt_mlp_report;

kpt_status kpt_mlp_default_train_opts(kpt_mlp_train_opts* opts);
/* Train on the GPU: exact-QLT samples from a Philox stream, forward / backward /
 * Adam kernels, then build the BF16 MFMA image used by the push. */
kpt_status kpt_mlp_train(kpt_ctx* ctx, const kpt_mlp_train_opts* opts, kpt_mlp_report* report);
/* Load / read the FP32 parameters (KPT_MLP_PARAMS floats) and the hull. */
kpt_status kpt_mlp_set(kpt_ctx* ctx, const float* params, size_t n, const kpt_mlp_hull* hull);
kpt_status kpt_mlp_get(kpt_ctx* ctx, float* params, size_t n, kpt_mlp_hull* hull);
/* Gradient of the mean squared error over n samples (features n x 3 and scaled
 * targets n x 2, host arrays) with respect to the parameters, from the training
 * kernels; n <= 65536. */
kpt_status kpt_mlp_gradient(kpt_ctx* ctx, const float* feats, const float* targets, uint64_t n, float* grad,
                            double* loss);
/* Network outputs for n host feature rows: mfma = 0 the FP32 training forward
 * kernel, mfma = 1 the BF16 MFMA wavefront evaluation of the push. */
kpt_status kpt_mlp_infer(kpt_ctx* ctx, const float* feats, uint64_t n, int32_t mfma, float* out);

/* ---- CSV outputs (host only) ------------------------------------------------ */

/* Per-frame timing and split counters: FrameTiming (src/main.cpp:355-363). */
typedef struct kpt_frame_timing {
    double read_seconds;
    double inject_seconds;
    double move_seconds;
    double split_seconds;
    uint64_t injected;
    uint64_t split;
    uint64_t skipped_split_capacity;
} kpt_frame_timing;

/* write_momentum_histogram's CSV (src/main.cpp:1944-1963): header
 * "bin,p_left,p_right,weight", edges exp(ln pmin + i dlnp), 16 significant digits. */
kpt_status kpt_write_histogram_csv(const char* path, int32_t bins, double pmin, double pmax,
                                   const double* weights);

/* write_summary_header / write_summary_row (src/main.cpp:2207-2235) into a
 * caller buffer (NUL-terminated, newline included). Returns KPT_ERR_ARG when the
 * buffer is too small. */
kpt_status kpt_format_summary_header(char* buf, size_t len);
kpt_status kpt_format_summary_row(char* buf, size_t len, int32_t frame, double time,
                                  const kpt_summary* summary, const kpt_frame_timing* timing,
                                  double total_frame_seconds, double total_elapsed_seconds);

#ifdef __cplusplus
}
#endif

#endif /* KPT_H */
