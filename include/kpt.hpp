/* kpt.hpp — header-only C++ facade over the C ABI of kpt.h with the names of the
 * reference's generic solver API, so code written against
 *   ParticleSystem        (include/ParticleSystem.hpp:307-1245),
 *   ParkerSolver          (include/ParkerSolver.hpp:64-1674),
 *   FocusTransportSolver  (include/FocusSolver.hpp:74-1890, alias FocusSolver)
 * keeps its call surface: set_coefficient_fields, compute_adaptive_time_step,
 * advance_deterministic, advance_stochastic (advance_sde),
 * set_maximum_pitch_angle_cosine, capture_previous_positions,
 * split_particles_on_energy_growth, reset_split_baselines,
 * apply_grid_boundary_conditions, sort_by_key, compact_active_particles,
 * write_binary_snapshot and read_binary_snapshot.
 *
 * Differences a caller sees:
 *   - one ParticleSystem owns one device context (kpt_ctx); the solvers work on
 *     field slots of that context and keep no particle copies;
 *   - advance_stochastic takes no RandomManager: every particle draws from its
 *     own counter-based Philox stream keyed by its id and the solver's step
 *     counter (ParticleSystem::set_injected_noise supplies a noise table
 *     instead, for step-level comparisons);
 *   - errors throw kpt::Error (a std::runtime_error carrying the kpt_status)
 *     with the reference's message where the reference has one.
 * Link with libkpt.so. */
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "kpt.h"

namespace kpt {

class Error : public std::runtime_error {
public:
    Error(kpt_status s, const std::string& what) : std::runtime_error(what), status(s) {}
    kpt_status status;
};

inline void check(const kpt_ctx* ctx, kpt_status st) {
    if (st == KPT_OK) return;
    const char* m = ctx ? kpt_last_error(ctx) : kpt_last_create_error();
    throw Error(st, m && *m ? m : "kpt: call failed");
}

inline kpt_particle_properties default_particle_properties() {
    kpt_particle_properties p{};
    kpt_default_particle_properties(&p);
    return p;
}

inline kpt_energy_split_policy default_splitting_policy() {
    kpt_energy_split_policy p{};
    kpt_default_energy_split_policy(&p);
    return p;
}

/* ParticleSystem: the particle set of one device context, with its properties
 * and splitting policy. */
class ParticleSystem {
public:
    explicit ParticleSystem(uint64_t capacity, const kpt_particle_properties& properties = default_particle_properties(),
                            const kpt_energy_split_policy& policy = default_splitting_policy(), uint64_t seed = 114514,
                            int device = 0, int32_t species = 3)
        : properties_(properties), policy_(policy), species_(species) {
        check(nullptr, kpt_create(&ctx_, device, capacity, seed));
    }
    ~ParticleSystem() {
        if (ctx_) kpt_destroy(ctx_);
    }
    ParticleSystem(const ParticleSystem&) = delete;
    ParticleSystem& operator=(const ParticleSystem&) = delete;
    ParticleSystem(ParticleSystem&& o) noexcept
        : ctx_(std::exchange(o.ctx_, nullptr)), properties_(o.properties_), policy_(o.policy_), species_(o.species_) {}

    kpt_ctx* handle() const { return ctx_; }
    const kpt_particle_properties& properties() const { return properties_; }
    const kpt_energy_split_policy& splitting_policy() const { return policy_; }
    int32_t species() const { return species_; }

    uint64_t particle_count() const {
        uint64_t n = 0;
        check(ctx_, kpt_particles_count(ctx_, &n));
        return n;
    }
    uint64_t particle_capacity() const {
        uint64_t n = 0;
        check(ctx_, kpt_particles_capacity(ctx_, &n));
        return n;
    }

    /* Replace the stored particles (host arrays are read during the call only). */
    void upload(const kpt_particles_host& h, uint64_t count) { check(ctx_, kpt_particles_upload(ctx_, &h, count)); }
    /* Copy the first `count` stored particles into the non-null host arrays. */
    void download(kpt_particles_host& h, uint64_t count) const {
        check(ctx_, kpt_particles_download(ctx_, &h, count));
    }

    /* Step-level comparisons: draws[particle][step][4] replace the Philox draws
     * of the solvers' advance_stochastic (clear_injected_noise returns to Philox). */
    void set_injected_noise(const double* draws, uint64_t particles, uint32_t steps) {
        check(ctx_, kpt_set_noise(ctx_, draws, particles, steps));
    }
    void clear_injected_noise() { check(ctx_, kpt_clear_noise(ctx_)); }

    void capture_previous_positions() { check(ctx_, kpt_capture_previous_positions(ctx_)); }

    kpt_split_statistics split_particles_on_energy_growth() {
        kpt_split_statistics st{};
        check(ctx_, kpt_split_energy_growth(ctx_, &properties_, &policy_, &st));
        return st;
    }
    void reset_split_baselines() { check(ctx_, kpt_reset_split_baselines(ctx_, &properties_)); }
    /* Wrap alive particles on the periodic axes of field slot `slot` (bit 0 x,
     * bit 1 y of periodic_mask) and mark escapes along the others; returns the
     * escaped count. */
    uint64_t apply_grid_boundary_conditions(int32_t slot, int32_t periodic_mask = 3) {
        uint64_t escaped = 0;
        check(ctx_, kpt_apply_boundaries(ctx_, slot, periodic_mask, &escaped));
        return escaped;
    }
    void sort_by_key(int32_t minimum_key, int32_t maximum_key) {
        check(ctx_, kpt_sort_by_key(ctx_, minimum_key, maximum_key));
    }
    uint64_t compact_active_particles() {
        uint64_t n = 0;
        check(ctx_, kpt_compact_active_particles(ctx_, &n));
        return n;
    }

    void write_binary_snapshot(const std::string& path) const {
        check(ctx_, kpt_particle_system_write_snapshot(ctx_, path.c_str(), &properties_, species_, &policy_));
    }
    /* The reference builds a new system from the file; here the file replaces
     * this system's particles, properties, species and policy. */
    void read_binary_snapshot(const std::string& path) {
        check(ctx_, kpt_particle_system_read_snapshot(ctx_, path.c_str(), &properties_, &species_, &policy_));
    }

private:
    kpt_ctx* ctx_ = nullptr;
    kpt_particle_properties properties_;
    kpt_energy_split_policy policy_;
    int32_t species_;
};

/* The B and V planes of one field, cell-centred values [ny][nx] with x fastest,
 * and its nx + 1, ny + 1 face coordinates (the AthenaFieldReader layout). */
struct FieldPlanes {
    int32_t nx = 0, ny = 0;
    const double* x_edges = nullptr;
    const double* y_edges = nullptr;
    const double* b[3] = {nullptr, nullptr, nullptr};
    const double* v[3] = {nullptr, nullptr, nullptr};
};

inline kpt_generic_parker_params default_parker_parameters() {
    kpt_generic_parker_params p{};
    kpt_generic_parker_default_params(&p);
    return p;
}

inline kpt_generic_focus_params default_focus_parameters() {
    kpt_generic_focus_params p{};
    kpt_generic_focus_default_params(&p);
    return p;
}

/* ParkerSolver: global-dt Parker transport of a ParticleSystem on one field slot. */
class ParkerSolver {
public:
    explicit ParkerSolver(ParticleSystem& particles,
                          const kpt_generic_parker_params& parameters = default_parker_parameters(),
                          int32_t slot = KPT_SLOT_LOWER)
        : particles_(particles), params_(parameters), slot_(slot) {}

    /* Upload B and V to the solver's slot and build the packed coefficient basis
     * (set_coefficient_fields, include/ParkerSolver.hpp:159-164, with the fields
     * of ParkerCoefficient::make_legacy_transport_coefficient_fields). */
    void set_coefficient_fields(const FieldPlanes& f) {
        kpt_ctx* c = particles_.handle();
        check(c, kpt_field_upload(c, slot_, KPT_MODEL_PARKER16, f.nx, f.ny, f.x_edges, f.y_edges, f.b[0], f.b[1],
                                  f.b[2], f.v[0], f.v[1], f.v[2]));
        check(c, kpt_generic_parker_build(c, &params_, slot_));
    }
    double compute_adaptive_time_step() const {
        double dt = 0.0;
        check(particles_.handle(), kpt_generic_parker_time_step(particles_.handle(), &params_, slot_, &dt));
        return dt;
    }
    double advance_deterministic() {
        double dt = 0.0;
        check(particles_.handle(),
              kpt_generic_parker_advance_deterministic(particles_.handle(), &params_, slot_, &dt, &failures_));
        return dt;
    }
    double advance_stochastic() {
        double dt = 0.0;
        check(particles_.handle(),
              kpt_generic_parker_step(particles_.handle(), &params_, slot_, step_, &dt, &failures_));
        ++step_;
        return dt;
    }
    double advance_sde() { return advance_stochastic(); }

    const kpt_generic_parker_params& parameters() const { return params_; }
    uint64_t last_failures() const { return failures_; }  // particles marked failed by the last advance
    uint32_t stochastic_steps() const { return step_; }

private:
    ParticleSystem& particles_;
    kpt_generic_parker_params params_;
    int32_t slot_;
    uint32_t step_ = 0;
    uint64_t failures_ = 0;
};

/* FocusTransportSolver: global-dt focused transport with slab-QLT pitch-angle
 * scattering of a ParticleSystem on one field slot. */
class FocusTransportSolver {
public:
    explicit FocusTransportSolver(ParticleSystem& particles,
                                  const kpt_generic_focus_params& parameters = default_focus_parameters(),
                                  int32_t slot = KPT_SLOT_LOWER)
        : particles_(particles), params_(parameters), slot_(slot) {}

    /* set_coefficient_fields (include/FocusSolver.hpp:184-190) with the fields of
     * FocusCoefficient::make_legacy_transport_coefficient_fields(_with_time_
     * derivative): dvdt is NULL or three planes [3][ny][nx] of dV/dt. */
    void set_coefficient_fields(const FieldPlanes& f, const double* dvdt = nullptr) {
        kpt_ctx* c = particles_.handle();
        check(c, kpt_field_upload(c, slot_, KPT_MODEL_PARKER16, f.nx, f.ny, f.x_edges, f.y_edges, f.b[0], f.b[1],
                                  f.b[2], f.v[0], f.v[1], f.v[2]));
        check(c, kpt_generic_focus_build(c, &params_, slot_, dvdt));
    }
    /* set_maximum_pitch_angle_cosine (include/FocusSolver.hpp:219-226): the
     * reflecting bound of |mu| in (0, 1]. */
    void set_maximum_pitch_angle_cosine(double mu_max) {
        if (!(mu_max > 0.0 && mu_max <= 1.0))  // false for NaN too
            throw Error(KPT_ERR_ARG, "FocusTransportSolver: maximum |mu| must satisfy 0 < value <= 1.");
        params_.maximum_pitch_angle_cosine = mu_max;
    }
    double compute_adaptive_time_step() const {
        double dt = 0.0;
        check(particles_.handle(), kpt_generic_focus_time_step(particles_.handle(), &params_, slot_, &dt));
        return dt;
    }
    double advance_deterministic() {
        double dt = 0.0;
        check(particles_.handle(),
              kpt_generic_focus_advance_deterministic(particles_.handle(), &params_, slot_, &dt, &failures_));
        return dt;
    }
    double advance_stochastic() {
        double dt = 0.0;
        check(particles_.handle(), kpt_generic_focus_step(particles_.handle(), &params_, slot_, step_, &dt, &failures_));
        ++step_;
        return dt;
    }
    double advance_sde() { return advance_stochastic(); }

    const kpt_generic_focus_params& parameters() const { return params_; }
    uint64_t last_failures() const { return failures_; }
    uint32_t stochastic_steps() const { return step_; }

private:
    ParticleSystem& particles_;
    kpt_generic_focus_params params_;
    int32_t slot_;
    uint32_t step_ = 0;
    uint64_t failures_ = 0;
};

using FocusSolver = FocusTransportSolver;

}  // namespace kpt
