// kpt_generator_capi.hip — host side of the generic particle generator C ABI
// (kpt_candidates_*, kpt_inject_*, kpt_recycle_*; include/ParticleGenerator.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "kpt_capi_internal.h"

// ---- generic particle generator (include/ParticleGenerator.hpp) -----------------

struct kpt_candidates {
    int device = 0;
    kpt::CandidateView view{};
    uint32_t* cells = nullptr;    // accepted cells (field candidates)
    double* cumulative = nullptr;
    double* face[2] = {nullptr, nullptr};         // copies of the slot's stretched tables,
    kpt::FaceBin* bins[2] = {nullptr, nullptr};   // so the set outlives later uploads
};

namespace {

kpt_status check_kinematics(kpt_ctx* ctx, const kpt_injection_kinematics* k) {
    if (!k) return fail(ctx, KPT_ERR_ARG, "injection kinematics pointer is null.");
    // validate_particle_injection_kinematics (include/ParticleGenerator.hpp:89-102)
    if (!std::isfinite(k->mu) || !std::isfinite(k->weight) || k->weight <= 0.0)
        return fail(ctx, KPT_ERR_ARG, "ParticleInjectionKinematics: mu must be finite and weight must be positive.");
    if (!std::isfinite(k->momentum) || k->momentum < 0.0)
        return fail(ctx, KPT_ERR_ARG,
                    "ParticleInjectionKinematics: momentum magnitude must be finite and non-negative.");
    if (k->status < 0 || k->status > 2) return fail(ctx, KPT_ERR_ARG, "particle status must be 0, 1 or 2.");
    if (!std::isfinite(k->time) || !std::isfinite(k->step_dt) || k->step_dt < 0.0)
        return fail(ctx, KPT_ERR_ARG, "injection time must be finite and the stored step finite and >= 0.");
    return KPT_OK;
}

// e0 = ParticleProperties::kinetic_energy_from_momentum_magnitude (include/
// ParticleSystem.hpp:170-193) of the injected momentum.
kpt_status injection_energy(kpt_ctx* ctx, const kpt_particle_properties* props, double momentum, double& e0) {
    kpt_particle_properties pr;
    kpt_default_particle_properties(&pr);
    if (props) pr = *props;
    const kpt_status st = check_properties(ctx, &pr);
    if (st != KPT_OK) return st;
    const double nm = momentum / (pr.rest_mass * pr.speed_of_light);
    e0 = (std::sqrt(1.0 + nm * nm) - 1.0) * pr.rest_mass * pr.speed_of_light * pr.speed_of_light;
    return KPT_OK;
}

// The slot, region and predicate of a generator call.
kpt_status make_source(kpt_ctx* ctx, int32_t slot, const kpt_region* region, const kpt_field_predicate* pred,
                       int32_t max_attempts, kpt::GeneratorSource& src) {
    if (slot != KPT_SLOT_LOWER && slot != KPT_SLOT_UPPER) return fail(ctx, KPT_ERR_ARG, "invalid field slot.");
    const FieldSlot& f = ctx->slot[slot];
    if (!f.valid) return fail(ctx, KPT_ERR_STATE, "field slot is empty.");
    src = kpt::GeneratorSource{};
    src.g = f.g;
    src.raw = f.raw;
    if (region) {
        const double b[4] = {region->x_lo, region->x_hi, region->y_lo, region->y_hi};
        for (int d = 0; d < 2; ++d)
            if (!std::isfinite(b[2 * d]) || !std::isfinite(b[2 * d + 1]) || !(b[2 * d + 1] > b[2 * d]))
                return fail(ctx, KPT_ERR_ARG, "AxisAlignedParticleRegion: bounds must be finite and ascending.");
        for (int q = 0; q < 4; ++q) src.region[q] = b[q];
    } else {
        // make_particle_region_from_grid (include/ParticleGenerator.hpp:208-228)
        src.region[0] = f.g.phys_lo[0];
        src.region[1] = f.g.phys_hi[0];
        src.region[2] = f.g.phys_lo[1];
        src.region[3] = f.g.phys_hi[1];
    }
    if (pred) {
        if (pred->component < 0 || pred->component > 5)
            return fail(ctx, KPT_ERR_ARG, "Field particle injection: component index is out of range.");
        if (pred->comparison < KPT_CMP_GREATER || pred->comparison > KPT_CMP_ABS_LESS)
            return fail(ctx, KPT_ERR_ARG, "unknown field predicate comparison.");
        if (!std::isfinite(pred->threshold)) return fail(ctx, KPT_ERR_ARG, "field predicate threshold must be finite.");
        src.use_predicate = 1;
        src.component = pred->component;
        src.comparison = pred->comparison;
        src.threshold = pred->threshold;
    }
    src.max_attempts = max_attempts;
    return KPT_OK;
}

void destroy_candidates(kpt_candidates* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    dev_free(c->cells);
    dev_free(c->cumulative);
    for (int d = 0; d < 2; ++d) {
        dev_free(c->face[d]);
        dev_free(c->bins[d]);
    }
    delete c;
}

// Candidate set over the cells of `src` (all of them without a predicate):
// flags and areas, scans, and private copies of the stretched-axis tables.
kpt_status build_candidates(kpt_ctx* ctx, const kpt::GeneratorSource& src, int field_kind, kpt_candidates** out) {
    const uint64_t cells = static_cast<uint64_t>(src.g.nx) * src.g.ny;
    if (cells > 0x7fffffffull) return fail(ctx, KPT_ERR_ARG, "field has too many cells for a candidate set.");
    (void)hipSetDevice(ctx->device);
    kpt_candidates* c = new kpt_candidates();
    c->device = ctx->device;
    kpt::CandidateBuild w{};
    size_t scan_bytes = 0;
    auto cleanup = [&](kpt_status st) {
        dev_free(w.flags);
        dev_free(w.offsets);
        dev_free(w.area);
        dev_free(w.out_area);
        dev_free(w.temp);
        if (st != KPT_OK) destroy_candidates(c);
        return st;
    };
    hipError_t e = kpt::candidate_scratch_bytes(cells, &scan_bytes);
    if (e == hipSuccess) e = dev_alloc(&w.flags, cells);
    if (e == hipSuccess) e = dev_alloc(&w.offsets, cells);
    if (e == hipSuccess) e = dev_alloc(&w.area, cells);
    if (e == hipSuccess) e = dev_alloc(&w.out_area, cells);
    if (e == hipSuccess) e = dev_alloc(&c->cells, cells);
    if (e == hipSuccess) e = dev_alloc(&c->cumulative, cells);
    if (e == hipSuccess) e = hipMalloc(&w.temp, std::max<size_t>(scan_bytes, 1));
    if (e != hipSuccess) return cleanup(hip_fail(ctx, e, "candidate set buffers"));
    w.temp_bytes = scan_bytes;
    w.out_cells = c->cells;
    kpt::GridDesc g = src.g;
    for (int d = 0; d < 2; ++d) {
        if (!src.g.face[d]) continue;
        const size_t nf = static_cast<size_t>(d == 0 ? g.nx : g.ny) + 1;
        if ((e = dev_alloc(&c->face[d], nf)) != hipSuccess ||
            (e = dev_alloc(&c->bins[d], static_cast<size_t>(g.bins_n[d]))) != hipSuccess ||
            (e = hipMemcpyAsync(c->face[d], src.g.face[d], nf * sizeof(double), hipMemcpyDeviceToDevice,
                                ctx->stream)) != hipSuccess ||
            (e = hipMemcpyAsync(c->bins[d], src.g.bins[d], static_cast<size_t>(g.bins_n[d]) * sizeof(kpt::FaceBin),
                                hipMemcpyDeviceToDevice, ctx->stream)) != hipSuccess)
            return cleanup(hip_fail(ctx, e, "candidate set grid tables"));
        g.face[d] = c->face[d];
        g.bins[d] = c->bins[d];
    }
    kpt::GeneratorSource s2 = src;
    s2.g = g;
    uint64_t count = cells;
    if (field_kind) {
        if ((e = kpt::launch_candidates(s2, 1, w, ctx->stream)) != hipSuccess)
            return cleanup(hip_fail(ctx, e, "candidate flags"));
        uint32_t last[2] = {0, 0};
        if ((e = hipMemcpyAsync(&last[0], w.offsets + cells - 1, 4, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
            (e = hipMemcpyAsync(&last[1], w.flags + cells - 1, 4, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
            (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
            return cleanup(hip_fail(ctx, e, "candidate count"));
        count = static_cast<uint64_t>(last[0]) + last[1];
        if ((e = kpt::launch_cumulative(w.out_area, c->cumulative, count, w.temp, w.temp_bytes, ctx->stream)) !=
            hipSuccess)
            return cleanup(hip_fail(ctx, e, "candidate weights"));
    } else {
        if ((e = kpt::launch_candidates(s2, 0, w, ctx->stream)) != hipSuccess ||
            (e = kpt::launch_cumulative(w.area, c->cumulative, count, w.temp, w.temp_bytes, ctx->stream)) !=
                hipSuccess)
            return cleanup(hip_fail(ctx, e, "grid candidate weights"));
        dev_free(c->cells);
        c->cells = nullptr;
    }
    double total = 0.0;
    if (count > 0 &&
        ((e = hipMemcpyAsync(&total, c->cumulative + count - 1, 8, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess))
        return cleanup(hip_fail(ctx, e, "candidate total"));
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return cleanup(hip_fail(ctx, e, "candidate set"));
    if (count > 0 && (!std::isfinite(total) || total <= 0.0))
        return cleanup(fail(ctx, KPT_ERR_ARG,
                            field_kind ? "Field particle candidate set: total physical volume must be positive."
                                       : "Grid particle candidate set: total physical volume must be positive."));
    c->view.g = g;
    c->view.cells = c->cells;
    c->view.cumulative = c->cumulative;
    c->view.count = count;
    c->view.total = total;
    c->view.kind = field_kind ? 0 : 1;
    *out = c;
    return cleanup(KPT_OK);
}

kpt_status read_counters(kpt_ctx* ctx, unsigned long long out[2]) {
    KPT_HIP(ctx, hipMemcpyAsync(out, ctx->counters, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KPT_OK;
}

}  // namespace

extern "C" {

kpt_status kpt_candidates_from_field(kpt_ctx* ctx, int32_t slot, const kpt_region* region,
                                     const kpt_field_predicate* pred, kpt_candidates** out) {
    if (!ctx || !pred || !out) return KPT_ERR_ARG;
    *out = nullptr;
    kpt::GeneratorSource src;
    const kpt_status st = make_source(ctx, slot, region, pred, 1, src);
    if (st != KPT_OK) return st;
    return build_candidates(ctx, src, 1, out);
}

kpt_status kpt_candidates_from_grid(kpt_ctx* ctx, int32_t slot, kpt_candidates** out) {
    if (!ctx || !out) return KPT_ERR_ARG;
    *out = nullptr;
    kpt::GeneratorSource src;
    const kpt_status st = make_source(ctx, slot, nullptr, nullptr, 1, src);
    if (st != KPT_OK) return st;
    if (src.g.nx <= 0 || src.g.ny <= 0) return fail(ctx, KPT_ERR_ARG, "Grid particle candidate set: grid domain has no cells.");
    return build_candidates(ctx, src, 0, out);
}

kpt_status kpt_candidates_info(const kpt_candidates* c, uint64_t* count, double* total_weight) {
    if (!c) return KPT_ERR_ARG;
    if (count) *count = c->view.count;
    if (total_weight) *total_weight = c->view.total;
    return KPT_OK;
}

kpt_status kpt_candidates_download(kpt_ctx* ctx, const kpt_candidates* c, uint32_t* cells, double* cumulative) {
    if (!ctx || !c) return KPT_ERR_ARG;
    const uint64_t n = c->view.count;
    if (n == 0) return KPT_OK;
    (void)hipSetDevice(c->device);
    if (cells) {
        if (c->cells) {
            KPT_HIP(ctx, hipMemcpy(cells, c->cells, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        } else {
            for (uint64_t k = 0; k < n; ++k) cells[k] = static_cast<uint32_t>(k);
        }
    }
    if (cumulative) KPT_HIP(ctx, hipMemcpy(cumulative, c->cumulative, n * sizeof(double), hipMemcpyDeviceToHost));
    return KPT_OK;
}

kpt_status kpt_candidates_destroy(kpt_candidates* c) {
    destroy_candidates(c);
    return KPT_OK;
}

kpt_status kpt_inject_from_candidates(kpt_ctx* ctx, const kpt_candidates* c, uint64_t n,
                                      const kpt_injection_kinematics* k, const kpt_particle_properties* props,
                                      uint64_t* injected) {
    if (!ctx || !c) return KPT_ERR_ARG;
    if (injected) *injected = 0;
    kpt_status st = check_kinematics(ctx, k);
    if (st != KPT_OK) return st;
    double e0 = 0.0;
    if ((st = injection_energy(ctx, props, k->momentum, e0)) != KPT_OK) return st;
    // validate_particle_injection_plan (include/ParticleGenerator.hpp:846-870)
    if (n > ctx->capacity - ctx->count)
        return fail(ctx, KPT_ERR_ARG, "ParticleInjectionPlan: requested range exceeds capacity.");
    if (c->view.count == 0)
        return fail(ctx, KPT_ERR_STATE, c->view.kind == 0
                                            ? "ParticleGenerator: field-conditioned injection has no candidate grid points."
                                            : "ParticleGenerator: grid injection has no candidate cells.");
    (void)hipSetDevice(ctx->device);
    const uint32_t key[2] = {static_cast<uint32_t>(ctx->seed), static_cast<uint32_t>(ctx->seed >> 32)};
    KPT_HIP(ctx, kpt::launch_inject_candidates(ctx->P, ctx->count, n, ctx->next_root, c->view, *k, e0,
                                               ctx->generator_calls, key, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->generator_calls += 1;
    ctx->count += n;
    ctx->next_root += n;
    if (injected) *injected = n;
    return KPT_OK;
}

kpt_status kpt_inject_uniform_if(kpt_ctx* ctx, int32_t slot, const kpt_region* region,
                                 const kpt_field_predicate* pred, int32_t max_attempts, uint64_t n,
                                 const kpt_injection_kinematics* k, const kpt_particle_properties* props,
                                 uint64_t* injected, uint64_t* rejected) {
    if (!ctx) return KPT_ERR_ARG;
    if (injected) *injected = 0;
    if (rejected) *rejected = 0;
    kpt_status st = check_kinematics(ctx, k);
    if (st != KPT_OK) return st;
    double e0 = 0.0;
    if ((st = injection_energy(ctx, props, k->momentum, e0)) != KPT_OK) return st;
    if (max_attempts <= 0 || max_attempts > 65535)
        return fail(ctx, KPT_ERR_ARG, "ParticleInjectionPlan: max_attempts must be positive (and below 65536).");
    if (n > ctx->capacity - ctx->count)
        return fail(ctx, KPT_ERR_ARG, "ParticleInjectionPlan: requested range exceeds capacity.");
    kpt::GeneratorSource src;
    if ((st = make_source(ctx, slot, region, pred, max_attempts, src)) != KPT_OK) return st;
    (void)hipSetDevice(ctx->device);
    const uint32_t key[2] = {static_cast<uint32_t>(ctx->seed), static_cast<uint32_t>(ctx->seed >> 32)};
    KPT_HIP(ctx, hipMemsetAsync(ctx->counters, 0, 2 * sizeof(unsigned long long), ctx->stream));
    KPT_HIP(ctx, kpt::launch_inject_uniform(ctx->P, ctx->count, n, ctx->next_root, src, *k, e0, ctx->generator_calls,
                                            key, ctx->counters, ctx->stream));
    unsigned long long cnt[2] = {0, 0};
    if ((st = read_counters(ctx, cnt)) != KPT_OK) return st;
    ctx->generator_calls += 1;
    ctx->count += n;
    ctx->next_root += n;
    if (injected) *injected = n;
    if (rejected) *rejected = cnt[0];
    return KPT_OK;
}

kpt_status kpt_recycle_uniform_if(kpt_ctx* ctx, int32_t slot, const kpt_region* region,
                                  const kpt_field_predicate* pred, int32_t max_attempts, int32_t selection,
                                  const kpt_injection_kinematics* k, const kpt_particle_properties* props,
                                  uint64_t* recycled, uint64_t* rejected) {
    if (!ctx) return KPT_ERR_ARG;
    if (recycled) *recycled = 0;
    if (rejected) *rejected = 0;
    kpt_status st = check_kinematics(ctx, k);
    if (st != KPT_OK) return st;
    double e0 = 0.0;
    if ((st = injection_energy(ctx, props, k->momentum, e0)) != KPT_OK) return st;
    if (max_attempts <= 0 || max_attempts > 65535)
        return fail(ctx, KPT_ERR_ARG, "ParticleGenerator: recycle max_attempts must be positive (and below 65536).");
    if (selection < KPT_RECYCLE_ESCAPED || selection > KPT_RECYCLE_ANY)
        return fail(ctx, KPT_ERR_ARG, "unknown recycle selection.");
    kpt::GeneratorSource src;
    if ((st = make_source(ctx, slot, region, pred, max_attempts, src)) != KPT_OK) return st;
    (void)hipSetDevice(ctx->device);
    const uint32_t key[2] = {static_cast<uint32_t>(ctx->seed), static_cast<uint32_t>(ctx->seed >> 32)};
    KPT_HIP(ctx, hipMemsetAsync(ctx->counters, 0, 2 * sizeof(unsigned long long), ctx->stream));
    KPT_HIP(ctx, kpt::launch_recycle(ctx->P, ctx->count, selection, 1, kpt::CandidateView{}, src, *k, e0,
                                     ctx->generator_calls, key, ctx->counters, ctx->stream));
    unsigned long long cnt[2] = {0, 0};
    if ((st = read_counters(ctx, cnt)) != KPT_OK) return st;
    ctx->generator_calls += 1;
    if (recycled) *recycled = cnt[0];
    if (rejected) *rejected = cnt[1];
    return KPT_OK;
}

kpt_status kpt_recycle_from_candidates(kpt_ctx* ctx, const kpt_candidates* c, int32_t selection,
                                       const kpt_injection_kinematics* k, const kpt_particle_properties* props,
                                       uint64_t* recycled) {
    if (!ctx || !c) return KPT_ERR_ARG;
    if (recycled) *recycled = 0;
    kpt_status st = check_kinematics(ctx, k);
    if (st != KPT_OK) return st;
    double e0 = 0.0;
    if ((st = injection_energy(ctx, props, k->momentum, e0)) != KPT_OK) return st;
    if (selection < KPT_RECYCLE_ESCAPED || selection > KPT_RECYCLE_ANY)
        return fail(ctx, KPT_ERR_ARG, "unknown recycle selection.");
    if (c->view.count == 0)
        return fail(ctx, KPT_ERR_STATE, c->view.kind == 0
                                            ? "ParticleGenerator: field-conditioned recycling has no candidate grid points."
                                            : "ParticleGenerator: grid recycling has no candidate cells.");
    (void)hipSetDevice(ctx->device);
    const uint32_t key[2] = {static_cast<uint32_t>(ctx->seed), static_cast<uint32_t>(ctx->seed >> 32)};
    KPT_HIP(ctx, hipMemsetAsync(ctx->counters, 0, 2 * sizeof(unsigned long long), ctx->stream));
    kpt::GeneratorSource none{};
    KPT_HIP(ctx, kpt::launch_recycle(ctx->P, ctx->count, selection, 0, c->view, none, *k, e0, ctx->generator_calls,
                                     key, ctx->counters, ctx->stream));
    unsigned long long cnt[2] = {0, 0};
    if ((st = read_counters(ctx, cnt)) != KPT_OK) return st;
    ctx->generator_calls += 1;
    if (recycled) *recycled = cnt[0];
    return KPT_OK;
}

kpt_status kpt_generator_calls(const kpt_ctx* ctx, uint32_t* calls) {
    if (!ctx || !calls) return KPT_ERR_ARG;
    *calls = ctx->generator_calls;
    return KPT_OK;
}

}  // extern "C"
