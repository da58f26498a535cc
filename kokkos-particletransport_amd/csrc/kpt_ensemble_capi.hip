// kpt_ensemble_capi.hip — host side of the batched-ensemble C ABI
// (kpt_ensemble_*): run table, tiles and per-XCD queues, the graph-captured
// frame loop, the launch-shape autotuner and the batched cell sort.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "kpt_capi_internal.h"

// ---- batched ensemble -----------------------------------------------------------

struct kpt_ensemble {
    kpt_ctx* ctx = nullptr;
    int runs = 0;
    int cu_count = 256;
    std::vector<uint64_t> cap, off, count, next_root, seed;
    std::vector<kpt_device_settings> s;
    std::vector<char> set;
    std::vector<FieldSlot> field;
    uint64_t total_cap = 0;
    kpt::ParticleArrays P{};
    kpt::ParticleArrays B{};  // autotune backup, allocated on first use
    bool backup_allocated = false;
    kpt::ParticleArrays P2{};  // cell sort target, allocated on first use
    bool p2_allocated = false;
    kpt::SortScratch sort{};       // sized for all slots (one batched sort)
    kpt::SortRun* d_sort_runs = nullptr;
    kpt::EnsembleRun* d_runs = nullptr;
    unsigned long long* d_counters = nullptr;  // 4 per run
    uint2* d_tiles = nullptr;
    size_t tiles_cap = 0;
    uint32_t* d_gstart = nullptr;  // 9 entries
    unsigned int* d_queues = nullptr;  // 8 entries
    kpt::EnsembleLoop* d_loop = nullptr;
    bool dirty = true;       // run table or tiles need a rebuild
    uint64_t mlp_version = 0;  // the context's network install the run table holds
    bool has_mlp = false;
    uint32_t n_tiles = 0;
    kpt_ensemble_tuning tune{0, 8, 0.0};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // kernel arguments the captured graph was recorded with
    const void* g_tiles = nullptr;
    int g_groups = 0, g_grid = 0, g_mlp = -1;
};

namespace {

using kpt::alloc_arrays;
using kpt::arrays_at;
using kpt::copy_arrays;
using kpt::free_arrays;

kpt_status ens_run_ok(kpt_ensemble* e, int32_t run) {
    if (!e) return KPT_ERR_ARG;
    if (run < 0 || run >= e->runs) return fail(e->ctx, KPT_ERR_ARG, "ensemble run index out of range.");
    return KPT_OK;
}

void drop_graph(kpt_ensemble* e) {
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->graph_exec = nullptr;
    e->graph = nullptr;
}

// Run table and tile lists from the runs' fields, settings and counts.
kpt_status ens_sync(kpt_ensemble* e) {
    kpt_ctx* ctx = e->ctx;
    // the table copies the network's activation, feature scaling and on/off flag:
    // a kpt_mlp_set or kpt_mlp_train since the last build invalidates it
    if (!e->dirty && e->mlp_version == ctx->mlp_version) return KPT_OK;
    std::vector<kpt::EnsembleRun> table(static_cast<size_t>(e->runs));
    e->has_mlp = false;
    for (int r = 0; r < e->runs; ++r) {
        const std::string tag = "ensemble run " + std::to_string(r) + ": ";
        if (!e->set[r]) return fail(ctx, KPT_ERR_STATE, tag + "settings are not set.");
        const FieldSlot& f = e->field[r];
        if (!f.valid) return fail(ctx, KPT_ERR_STATE, tag + "field is not loaded.");
        const kpt_device_settings& s = e->s[r];
        kpt::EnsembleRun& R = table[r];
        kpt::PushArgs& a = R.a;
        if (f.model == KPT_MODEL_PARKER16) {
            if (s.focused_transport != 0) return fail(ctx, KPT_ERR_ARG, tag + "focused transport needs a FOCUSED21 field.");
            R.kind = kpt::kEnsParker32;
        } else {
            if (s.focused_transport == 0 || (s.closure != KPT_CLOSURE_QLT && s.closure != KPT_CLOSURE_MLP))
                return fail(ctx, KPT_ERR_ARG,
                            tag + "a FOCUSED21 run must be focused transport with the QLT or learned closure.");
            R.kind = s.closure == KPT_CLOSURE_MLP ? kpt::kEnsFocusedMlp32 : kpt::kEnsFocusedQlt32;
            if (s.closure == KPT_CLOSURE_MLP && s.mlp_passes == 3)
                return fail(ctx, KPT_ERR_ARG,
                            tag + "the batched ensemble evaluates the network once per sub-step (mlp_passes 0 or 1).");
        }
        a.P = arrays_at(e->P, e->off[r]);
        a.count = e->count[r];
        a.field_lo = a.field_hi = f.c32;
        // one frame per run: its node-major single-frame records (read through L1/L2;
        // LDS staging of them is compiled out by default, KPT_ENSEMBLE_STAGE_BYTES)
        a.field_pair = nullptr;
        a.field_pair64 = nullptr;
        {
            const kpt_status rs = ensure_rec32(ctx, e->field[r]);
            if (rs != KPT_OK) return rs;
        }
        a.field_rec = f.rec32;
        a.stage = ctx->push_stage ? 1 : 0;
        a.field_lo64 = a.field_hi64 = f.c64;
        a.plane = kpt::coefficient_plane(f.g.nx, f.g.ny);
        a.g = f.g;
        a.s = s;
        a.t0 = a.t_end = a.t_stop = 0.0;  // set on the device per interval
        a.first_chunk = 1;
        a.key[0] = static_cast<uint32_t>(e->seed[r]);
        a.key[1] = static_cast<uint32_t>(e->seed[r] >> 32);
        a.noise = nullptr;
        a.noise_steps = 0;
        a.qlt_norm = s.qlt_spectral_index * std::sin(M_PI / s.qlt_spectral_index) / M_PI;
        a.k64 = kpt::make_step_const<double>(s, a.qlt_norm);
        a.k32 = kpt::make_step_const<float>(s, a.qlt_norm);
        a.mlp_image = nullptr;
        a.mlp_on = 0;
        a.mlp_act = kpt::kActSilu;
        if (R.kind == kpt::kEnsFocusedMlp32) {
            if (!ctx->mlp_valid)
                return fail(ctx, KPT_ERR_STATE, tag + "no pitch-angle network is loaded (kpt_mlp_train / kpt_mlp_set).");
            const kpt_mlp_hull& h = ctx->mlp_hull;
            a.mlp_image = ctx->mlp_image;
            a.mlp_scale = kpt::mlp_scale(h);
            a.mlp_act = h.activation;
            a.mlp_on = (s.qlt_spectral_index >= h.gamma_lo && s.qlt_spectral_index <= h.gamma_hi && s.qlt_h0 == h.h0)
                           ? 1 : 0;
        }
        if (R.kind == kpt::kEnsFocusedMlp32) e->has_mlp = true;
        a.counters = e->d_counters + 4 * static_cast<size_t>(r);
        a.self = &e->d_runs[r].a;  // the run table is the arguments' device copy
        R.pad = 0;
    }
    // tiles of 256 particles; groups: one queue, or one per XCD with the runs
    // assigned longest-first to the least-loaded queue
    const int G = e->tune.groups;
    std::vector<std::vector<uint2>> per(G);
    std::vector<int> order(e->runs);
    for (int r = 0; r < e->runs; ++r) order[r] = r;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return e->count[a] > e->count[b]; });
    std::vector<uint64_t> load(G, 0);
    for (int r : order) {
        const uint64_t nt = (e->count[r] + 255) / 256;
        if (nt == 0) continue;
        int gmin = 0;
        for (int g = 1; g < G; ++g)
            if (load[g] < load[gmin]) gmin = g;
        load[gmin] += nt;
        for (uint64_t k = 0; k < nt; ++k) per[gmin].push_back(make_uint2(static_cast<uint32_t>(r), static_cast<uint32_t>(256 * k)));
    }
    std::vector<uint2> tiles;
    std::vector<uint32_t> gstart(9, 0);
    for (int g = 0; g < G; ++g) {
        gstart[g] = static_cast<uint32_t>(tiles.size());
        tiles.insert(tiles.end(), per[g].begin(), per[g].end());
    }
    for (int g = G; g <= 8; ++g) gstart[g] = static_cast<uint32_t>(tiles.size());
    e->n_tiles = static_cast<uint32_t>(tiles.size());
    if (tiles.size() > e->tiles_cap) {
        drop_graph(e);
        dev_free(e->d_tiles);
        e->d_tiles = nullptr;
        e->tiles_cap = 0;
        KPT_HIP(ctx, dev_alloc(&e->d_tiles, std::max<size_t>(tiles.size(), 1)));
        e->tiles_cap = std::max<size_t>(tiles.size(), 1);
    }
    if (!tiles.empty())
        KPT_HIP(ctx, hipMemcpy(e->d_tiles, tiles.data(), tiles.size() * sizeof(uint2), hipMemcpyHostToDevice));
    KPT_HIP(ctx, hipMemcpy(e->d_gstart, gstart.data(), 9 * sizeof(uint32_t), hipMemcpyHostToDevice));
    KPT_HIP(ctx, hipMemcpy(e->d_runs, table.data(), table.size() * sizeof(kpt::EnsembleRun), hipMemcpyHostToDevice));
    e->dirty = false;
    e->mlp_version = ctx->mlp_version;
    return KPT_OK;
}

int ens_grid(const kpt_ensemble* e) { return e->tune.grid_blocks > 0 ? e->tune.grid_blocks : 2 * e->cu_count; }

// One interval of every run: advance the run table, then the persistent push.
hipError_t ens_launch(kpt_ensemble* e) {
    kpt_ctx* ctx = e->ctx;
    hipError_t st = kpt::launch_ensemble_advance(e->d_runs, e->runs, e->d_loop, e->d_queues, e->tune.groups, ctx->stream);
    if (st != hipSuccess) return st;
    return kpt::launch_ensemble(e->d_runs, e->d_tiles, e->d_gstart, e->tune.groups, e->d_queues, ens_grid(e),
                                ctx->mlp_image, e->has_mlp ? 1 : 0, ctx->stream);
}

kpt_status ens_set_loop(kpt_ensemble* e, uint32_t first_frame, double window) {
    kpt::EnsembleLoop L{first_frame, 0u, window};
    KPT_HIP(e->ctx, hipMemcpy(e->d_loop, &L, sizeof(L), hipMemcpyHostToDevice));
    return KPT_OK;
}

}  // namespace

extern "C" {

kpt_status kpt_ensemble_create(kpt_ctx* ctx, int32_t runs, const uint64_t* capacities, kpt_ensemble** out) {
    if (!ctx || !out || !capacities) return KPT_ERR_ARG;
    *out = nullptr;
    if (runs <= 0 || runs > 65536) return fail(ctx, KPT_ERR_ARG, "ensemble run count must be in [1, 65536].");
    (void)hipSetDevice(ctx->device);
    kpt_ensemble* e = new kpt_ensemble();
    e->ctx = ctx;
    e->runs = runs;
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, ctx->device) == hipSuccess && cu > 0)
        e->cu_count = cu;
    e->cap.assign(capacities, capacities + runs);
    e->off.resize(runs);
    e->count.assign(runs, 0);
    e->next_root.assign(runs, 0);
    e->seed.assign(runs, 0);
    e->s.resize(runs);
    e->set.assign(runs, 0);
    e->field.resize(runs);
    uint64_t total = 0;
    for (int r = 0; r < runs; ++r) {
        if (e->cap[r] > (1ull << 32)) {
            delete e;
            return fail(ctx, KPT_ERR_ARG, "ensemble run capacity must be at most 2^32.");
        }
        e->off[r] = total;
        total += e->cap[r];
    }
    e->total_cap = total;
    hipError_t er = alloc_arrays(e->P, total);
    if (er == hipSuccess) er = dev_alloc(&e->d_runs, static_cast<size_t>(runs));
    if (er == hipSuccess) er = dev_alloc(&e->d_counters, 4 * static_cast<size_t>(runs));
    if (er == hipSuccess) er = dev_alloc(&e->d_gstart, 9);
    if (er == hipSuccess) er = dev_alloc(&e->d_queues, 8);
    if (er == hipSuccess) er = dev_alloc(&e->d_loop, 1);
    if (er == hipSuccess && total) er = hipMemset(e->P.e0, 0, total * 8);
    if (er == hipSuccess && total) er = hipMemset(e->P.substep, 0, total * 4);
    if (er == hipSuccess && total) er = hipMemset(e->P.sort_key, 0, total * 4);
    if (er != hipSuccess) {
        kpt_ensemble_destroy(e);
        return hip_fail(ctx, er, "ensemble allocation");
    }
    *out = e;
    return KPT_OK;
}

kpt_status kpt_ensemble_destroy(kpt_ensemble* e) {
    if (!e) return KPT_OK;
    (void)hipSetDevice(e->ctx->device);
    (void)hipStreamSynchronize(e->ctx->stream);
    drop_graph(e);
    free_arrays(e->P);
    if (e->backup_allocated) free_arrays(e->B);
    if (e->p2_allocated) free_arrays(e->P2);
    for (int k = 0; k < 2; ++k) {
        dev_free(e->sort.keys[k]);
        dev_free(e->sort.vals[k]);
    }
    dev_free(e->sort.temp);
    dev_free(e->d_sort_runs);
    for (FieldSlot& f : e->field) {
        dev_free(f.raw);
        dev_free(f.c64);
        dev_free(f.c32);
        dev_free(f.rec32);
        dev_free(f.gen);
        dev_free(f.foc);
    }
    for (void* q : {static_cast<void*>(e->d_runs), static_cast<void*>(e->d_counters), static_cast<void*>(e->d_tiles),
                    static_cast<void*>(e->d_gstart), static_cast<void*>(e->d_queues), static_cast<void*>(e->d_loop)})
        dev_free(q);
    delete e;
    return KPT_OK;
}

kpt_status kpt_ensemble_field_upload(kpt_ensemble* e, int32_t run, int32_t model, int32_t nx, int32_t ny,
                                     const double* xe, const double* ye, const double* bx, const double* by,
                                     const double* bz, const double* vx, const double* vy, const double* vz) {
    kpt_status st = ens_run_ok(e, run);
    if (st != KPT_OK) return st;
    kpt_ctx* ctx = e->ctx;
    if (!xe || !ye || !bx || !by || !bz || !vx || !vy || !vz) return fail(ctx, KPT_ERR_ARG, "field array pointer is null.");
    const size_t cells = static_cast<size_t>(std::max(nx, 0)) * std::max(ny, 0);
    for (const double* plane : {bx, by, bz, vx, vy, vz})
        for (size_t k = 0; k < cells; ++k)
            if (!std::isfinite(plane[k])) return fail(ctx, KPT_ERR_ARG, "AthenaFieldReader: non-finite field value.");
    for (const double* ed : {xe, ye}) {
        const int n = ed == xe ? nx : ny;
        if (n > 1 && !faces_uniform(ed, static_cast<size_t>(n) + 1))
            return fail(ctx, KPT_ERR_ARG, "ensemble runs use uniform grids.");
    }
    const double* planes[6] = {bx, by, bz, vx, vy, vz};
    (void)hipSetDevice(ctx->device);
    e->dirty = true;
    return upload_into(ctx, e->field[run], model, nx, ny, xe, ye, planes);
}

kpt_status kpt_ensemble_set_run(kpt_ensemble* e, int32_t run, const kpt_device_settings* s, uint64_t seed) {
    kpt_status st = ens_run_ok(e, run);
    if (st != KPT_OK) return st;
    st = check_settings(e->ctx, s);
    if (st != KPT_OK) return st;
    if (s->precision != KPT_PREC_FP32) return fail(e->ctx, KPT_ERR_ARG, "ensemble runs use the FP32 push.");
    e->s[run] = *s;
    e->s[run].time_interpolation = 0;  // one frame per run
    e->s[run].push_chunks = 1;
    e->seed[run] = seed;
    e->set[run] = 1;
    e->dirty = true;
    return KPT_OK;
}

kpt_status kpt_ensemble_particles_upload(kpt_ensemble* e, int32_t run, const kpt_particles_host* h, uint64_t count) {
    kpt_status st = ens_run_ok(e, run);
    if (st != KPT_OK) return st;
    kpt_ctx* ctx = e->ctx;
    if (!h) return KPT_ERR_ARG;
    if (count > e->cap[run]) return fail(ctx, KPT_ERR_CAPACITY, "particle count exceeds the run's capacity.");
    if (count > 0 && (!h->x || !h->y || !h->p || !h->t || !h->step_dt || !h->weight || !h->mu || !h->status ||
                      !h->split_level))
        return fail(ctx, KPT_ERR_ARG, "particle array pointer is null.");
    (void)hipSetDevice(ctx->device);
    const kpt::ParticleArrays P = arrays_at(e->P, e->off[run]);
    uint64_t max_root = 0;
    if (count > 0) {
        KPT_HIP(ctx, hipMemcpy(P.x, h->x, count * 8, hipMemcpyHostToDevice));
        KPT_HIP(ctx, hipMemcpy(P.y, h->y, count * 8, hipMemcpyHostToDevice));
        KPT_HIP(ctx, hipMemcpy(P.p, h->p, count * 8, hipMemcpyHostToDevice));
        KPT_HIP(ctx, hipMemcpy(P.t, h->t, count * 8, hipMemcpyHostToDevice));
        KPT_HIP(ctx, hipMemcpy(P.step_dt, h->step_dt, count * 8, hipMemcpyHostToDevice));
        KPT_HIP(ctx, hipMemcpy(P.weight, h->weight, count * 8, hipMemcpyHostToDevice));
        KPT_HIP(ctx, hipMemcpy(P.mu, h->mu, count * 8, hipMemcpyHostToDevice));
        KPT_HIP(ctx, hipMemcpy(P.status, h->status, count * 4, hipMemcpyHostToDevice));
        KPT_HIP(ctx, hipMemcpy(P.split_level, h->split_level, count * 4, hipMemcpyHostToDevice));
        if (h->e0)
            KPT_HIP(ctx, hipMemcpy(P.e0, h->e0, count * 8, hipMemcpyHostToDevice));
        else
            KPT_HIP(ctx, hipMemset(P.e0, 0, count * 8));
        KPT_HIP(ctx, upload_system_arrays(P, h, count));
        std::vector<uint64_t> ids(count);
        for (uint64_t i = 0; i < count; ++i) ids[i] = h->id ? h->id[i] : i;
        for (uint64_t i = 0; i < count; ++i) max_root = std::max<uint64_t>(max_root, (ids[i] & 0xFFFFFFFFull) + 1);
        KPT_HIP(ctx, hipMemcpy(P.id, ids.data(), count * 8, hipMemcpyHostToDevice));
    }
    e->count[run] = count;
    e->next_root[run] = max_root;
    e->dirty = true;
    return KPT_OK;
}

kpt_status kpt_ensemble_particles_download(kpt_ensemble* e, int32_t run, kpt_particles_host* h, uint64_t count) {
    kpt_status st = ens_run_ok(e, run);
    if (st != KPT_OK) return st;
    kpt_ctx* ctx = e->ctx;
    if (!h) return KPT_ERR_ARG;
    if (count > e->count[run]) return fail(ctx, KPT_ERR_ARG, "download count exceeds stored particles.");
    if (count == 0) return KPT_OK;
    (void)hipSetDevice(ctx->device);
    const kpt::ParticleArrays P = arrays_at(e->P, e->off[run]);
    if (h->x) KPT_HIP(ctx, hipMemcpy(h->x, P.x, count * 8, hipMemcpyDeviceToHost));
    if (h->y) KPT_HIP(ctx, hipMemcpy(h->y, P.y, count * 8, hipMemcpyDeviceToHost));
    if (h->p) KPT_HIP(ctx, hipMemcpy(h->p, P.p, count * 8, hipMemcpyDeviceToHost));
    if (h->t) KPT_HIP(ctx, hipMemcpy(h->t, P.t, count * 8, hipMemcpyDeviceToHost));
    if (h->step_dt) KPT_HIP(ctx, hipMemcpy(h->step_dt, P.step_dt, count * 8, hipMemcpyDeviceToHost));
    if (h->weight) KPT_HIP(ctx, hipMemcpy(h->weight, P.weight, count * 8, hipMemcpyDeviceToHost));
    if (h->mu) KPT_HIP(ctx, hipMemcpy(h->mu, P.mu, count * 8, hipMemcpyDeviceToHost));
    if (h->status) KPT_HIP(ctx, hipMemcpy(h->status, P.status, count * 4, hipMemcpyDeviceToHost));
    if (h->split_level) KPT_HIP(ctx, hipMemcpy(h->split_level, P.split_level, count * 4, hipMemcpyDeviceToHost));
    if (h->id) KPT_HIP(ctx, hipMemcpy(h->id, P.id, count * 8, hipMemcpyDeviceToHost));
    if (h->e0) KPT_HIP(ctx, hipMemcpy(h->e0, P.e0, count * 8, hipMemcpyDeviceToHost));
    KPT_HIP(ctx, download_system_arrays(h, P, count));
    return KPT_OK;
}

kpt_status kpt_ensemble_count(kpt_ensemble* e, int32_t run, uint64_t* count) {
    kpt_status st = ens_run_ok(e, run);
    if (st != KPT_OK) return st;
    if (!count) return KPT_ERR_ARG;
    *count = e->count[run];
    return KPT_OK;
}

kpt_status kpt_ensemble_inject(kpt_ensemble* e, int32_t run, uint64_t n, double t0, uint32_t frame,
                               uint64_t* injected) {
    kpt_status st = ens_run_ok(e, run);
    if (st != KPT_OK) return st;
    kpt_ctx* ctx = e->ctx;
    if (!e->set[run]) return fail(ctx, KPT_ERR_STATE, "ensemble run settings are not set.");
    if (injected) *injected = 0;
    if (e->count[run] >= e->cap[run]) return KPT_OK;
    const uint64_t m = std::min(n, e->cap[run] - e->count[run]);
    (void)hipSetDevice(ctx->device);
    const uint32_t key[2] = {static_cast<uint32_t>(e->seed[run]), static_cast<uint32_t>(e->seed[run] >> 32)};
    const kpt_inject_opts o{};
    KPT_HIP(ctx, kpt::launch_inject(arrays_at(e->P, e->off[run]), e->count[run], m, e->next_root[run], e->s[run], o, t0,
                                    frame, key, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    e->count[run] += m;
    e->next_root[run] += m;
    e->dirty = true;
    if (injected) *injected = m;
    return KPT_OK;
}

kpt_status kpt_ensemble_set_tuning(kpt_ensemble* e, const kpt_ensemble_tuning* t) {
    if (!e || !t) return KPT_ERR_ARG;
    if (t->groups != 1 && t->groups != 8) return fail(e->ctx, KPT_ERR_ARG, "ensemble queue groups must be 1 or 8.");
    if (t->grid_blocks < 0 || t->grid_blocks > (1 << 20))
        return fail(e->ctx, KPT_ERR_ARG, "ensemble grid must be in [0, 2^20] workgroups.");
    if (t->groups != e->tune.groups) e->dirty = true;
    e->tune.grid_blocks = t->grid_blocks;
    e->tune.groups = t->groups;
    e->tune.ms = t->ms;
    return KPT_OK;
}

kpt_status kpt_ensemble_push(kpt_ensemble* e, uint32_t first_frame, uint32_t frames, int32_t use_graph,
                             kpt_ensemble_run_stats* per_run, double* kernel_ms) {
    if (!e) return KPT_ERR_ARG;
    kpt_ctx* ctx = e->ctx;
    (void)hipSetDevice(ctx->device);
    kpt_status st = ens_sync(e);
    if (st != KPT_OK) return st;
    KPT_HIP(ctx, hipMemsetAsync(e->d_counters, 0, 4 * sizeof(unsigned long long) * e->runs, ctx->stream));
    st = ens_set_loop(e, first_frame, 1.0);
    if (st != KPT_OK) return st;
    if (use_graph && frames > 0) {
        const int grid = ens_grid(e), mlp = e->has_mlp ? 1 : 0;
        if (!e->graph_exec || e->g_tiles != e->d_tiles || e->g_groups != e->tune.groups || e->g_grid != grid ||
            e->g_mlp != mlp) {
            drop_graph(e);
            KPT_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            const hipError_t le = ens_launch(e);
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
            if (le != hipSuccess || ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return hip_fail(ctx, le != hipSuccess ? le : ce, "ensemble graph capture");
            }
            e->graph = g;
            KPT_HIP(ctx, hipGraphInstantiate(&e->graph_exec, e->graph, nullptr, nullptr, 0));
            e->g_tiles = e->d_tiles;
            e->g_groups = e->tune.groups;
            e->g_grid = grid;
            e->g_mlp = mlp;
        }
    }
    KPT_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (uint32_t k = 0; k < frames; ++k) {
        if (use_graph)
            KPT_HIP(ctx, hipGraphLaunch(e->graph_exec, ctx->stream));
        else
            KPT_HIP(ctx, ens_launch(e));
    }
    KPT_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<unsigned long long> c(4 * static_cast<size_t>(e->runs));
    KPT_HIP(ctx, hipMemcpyAsync(c.data(), e->d_counters, c.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    KPT_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (kernel_ms) *kernel_ms = ms;
    if (per_run)
        for (int r = 0; r < e->runs; ++r) {
            per_run[r].particle_steps = c[4 * r];
            per_run[r].failures = c[4 * r + 1];
            per_run[r].active_in = c[4 * r + 2];
        }
    return KPT_OK;
}

kpt_status kpt_ensemble_autotune(kpt_ensemble* e, uint32_t frame, double window, kpt_ensemble_tuning* chosen,
                                 kpt_ensemble_tuning* table, int32_t table_len, int32_t* tried) {
    if (!e) return KPT_ERR_ARG;
    kpt_ctx* ctx = e->ctx;
    if (!(window > 0.0 && window <= 1.0)) return fail(ctx, KPT_ERR_ARG, "autotune window must be in (0, 1].");
    (void)hipSetDevice(ctx->device);
    kpt_status st = ens_sync(e);
    if (st != KPT_OK) return st;
    if (!e->backup_allocated) {
        KPT_HIP(ctx, alloc_arrays(e->B, e->total_cap));
        e->backup_allocated = true;
    }
    KPT_HIP(ctx, copy_arrays(e->B, e->P, e->total_cap, ctx->stream));
    const kpt_ensemble_tuning saved = e->tune;
    std::vector<kpt_ensemble_tuning> cands;
    for (int groups : {8, 1})
        for (int mult : {1, 2, 3, 4}) cands.push_back(kpt_ensemble_tuning{mult * e->cu_count, groups, 0.0});
    int best = -1;
    for (size_t c = 0; c <= cands.size(); ++c) {
        // candidate 0 runs twice: the first pass warms the caches and is not kept
        const size_t ci = c == 0 ? 0 : c - 1;
        st = kpt_ensemble_set_tuning(e, &cands[ci]);
        if (st == KPT_OK) st = ens_sync(e);
        if (st == KPT_OK) st = ens_set_loop(e, frame, window);
        if (st != KPT_OK) break;
        KPT_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        KPT_HIP(ctx, ens_launch(e));
        KPT_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
        float ms = 0.0f;
        KPT_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        KPT_HIP(ctx, copy_arrays(e->P, e->B, e->total_cap, ctx->stream));
        if (c == 0) continue;
        cands[ci].ms = ms;
        if (best < 0 || ms < cands[best].ms) best = static_cast<int>(ci);
    }
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (st != KPT_OK) {
        (void)kpt_ensemble_set_tuning(e, &saved);
        return st;
    }
    st = kpt_ensemble_set_tuning(e, &cands[best]);
    if (st != KPT_OK) return st;
    if (chosen) *chosen = cands[best];
    const int n = std::min<int>(static_cast<int>(cands.size()), std::max(table_len, 0));
    for (int k = 0; table && k < n; ++k) table[k] = cands[k];
    if (tried) *tried = static_cast<int32_t>(cands.size());
    return KPT_OK;
}

kpt_status kpt_ensemble_histogram_logp(kpt_ensemble* e, int32_t run, int32_t bins, double pmin, double pmax,
                                       double* out) {
    kpt_status st = ens_run_ok(e, run);
    if (st != KPT_OK) return st;
    kpt_ctx* ctx = e->ctx;
    if (!out) return KPT_ERR_ARG;
    if (bins <= 0 || bins > 4096) return fail(ctx, KPT_ERR_ARG, "histogram bins must be in [1, 4096].");
    if (!(pmin > 0.0) || !(pmax > pmin) || !std::isfinite(pmax))
        return fail(ctx, KPT_ERR_ARG, "histogram range must satisfy 0 < pmin < pmax.");
    const double log_pmin = std::log(pmin);
    const double inv_dlogp = static_cast<double>(bins) / (std::log(pmax) - log_pmin);
    (void)hipSetDevice(ctx->device);
    KPT_HIP(ctx, kpt::launch_hist_logp(arrays_at(e->P, e->off[run]), e->count[run], bins, log_pmin, inv_dlogp, pmin,
                                       pmax, ctx->hist_partials, ctx->hist_out, ctx->stream));
    KPT_HIP(ctx, hipMemcpyAsync(out, ctx->hist_out, bins * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KPT_OK;
}

}  // extern "C"

extern "C" kpt_status kpt_ensemble_sort(kpt_ensemble* e, uint64_t* active) {
    if (!e) return KPT_ERR_ARG;
    kpt_ctx* ctx = e->ctx;
    (void)hipSetDevice(ctx->device);
    for (int r = 0; r < e->runs; ++r) {
        if (!e->field[r].valid) return fail(ctx, KPT_ERR_STATE, "ensemble run " + std::to_string(r) + ": field is not loaded.");
        if (!e->set[r]) return fail(ctx, KPT_ERR_STATE, "ensemble run " + std::to_string(r) + ": settings are not set.");
    }
    if (!e->p2_allocated) {
        const uint64_t n = std::max<uint64_t>(e->total_cap, 1);
        KPT_HIP(ctx, alloc_arrays(e->P2, e->total_cap));
        for (int k = 0; k < 2; ++k) {
            KPT_HIP(ctx, dev_alloc(&e->sort.keys[k], n));
            KPT_HIP(ctx, dev_alloc(&e->sort.vals[k], n));
        }
        e->sort.temp_bytes = kpt::sort_temp_bytes(n);
        KPT_HIP(ctx, hipMalloc(&e->sort.temp, std::max<size_t>(e->sort.temp_bytes, 1)));
        KPT_HIP(ctx, dev_alloc(&e->d_sort_runs, static_cast<size_t>(e->runs)));
        e->p2_allocated = true;
    }
    unsigned long long* act_dev = ctx->counters + 4;
    KPT_HIP(ctx, hipMemsetAsync(act_dev, 0, sizeof(unsigned long long), ctx->stream));
    // One batched sort when (run, node key) fits 32 bits: key bits of the largest
    // grid plus the two classes after its nodes, above them the run index.
    uint32_t limit = 0;
    for (int r = 0; r < e->runs; ++r) limit = std::max(limit, kpt::sort_key_limit(e->field[r].g));
    uint32_t run_shift = 1;
    while (run_shift < 32 && (1ull << run_shift) < static_cast<uint64_t>(limit) + 2) ++run_shift;
    uint32_t run_bits = 0;
    while ((1ull << run_bits) < static_cast<uint64_t>(e->runs)) ++run_bits;
    if (run_shift + run_bits <= 32 && e->total_cap <= 0x7fffffffull) {
        std::vector<kpt::SortRun> sr(static_cast<size_t>(e->runs));
        for (int r = 0; r < e->runs; ++r) {
            sr[r].g = e->field[r].g;
            sr[r].shift_x = e->s[r].interpolation_shift_x;
            sr[r].shift_y = e->s[r].interpolation_shift_y;
            sr[r].off = e->off[r];
            sr[r].count = e->count[r];
        }
        KPT_HIP(ctx, hipMemcpyAsync(e->d_sort_runs, sr.data(), sr.size() * sizeof(kpt::SortRun),
                                    hipMemcpyHostToDevice, ctx->stream));
        KPT_HIP(ctx, kpt::launch_ensemble_sort(e->d_sort_runs, e->runs, e->total_cap, run_shift, e->P, e->P2,
                                               e->sort, act_dev, ctx->stream));
        KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));  // `sr` is read by the copy
    } else {
        for (int r = 0; r < e->runs; ++r) {
            if (e->count[r] == 0) continue;
            const kpt_device_settings& s = e->s[r];
            KPT_HIP(ctx, kpt::launch_sort(arrays_at(e->P, e->off[r]), arrays_at(e->P2, e->off[r]), e->count[r],
                                          e->field[r].g, s.interpolation_shift_x, s.interpolation_shift_y,
                                          std::numeric_limits<double>::infinity(), e->sort, act_dev, ctx->stream));
        }
    }
    unsigned long long act = 0;
    KPT_HIP(ctx, hipMemcpyAsync(&act, act_dev, sizeof(act), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::swap(e->P, e->P2);
    e->dirty = true;  // the run table points into the other particle set now
    if (active) *active = act;
    return KPT_OK;
}

