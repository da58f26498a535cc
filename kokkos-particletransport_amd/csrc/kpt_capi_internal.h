// kpt_capi_internal.h — the context and field-slot types and the small helpers
// shared by the C ABI's translation units: kpt_capi.hip (context, fields,
// particles, push, diagnostics, solvers), kpt_ensemble_capi.hip (batched
// ensembles) and kpt_generator_capi.hip (the generic particle generator).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "kpt_internal.h"

namespace kpt_capi {

struct FieldSlot {
    bool valid = false;
    int model = 0;
    kpt::GridDesc g{};
    double* raw = nullptr;  // 6 planes of nx*ny (Bx, By, Bz, Vx, Vy, Vz)
    double* c64 = nullptr;
    float* c32 = nullptr;
    // node-major records of the FP32 frame (static-field FP32 pushes), built
    // from c32 when a push first needs them after an upload
    float* rec32 = nullptr;
    size_t rec32_cap = 0;
    uint64_t rec32_version = 0;
    size_t nodes_cap = 0;
    size_t raw_cap = 0;
    uint64_t version = 0;  // changes with every upload into the slot
    // generic Parker basis built from raw (kpt_generic_parker_build)
    double* gen = nullptr;
    size_t gen_cap = 0;
    bool gen_valid = false;
    // generic focused basis built from raw (kpt_generic_focus_build)
    double* foc = nullptr;
    size_t foc_cap = 0;
    bool foc_valid = false;
    // stretched axes: face coordinates and bin table on the device
    double* face_dev[2] = {nullptr, nullptr};
    kpt::FaceBin* bins_dev[2] = {nullptr, nullptr};
    size_t face_cap[2] = {0, 0};
    size_t bins_cap[2] = {0, 0};
    // host copies of the tables, alive until the next fill of this slot (the
    // asynchronous copies read them)
    std::vector<double> host_face[2];
    std::vector<kpt::FaceBin> host_bins[2];
};

// Background load of the next frame (kpt_field_stage_compact): a host thread
// reads and validates the file into pinned memory and copies it to `slot` on
// `stream`, while the caller's pushes run on the context's stream; the advance
// packs it on the context's stream.
struct FieldStage {
    FieldSlot slot;
    std::thread worker;
    bool busy = false;              // a stage was started and not yet finished
    hipStream_t stream = nullptr;   // default priority, like the context's stream
    hipEvent_t ready = nullptr;     // context-stream work queued before the stage
    hipEvent_t done = nullptr;      // the staged frame is packed
    double* pinned = nullptr;       // six planes, page-locked
    size_t pinned_cap = 0;          // doubles
    kpt_status code = KPT_OK;
    std::string err;
    double read_seconds = 0.0;
};

}  // namespace kpt_capi

struct kpt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t capacity = 0;
    uint64_t count = 0;
    uint64_t seed = 0;
    uint64_t next_root = 0;
    uint32_t generator_calls = 0;  // Philox counter word of the generic generator's draws
    kpt::ParticleArrays P{};
    kpt_capi::FieldSlot slot[2];
    kpt_capi::FieldStage stage;
    uint64_t field_versions = 0;
    // FP32 frame-interleaved pair of the lower and upper frames (built by the
    // FP32 time-blended push when either slot changed)
    float* pair32 = nullptr;
    size_t pair32_cap = 0;
    uint64_t pair_versions[2] = {0, 0};
    double* pair64 = nullptr;
    size_t pair64_cap = 0;
    uint64_t pair64_versions[2] = {0, 0};
    // learned closure: FP32 parameters (device and host copies), hull, MFMA image;
    // mlp_version counts installs (ensembles rebuild their run tables on a change)
    bool mlp_valid = false;
    uint64_t mlp_version = 0;
    // FP32 pushes stage their stencil records in LDS (environment KPT_PUSH_STAGE=0
    // at context creation turns it off, for A/B measurements and tests)
    bool push_stage = true;
    kpt_mlp_hull mlp_hull{};
    float* mlp_params = nullptr;
    void* mlp_image = nullptr;
    std::vector<float> mlp_host;
    double* noise = nullptr;
    uint64_t noise_particles = 0;
    uint32_t noise_steps = 0;
    unsigned long long* counters = nullptr;  // 8 x u64
    kpt::PushArgs* d_args = nullptr;         // device copy of the current push launch's arguments
    uint32_t* split_scratch = nullptr;
    size_t split_scratch_cap = 0;
    double* hist_partials = nullptr;
    double* hist_out = nullptr;
    void* summary_partials = nullptr;
    kpt_summary* summary_out = nullptr;
    // cell sort: second particle set (allocated on first sort) and scratch
    kpt::ParticleArrays P2{};
    bool p2_allocated = false;
    kpt::SortScratch sort{};
    std::string err;
    char device_name[64] = {0};
};

namespace kpt_capi {

inline kpt_status fail(kpt_ctx* ctx, kpt_status code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

inline kpt_status hip_fail(kpt_ctx* ctx, hipError_t e, const char* what) {
    return fail(ctx, KPT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define KPT_HIP(ctx, call)                                  \
    do {                                                    \
        const hipError_t e_ = (call);                       \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
    } while (0)

template <typename T>
inline hipError_t dev_alloc(T** p, size_t elems) {
    *p = nullptr;
    if (elems == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void**>(p), elems * sizeof(T));
}

inline void dev_free(void* p) {
    if (p) (void)hipFree(p);
}

// The ParticleSystem arrays of n uploaded particles: previous positions default
// to the uploaded position, sort keys to 0.
inline hipError_t upload_system_arrays(const kpt::ParticleArrays& P, const kpt_particles_host* h, uint64_t n) {
    if (n == 0) return hipSuccess;
    hipError_t e = hipSuccess;
    auto put = [&](void* d, const void* src) {
        if (e == hipSuccess) e = hipMemcpy(d, src, n * 8, hipMemcpyHostToDevice);
    };
    put(P.prev_x, h->prev_x ? h->prev_x : h->x);
    put(P.prev_y, h->prev_y ? h->prev_y : h->y);
    put(P.prev_step_x, h->prev_step_x ? h->prev_step_x : h->x);
    put(P.prev_step_y, h->prev_step_y ? h->prev_step_y : h->y);
    if (e == hipSuccess)
        e = h->sort_key ? hipMemcpy(P.sort_key, h->sort_key, n * 4, hipMemcpyHostToDevice)
                        : hipMemset(P.sort_key, 0, n * 4);
    return e;
}

inline hipError_t download_system_arrays(kpt_particles_host* h, const kpt::ParticleArrays& P, uint64_t n) {
    hipError_t e = hipSuccess;
    auto get = [&](void* d, const void* src, size_t es) {
        if (e == hipSuccess && d && n) e = hipMemcpy(d, src, n * es, hipMemcpyDeviceToHost);
    };
    get(h->prev_x, P.prev_x, 8);
    get(h->prev_y, P.prev_y, 8);
    get(h->prev_step_x, P.prev_step_x, 8);
    get(h->prev_step_y, P.prev_step_y, 8);
    get(h->sort_key, P.sort_key, 4);
    return e;
}

// Helpers defined in kpt_capi.hip.
kpt_status check_settings(kpt_ctx* ctx, const kpt_device_settings* s);
kpt_status check_properties(kpt_ctx* ctx, const kpt_particle_properties* pr);
kpt_status upload_into(kpt_ctx* ctx, FieldSlot& f, int32_t model, int32_t nx, int32_t ny, const double* xe,
                       const double* ye, const double* const planes[6]);
bool faces_uniform(const double* edges, size_t count);
// The slot's node-major FP32 records (rec32), rebuilt when the slot changed.
kpt_status ensure_rec32(kpt_ctx* ctx, FieldSlot& f);
kpt_status push_args(kpt_ctx* ctx, const kpt_device_settings* s, double t0, uint32_t frame, bool paired,
                     kpt::PushArgs& a, bool& focused);

// Defined in kpt_noalloc.hip: looks up, once per device and with that device
// current, the address of the no-allocation push's device scratch array.
hipError_t noalloc_setup(int device);

}  // namespace kpt_capi

using namespace kpt_capi;
