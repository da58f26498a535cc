// kpt_capi.hip — host side of the C ABI declared in include/kpt.h.
//
// Owns the device buffers (particles in SoA, two coefficient frames in FP64 and
// FP32, noise table, scratch for split / histogram / summary), validates
// arguments the way the reference driver does (src/main.cpp:187-240,
// include/AthenaFieldReader.hpp:86-270) and launches the kernels of
// kpt_kernels.hip on the context's stream. Every call is synchronous. The
// batched ensemble and the generic particle generator have their own files
// (kpt_ensemble_capi.hip, kpt_generator_capi.hip); kpt_capi_internal.h holds
// the context type they share.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <chrono>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "kpt_capi_internal.h"

namespace {

thread_local std::string g_create_error;

}  // namespace

// The spacing test of validate_uniform_face_edges (include/AthenaFieldReader.hpp:
// 167-193): all intervals equal the mean interval to 1024 ulp of the scale.
bool kpt_capi::faces_uniform(const double* edges, size_t count) {
    const double span = edges[count - 1] - edges[0];
    const double step = span / static_cast<double>(count - 1);
    for (size_t i = 1; i < count; ++i) {
        const double cur = edges[i] - edges[i - 1];
        const double mag = std::max({std::fabs(span), std::fabs(step), std::fabs(cur), 1.0});
        if (std::fabs(cur - step) > 1024.0 * std::numeric_limits<double>::epsilon() * mag) return false;
    }
    return true;
}

namespace {


// A stretched axis (non-uniform faces) spans exactly [e_0, e_n]; a uniform one
// keeps the analytic bounds start + k step.
kpt::GridDesc make_grid_desc(int32_t nx, int32_t ny, const double* xe, const double* ye) {
    kpt::GridDesc g{};
    g.nx = nx;
    g.ny = ny;
    const double* e[2] = {xe, ye};
    const int32_t n[2] = {nx, ny};
    for (int d = 0; d < 2; ++d) {
        const bool stretched = n[d] > 1 && !faces_uniform(e[d], static_cast<size_t>(n[d]) + 1);
        g.start[d] = e[d][0];
        g.step[d] = (e[d][n[d]] - e[d][0]) / static_cast<double>(n[d]);
        g.inv_step[d] = 1.0 / g.step[d];
        g.phys_lo[d] = stretched ? e[d][0] : g.start[d] + 0.0 * g.step[d];
        g.phys_hi[d] = stretched ? e[d][n[d]] : g.start[d] + static_cast<double>(n[d]) * g.step[d];
        g.width[d] = g.phys_hi[d] - g.phys_lo[d];
        g.inv_width[d] = g.width[d] != 0.0 ? 1.0 / g.width[d] : 0.0;
        g.face[d] = nullptr;
        g.bins[d] = nullptr;
        g.bins_n[d] = 0;
        g.bins_scale[d] = 0.0;
    }
    return g;
}

// Bin table of a stretched axis: m equal bins over [e_0, e_n). Bin b records the
// interval k that contains its lower end less a guard band g and the faces e_k,
// e_{k+1}, e_{k+2}. m starts at the axis length over the narrowest interval and
// doubles until every bin, widened by g on both sides, lies in [e_k, e_{k+2})
// (in [e_k, e_n) for the last two intervals, as queries stay below e_n):
// g (1e-9 of the axis) exceeds by far the rounding of the device's bin index, so
// a query always finds its interval among k and k + 1. Fails past 2^22 bins.
bool stretched_table(const double* e, int n, std::vector<kpt::FaceBin>& bins, double& scale) {
    const double w = e[n] - e[0];
    const double g = 1e-9 * w;
    double hmin = w;
    for (int k = 0; k < n; ++k) hmin = std::min(hmin, e[k + 1] - e[k]);
    const auto face = [&](int k) { return e[std::min(k, n)]; };
    for (double m0 = std::max(static_cast<double>(n), std::ceil(w / hmin)); m0 <= static_cast<double>(1 << 22);
         m0 *= 2.0) {
        const int m = static_cast<int>(m0);
        bins.assign(static_cast<size_t>(m), kpt::FaceBin{});
        bool ok = true;
        for (int b = 0; b < m && ok; ++b) {
            const double lo = e[0] + w * static_cast<double>(b) / m - g;
            const double hi = e[0] + w * static_cast<double>(b + 1) / m + g;
            int k = static_cast<int>(std::upper_bound(e, e + n + 1, lo) - e) - 1;
            k = std::min(std::max(k, 0), n - 1);
            bins[b] = kpt::FaceBin{face(k), face(k + 1), face(k + 2), k, 0};
            ok = (k == 0 || e[k] <= lo) && (k + 2 >= n || hi < e[k + 2]);  // queries lie in [e_0, e_n)
        }
        if (ok) {
            scale = static_cast<double>(m) / w;
            return true;
        }
    }
    return false;
}

// validate_face_edges / validate_uniform_face_edges
// (include/AthenaFieldReader.hpp:147-193).
bool check_edges(const double* edges, size_t count, const char* axis, bool require_uniform,
                 std::string& err) {
    if (count < 2) {
        err = std::string("AthenaFieldReader: ") + axis + " requires at least two face coordinates.";
        return false;
    }
    for (size_t i = 0; i < count; ++i) {
        if (!std::isfinite(edges[i])) {
            err = std::string("AthenaFieldReader: non-finite ") + axis + " face coordinate.";
            return false;
        }
        if (i > 0 && !(edges[i] > edges[i - 1])) {
            err = std::string("AthenaFieldReader: ") + axis + " face coordinates must be strictly ascending.";
            return false;
        }
    }
    if (!require_uniform) return true;
    const double intervals = static_cast<double>(count - 1);
    const double span = edges[count - 1] - edges[0];
    const double step = span / intervals;
    if (!std::isfinite(step) || step <= 0.0) {
        err = std::string("AthenaFieldReader: ") + axis + " face spacing must be positive and finite.";
        return false;
    }
    for (size_t i = 1; i < count; ++i) {
        const double cur = edges[i] - edges[i - 1];
        const double mag = std::max({std::fabs(span), std::fabs(step), std::fabs(cur), 1.0});
        const double tol = 1024.0 * std::numeric_limits<double>::epsilon() * mag;
        if (std::fabs(cur - step) > tol) {
            err = std::string("AthenaFieldReader: ") + axis +
                  " face coordinates must be uniformly spaced for analytic grid loading.";
            return false;
        }
    }
    return true;
}

}  // namespace

kpt_status kpt_capi::check_settings(kpt_ctx* ctx, const kpt_device_settings* s) {
    if (!s) return fail(ctx, KPT_ERR_ARG, "settings pointer is null.");
    const bool ok = std::isfinite(s->mhd_dt) && s->mhd_dt > 0.0 && std::isfinite(s->p0) && s->p0 > 0.0 &&
                    std::isfinite(s->dt_min) && s->dt_min > 0.0 && std::isfinite(s->dt_max) &&
                    s->dt_max >= s->dt_min && std::isfinite(s->x_width) && s->x_width > 0.0 &&
                    std::isfinite(s->y_width) && s->y_width > 0.0 && std::isfinite(s->dx) && s->dx > 0.0 &&
                    std::isfinite(s->dy) && s->dy > 0.0 && std::isfinite(s->charge) && s->charge != 0.0 &&
                    std::isfinite(s->kpara0) && s->kpara0 > 0.0 && std::isfinite(s->kperp_over_kpara) &&
                    s->kperp_over_kpara >= 0.0 && s->max_particle_steps_per_interval > 0 &&
                    (s->precision == KPT_PREC_FP64 || s->precision == KPT_PREC_FP32) &&
                    s->push_chunks >= 0 && s->push_chunks <= 4096 && std::isfinite(s->fp64_null_b) &&
                    s->fp64_null_b >= 0.0;
    if (!ok) return fail(ctx, KPT_ERR_ARG, "Reconnection2DSettings: invalid physical parameters.");
    if (s->closure != KPT_CLOSURE_REDUCED) {
        // PitchAngleScatteringParameters::validate (LegencyModel.hpp:207-222) plus
        // the positive sigma2 and Lc that evaluate_pitch_angle_scattering requires
        const bool qok = (s->closure == KPT_CLOSURE_QLT || s->closure == KPT_CLOSURE_MLP) &&
                         std::isfinite(s->qlt_spectral_index) && s->qlt_spectral_index > 1.0 &&
                         std::isfinite(s->qlt_h0) && s->qlt_h0 > 0.0 && std::isfinite(s->qlt_sigma2) &&
                         s->qlt_sigma2 > 0.0 && std::isfinite(s->qlt_correlation_length) &&
                         s->qlt_correlation_length > 0.0 && s->mu_max > 0.0 && s->mu_max <= 1.0;
        if (!qok)
            return fail(ctx, KPT_ERR_ARG,
                        "LegencyModel::PitchAngleScatteringParameters: coefficients and floors must be finite, "
                        "with spectral_index > 1, resonance_regularization_h0 > 0, sigma2 > 0, Lc > 0 and "
                        "0 < mu_max <= 1.");
        if (s->closure == KPT_CLOSURE_MLP && s->mlp_passes != 0 && s->mlp_passes != 1 && s->mlp_passes != 3)
            return fail(ctx, KPT_ERR_ARG,
                        "mlp_passes must be 0 or 1 (one network evaluation per sub-step) or 3 (one per Picard "
                        "pass).");
    }
    return KPT_OK;
}

namespace {


kpt_status sort_in_place(kpt_ctx* ctx, const kpt_device_settings& s, double t_done, bool compact,
                         uint64_t* active);

// Device buffers of a slot for an nx x ny field of `model` components; reused
// when large enough, so a run of equal frames allocates once.
hipError_t alloc_field_buffers(FieldSlot& f, int32_t nx, int32_t ny, int model) {
    const size_t raw = static_cast<size_t>(nx) * ny * 6;
    const size_t nodes = kpt::coefficient_plane(nx, ny) * static_cast<size_t>(model);
    hipError_t e = hipSuccess;
    if (raw > f.raw_cap) {
        dev_free(f.raw);
        f.raw = nullptr;
        f.raw_cap = 0;
        if ((e = dev_alloc(&f.raw, raw)) != hipSuccess) return e;
        f.raw_cap = raw;
    }
    if (nodes > f.nodes_cap) {
        dev_free(f.c64);
        dev_free(f.c32);
        f.c64 = nullptr;
        f.c32 = nullptr;
        f.nodes_cap = 0;
        if ((e = dev_alloc(&f.c64, nodes)) != hipSuccess) return e;
        if ((e = dev_alloc(&f.c32, nodes)) != hipSuccess) return e;
        f.nodes_cap = nodes;
    }
    return hipSuccess;
}

// Compact-field parser shared by kpt_compact_field_read, kpt_field_load_compact
// and the staging thread. `planes(cells)` returns room for the six planes of
// cells values each (nullptr: out of memory).
bool read_compact_into(const char* path, int32_t& nx, int32_t& ny, std::vector<double>& xe,
                       std::vector<double>& ye, const std::function<double*(size_t)>& planes,
                       bool allow_nonuniform, std::string& err) {
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        err = "AthenaFieldReader: failed to open Athena binary file.";
        return false;
    }
    in.read(reinterpret_cast<char*>(&nx), 4);
    if (!in) {
        err = "AthenaFieldReader: failed to read nx.";
        return false;
    }
    in.read(reinterpret_cast<char*>(&ny), 4);
    if (!in) {
        err = "AthenaFieldReader: failed to read ny.";
        return false;
    }
    if (nx <= 0 || ny <= 0) {
        err = "AthenaFieldReader: Athena binary dimensions must be positive.";
        return false;
    }
    const size_t cells = static_cast<size_t>(nx) * static_cast<size_t>(ny);
    xe.resize(static_cast<size_t>(nx) + 1);
    ye.resize(static_cast<size_t>(ny) + 1);
    in.read(reinterpret_cast<char*>(xe.data()), static_cast<std::streamsize>(xe.size() * 8));
    if (!in) {
        err = "AthenaFieldReader: failed to read x_edges.";
        return false;
    }
    in.read(reinterpret_cast<char*>(ye.data()), static_cast<std::streamsize>(ye.size() * 8));
    if (!in) {
        err = "AthenaFieldReader: failed to read y_edges.";
        return false;
    }
    if (!check_edges(xe.data(), xe.size(), "x", !allow_nonuniform, err)) return false;
    if (!check_edges(ye.data(), ye.size(), "y", !allow_nonuniform, err)) return false;
    double* comps = planes(cells);
    if (comps == nullptr) {
        err = "AthenaFieldReader: out of host memory for the field planes.";
        return false;
    }
    static const char* names[6] = {"Bx", "By", "Bz", "Vx", "Vy", "Vz"};
    for (int c = 0; c < 6; ++c) {
        double* dst = comps + static_cast<size_t>(c) * cells;
        in.read(reinterpret_cast<char*>(dst), static_cast<std::streamsize>(cells * 8));
        if (!in) {
            err = std::string("AthenaFieldReader: failed to read ") + names[c] + ".";
            return false;
        }
        for (size_t k = 0; k < cells; ++k) {
            if (!std::isfinite(dst[k])) {
                err = std::string("AthenaFieldReader: non-finite ") + names[c] + " value.";
                return false;
            }
        }
    }
    if (in.peek() != std::char_traits<char>::eof()) {
        err = "AthenaFieldReader: Athena binary file has trailing payload.";
        return false;
    }
    return true;
}

bool read_compact(const char* path, int32_t& nx, int32_t& ny, std::vector<double>& xe,
                  std::vector<double>& ye, std::vector<double>& comps, bool allow_nonuniform,
                  std::string& err) {
    return read_compact_into(
        path, nx, ny, xe, ye,
        [&](size_t cells) {
            comps.resize(cells * 6);
            return comps.data();
        },
        allow_nonuniform, err);
}

// Raw planes, stretched-axis tables and (with `pack`) the packed FP64 / FP32
// frames of one field into `f`, by work queued on `stream` (not waited for).
// Touches no context state, so the staging thread uses it too. On failure
// returns the status and sets `err`.
kpt_status fill_slot(FieldSlot& f, int32_t model, int32_t nx, int32_t ny, const double* xe, const double* ye,
                     const double* const planes[6], hipStream_t stream, bool pack, std::string& err) {
    if (model != KPT_MODEL_PARKER16 && model != KPT_MODEL_FOCUSED21) {
        err = "invalid coefficient model (16 or 21).";
        return KPT_ERR_ARG;
    }
    if (nx <= 0 || ny <= 0) {
        err = "AthenaFieldReader: Athena binary dimensions must be positive.";
        return KPT_ERR_ARG;
    }
    if (!check_edges(xe, static_cast<size_t>(nx) + 1, "x", false, err) ||
        !check_edges(ye, static_cast<size_t>(ny) + 1, "y", false, err))
        return KPT_ERR_ARG;
    f.valid = false;
    const auto hip_err = [&](hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return KPT_ERR_DEVICE;
    };
    hipError_t e = alloc_field_buffers(f, nx, ny, model);
    if (e != hipSuccess) return hip_err(e, "field buffers");
    const size_t cells = static_cast<size_t>(nx) * ny;
    for (int c = 0; c < 6; ++c) {
        e = hipMemcpyAsync(f.raw + c * cells, planes[c], cells * sizeof(double), hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) return hip_err(e, "field upload");
    }
    f.g = make_grid_desc(nx, ny, xe, ye);
    const double* faces[2] = {xe, ye};
    const int32_t n[2] = {nx, ny};
    for (int d = 0; d < 2; ++d) {
        if (n[d] <= 1 || faces_uniform(faces[d], static_cast<size_t>(n[d]) + 1)) continue;
        std::vector<kpt::FaceBin>& bins = f.host_bins[d];
        double scale = 0.0;
        if (!stretched_table(faces[d], n[d], bins, scale)) {
            err = std::string(d == 0 ? "x" : "y") +
                  " faces are too strongly graded: the narrowest interval is below 2^-22 of the axis.";
            return KPT_ERR_ARG;
        }
        const size_t nf = static_cast<size_t>(n[d]) + 1;
        if (nf > f.face_cap[d]) {
            dev_free(f.face_dev[d]);
            f.face_dev[d] = nullptr;
            f.face_cap[d] = 0;
            if ((e = dev_alloc(&f.face_dev[d], nf)) != hipSuccess) return hip_err(e, "face table");
            f.face_cap[d] = nf;
        }
        if (bins.size() > f.bins_cap[d]) {
            dev_free(f.bins_dev[d]);
            f.bins_dev[d] = nullptr;
            f.bins_cap[d] = 0;
            if ((e = dev_alloc(&f.bins_dev[d], bins.size())) != hipSuccess) return hip_err(e, "bin table");
            f.bins_cap[d] = bins.size();
        }
        f.host_face[d].assign(faces[d], faces[d] + nf);
        if ((e = hipMemcpyAsync(f.face_dev[d], f.host_face[d].data(), nf * sizeof(double), hipMemcpyHostToDevice,
                                stream)) != hipSuccess ||
            (e = hipMemcpyAsync(f.bins_dev[d], bins.data(), bins.size() * sizeof(kpt::FaceBin),
                                hipMemcpyHostToDevice, stream)) != hipSuccess)
            return hip_err(e, "stretched-axis tables");
        f.g.face[d] = f.face_dev[d];
        f.g.bins[d] = f.bins_dev[d];
        f.g.bins_n[d] = static_cast<int32_t>(bins.size());
        f.g.bins_scale[d] = scale;
    }
    f.model = model;
    if (pack && (e = kpt::launch_pack(f.raw, f.g, model, f.c64, f.c32, stream)) != hipSuccess)
        return hip_err(e, "pack");
    f.gen_valid = false;
    f.foc_valid = false;
    return KPT_OK;
}

}  // namespace

// fill_slot on the context's stream, waited for: the slot is ready on return.
kpt_status kpt_capi::upload_into(kpt_ctx* ctx, FieldSlot& f, int32_t model, int32_t nx, int32_t ny, const double* xe,
                       const double* ye, const double* const planes[6]) {
    std::string err;
    const kpt_status st = fill_slot(f, model, nx, ny, xe, ye, planes, ctx->stream, true, err);
    if (st != KPT_OK) return fail(ctx, st, err);
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    f.version = ++ctx->field_versions;
    f.valid = true;
    return KPT_OK;
}

namespace {


kpt_status upload_field(kpt_ctx* ctx, int32_t slot, int32_t model, int32_t nx, int32_t ny,
                        const double* xe, const double* ye, const double* const planes[6]) {
    if (slot != KPT_SLOT_LOWER && slot != KPT_SLOT_UPPER) return fail(ctx, KPT_ERR_ARG, "invalid field slot.");
    return upload_into(ctx, ctx->slot[slot], model, nx, ny, xe, ye, planes);
}

// Staging thread: read and validate the file into pinned memory, then copy it
// into the staging slot on the staging stream. Reports through the
// stage's code and message only (the caller's thread owns the context).
void stage_run(kpt_ctx* ctx, std::string path, int32_t model, bool allow_stretched) {
    FieldStage& S = ctx->stage;
    const auto t0 = std::chrono::steady_clock::now();
    hipError_t e = hipSetDevice(ctx->device);
    // the previous stage's copies out of the pinned buffer have finished
    if (e == hipSuccess) e = hipEventSynchronize(S.done);
    if (e != hipSuccess) {
        S.code = KPT_ERR_DEVICE;
        S.err = std::string("field stage: ") + hipGetErrorString(e);
        return;
    }
    int32_t nx = 0, ny = 0;
    std::vector<double> xe, ye;
    std::string err;
    const bool ok = read_compact_into(
        path.c_str(), nx, ny, xe, ye,
        [&](size_t cells) -> double* {
            if (cells * 6 > S.pinned_cap) {
                if (S.pinned) (void)hipHostFree(S.pinned);
                S.pinned = nullptr;
                S.pinned_cap = 0;
                if (hipHostMalloc(reinterpret_cast<void**>(&S.pinned), cells * 6 * sizeof(double)) != hipSuccess) {
                    S.pinned = nullptr;
                    return nullptr;
                }
                S.pinned_cap = cells * 6;
            }
            return S.pinned;
        },
        allow_stretched, err);
    S.read_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!ok) {
        S.code = KPT_ERR_IO;
        S.err = err;
        return;
    }
    // the slot's buffers held the frame that work queued before the stage reads
    if ((e = hipStreamWaitEvent(S.stream, S.ready, 0)) != hipSuccess) {
        S.code = KPT_ERR_DEVICE;
        S.err = std::string("field stage: ") + hipGetErrorString(e);
        return;
    }
    const size_t cells = static_cast<size_t>(nx) * ny;
    const double* planes[6];
    for (int c = 0; c < 6; ++c) planes[c] = S.pinned + c * cells;
    // the pack waits for the advance: run beside a push, its 130 MB of frame
    // writes (1024 x 512, 21 components) evict the push's frames from the caches
    S.code = fill_slot(S.slot, model, nx, ny, xe.data(), ye.data(), planes, S.stream, false, err);
    if (S.code != KPT_OK) {
        S.err = err;
        return;
    }
    if ((e = hipEventRecord(S.done, S.stream)) != hipSuccess) {
        S.code = KPT_ERR_DEVICE;
        S.err = std::string("field stage: ") + hipGetErrorString(e);
    }
}

}  // namespace

extern "C" {

const char* kpt_version(void) { return "kpt-mi355x 0.1 (gfx950)"; }

const char* kpt_last_create_error(void) { return g_create_error.c_str(); }

kpt_status kpt_create(kpt_ctx** out, int device, uint64_t capacity, uint64_t seed) {
    g_create_error.clear();
    if (!out) {
        g_create_error = "output pointer is null.";
        return KPT_ERR_ARG;
    }
    *out = nullptr;
    if (capacity == 0) {
        g_create_error = "Reconnection2DSettings: particle counts and capacity must be positive.";
        return KPT_ERR_ARG;
    }
    if (capacity >= (1ull << 32)) {
        g_create_error = "capacity must be below 2^32 particles.";
        return KPT_ERR_ARG;
    }
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        g_create_error = "no HIP device is available (this library runs only on gfx950).";
        return KPT_ERR_DEVICE;
    }
    if (device < 0 || device >= n) {
        g_create_error = "device index out of range.";
        return KPT_ERR_ARG;
    }
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) {
        g_create_error = std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
        return KPT_ERR_DEVICE;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        g_create_error = std::string("device architecture ") + prop.gcnArchName + " is not gfx950.";
        return KPT_ERR_DEVICE;
    }
    if ((e = hipSetDevice(device)) != hipSuccess) {
        g_create_error = std::string("hipSetDevice: ") + hipGetErrorString(e);
        return KPT_ERR_DEVICE;
    }
    kpt_ctx* ctx = new kpt_ctx();
    ctx->device = device;
    ctx->capacity = capacity;
    ctx->seed = seed;
    if (const char* st = std::getenv("KPT_PUSH_STAGE")) ctx->push_stage = std::strcmp(st, "0") != 0;
    std::snprintf(ctx->device_name, sizeof(ctx->device_name), "%s", prop.gcnArchName);
    auto bail = [&](hipError_t err, const char* what) {
        g_create_error = std::string(what) + ": " + hipGetErrorString(err);
        kpt_destroy(ctx);
        return KPT_ERR_DEVICE;
    };
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess) return bail(e, "hipStreamCreate");
    if ((e = hipEventCreate(&ctx->ev0)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreate(&ctx->ev1)) != hipSuccess) return bail(e, "hipEventCreate");
    kpt::ParticleArrays& P = ctx->P;
    if ((e = kpt::alloc_arrays(P, capacity)) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = dev_alloc(&ctx->counters, 8)) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = dev_alloc(&ctx->d_args, 1)) != hipSuccess) return bail(e, "hipMalloc");
    // the no-allocation push's scratch address on this device; a failure here is
    // reported by that entry point, not by every context
    (void)noalloc_setup(device);
    ctx->split_scratch_cap = kpt::split_scratch_elems(capacity);
    if ((e = dev_alloc(&ctx->split_scratch, ctx->split_scratch_cap)) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = dev_alloc(&ctx->hist_partials, kpt::hist_partial_elems(4096))) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = dev_alloc(&ctx->hist_out, 4096)) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = hipMalloc(&ctx->summary_partials, kpt::summary_partial_bytes())) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = dev_alloc(&ctx->summary_out, 1)) != hipSuccess) return bail(e, "hipMalloc");
    // ReconnectionParticleStorage: status = 2, split level = 0, weight = 0
    // (src/main.cpp:329-331).
    std::vector<int32_t> two(static_cast<size_t>(std::min<uint64_t>(capacity, 1u << 20)), 2);
    for (uint64_t off = 0; off < capacity; off += two.size()) {
        const size_t n2 = static_cast<size_t>(std::min<uint64_t>(two.size(), capacity - off));
        if ((e = hipMemcpy(P.status + off, two.data(), n2 * 4, hipMemcpyHostToDevice)) != hipSuccess)
            return bail(e, "hipMemcpy");
    }
    if ((e = hipMemset(P.split_level, 0, capacity * 4)) != hipSuccess) return bail(e, "hipMemset");
    if ((e = hipMemset(P.weight, 0, capacity * 8)) != hipSuccess) return bail(e, "hipMemset");
    if ((e = hipMemset(P.e0, 0, capacity * 8)) != hipSuccess) return bail(e, "hipMemset");
    if ((e = hipMemset(P.sort_key, 0, capacity * 4)) != hipSuccess) return bail(e, "hipMemset");
    *out = ctx;
    return KPT_OK;
}

kpt_status kpt_destroy(kpt_ctx* ctx) {
    if (!ctx) return KPT_OK;
    (void)hipSetDevice(ctx->device);
    FieldStage& S = ctx->stage;
    if (S.busy) S.worker.join();
    if (S.stream) (void)hipStreamSynchronize(S.stream);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    kpt::free_arrays(ctx->P);
    for (void* p : {static_cast<void*>(ctx->counters), static_cast<void*>(ctx->d_args),
                    static_cast<void*>(ctx->split_scratch), static_cast<void*>(ctx->hist_partials),
                    static_cast<void*>(ctx->hist_out), ctx->summary_partials,
                    static_cast<void*>(ctx->summary_out), static_cast<void*>(ctx->noise)})
        dev_free(p);
    for (FieldSlot* fp : {&ctx->slot[0], &ctx->slot[1], &S.slot}) {
        FieldSlot& f = *fp;
        dev_free(f.raw);
        dev_free(f.c64);
        dev_free(f.c32);
        dev_free(f.rec32);
        dev_free(f.gen);
        dev_free(f.foc);
        for (int d = 0; d < 2; ++d) {
            dev_free(f.face_dev[d]);
            dev_free(f.bins_dev[d]);
        }
    }
    dev_free(ctx->pair32);
    dev_free(ctx->pair64);
    dev_free(ctx->mlp_params);
    dev_free(ctx->mlp_image);
    if (ctx->p2_allocated) kpt::free_arrays(ctx->P2);
    for (int k = 0; k < 2; ++k) {
        dev_free(ctx->sort.keys[k]);
        dev_free(ctx->sort.vals[k]);
    }
    dev_free(ctx->sort.temp);
    if (S.pinned) (void)hipHostFree(S.pinned);
    if (S.ready) (void)hipEventDestroy(S.ready);
    if (S.done) (void)hipEventDestroy(S.done);
    if (S.stream) (void)hipStreamDestroy(S.stream);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return KPT_OK;
}

const char* kpt_last_error(const kpt_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

kpt_status kpt_device_name(kpt_ctx* ctx, char* buf, size_t len) {
    if (!ctx || !buf || len == 0) return KPT_ERR_ARG;
    std::snprintf(buf, len, "%s", ctx->device_name);
    return KPT_OK;
}

kpt_status kpt_default_host_settings(kpt_host_settings* o) {
    if (!o) return KPT_ERR_ARG;
    o->max_particle_steps_per_interval = 200000;
    o->momentum_dependency = 1;
    o->magnetic_dependency = 1;
    o->split_particles = 1;
    o->time_interpolation = 1;
    o->focused_transport = 0;
    o->mhd_dt = 0.1;
    o->p0 = 0.1;
    o->pmin = 1.0e-2;
    o->pmax = 1.0e1;
    o->gamma_turb = 1.6666667;
    o->kpara0 = 0.00743592;
    o->kperp_over_kpara = 0.01;
    o->dt_min_rel = 1.0e-7;
    o->dt_max_rel = 1.0e-2;
    o->drift_param1 = 850964.408;
    o->drift_param2 = 13575468.975;
    o->charge = -1.0;
    o->particle_v0 = 17.20195;
    o->duu0 = 5578.445;
    o->split_ratio = 2.0;
    o->pmin_split_over_p0 = 2.0;
    o->mu_max = 0.99;
    return KPT_OK;
}

kpt_status kpt_make_device_settings(const kpt_host_settings* h, int32_t nx, int32_t ny, const double* xe,
                                    const double* ye, kpt_device_settings* o) {
    if (!h || !o || !xe || !ye || nx <= 0 || ny <= 0) return KPT_ERR_ARG;
    const kpt::GridDesc g = make_grid_desc(nx, ny, xe, ye);
    std::memset(o, 0, sizeof(*o));
    o->max_particle_steps_per_interval = h->max_particle_steps_per_interval;
    o->momentum_dependency = h->momentum_dependency;
    o->magnetic_dependency = h->magnetic_dependency;
    o->split_particles = h->split_particles ? 1 : 0;
    o->time_interpolation = h->time_interpolation ? 1 : 0;
    o->focused_transport = h->focused_transport ? 1 : 0;
    o->x_lower = g.phys_lo[0];
    o->x_upper = g.phys_hi[0];
    o->y_lower = g.phys_lo[1];
    o->y_upper = g.phys_hi[1];
    o->x_width = o->x_upper - o->x_lower;
    o->y_width = o->y_upper - o->y_lower;
    o->dx = o->x_width / static_cast<double>(nx);
    o->dy = o->y_width / static_cast<double>(ny);
    // the driver's node convention: a value sits half a cell before the
    // cell-centred sample point; on a stretched axis values sit at the faces
    o->interpolation_shift_x = nx > 1 && !faces_uniform(xe, static_cast<size_t>(nx) + 1) ? 0.0 : 0.5 * o->dx;
    o->interpolation_shift_y = ny > 1 && !faces_uniform(ye, static_cast<size_t>(ny) + 1) ? 0.0 : 0.5 * o->dy;
    o->mhd_dt = h->mhd_dt;
    o->p0 = h->p0;
    o->pmax = h->pmax;
    o->gamma_turb = h->gamma_turb;
    o->pindex = 3.0 - h->gamma_turb;
    o->kpara0 = h->kpara0;
    o->kperp_over_kpara = h->kperp_over_kpara;
    o->dt_min = h->dt_min_rel * h->mhd_dt;
    o->dt_max = h->dt_max_rel * h->mhd_dt;
    o->drift_param1 = h->drift_param1;
    o->drift_param2 = h->drift_param2;
    o->charge = h->charge;
    o->particle_v0 = h->particle_v0;
    o->duu0 = h->duu0;
    o->pitch_angle_scattering_h0 = 0.2;  // member default, src/main.cpp:279
    o->split_ratio = h->split_ratio;
    o->pmin_split = h->pmin_split_over_p0 * h->p0;
    o->mu_max = h->mu_max;
    o->precision = KPT_PREC_FP64;
    o->push_chunks = 1;
    o->closure = KPT_CLOSURE_REDUCED;
    o->qlt_sigma2 = 0.0;
    o->qlt_correlation_length = 0.0;
    o->qlt_spectral_index = 5.0 / 3.0;  // PitchAngleScatteringParameters (LegencyModel.hpp:196-198)
    o->qlt_h0 = 1.0e-6;
    o->fp64_null_b = 0.02;
    return KPT_OK;
}

kpt_status kpt_set_seed(kpt_ctx* ctx, uint64_t seed) {
    if (!ctx) return KPT_ERR_ARG;
    ctx->seed = seed;
    return KPT_OK;
}

// ---- fields ----------------------------------------------------------------

kpt_status kpt_field_upload(kpt_ctx* ctx, int32_t slot, int32_t model, int32_t nx, int32_t ny,
                            const double* xe, const double* ye, const double* bx, const double* by,
                            const double* bz, const double* vx, const double* vy, const double* vz) {
    if (!ctx) return KPT_ERR_ARG;
    if (!xe || !ye || !bx || !by || !bz || !vx || !vy || !vz)
        return fail(ctx, KPT_ERR_ARG, "field array pointer is null.");
    for (const double* plane : {bx, by, bz, vx, vy, vz}) {
        const size_t cells = static_cast<size_t>(std::max(nx, 0)) * std::max(ny, 0);
        for (size_t k = 0; k < cells; ++k)
            if (!std::isfinite(plane[k])) return fail(ctx, KPT_ERR_ARG, "AthenaFieldReader: non-finite field value.");
    }
    const double* planes[6] = {bx, by, bz, vx, vy, vz};
    (void)hipSetDevice(ctx->device);
    return upload_field(ctx, slot, model, nx, ny, xe, ye, planes);
}

kpt_status kpt_field_load_compact(kpt_ctx* ctx, const char* path, int32_t slot, int32_t model) {
    return kpt_field_load_compact_ex(ctx, path, slot, model, 0);
}

kpt_status kpt_field_load_compact_ex(kpt_ctx* ctx, const char* path, int32_t slot, int32_t model,
                                     int32_t allow_stretched) {
    if (!ctx || !path) return KPT_ERR_ARG;
    int32_t nx = 0, ny = 0;
    std::vector<double> xe, ye, comps;
    std::string err;
    if (!read_compact(path, nx, ny, xe, ye, comps, allow_stretched != 0, err)) return fail(ctx, KPT_ERR_IO, err);
    const size_t cells = static_cast<size_t>(nx) * ny;
    const double* planes[6];
    for (int c = 0; c < 6; ++c) planes[c] = comps.data() + c * cells;
    (void)hipSetDevice(ctx->device);
    return upload_field(ctx, slot, model, nx, ny, xe.data(), ye.data(), planes);
}

kpt_status kpt_field_advance(kpt_ctx* ctx) {
    if (!ctx) return KPT_ERR_ARG;
    if (!ctx->slot[KPT_SLOT_UPPER].valid) return fail(ctx, KPT_ERR_STATE, "upper field slot is empty.");
    std::swap(ctx->slot[KPT_SLOT_LOWER], ctx->slot[KPT_SLOT_UPPER]);
    ctx->slot[KPT_SLOT_UPPER].valid = false;
    return KPT_OK;
}

kpt_status kpt_field_stage_compact(kpt_ctx* ctx, const char* path, int32_t model, int32_t allow_stretched) {
    if (!ctx || !path) return KPT_ERR_ARG;
    if (model != KPT_MODEL_PARKER16 && model != KPT_MODEL_FOCUSED21)
        return fail(ctx, KPT_ERR_ARG, "invalid coefficient model (16 or 21).");
    FieldStage& S = ctx->stage;
    if (S.busy)
        return fail(ctx, KPT_ERR_STATE, "a field stage is in flight: finish it with kpt_field_advance_staged.");
    (void)hipSetDevice(ctx->device);
    if (!S.stream) {
        KPT_HIP(ctx, hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
        KPT_HIP(ctx, hipEventCreateWithFlags(&S.ready, hipEventDisableTiming));
        KPT_HIP(ctx, hipEventCreateWithFlags(&S.done, hipEventDisableTiming));
    }
    KPT_HIP(ctx, hipEventRecord(S.ready, ctx->stream));
    S.code = KPT_OK;
    S.err.clear();
    S.read_seconds = 0.0;
    S.slot.valid = false;
    try {
        S.worker = std::thread(stage_run, ctx, std::string(path), model, allow_stretched != 0);
    } catch (const std::exception& ex) {
        return fail(ctx, KPT_ERR_STATE, std::string("field stage thread: ") + ex.what());
    }
    S.busy = true;
    return KPT_OK;
}

kpt_status kpt_field_advance_staged(kpt_ctx* ctx, double* read_seconds) {
    if (!ctx) return KPT_ERR_ARG;
    FieldStage& S = ctx->stage;
    if (!S.busy) return fail(ctx, KPT_ERR_STATE, "no field stage in flight.");
    S.worker.join();
    S.busy = false;
    if (read_seconds) *read_seconds = S.read_seconds;
    if (S.code != KPT_OK) return fail(ctx, S.code, S.err);
    if (!ctx->slot[KPT_SLOT_UPPER].valid) return fail(ctx, KPT_ERR_STATE, "upper field slot is empty.");
    (void)hipSetDevice(ctx->device);
    KPT_HIP(ctx, hipStreamWaitEvent(ctx->stream, S.done, 0));
    KPT_HIP(ctx, kpt::launch_pack(S.slot.raw, S.slot.g, S.slot.model, S.slot.c64, S.slot.c32, ctx->stream));
    std::swap(ctx->slot[KPT_SLOT_LOWER], ctx->slot[KPT_SLOT_UPPER]);
    std::swap(ctx->slot[KPT_SLOT_UPPER], S.slot);
    FieldSlot& up = ctx->slot[KPT_SLOT_UPPER];
    up.version = ++ctx->field_versions;
    up.valid = true;
    S.slot.valid = false;
    return KPT_OK;
}

kpt_status kpt_field_download_coefficients(kpt_ctx* ctx, int32_t slot, double* out, size_t out_len) {
    if (!ctx || !out || (slot != 0 && slot != 1)) return KPT_ERR_ARG;
    const FieldSlot& f = ctx->slot[slot];
    if (!f.valid) return fail(ctx, KPT_ERR_STATE, "field slot is empty.");
    const size_t nodes = kpt::coefficient_plane(f.g.nx, f.g.ny);
    if (out_len < nodes * f.model) return fail(ctx, KPT_ERR_ARG, "output buffer too small.");
    std::vector<double> tmp(nodes * f.model);
    (void)hipSetDevice(ctx->device);
    KPT_HIP(ctx, hipMemcpy(tmp.data(), f.c64, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t n = 0; n < nodes; ++n)
        for (int c = 0; c < f.model; ++c) out[n * f.model + c] = tmp[c * nodes + n];
    return KPT_OK;
}

kpt_status kpt_compact_field_read(const char* path, int32_t* nx, int32_t* ny, double* x_edges, double* y_edges,
                                  double* bx, double* by, double* bz, double* vx, double* vy, double* vz,
                                  int32_t allow_nonuniform, char* err, size_t err_len) {
    int32_t n0 = 0, n1 = 0;
    std::vector<double> xe, ye, comps;
    std::string e;
    if (!path || !read_compact(path, n0, n1, xe, ye, comps, allow_nonuniform != 0, e)) {
        if (err && err_len) std::snprintf(err, err_len, "%s", path ? e.c_str() : "path is null.");
        return path ? KPT_ERR_IO : KPT_ERR_ARG;
    }
    if (nx) *nx = n0;
    if (ny) *ny = n1;
    const size_t cells = static_cast<size_t>(n0) * n1;
    if (x_edges) std::memcpy(x_edges, xe.data(), xe.size() * 8);
    if (y_edges) std::memcpy(y_edges, ye.data(), ye.size() * 8);
    double* outs[6] = {bx, by, bz, vx, vy, vz};
    for (int c = 0; c < 6; ++c)
        if (outs[c]) std::memcpy(outs[c], comps.data() + c * cells, cells * 8);
    return KPT_OK;
}

kpt_status kpt_compact_field_write(const char* path, int32_t nx, int32_t ny, const double* xe, const double* ye,
                                   const double* bx, const double* by, const double* bz, const double* vx,
                                   const double* vy, const double* vz) {
    if (!path || nx <= 0 || ny <= 0 || !xe || !ye || !bx || !by || !bz || !vx || !vy || !vz) return KPT_ERR_ARG;
    std::ofstream out(path, std::ios::binary);
    if (!out) return KPT_ERR_IO;
    const size_t cells = static_cast<size_t>(nx) * ny;
    out.write(reinterpret_cast<const char*>(&nx), 4);
    out.write(reinterpret_cast<const char*>(&ny), 4);
    out.write(reinterpret_cast<const char*>(xe), static_cast<std::streamsize>((static_cast<size_t>(nx) + 1) * 8));
    out.write(reinterpret_cast<const char*>(ye), static_cast<std::streamsize>((static_cast<size_t>(ny) + 1) * 8));
    for (const double* plane : {bx, by, bz, vx, vy, vz})
        out.write(reinterpret_cast<const char*>(plane), static_cast<std::streamsize>(cells * 8));
    return out ? KPT_OK : KPT_ERR_IO;
}

// ---- particles -------------------------------------------------------------

kpt_status kpt_particles_count(kpt_ctx* ctx, uint64_t* count) {
    if (!ctx || !count) return KPT_ERR_ARG;
    *count = ctx->count;
    return KPT_OK;
}

kpt_status kpt_particles_capacity(kpt_ctx* ctx, uint64_t* capacity) {
    if (!ctx || !capacity) return KPT_ERR_ARG;
    *capacity = ctx->capacity;
    return KPT_OK;
}

kpt_status kpt_particles_upload(kpt_ctx* ctx, const kpt_particles_host* h, uint64_t count) {
    if (!ctx || !h) return KPT_ERR_ARG;
    if (count > ctx->capacity) return fail(ctx, KPT_ERR_CAPACITY, "particle count exceeds capacity.");
    if (count > 0 && (!h->x || !h->y || !h->p || !h->t || !h->step_dt || !h->weight || !h->mu || !h->status ||
                      !h->split_level))
        return fail(ctx, KPT_ERR_ARG, "particle array pointer is null.");
    (void)hipSetDevice(ctx->device);
    kpt::ParticleArrays& P = ctx->P;
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
    }
    uint64_t max_root = 0;
    if (h->id) {
        if (count > 0) KPT_HIP(ctx, hipMemcpy(P.id, h->id, count * 8, hipMemcpyHostToDevice));
        for (uint64_t i = 0; i < count; ++i) max_root = std::max<uint64_t>(max_root, (h->id[i] & 0xFFFFFFFFull) + 1);
    } else if (count > 0) {
        std::vector<uint64_t> ids(count);
        for (uint64_t i = 0; i < count; ++i) ids[i] = i;
        KPT_HIP(ctx, hipMemcpy(P.id, ids.data(), count * 8, hipMemcpyHostToDevice));
        max_root = count;
    }
    // Slots past the stored range return to the storage defaults.
    if (ctx->count > count) {
        std::vector<int32_t> two(ctx->count - count, 2);
        KPT_HIP(ctx, hipMemcpy(P.status + count, two.data(), two.size() * 4, hipMemcpyHostToDevice));
    }
    ctx->count = count;
    ctx->next_root = max_root;
    return KPT_OK;
}

kpt_status kpt_particles_download(kpt_ctx* ctx, kpt_particles_host* h, uint64_t count) {
    if (!ctx || !h) return KPT_ERR_ARG;
    if (count > ctx->count) return fail(ctx, KPT_ERR_ARG, "download count exceeds stored particles.");
    if (count == 0) return KPT_OK;
    (void)hipSetDevice(ctx->device);
    const kpt::ParticleArrays& P = ctx->P;
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

// ---- hot path ----------------------------------------------------------------

kpt_status kpt_inject(kpt_ctx* ctx, const kpt_device_settings* s, const kpt_inject_opts* opts, uint64_t n,
                      double t0, uint32_t frame, uint64_t* injected) {
    if (!ctx) return KPT_ERR_ARG;
    kpt_status st = check_settings(ctx, s);
    if (st != KPT_OK) return st;
    kpt_inject_opts o{};
    if (opts) o = *opts;
    if (o.use_region && !(o.region_x_hi > o.region_x_lo && o.region_y_hi > o.region_y_lo))
        return fail(ctx, KPT_ERR_ARG, "injection region must have positive extent.");
    if (injected) *injected = 0;
    if (ctx->count >= ctx->capacity) return KPT_OK;
    const uint64_t m = std::min(n, ctx->capacity - ctx->count);
    (void)hipSetDevice(ctx->device);
    const uint32_t key[2] = {static_cast<uint32_t>(ctx->seed), static_cast<uint32_t>(ctx->seed >> 32)};
    KPT_HIP(ctx, kpt::launch_inject(ctx->P, ctx->count, m, ctx->next_root, *s, o, t0, frame, key, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->count += m;
    ctx->next_root += m;
    if (injected) *injected = m;
    return KPT_OK;
}

extern "C++" kpt_status kpt_capi::ensure_rec32(kpt_ctx* ctx, FieldSlot& f) {
    if (f.rec32 != nullptr && f.rec32_version == f.version) return KPT_OK;
    const uint64_t plane = kpt::coefficient_plane(f.g.nx, f.g.ny);
    const size_t elems = kpt::rec_elems(plane, f.model);
    if (elems > f.rec32_cap) {
        dev_free(f.rec32);
        f.rec32 = nullptr;
        f.rec32_cap = 0;
        KPT_HIP(ctx, dev_alloc(&f.rec32, elems));
        f.rec32_cap = elems;
    }
    KPT_HIP(ctx, kpt::launch_rec(f.c32, plane, f.model, f.rec32, ctx->stream));
    f.rec32_version = f.version;
    return KPT_OK;
}

// The push's launch arguments for the interval [t0, t0 + mhd_dt) of `frame`
// (per-chunk fields t_stop, first_chunk, counters and self are the caller's).
// `paired`: a time-blended push reads the frames from the node-major pair
// buffer of its precision, built (and grown) here; otherwise the kernel reads
// both component-planar frames.
extern "C++" kpt_status kpt_capi::push_args(kpt_ctx* ctx, const kpt_device_settings* s, double t0, uint32_t frame, bool paired,
                               kpt::PushArgs& a, bool& focused) {
    if (!ctx) return KPT_ERR_ARG;
    kpt_status st = check_settings(ctx, s);
    if (st != KPT_OK) return st;
    const FieldSlot& lo = ctx->slot[KPT_SLOT_LOWER];
    const FieldSlot& hi = ctx->slot[KPT_SLOT_UPPER];
    if (!lo.valid) return fail(ctx, KPT_ERR_STATE, "lower coefficient frame is not loaded.");
    if (s->time_interpolation != 0) {
        if (!hi.valid) return fail(ctx, KPT_ERR_STATE, "upper coefficient frame is not loaded.");
        if (hi.model != lo.model || hi.g.nx != lo.g.nx || hi.g.ny != lo.g.ny || hi.g.start[0] != lo.g.start[0] ||
            hi.g.start[1] != lo.g.start[1] || hi.g.step[0] != lo.g.step[0] || hi.g.step[1] != lo.g.step[1])
            return fail(ctx, KPT_ERR_STATE, "lower and upper coefficient frames have different grids or models.");
    }
    if (ctx->noise) {
        if (ctx->noise_particles < ctx->count)
            return fail(ctx, KPT_ERR_STATE, "noise table covers fewer particles than are stored.");
        if (static_cast<uint64_t>(s->max_particle_steps_per_interval) > ctx->noise_steps)
            return fail(ctx, KPT_ERR_STATE, "max_particle_steps_per_interval exceeds the noise table's steps.");
    }
    a = kpt::PushArgs{};
    a.count = ctx->count;
    const bool fp32 = s->precision == KPT_PREC_FP32;
    a.field_lo = fp32 ? static_cast<const void*>(lo.c32) : static_cast<const void*>(lo.c64);
    a.field_hi = s->time_interpolation != 0 ? (fp32 ? static_cast<const void*>(hi.c32) : static_cast<const void*>(hi.c64))
                                            : a.field_lo;
    a.field_lo64 = lo.c64;
    a.field_hi64 = s->time_interpolation != 0 ? hi.c64 : lo.c64;
    a.field_pair = nullptr;
    if (paired && fp32 && s->time_interpolation != 0) {
        const uint64_t plane = kpt::coefficient_plane(lo.g.nx, lo.g.ny);
        const size_t elems = kpt::pair_elems(plane, lo.model);
        if (elems > ctx->pair32_cap) {
            dev_free(ctx->pair32);
            ctx->pair32 = nullptr;
            ctx->pair32_cap = 0;
            ctx->pair_versions[0] = ctx->pair_versions[1] = 0;
            KPT_HIP(ctx, dev_alloc(&ctx->pair32, elems));
            ctx->pair32_cap = elems;
        }
        if (ctx->pair_versions[0] != lo.version || ctx->pair_versions[1] != hi.version) {
            KPT_HIP(ctx, kpt::launch_pair(lo.c32, hi.c32, plane, lo.model, ctx->pair32, ctx->stream));
            ctx->pair_versions[0] = lo.version;
            ctx->pair_versions[1] = hi.version;
        }
        a.field_pair = ctx->pair32;
    }
    a.field_rec = nullptr;
    if (paired && fp32 && s->time_interpolation == 0) {
        st = ensure_rec32(ctx, ctx->slot[KPT_SLOT_LOWER]);
        if (st != KPT_OK) return st;
        a.field_rec = lo.rec32;
    }
    a.field_pair64 = nullptr;
    if (paired && !fp32 && s->time_interpolation != 0) {
        const uint64_t plane = kpt::coefficient_plane(lo.g.nx, lo.g.ny);
        const size_t elems = kpt::pair64_elems(plane, lo.model);
        if (elems > ctx->pair64_cap) {
            dev_free(ctx->pair64);
            ctx->pair64 = nullptr;
            ctx->pair64_cap = 0;
            ctx->pair64_versions[0] = ctx->pair64_versions[1] = 0;
            KPT_HIP(ctx, dev_alloc(&ctx->pair64, elems));
            ctx->pair64_cap = elems;
        }
        if (ctx->pair64_versions[0] != lo.version || ctx->pair64_versions[1] != hi.version) {
            KPT_HIP(ctx, kpt::launch_pair64(lo.c64, hi.c64, plane, lo.model, ctx->pair64, ctx->stream));
            ctx->pair64_versions[0] = lo.version;
            ctx->pair64_versions[1] = hi.version;
        }
        a.field_pair64 = ctx->pair64;
    }
    a.stage = ctx->push_stage ? 1 : 0;
    a.plane = kpt::coefficient_plane(lo.g.nx, lo.g.ny);
    a.g = lo.g;
    a.s = *s;
    a.t0 = t0;
    a.t_end = t0 + s->mhd_dt;
    a.t_stop = a.t_end;
    a.first_chunk = 1;
    a.key[0] = static_cast<uint32_t>(ctx->seed);
    a.key[1] = static_cast<uint32_t>(ctx->seed >> 32);
    a.frame = frame;
    a.noise = ctx->noise;
    a.noise_steps = ctx->noise_steps;
    a.qlt_norm = s->qlt_spectral_index * std::sin(M_PI / s->qlt_spectral_index) / M_PI;
    a.k64 = kpt::make_step_const<double>(*s, a.qlt_norm);
    a.k32 = kpt::make_step_const<float>(*s, a.qlt_norm);
    a.mlp_image = nullptr;
    a.mlp_on = 0;
    a.mlp_act = kpt::kActSilu;
    if (s->closure == KPT_CLOSURE_MLP) {
        if (!fp32) return fail(ctx, KPT_ERR_ARG, "the learned pitch-angle closure runs in the FP32 push.");
        if (!ctx->mlp_valid) return fail(ctx, KPT_ERR_STATE, "no pitch-angle network is loaded (kpt_mlp_train / kpt_mlp_set).");
        const kpt_mlp_hull& h = ctx->mlp_hull;
        a.mlp_image = ctx->mlp_image;
        a.mlp_scale = kpt::mlp_scale(h);
        a.mlp_act = h.activation;
        a.mlp_on = (s->qlt_spectral_index >= h.gamma_lo && s->qlt_spectral_index <= h.gamma_hi &&
                    s->qlt_h0 == h.h0) ? 1 : 0;
    }
    focused = lo.model == KPT_MODEL_FOCUSED21 && s->focused_transport != 0;
    return KPT_OK;
}

kpt_status kpt_push_interval(kpt_ctx* ctx, const kpt_device_settings* s, double t0, uint32_t frame,
                             kpt_push_stats* stats) {
    kpt::PushArgs a;
    bool focused = false;
    kpt_status st = push_args(ctx, s, t0, frame, true, a, focused);
    if (st != KPT_OK) return st;
    const FieldSlot& lo = ctx->slot[KPT_SLOT_LOWER];
    const bool fp32 = s->precision == KPT_PREC_FP32;
    a.counters = ctx->counters;
    (void)hipSetDevice(ctx->device);
    const int chunks = s->push_chunks > 1 ? s->push_chunks : 1;
    const bool resort = chunks > 1 && ctx->noise == nullptr;
    unsigned long long total[3] = {0, 0, 0};
    double ms_total = 0.0;
    for (int k = 0; k < chunks; ++k) {
        if (resort) {
            st = sort_in_place(ctx, *s, a.t_end, false, nullptr);
            if (st != KPT_OK) return st;
        }
        a.P = ctx->P;
        a.first_chunk = k == 0 ? 1 : 0;
        a.t_stop = k == chunks - 1 ? a.t_end : t0 + s->mhd_dt * static_cast<double>(k + 1) / chunks;
        KPT_HIP(ctx, hipMemsetAsync(ctx->counters, 0, 4 * sizeof(unsigned long long), ctx->stream));
        // the arguments in device memory for the push's near-null sub-steps; `a`
        // stays unchanged until the stream is synchronised below
        a.self = ctx->d_args;
        KPT_HIP(ctx, hipMemcpyAsync(ctx->d_args, &a, sizeof(a), hipMemcpyHostToDevice, ctx->stream));
        KPT_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        KPT_HIP(ctx, kpt::launch_push(a, lo.model, focused, fp32, s->closure, ctx->stream));
        KPT_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        unsigned long long c[4];
        KPT_HIP(ctx, hipMemcpyAsync(c, ctx->counters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
        KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
        float ms = 0.0f;
        KPT_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ms_total += ms;
        total[0] += c[0];
        total[1] += c[1];
        if (k == 0) total[2] = c[2];
    }
    if (stats) {
        stats->particle_steps = total[0];
        stats->failures = total[1];
        stats->active_in = total[2];
        stats->kernel_ms = ms_total;
        stats->launches = ctx->count > 0 ? static_cast<uint32_t>(chunks) : 0u;
        stats->reserved = 0;
    }
    return KPT_OK;
}

kpt_status kpt_split(kpt_ctx* ctx, const kpt_device_settings* s, uint64_t* split, uint64_t* skipped) {
    if (!ctx) return KPT_ERR_ARG;
    kpt_status st = check_settings(ctx, s);
    if (st != KPT_OK) return st;
    if (split) *split = 0;
    if (skipped) *skipped = 0;
    if (ctx->count >= ctx->capacity || s->split_particles == 0 || ctx->count == 0) return KPT_OK;
    (void)hipSetDevice(ctx->device);
    KPT_HIP(ctx, hipMemsetAsync(ctx->counters, 0, 8 * sizeof(unsigned long long), ctx->stream));
    kpt::SplitTable tab;
    for (int L = 0; L < kpt::SplitTable::kLevels; ++L)
        tab.thr[L] = s->pmin_split * std::pow(s->split_ratio, static_cast<double>(L));
    KPT_HIP(ctx, kpt::launch_split(ctx->P, ctx->count, ctx->capacity, *s, tab, ctx->split_scratch,
                                   ctx->counters, ctx->stream));
    unsigned long long c[2];
    uint32_t total = 0;
    const unsigned tiles = static_cast<unsigned>(kpt::split_scratch_elems(ctx->count) - 1);
    KPT_HIP(ctx, hipMemcpyAsync(c, ctx->counters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipMemcpyAsync(&total, ctx->split_scratch + tiles, 4, hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->count = std::min<uint64_t>(ctx->count + total, ctx->capacity);
    if (split) *split = c[0];
    if (skipped) *skipped = c[1];
    return KPT_OK;
}

kpt_status kpt_histogram_logp(kpt_ctx* ctx, int32_t bins, double pmin, double pmax, double* out) {
    if (!ctx || !out) return KPT_ERR_ARG;
    if (bins <= 0 || bins > 4096) return fail(ctx, KPT_ERR_ARG, "histogram bins must be in [1, 4096].");
    if (!(pmin > 0.0) || !(pmax > pmin) || !std::isfinite(pmax))
        return fail(ctx, KPT_ERR_ARG, "histogram range must satisfy 0 < pmin < pmax.");
    const double log_pmin = std::log(pmin);
    const double inv_dlogp = static_cast<double>(bins) / (std::log(pmax) - log_pmin);
    (void)hipSetDevice(ctx->device);
    KPT_HIP(ctx, kpt::launch_hist_logp(ctx->P, ctx->count, bins, log_pmin, inv_dlogp, pmin, pmax, ctx->hist_partials,
                                       ctx->hist_out, ctx->stream));
    KPT_HIP(ctx, hipMemcpyAsync(out, ctx->hist_out, bins * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KPT_OK;
}

kpt_status kpt_histogram_xy(kpt_ctx* ctx, int32_t bx, int32_t by, double x_lo, double x_hi, double y_lo,
                            double y_hi, double* out) {
    if (!ctx || !out) return KPT_ERR_ARG;
    if (bx <= 0 || by <= 0 || static_cast<int64_t>(bx) * by > 4096)
        return fail(ctx, KPT_ERR_ARG, "spatial histogram must have 1..4096 bins.");
    if (!(x_hi > x_lo) || !(y_hi > y_lo)) return fail(ctx, KPT_ERR_ARG, "spatial histogram range is empty.");
    (void)hipSetDevice(ctx->device);
    KPT_HIP(ctx, kpt::launch_hist_xy(ctx->P, ctx->count, bx, by, x_lo, x_hi, y_lo, y_hi, ctx->hist_partials,
                                     ctx->hist_out, ctx->stream));
    KPT_HIP(ctx, hipMemcpyAsync(out, ctx->hist_out, static_cast<size_t>(bx) * by * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KPT_OK;
}

kpt_status kpt_sort_particles(kpt_ctx* ctx, const kpt_device_settings* s, int32_t compact, uint64_t* active) {
    if (!ctx) return KPT_ERR_ARG;
    kpt_status st = check_settings(ctx, s);
    if (st != KPT_OK) return st;
    if (ctx->noise) return fail(ctx, KPT_ERR_STATE, "the injected-noise table is keyed by slot: clear it before sorting.");
    if (!ctx->slot[KPT_SLOT_LOWER].valid) return fail(ctx, KPT_ERR_STATE, "lower coefficient frame is not loaded.");
    (void)hipSetDevice(ctx->device);
    return sort_in_place(ctx, *s, std::numeric_limits<double>::infinity(), compact != 0, active);
}

}  // extern "C"

namespace {

// Cell sort of the stored particles into the second particle set, then swap.
// Particles with t >= t_done sort after the moving ones (see launch_sort).
kpt_status sort_in_place(kpt_ctx* ctx, const kpt_device_settings& s, double t_done, bool compact,
                         uint64_t* active) {
    const FieldSlot& lo = ctx->slot[KPT_SLOT_LOWER];
    if (!ctx->p2_allocated) {
        kpt::ParticleArrays& Q = ctx->P2;
        const uint64_t cap = ctx->capacity;
        KPT_HIP(ctx, kpt::alloc_arrays(Q, cap));
        for (int k = 0; k < 2; ++k) {
            KPT_HIP(ctx, dev_alloc(&ctx->sort.keys[k], cap));
            KPT_HIP(ctx, dev_alloc(&ctx->sort.vals[k], cap));
        }
        ctx->sort.temp_bytes = kpt::sort_temp_bytes(cap);
        KPT_HIP(ctx, hipMalloc(&ctx->sort.temp, std::max<size_t>(ctx->sort.temp_bytes, 1)));
        // the slots past `count` keep the storage default status 2
        std::vector<int32_t> two(static_cast<size_t>(std::min<uint64_t>(cap, 1u << 20)), 2);
        for (uint64_t off = 0; off < cap; off += two.size()) {
            const size_t n2 = static_cast<size_t>(std::min<uint64_t>(two.size(), cap - off));
            KPT_HIP(ctx, hipMemcpy(Q.status + off, two.data(), n2 * 4, hipMemcpyHostToDevice));
        }
        ctx->p2_allocated = true;
    }
    unsigned long long* act_dev = ctx->counters + 4;
    KPT_HIP(ctx, hipMemsetAsync(act_dev, 0, sizeof(unsigned long long), ctx->stream));
    KPT_HIP(ctx, kpt::launch_sort(ctx->P, ctx->P2, ctx->count, lo.g, s.interpolation_shift_x,
                                  s.interpolation_shift_y, t_done, ctx->sort, act_dev, ctx->stream));
    unsigned long long act = 0;
    KPT_HIP(ctx, hipMemcpyAsync(&act, act_dev, sizeof(act), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::swap(ctx->P, ctx->P2);
    if (compact) {
        // dropped slots return to the storage default (status 2)
        if (act < ctx->count) {
            std::vector<int32_t> two(ctx->count - act, 2);
            KPT_HIP(ctx, hipMemcpy(ctx->P.status + act, two.data(), two.size() * 4, hipMemcpyHostToDevice));
        }
        ctx->count = act;
    }
    if (active) *active = act;
    return KPT_OK;
}

}  // namespace

extern "C" {

kpt_status kpt_summary_compute(kpt_ctx* ctx, kpt_summary* out) {
    if (!ctx || !out) return KPT_ERR_ARG;
    if (ctx->count == 0) {
        std::memset(out, 0, sizeof(*out));
        return KPT_OK;
    }
    (void)hipSetDevice(ctx->device);
    KPT_HIP(ctx, kpt::launch_summary(ctx->P, ctx->count, ctx->summary_partials, ctx->summary_out, ctx->stream));
    KPT_HIP(ctx, hipMemcpyAsync(out, ctx->summary_out, sizeof(kpt_summary), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KPT_OK;
}

// ---- noise -------------------------------------------------------------------

kpt_status kpt_set_noise(kpt_ctx* ctx, const double* draws, uint64_t n_particles, uint32_t steps) {
    if (!ctx || !draws || n_particles == 0 || steps == 0)
        return fail(ctx, KPT_ERR_ARG, "noise table must be non-empty.");
    (void)hipSetDevice(ctx->device);
    dev_free(ctx->noise);
    ctx->noise = nullptr;
    const size_t elems = static_cast<size_t>(n_particles) * steps * 4;
    KPT_HIP(ctx, dev_alloc(&ctx->noise, elems));
    KPT_HIP(ctx, hipMemcpy(ctx->noise, draws, elems * sizeof(double), hipMemcpyHostToDevice));
    ctx->noise_particles = n_particles;
    ctx->noise_steps = steps;
    return KPT_OK;
}

kpt_status kpt_clear_noise(kpt_ctx* ctx) {
    if (!ctx) return KPT_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    dev_free(ctx->noise);
    ctx->noise = nullptr;
    ctx->noise_particles = 0;
    ctx->noise_steps = 0;
    return KPT_OK;
}

void kpt_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t m0 = static_cast<uint64_t>(0xD2511F53u) * c0;
        const uint64_t m1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
        const uint32_t n0 = static_cast<uint32_t>(m1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = static_cast<uint32_t>(m0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = static_cast<uint32_t>(m1);
        c2 = n2;
        c3 = static_cast<uint32_t>(m0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// ---- v4 snapshots --------------------------------------------------------------

}  // extern "C"

namespace {

// v4 writer. `resumable` stores the particle ids and split baselines in the
// records' id and E0 fields instead of the driver's slot index and constant E0.
kpt_status write_snapshot(const char* path, const kpt_particles_host* h, uint64_t count, uint64_t capacity,
                          double charge, double split_ratio, double p0, bool resumable) {
    if (!path || !h) return KPT_ERR_ARG;
    if (resumable && count > 0 && (!h->id || !h->e0)) return KPT_ERR_ARG;
    if (count > 0 && (!h->x || !h->y || !h->p || !h->mu || !h->weight || !h->status || !h->split_level))
        return KPT_ERR_ARG;
    std::ofstream out(path, std::ios::binary);
    if (!out) return KPT_ERR_IO;
    // Header (src/main.cpp:2155-2181): 88 bytes. Records (:2183-2201): 100 bytes.
    std::vector<char> buf;
    buf.reserve(88 + static_cast<size_t>(count) * 100);
    auto put = [&buf](const void* v, size_t bytes) {
        const char* c = static_cast<const char*>(v);
        buf.insert(buf.end(), c, c + bytes);
    };
    const char magic[8] = {'K', 'P', 'T', 'P', 'R', 'T', '\0', '\0'};
    const uint32_t version = 4u, endian = 0x01020304u;
    const int32_t dim = 2, species = 3;
    const double rest_mass = 1.0, c_light = 1.0, energy_scale = 1.0, min_child_weight = 0.0;
    put(magic, 8);
    put(&version, 4);
    put(&endian, 4);
    put(&dim, 4);
    put(&species, 4);
    put(&count, 8);
    put(&capacity, 8);
    put(&rest_mass, 8);
    put(&charge, 8);
    put(&c_light, 8);
    put(&energy_scale, 8);
    put(&split_ratio, 8);
    put(&min_child_weight, 8);
    const double e0 = std::sqrt(1.0 + p0 * p0) - 1.0;
    const int32_t sort_key = 0;
    for (uint64_t i = 0; i < count; ++i) {
        put(resumable ? static_cast<const void*>(&h->id[i]) : static_cast<const void*>(&i), 8);
        put(&h->status[i], 4);
        put(&h->split_level[i], 4);
        put(&sort_key, 4);
        for (int rep = 0; rep < 3; ++rep) {  // position, previous, previous-step
            put(&h->x[i], 8);
            put(&h->y[i], 8);
        }
        put(&h->p[i], 8);
        put(&h->mu[i], 8);
        put(&h->weight[i], 8);
        put(resumable ? &h->e0[i] : &e0, 8);
    }
    out.write(buf.data(), static_cast<std::streamsize>(buf.size()));
    return out ? KPT_OK : KPT_ERR_IO;
}

}  // namespace

extern "C" {

kpt_status kpt_snapshot_write_host(const char* path, const kpt_particles_host* h, uint64_t count,
                                   uint64_t capacity, double charge, double split_ratio, double p0) {
    return write_snapshot(path, h, count, capacity, charge, split_ratio, p0, false);
}

kpt_status kpt_snapshot_read_host(const char* path, kpt_particles_host* h, uint64_t* count_io, double charge,
                                  double split_ratio, char* err, size_t err_len) {
    auto bad = [&](kpt_status code, const std::string& m) {
        if (err && err_len) std::snprintf(err, err_len, "%s", m.c_str());
        return code;
    };
    if (!path || !count_io) return bad(KPT_ERR_ARG, "path or count pointer is null.");
    std::ifstream in(path, std::ios::binary);
    if (!in) return bad(KPT_ERR_IO, "Failed to open restart particle snapshot.");
    char magic[8];
    in.read(magic, 8);
    const char expect[8] = {'K', 'P', 'T', 'P', 'R', 'T', '\0', '\0'};
    if (!in || std::memcmp(magic, expect, 8) != 0)
        return bad(KPT_ERR_IO, "Restart particle snapshot is not a repository KPTPRT file.");
    auto get = [&in](void* v, size_t bytes) {
        in.read(static_cast<char*>(v), static_cast<std::streamsize>(bytes));
        return static_cast<bool>(in);
    };
    uint32_t version = 0, endian = 0;
    int32_t dim = 0, species = 0;
    uint64_t count = 0, capacity = 0;
    double hdr[6];
    if (!get(&version, 4) || !get(&endian, 4)) return bad(KPT_ERR_IO, "Failed to read particle binary scalar.");
    if (version != 4u) return bad(KPT_ERR_IO, "Restart particle snapshot must use particle binary version 4.");
    if (endian != 0x01020304u) return bad(KPT_ERR_IO, "Restart particle snapshot endian does not match this host.");
    if (!get(&dim, 4) || !get(&species, 4) || !get(&count, 8) || !get(&capacity, 8) || !get(hdr, 48))
        return bad(KPT_ERR_IO, "Failed to read particle binary scalar.");
    if (dim != 2)
        return bad(KPT_ERR_IO, "Restart particle snapshot dimensionality does not match reconnection storage.");
    if (h == nullptr) {
        *count_io = count;
        return KPT_OK;
    }
    if (count > *count_io) {
        return bad(KPT_ERR_CAPACITY, "Restart particle snapshot stores " + std::to_string(count) +
                                         " particles, but configured capacity is " + std::to_string(*count_io) +
                                         ". Snapshot capacity was " + std::to_string(capacity) + ".");
    }
    auto matches = [](double a, double b) {
        if (!std::isfinite(a) || !std::isfinite(b)) return false;
        const double scale = std::max({1.0, std::fabs(a), std::fabs(b)});
        return std::fabs(a - b) <= 1.0e-12 * scale;
    };
    if (!matches(hdr[1], charge))
        return bad(KPT_ERR_ARG, "Restart particle snapshot charge does not match current settings.");
    if (!matches(hdr[4], split_ratio))
        return bad(KPT_ERR_ARG, "Restart particle snapshot split ratio does not match current settings.");
    std::vector<char> rec(static_cast<size_t>(count) * 100);
    if (count > 0 && !get(rec.data(), rec.size())) return bad(KPT_ERR_IO, "Failed to read particle binary scalar.");
    for (uint64_t i = 0; i < count; ++i) {
        const char* r = rec.data() + i * 100;
        if (h->id) std::memcpy(&h->id[i], r + 0, 8);
        if (h->status) std::memcpy(&h->status[i], r + 8, 4);
        if (h->split_level) std::memcpy(&h->split_level[i], r + 12, 4);
        if (h->x) std::memcpy(&h->x[i], r + 20, 8);
        if (h->y) std::memcpy(&h->y[i], r + 28, 8);
        if (h->p) std::memcpy(&h->p[i], r + 68, 8);
        if (h->mu) std::memcpy(&h->mu[i], r + 76, 8);
        if (h->weight) std::memcpy(&h->weight[i], r + 84, 8);
        if (h->e0) std::memcpy(&h->e0[i], r + 92, 8);
    }
    *count_io = count;
    return KPT_OK;
}

kpt_status kpt_snapshot_write_resumable(kpt_ctx* ctx, const char* path, double charge, double split_ratio,
                                        double p0) {
    if (!ctx || !path) return KPT_ERR_ARG;
    const uint64_t n = ctx->count;
    std::vector<double> x(n), y(n), p(n), mu(n), w(n), e0(n);
    std::vector<int32_t> status(n), level(n);
    std::vector<uint64_t> id(n);
    kpt_particles_host h{x.data(), y.data(), p.data(), nullptr, nullptr, w.data(), mu.data(),
                         status.data(), level.data(), id.data(), e0.data()};
    kpt_status st = kpt_particles_download(ctx, &h, n);
    if (st != KPT_OK) return st;
    st = write_snapshot(path, &h, n, ctx->capacity, charge, split_ratio, p0, true);
    if (st != KPT_OK) return fail(ctx, st, "Failed to write particle snapshot.");
    return KPT_OK;
}

kpt_status kpt_snapshot_write(kpt_ctx* ctx, const char* path, double charge, double split_ratio, double p0) {
    if (!ctx || !path) return KPT_ERR_ARG;
    const uint64_t n = ctx->count;
    std::vector<double> x(n), y(n), p(n), mu(n), w(n);
    std::vector<int32_t> status(n), level(n);
    kpt_particles_host h{x.data(), y.data(), p.data(), nullptr, nullptr, w.data(), mu.data(),
                         status.data(), level.data(), nullptr};
    kpt_status st = kpt_particles_download(ctx, &h, n);
    if (st != KPT_OK) return st;
    st = kpt_snapshot_write_host(path, &h, n, ctx->capacity, charge, split_ratio, p0);
    if (st != KPT_OK) return fail(ctx, st, "Failed to write particle snapshot.");
    return KPT_OK;
}

kpt_status kpt_snapshot_read(kpt_ctx* ctx, const char* path, double charge, double split_ratio,
                             double restart_time, double dt_min) {
    if (!ctx || !path) return KPT_ERR_ARG;
    char err[512] = {0};
    uint64_t count = 0;
    kpt_status st = kpt_snapshot_read_host(path, nullptr, &count, charge, split_ratio, err, sizeof(err));
    if (st != KPT_OK) return fail(ctx, st, err);
    if (count > ctx->capacity) {
        return fail(ctx, KPT_ERR_CAPACITY, "Restart particle snapshot stores " + std::to_string(count) +
                                               " particles, but configured capacity is " +
                                               std::to_string(ctx->capacity) + ".");
    }
    std::vector<double> x(count), y(count), p(count), t(count, restart_time), dt(count, dt_min), w(count), mu(count),
        e0(count);
    std::vector<int32_t> status(count), level(count);
    std::vector<uint64_t> id(count);
    kpt_particles_host h{x.data(), y.data(), p.data(), t.data(), dt.data(), w.data(), mu.data(),
                         status.data(), level.data(), id.data(), e0.data()};
    st = kpt_snapshot_read_host(path, &h, &count, charge, split_ratio, err, sizeof(err));
    if (st != KPT_OK) return fail(ctx, st, err);
    return kpt_particles_upload(ctx, &h, count);
}

// ---- ParticleSystem storage -----------------------------------------------------

kpt_status kpt_capture_previous_positions(kpt_ctx* ctx) {
    if (!ctx) return KPT_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    KPT_HIP(ctx, kpt::launch_capture_previous(ctx->P, ctx->count, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KPT_OK;
}

}  // extern "C"

namespace {

kpt_status sort_stored(kpt_ctx* ctx, int32_t min_key, int32_t max_key, bool compact, uint64_t* active) {
    if (!ctx) return KPT_ERR_ARG;
    if (max_key < min_key) return fail(ctx, KPT_ERR_ARG, "ParticleSystem: invalid sort key range.");
    (void)hipSetDevice(ctx->device);
    if (!ctx->p2_allocated) {
        kpt::ParticleArrays& Q = ctx->P2;
        const uint64_t cap = ctx->capacity;
        KPT_HIP(ctx, kpt::alloc_arrays(Q, cap));
        for (int k = 0; k < 2; ++k) {
            KPT_HIP(ctx, dev_alloc(&ctx->sort.keys[k], cap));
            KPT_HIP(ctx, dev_alloc(&ctx->sort.vals[k], cap));
        }
        ctx->sort.temp_bytes = kpt::sort_temp_bytes(cap);
        KPT_HIP(ctx, hipMalloc(&ctx->sort.temp, std::max<size_t>(ctx->sort.temp_bytes, 1)));
        std::vector<int32_t> two(static_cast<size_t>(std::min<uint64_t>(cap, 1u << 20)), 2);
        for (uint64_t off = 0; off < cap; off += two.size()) {
            const size_t n2 = static_cast<size_t>(std::min<uint64_t>(two.size(), cap - off));
            KPT_HIP(ctx, hipMemcpy(Q.status + off, two.data(), n2 * 4, hipMemcpyHostToDevice));
        }
        ctx->p2_allocated = true;
    }
    unsigned long long* act_dev = ctx->counters + 4;
    KPT_HIP(ctx, hipMemsetAsync(act_dev, 0, sizeof(unsigned long long), ctx->stream));
    KPT_HIP(ctx, kpt::launch_sort_stored(ctx->P, ctx->P2, ctx->count, min_key, max_key, compact ? 1 : 0, ctx->sort,
                                         act_dev, ctx->stream));
    unsigned long long act = 0;
    KPT_HIP(ctx, hipMemcpyAsync(&act, act_dev, sizeof(act), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::swap(ctx->P, ctx->P2);
    if (compact && act < ctx->count) {
        // the dropped slots return to the storage default (status 2)
        std::vector<int32_t> two(static_cast<size_t>(ctx->count - act), 2);
        KPT_HIP(ctx, hipMemcpy(ctx->P.status + act, two.data(), two.size() * 4, hipMemcpyHostToDevice));
        ctx->count = act;
    }
    if (active) *active = act;
    return KPT_OK;
}

// ParticleProperties::validate (include/ParticleSystem.hpp:155-165) and
// ParticleSplittingPolicy::validate (:287-293) of a snapshot header.
bool properties_ok(const kpt_particle_properties& pr) {
    return std::isfinite(pr.rest_mass) && pr.rest_mass > 0.0 && std::isfinite(pr.charge) &&
           std::isfinite(pr.speed_of_light) && pr.speed_of_light > 0.0 && std::isfinite(pr.energy_scale_erg) &&
           pr.energy_scale_erg > 0.0;
}
bool policy_ok(const kpt_energy_split_policy& po) {
    return std::isfinite(po.energy_split_ratio) && po.energy_split_ratio > 1.0 &&
           std::isfinite(po.minimum_child_weight) && po.minimum_child_weight >= 0.0;
}

}  // namespace

extern "C" {

kpt_status kpt_sort_by_key(kpt_ctx* ctx, int32_t min_key, int32_t max_key) {
    if (ctx && (ctx->count <= 1 || min_key == max_key)) return KPT_OK;  // nothing to order (:730-732)
    return sort_stored(ctx, min_key, max_key, false, nullptr);
}

kpt_status kpt_compact_active_particles(kpt_ctx* ctx, uint64_t* active) {
    if (!ctx) return KPT_ERR_ARG;
    if (ctx->count == 0) {
        if (active) *active = 0;
        return KPT_OK;
    }
    return sort_stored(ctx, 0, 1, true, active);
}

kpt_status kpt_particle_system_write_snapshot(kpt_ctx* ctx, const char* path, const kpt_particle_properties* props,
                                              int32_t species, const kpt_energy_split_policy* policy) {
    if (!ctx || !path || !props || !policy) return KPT_ERR_ARG;
    if (!properties_ok(*props) || species < 0 || species > 3)
        return fail(ctx, KPT_ERR_ARG, "ParticleProperties: invalid particle properties or species.");
    if (!policy_ok(*policy))
        return fail(ctx, KPT_ERR_ARG,
                    "ParticleSplittingPolicy: energy split ratio must be greater than one and minimum child weight "
                    "must be finite and non-negative.");
    const uint64_t n = ctx->count;
    std::vector<double> x(n), y(n), p(n), mu(n), w(n), e0(n), px(n), py(n), psx(n), psy(n);
    std::vector<int32_t> status(n), level(n), key(n);
    std::vector<uint64_t> id(n);
    kpt_particles_host h{x.data(),  y.data(),  p.data(),   nullptr,    nullptr,    w.data(),   mu.data(),
                         status.data(), level.data(), id.data(), e0.data(), px.data(), py.data(), psx.data(),
                         psy.data(), key.data()};
    kpt_status st = kpt_particles_download(ctx, &h, n);
    if (st != KPT_OK) return st;
    std::ofstream out(path, std::ios::binary);
    if (!out) return fail(ctx, KPT_ERR_IO, "ParticleSystem: failed to open binary particle output file.");
    // write_binary_snapshot (include/ParticleSystem.hpp:936-1006): 88-byte header,
    // 100-byte records
    std::vector<char> buf;
    buf.reserve(88 + static_cast<size_t>(n) * 100);
    auto put = [&buf](const void* v, size_t bytes) {
        const char* c = static_cast<const char*>(v);
        buf.insert(buf.end(), c, c + bytes);
    };
    const char magic[8] = {'K', 'P', 'T', 'P', 'R', 'T', '\0', '\0'};
    const uint32_t version = 4u, endian = 0x01020304u;
    const int32_t dim = 2;
    const uint64_t cap = ctx->capacity;
    put(magic, 8);
    put(&version, 4);
    put(&endian, 4);
    put(&dim, 4);
    put(&species, 4);
    put(&n, 8);
    put(&cap, 8);
    put(&props->rest_mass, 8);
    put(&props->charge, 8);
    put(&props->speed_of_light, 8);
    put(&props->energy_scale_erg, 8);
    put(&policy->energy_split_ratio, 8);
    put(&policy->minimum_child_weight, 8);
    for (uint64_t i = 0; i < n; ++i) {
        put(&id[i], 8);
        put(&status[i], 4);
        put(&level[i], 4);
        put(&key[i], 4);
        put(&x[i], 8);
        put(&y[i], 8);
        put(&px[i], 8);
        put(&py[i], 8);
        put(&psx[i], 8);
        put(&psy[i], 8);
        put(&p[i], 8);
        put(&mu[i], 8);
        put(&w[i], 8);
        put(&e0[i], 8);
    }
    out.write(buf.data(), static_cast<std::streamsize>(buf.size()));
    if (!out) return fail(ctx, KPT_ERR_IO, "ParticleSystem: failed to write binary particle output file.");
    return KPT_OK;
}

kpt_status kpt_particle_system_read_snapshot(kpt_ctx* ctx, const char* path, kpt_particle_properties* props,
                                             int32_t* species, kpt_energy_split_policy* policy) {
    if (!ctx || !path) return KPT_ERR_ARG;
    // read_binary_snapshot (include/ParticleSystem.hpp:1011-1145) with its checks
    // and messages; the set replaces the stored particles (time 0, step_dt 0)
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(ctx, KPT_ERR_IO, "ParticleSystem: failed to open binary particle input file.");
    auto get = [&in](void* v, size_t bytes) {
        in.read(static_cast<char*>(v), static_cast<std::streamsize>(bytes));
        return static_cast<bool>(in);
    };
    char magic[8];
    const char expect[8] = {'K', 'P', 'T', 'P', 'R', 'T', '\0', '\0'};
    if (!get(magic, 8) || std::memcmp(magic, expect, 8) != 0)
        return fail(ctx, KPT_ERR_IO, "ParticleSystem: unrecognized particle binary magic.");
    uint32_t version = 0, endian = 0;
    int32_t dim = 0, sp = 0;
    uint64_t count = 0, capacity = 0;
    kpt_particle_properties pr{};
    kpt_energy_split_policy po{};
    if (!get(&version, 4) || !get(&endian, 4))
        return fail(ctx, KPT_ERR_IO, "ParticleSystem: failed to read binary particle scalar.");
    if (version != 4u) return fail(ctx, KPT_ERR_IO, "ParticleSystem: only particle binary version 4 is restart-readable.");
    if (endian != 0x01020304u)
        return fail(ctx, KPT_ERR_IO, "ParticleSystem: particle binary endian does not match this host.");
    if (!get(&dim, 4) || !get(&sp, 4) || !get(&count, 8) || !get(&capacity, 8))
        return fail(ctx, KPT_ERR_IO, "ParticleSystem: failed to read binary particle scalar.");
    if (dim != 2) return fail(ctx, KPT_ERR_IO, "ParticleSystem: snapshot dimension does not match ParticleSystem.");
    if (sp < 0 || sp > 3) return fail(ctx, KPT_ERR_IO, "ParticleSystem: snapshot contains an invalid particle species.");
    if (!get(&pr.rest_mass, 8) || !get(&pr.charge, 8) || !get(&pr.speed_of_light, 8) || !get(&pr.energy_scale_erg, 8) ||
        !get(&po.energy_split_ratio, 8) || !get(&po.minimum_child_weight, 8))
        return fail(ctx, KPT_ERR_IO, "ParticleSystem: failed to read binary particle scalar.");
    if (!properties_ok(pr)) return fail(ctx, KPT_ERR_IO, "ParticleProperties: invalid particle properties.");
    if (!policy_ok(po))
        return fail(ctx, KPT_ERR_IO,
                    "ParticleSplittingPolicy: energy split ratio must be greater than one and minimum child weight "
                    "must be finite and non-negative.");
    if (count > ctx->capacity)
        return fail(ctx, KPT_ERR_CAPACITY, "ParticleSystem: snapshot stores " + std::to_string(count) +
                                               " particles, more than the capacity " + std::to_string(ctx->capacity) +
                                               ".");
    const uint64_t n = count;
    std::vector<double> x(n), y(n), p(n), t(n, 0.0), dt(n, 0.0), w(n), mu(n), e0(n), px(n), py(n), psx(n), psy(n);
    std::vector<int32_t> status(n), level(n), key(n);
    std::vector<uint64_t> id(n);
    std::vector<char> rec(static_cast<size_t>(n) * 100);
    if (n > 0 && !get(rec.data(), rec.size()))
        return fail(ctx, KPT_ERR_IO, "ParticleSystem: failed to read binary particle scalar.");
    for (uint64_t i = 0; i < n; ++i) {
        const char* r = rec.data() + i * 100;
        std::memcpy(&id[i], r + 0, 8);
        std::memcpy(&status[i], r + 8, 4);
        std::memcpy(&level[i], r + 12, 4);
        std::memcpy(&key[i], r + 16, 4);
        std::memcpy(&x[i], r + 20, 8);
        std::memcpy(&y[i], r + 28, 8);
        std::memcpy(&px[i], r + 36, 8);
        std::memcpy(&py[i], r + 44, 8);
        std::memcpy(&psx[i], r + 52, 8);
        std::memcpy(&psy[i], r + 60, 8);
        std::memcpy(&p[i], r + 68, 8);
        std::memcpy(&mu[i], r + 76, 8);
        std::memcpy(&w[i], r + 84, 8);
        std::memcpy(&e0[i], r + 92, 8);
    }
    kpt_particles_host h{x.data(),  y.data(),  p.data(),   t.data(),   dt.data(),  w.data(),   mu.data(),
                         status.data(), level.data(), id.data(), e0.data(), px.data(), py.data(), psx.data(),
                         psy.data(), key.data()};
    kpt_status st = kpt_particles_upload(ctx, &h, n);
    if (st != KPT_OK) return st;
    if (props) *props = pr;
    if (species) *species = sp;
    if (policy) *policy = po;
    return KPT_OK;
}

}  // extern "C"

namespace {

bool hull_ok(const kpt_mlp_hull& h) {
    return std::isfinite(h.mu_lo) && h.mu_lo > 0.0 && h.mu_lo < 1.0 && std::isfinite(h.log_xi_lo) &&
           std::isfinite(h.log_xi_hi) && h.log_xi_hi > h.log_xi_lo && std::isfinite(h.gamma_lo) &&
           std::isfinite(h.gamma_hi) && h.gamma_hi > h.gamma_lo && h.gamma_lo > 1.0 && std::isfinite(h.h0) &&
           h.h0 > 0.0 && std::isfinite(h.out_mean) && std::isfinite(h.out_scale) && h.out_scale > 0.0 &&
           kpt::mlp_activation_ok(h.activation);
}

// Install FP32 parameters: device copy, host copy and the BF16 MFMA image.
kpt_status mlp_install(kpt_ctx* ctx, const float* params, const kpt_mlp_hull& h) {
    if (!ctx->mlp_params) KPT_HIP(ctx, dev_alloc(&ctx->mlp_params, kpt::kMlpParams));
    if (!ctx->mlp_image) KPT_HIP(ctx, hipMalloc(&ctx->mlp_image, kpt::kMlpImageBytes));
    ctx->mlp_host.assign(params, params + kpt::kMlpParams);
    std::vector<unsigned char> image(kpt::kMlpImageBytes, 0);
    kpt::mlp_build_image(params, h.activation, image.data());
    KPT_HIP(ctx, hipMemcpy(ctx->mlp_params, params, kpt::kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    KPT_HIP(ctx, hipMemcpy(ctx->mlp_image, image.data(), image.size(), hipMemcpyHostToDevice));
    ctx->mlp_hull = h;
    ctx->mlp_valid = true;
    ++ctx->mlp_version;
    return KPT_OK;
}

// Scoped device buffers of the training and test calls.
struct DevBufs {
    std::vector<void*> ptrs;
    ~DevBufs() {
        for (void* p : ptrs) dev_free(p);
    }
    template <typename T>
    hipError_t get(T** p, size_t elems) {
        hipError_t e = dev_alloc(p, elems);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

kpt_status kpt_mlp_default_train_opts(kpt_mlp_train_opts* o) {
    if (!o) return KPT_ERR_ARG;
    std::memset(o, 0, sizeof(*o));
    o->samples = 1ull << 22;
    o->steps = 8000;
    o->batch = 4096;
    o->lr = 3.0e-3;
    o->seed = 0x4D4C50ull;
    o->hull.mu_lo = 1.0e-3;
    o->hull.log_xi_lo = std::log(1.0e-7);
    o->hull.log_xi_hi = 0.0;
    o->hull.gamma_lo = 1.5;
    o->hull.gamma_hi = 2.0;
    o->hull.h0 = 1.0e-6;
    o->hull.out_mean = 0.0;
    o->hull.out_scale = 1.0;
    o->hull.activation = KPT_MLP_SILU;
    return KPT_OK;
}

kpt_status kpt_mlp_set(kpt_ctx* ctx, const float* params, size_t n, const kpt_mlp_hull* hull) {
    if (!ctx || !params || !hull) return KPT_ERR_ARG;
    if (n != static_cast<size_t>(kpt::kMlpParams))
        return fail(ctx, KPT_ERR_ARG, "the pitch-angle network has KPT_MLP_PARAMS parameters.");
    if (!hull_ok(*hull)) return fail(ctx, KPT_ERR_ARG, "invalid pitch-angle network hull.");
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(params[i])) return fail(ctx, KPT_ERR_ARG, "non-finite network parameter.");
    (void)hipSetDevice(ctx->device);
    return mlp_install(ctx, params, *hull);
}

kpt_status kpt_mlp_get(kpt_ctx* ctx, float* params, size_t n, kpt_mlp_hull* hull) {
    if (!ctx) return KPT_ERR_ARG;
    if (!ctx->mlp_valid) return fail(ctx, KPT_ERR_STATE, "no pitch-angle network is loaded.");
    if (params) {
        if (n != static_cast<size_t>(kpt::kMlpParams))
            return fail(ctx, KPT_ERR_ARG, "the pitch-angle network has KPT_MLP_PARAMS parameters.");
        std::memcpy(params, ctx->mlp_host.data(), n * sizeof(float));
    }
    if (hull) *hull = ctx->mlp_hull;
    return KPT_OK;
}

kpt_status kpt_mlp_gradient(kpt_ctx* ctx, const float* feats, const float* targets, uint64_t n, float* grad,
                            double* loss) {
    if (!ctx || !feats || !targets || !grad) return KPT_ERR_ARG;
    if (!ctx->mlp_valid) return fail(ctx, KPT_ERR_STATE, "no pitch-angle network is loaded.");
    if (n == 0 || n > 65536) return fail(ctx, KPT_ERR_ARG, "gradient batch must hold 1..65536 samples.");
    (void)hipSetDevice(ctx->device);
    DevBufs b;
    float *x, *y, *z, *yhat, *dz, *dy, *g, *gp, *lp;
    const size_t parts = kpt::mlp_loss_partials(n);
    KPT_HIP(ctx, b.get(&gp, kpt::mlp_grad_partials(n)));
    KPT_HIP(ctx, b.get(&x, n * 3));
    KPT_HIP(ctx, b.get(&y, n * 2));
    KPT_HIP(ctx, b.get(&z, n * 3 * kpt::kMlpH));
    KPT_HIP(ctx, b.get(&yhat, n * 2));
    KPT_HIP(ctx, b.get(&dz, n * 3 * kpt::kMlpH));
    KPT_HIP(ctx, b.get(&dy, n * 2));
    KPT_HIP(ctx, b.get(&g, kpt::kMlpParams));
    KPT_HIP(ctx, b.get(&lp, parts));
    KPT_HIP(ctx, hipMemcpyAsync(x, feats, n * 3 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    KPT_HIP(ctx, hipMemcpyAsync(y, targets, n * 2 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    const int act = ctx->mlp_hull.activation;
    KPT_HIP(ctx, kpt::launch_mlp_forward(ctx->mlp_params, act, x, n, z, yhat, ctx->stream));
    KPT_HIP(ctx, kpt::launch_mlp_backward(ctx->mlp_params, act, x, z, yhat, y, n, dz, dy, gp, g, lp, ctx->stream));
    std::vector<float> lh(parts);
    KPT_HIP(ctx, hipMemcpyAsync(grad, g, kpt::kMlpParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipMemcpyAsync(lh.data(), lp, parts * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (loss) {
        double t = 0.0;
        for (float v : lh) t += v;
        *loss = t;
    }
    return KPT_OK;
}

kpt_status kpt_mlp_infer(kpt_ctx* ctx, const float* feats, uint64_t n, int32_t mfma, float* out) {
    if (!ctx || !feats || !out) return KPT_ERR_ARG;
    if (!ctx->mlp_valid) return fail(ctx, KPT_ERR_STATE, "no pitch-angle network is loaded.");
    if (n == 0) return KPT_OK;
    (void)hipSetDevice(ctx->device);
    DevBufs b;
    float *x, *o;
    KPT_HIP(ctx, b.get(&x, n * 3));
    KPT_HIP(ctx, b.get(&o, n * 2));
    KPT_HIP(ctx, hipMemcpyAsync(x, feats, n * 3 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    const int act = ctx->mlp_hull.activation;
    if (mfma)
        KPT_HIP(ctx, kpt::launch_mlp_infer(ctx->mlp_image, act, x, n, o, ctx->stream));
    else
        KPT_HIP(ctx, kpt::launch_mlp_forward(ctx->mlp_params, act, x, n, nullptr, o, ctx->stream));
    KPT_HIP(ctx, hipMemcpyAsync(out, o, n * 2 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KPT_OK;
}

kpt_status kpt_mlp_train(kpt_ctx* ctx, const kpt_mlp_train_opts* opts, kpt_mlp_report* report) {
    if (!ctx || !opts) return KPT_ERR_ARG;
    kpt_mlp_hull h = opts->hull;
    h.out_mean = 0.0;
    h.out_scale = 1.0;
    if (!hull_ok(h)) return fail(ctx, KPT_ERR_ARG, "invalid pitch-angle network hull.");
    const uint64_t N = opts->samples;
    const uint32_t B = opts->batch;
    if (B == 0 || B > 65536 || N < B || N > (1ull << 26) || opts->steps == 0 || !(opts->lr > 0.0))
        return fail(ctx, KPT_ERR_ARG, "training needs 1 <= batch <= 65536, batch <= samples <= 2^26, steps > 0, lr > 0.");
    (void)hipSetDevice(ctx->device);
    DevBufs b;
    float *X, *Y, *z, *yhat, *dz, *dy, *g, *gp, *m, *v, *lp, *params;
    double* raw;
    const size_t parts = kpt::mlp_loss_partials(B);
    KPT_HIP(ctx, b.get(&gp, kpt::mlp_grad_partials(B)));
    KPT_HIP(ctx, b.get(&X, N * 3));
    KPT_HIP(ctx, b.get(&Y, N * 2));
    KPT_HIP(ctx, b.get(&raw, N));
    KPT_HIP(ctx, b.get(&z, static_cast<size_t>(B) * 3 * kpt::kMlpH));
    KPT_HIP(ctx, b.get(&yhat, static_cast<size_t>(B) * 2));
    KPT_HIP(ctx, b.get(&dz, static_cast<size_t>(B) * 3 * kpt::kMlpH));
    KPT_HIP(ctx, b.get(&dy, static_cast<size_t>(B) * 2));
    KPT_HIP(ctx, b.get(&g, kpt::kMlpParams));
    KPT_HIP(ctx, b.get(&m, kpt::kMlpParams));
    KPT_HIP(ctx, b.get(&v, kpt::kMlpParams));
    KPT_HIP(ctx, b.get(&lp, parts));
    KPT_HIP(ctx, b.get(&params, kpt::kMlpParams));

    // training set: exact slab-QLT targets, ln R standardised over the set
    KPT_HIP(ctx, kpt::launch_mlp_generate(X, Y, raw, N, opts->seed, 0, h, ctx->stream));
    std::vector<double> lnr(N);
    KPT_HIP(ctx, hipMemcpyAsync(lnr.data(), raw, N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double mean = 0.0, var = 0.0;
    for (double r : lnr) mean += r;
    mean /= static_cast<double>(N);
    for (double r : lnr) var += (r - mean) * (r - mean);
    const double sd = std::sqrt(var / static_cast<double>(N));
    h.out_mean = mean;
    h.out_scale = sd > 0.0 ? sd : 1.0;
    KPT_HIP(ctx, kpt::launch_mlp_normalize(Y, N, static_cast<float>(h.out_mean),
                                           static_cast<float>(1.0 / h.out_scale), ctx->stream));

    // initial weights: uniform +-sqrt(3 / fan_in) from splitmix64(seed), zero biases
    std::vector<float> p0(kpt::kMlpParams, 0.0f);
    uint64_t st = opts->seed ^ 0xA5A5A5A5DEADBEEFull;
    auto init = [&](int off, int rows, int cols) {
        const double a = std::sqrt(3.0 / cols);
        for (int i = 0; i < rows * cols; ++i) {
            const double u = (static_cast<double>(splitmix64(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
            p0[off + i] = static_cast<float>(a * (2.0 * u - 1.0));
        }
    };
    init(kpt::kOffW1, kpt::kMlpH, kpt::kMlpIn);
    init(kpt::kOffW2, kpt::kMlpH, kpt::kMlpH);
    init(kpt::kOffW3, kpt::kMlpH, kpt::kMlpH);
    init(kpt::kOffW4, kpt::kMlpOut, kpt::kMlpH);
    KPT_HIP(ctx, hipMemcpyAsync(params, p0.data(), p0.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    KPT_HIP(ctx, hipMemsetAsync(m, 0, kpt::kMlpParams * sizeof(float), ctx->stream));
    KPT_HIP(ctx, hipMemsetAsync(v, 0, kpt::kMlpParams * sizeof(float), ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));

    KPT_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    const uint64_t batches = N / B;
    double b1c = 1.0, b2c = 1.0;
    for (uint32_t it = 1; it <= opts->steps; ++it) {
        const uint64_t off = ((it - 1) % batches) * B;
        const float lr = static_cast<float>(opts->lr * 0.5 * (1.0 + std::cos(M_PI * it / opts->steps)));
        b1c *= 0.9;
        b2c *= 0.999;
        KPT_HIP(ctx, kpt::launch_mlp_forward(params, h.activation, X + 3 * off, B, z, yhat, ctx->stream));
        KPT_HIP(ctx, kpt::launch_mlp_backward(params, h.activation, X + 3 * off, z, yhat, Y + 2 * off, B, dz, dy, gp,
                                              g, lp, ctx->stream));
        KPT_HIP(ctx, kpt::launch_mlp_adam(params, m, v, g, lr, static_cast<float>(1.0 - b1c),
                                          static_cast<float>(1.0 - b2c), ctx->stream));
    }
    KPT_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<float> lh(parts), pf(kpt::kMlpParams);
    KPT_HIP(ctx, hipMemcpyAsync(lh.data(), lp, parts * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipMemcpyAsync(pf.data(), params, pf.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float train_ms = 0.0f;
    KPT_HIP(ctx, hipEventElapsedTime(&train_ms, ctx->ev0, ctx->ev1));
    for (float w : pf)
        if (!std::isfinite(w)) return fail(ctx, KPT_ERR_DEVICE, "pitch-angle network training diverged.");
    kpt_status s = mlp_install(ctx, pf.data(), h);
    if (s != KPT_OK) return s;

    // held-out check of the installed BF16 MFMA evaluation against exact FP64 targets
    const uint64_t V = 65536;
    float *vx, *vy, *vo;
    double* vraw;
    KPT_HIP(ctx, b.get(&vx, V * 3));
    KPT_HIP(ctx, b.get(&vy, V * 2));
    KPT_HIP(ctx, b.get(&vo, V * 2));
    KPT_HIP(ctx, b.get(&vraw, V));
    KPT_HIP(ctx, kpt::launch_mlp_generate(vx, vy, vraw, V, opts->seed, N, h, ctx->stream));
    KPT_HIP(ctx, kpt::launch_mlp_infer(ctx->mlp_image, h.activation, vx, V, vo, ctx->stream));
    std::vector<double> vr(V);
    std::vector<float> ty(2 * V), to(2 * V);
    KPT_HIP(ctx, hipMemcpyAsync(vr.data(), vraw, V * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipMemcpyAsync(ty.data(), vy, 2 * V * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipMemcpyAsync(to.data(), vo, 2 * V * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (report) {
        std::memset(report, 0, sizeof(*report));
        double loss = 0.0;
        for (float w : lh) loss += w;
        double s_lnr = 0, m_lnr = 0, s_y1 = 0, m_y1 = 0, s_d = 0;
        for (uint64_t i = 0; i < V; ++i) {
            const double e0 = h.out_mean + h.out_scale * to[2 * i] - vr[i];
            const double e1 = static_cast<double>(to[2 * i + 1]) - ty[2 * i + 1];
            const double ed = std::expm1(e0);
            s_lnr += e0 * e0;
            m_lnr = std::max(m_lnr, std::fabs(e0));
            s_y1 += e1 * e1;
            m_y1 = std::max(m_y1, std::fabs(e1));
            s_d += ed * ed;
        }
        report->final_loss = loss;
        report->val_rms_lnr = std::sqrt(s_lnr / V);
        report->val_max_lnr = m_lnr;
        report->val_rms_y1 = std::sqrt(s_y1 / V);
        report->val_max_y1 = m_y1;
        report->val_rms_rel_d = std::sqrt(s_d / V);
        report->train_ms = train_ms;
        report->steps = opts->steps;
        report->validation_samples = static_cast<uint32_t>(V);
    }
    return KPT_OK;
}

}  // extern "C"

namespace {

kpt_status check_generic(kpt_ctx* ctx, const kpt_generic_parker_params* p) {
    if (!p) return fail(ctx, KPT_ERR_ARG, "generic Parker parameters pointer is null.");
    // LegencyModel::DiffusionModelParameters::validate (LegencyModel.hpp:80-105)
    if (!std::isfinite(p->parallel_prefactor) || p->parallel_prefactor <= 0.0 ||
        !std::isfinite(p->perpendicular_prefactor) || p->perpendicular_prefactor <= 0.0 ||
        !std::isfinite(p->perpendicular_parallel_ratio) || p->perpendicular_parallel_ratio < 0.0 ||
        !std::isfinite(p->minimum_magnetic_field_magnitude) || p->minimum_magnetic_field_magnitude < 0.0 ||
        !std::isfinite(p->minimum_solar_wind_speed) || p->minimum_solar_wind_speed < 0.0 ||
        !std::isfinite(p->minimum_sigma2) || p->minimum_sigma2 < 0.0)
        return fail(ctx, KPT_ERR_ARG,
                    "LegencyModel::DiffusionModelParameters: coefficients and floors must be finite positive or "
                    "non-negative values.");
    if (p->perpendicular_model != 0 && p->perpendicular_model != 1)
        return fail(ctx, KPT_ERR_ARG, "invalid perpendicular diffusion model (0 README formula, 1 constant ratio).");
    if (p->perpendicular_model == 1 && p->perpendicular_parallel_ratio <= 0.0)
        return fail(ctx, KPT_ERR_ARG,
                    "LegencyModel::DiffusionModelParameters: constant perpendicular ratio must be positive when "
                    "selected.");
    // TurbulencePropertyModel::validate and ParticleProperties::validate
    if (!std::isfinite(p->sigma2) || p->sigma2 <= 0.0 || !std::isfinite(p->l2d) || p->l2d <= 0.0 ||
        !std::isfinite(p->lslab) || p->lslab <= 0.0 || !std::isfinite(p->lc))
        return fail(ctx, KPT_ERR_ARG, "TurbulencePropertyModel: coefficients must be positive finite values.");
    if (!std::isfinite(p->rest_mass) || p->rest_mass <= 0.0 || !std::isfinite(p->charge) ||
        !std::isfinite(p->speed_of_light) || p->speed_of_light <= 0.0)
        return fail(ctx, KPT_ERR_ARG,
                    "ParticleProperties: rest mass and speed of light must be positive finite dimensionless values, "
                    "charge must be finite, and the energy scale must be positive and finite.");
    if (p->charge == 0.0)
        return fail(ctx, KPT_ERR_ARG, "ParkerSolver: Parker drift requires a finite nonzero particle charge.");
    if (!std::isfinite(p->courant) || p->courant <= 0.0)
        return fail(ctx, KPT_ERR_ARG, "ParkerSolver: Courant scale must be positive and finite.");
    if (!std::isfinite(p->minimum_magnetic_field_squared_for_drift) || p->minimum_magnetic_field_squared_for_drift < 0.0)
        return fail(ctx, KPT_ERR_ARG, "ParkerCoefficient: curl(B/B^2) floor must be finite and non-negative.");
    return KPT_OK;
}

kpt_status generic_slot(kpt_ctx* ctx, int32_t slot, const FieldSlot** out) {
    if (slot != KPT_SLOT_LOWER && slot != KPT_SLOT_UPPER) return fail(ctx, KPT_ERR_ARG, "invalid field slot.");
    const FieldSlot& f = ctx->slot[slot];
    if (!f.valid || !f.gen_valid)
        return fail(ctx, KPT_ERR_STATE, "ParkerSolver: coefficient fields are not configured.");
    *out = &f;
    return KPT_OK;
}

// courant * min over alive particles; 0 when none gives a finite bound.
kpt_status generic_dt(kpt_ctx* ctx, const kpt_generic_parker_params* p, const FieldSlot& f, double* dt) {
    const double large = 1.0e300;
    unsigned long long bits;
    std::memcpy(&bits, &large, 8);
    KPT_HIP(ctx, hipMemcpyAsync(ctx->counters + 6, &bits, 8, hipMemcpyHostToDevice, ctx->stream));
    KPT_HIP(ctx, kpt::launch_generic_min_dt(ctx->P, ctx->count, f.gen, f.g, *p, ctx->counters + 6, ctx->stream));
    KPT_HIP(ctx, hipMemcpyAsync(&bits, ctx->counters + 6, 8, hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double m;
    std::memcpy(&m, &bits, 8);
    *dt = (!std::isfinite(m) || m >= 0.5 * large) ? 0.0 : p->courant * m;
    return KPT_OK;
}

}  // namespace

extern "C" {

kpt_status kpt_generic_parker_default_params(kpt_generic_parker_params* p) {
    if (!p) return KPT_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    p->perpendicular_model = 0;
    p->parallel_prefactor = 1.622;   // DiffusionModelParameters (LegencyModel.hpp:27-35)
    p->perpendicular_prefactor = 0.198;
    p->sigma2 = 0.5;                 // TurbulencePropertyModel (TurbulenceProperties.hpp:59-70)
    p->l2d = 1.0;
    p->lslab = 1.0;
    p->lc = -1.0;
    p->rest_mass = 1.0;              // ParticleProperties (ParticleSystem.hpp:145-150)
    p->charge = 1.0;
    p->speed_of_light = 1.0;
    p->courant = 0.3;                // SolverBase (SolverBase.hpp:156)
    return KPT_OK;
}

kpt_status kpt_generic_parker_build(kpt_ctx* ctx, const kpt_generic_parker_params* p, int32_t slot) {
    if (!ctx) return KPT_ERR_ARG;
    kpt_status st = check_generic(ctx, p);
    if (st != KPT_OK) return st;
    if (slot != KPT_SLOT_LOWER && slot != KPT_SLOT_UPPER) return fail(ctx, KPT_ERR_ARG, "invalid field slot.");
    FieldSlot& f = ctx->slot[slot];
    if (!f.valid) return fail(ctx, KPT_ERR_STATE, "field slot is empty.");
    (void)hipSetDevice(ctx->device);
    const size_t elems = kpt::coefficient_plane(f.g.nx, f.g.ny) * KPT_GENERIC_PARKER_COMPONENTS;
    if (elems > f.gen_cap) {
        dev_free(f.gen);
        f.gen = nullptr;
        f.gen_cap = 0;
        KPT_HIP(ctx, dev_alloc(&f.gen, elems));
        f.gen_cap = elems;
    }
    f.gen_valid = false;
    int* bad = reinterpret_cast<int*>(ctx->counters + 7);
    KPT_HIP(ctx, hipMemsetAsync(bad, 0, sizeof(int), ctx->stream));
    KPT_HIP(ctx, kpt::launch_generic_pack(f.raw, f.g, *p, f.gen, bad, ctx->stream));
    int h = 0;
    KPT_HIP(ctx, hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (h)
        return fail(ctx, KPT_ERR_ARG,
                    "ParkerCoefficient / LegencyModel: |B| must be positive (or floored) and V, Lc, Omega and sigma2 "
                    "positive wherever the coefficients are evaluated.");
    f.gen_valid = true;
    return KPT_OK;
}

kpt_status kpt_generic_parker_download(kpt_ctx* ctx, int32_t slot, double* out, size_t out_len) {
    if (!ctx || !out) return KPT_ERR_ARG;
    const FieldSlot* f;
    kpt_status st = generic_slot(ctx, slot, &f);
    if (st != KPT_OK) return st;
    const size_t plane = kpt::coefficient_plane(f->g.nx, f->g.ny);
    const int C = KPT_GENERIC_PARKER_COMPONENTS;
    if (out_len < plane * C) return fail(ctx, KPT_ERR_ARG, "output buffer too small.");
    std::vector<double> tmp(plane * C);
    (void)hipSetDevice(ctx->device);
    KPT_HIP(ctx, hipMemcpy(tmp.data(), f->gen, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t n = 0; n < plane; ++n)
        for (int k = 0; k < C; ++k) out[n * C + k] = tmp[k * plane + n];
    return KPT_OK;
}

kpt_status kpt_generic_parker_time_step(kpt_ctx* ctx, const kpt_generic_parker_params* p, int32_t slot,
                                        double* dt) {
    if (!ctx || !dt) return KPT_ERR_ARG;
    kpt_status st = check_generic(ctx, p);
    if (st != KPT_OK) return st;
    const FieldSlot* f;
    st = generic_slot(ctx, slot, &f);
    if (st != KPT_OK) return st;
    (void)hipSetDevice(ctx->device);
    return generic_dt(ctx, p, *f, dt);
}

}  // extern "C"

namespace {

// ParkerSolver::advance_stochastic / advance_deterministic (include/
// ParkerSolver.hpp:228-299): dt, then boundary wrap, capture of the previous
// positions and the step in one kernel.
kpt_status generic_advance(kpt_ctx* ctx, const kpt_generic_parker_params* p, int32_t slot, uint32_t step,
                           bool deterministic, double* dt_used, uint64_t* failures) {
    if (!ctx) return KPT_ERR_ARG;
    kpt_status st = check_generic(ctx, p);
    if (st != KPT_OK) return st;
    const FieldSlot* f;
    st = generic_slot(ctx, slot, &f);
    if (st != KPT_OK) return st;
    if (!deterministic && ctx->noise && (ctx->noise_particles < ctx->count || step >= ctx->noise_steps))
        return fail(ctx, KPT_ERR_STATE, "noise table does not cover this step of every stored particle.");
    (void)hipSetDevice(ctx->device);
    double dt = 0.0;
    st = generic_dt(ctx, p, *f, &dt);
    if (st != KPT_OK) return st;
    if (!std::isfinite(dt) || dt <= 0.0)
        return fail(ctx, KPT_ERR_STATE, deterministic
                                            ? "ParkerSolver: adaptive deterministic time step is not positive and finite."
                                            : "ParkerSolver: adaptive stochastic time step is not positive and finite.");
    const uint32_t key[2] = {static_cast<uint32_t>(ctx->seed), static_cast<uint32_t>(ctx->seed >> 32)};
    KPT_HIP(ctx, hipMemsetAsync(ctx->counters + 5, 0, sizeof(unsigned long long), ctx->stream));
    KPT_HIP(ctx, kpt::launch_generic_step(ctx->P, ctx->count, f->gen, f->g, *p, dt, step, key, ctx->noise,
                                          ctx->noise_steps, deterministic ? 1 : 0, ctx->counters + 5, ctx->stream));
    unsigned long long nf = 0;
    KPT_HIP(ctx, hipMemcpyAsync(&nf, ctx->counters + 5, 8, hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (dt_used) *dt_used = dt;
    if (failures) *failures = nf;
    return KPT_OK;
}

}  // namespace

extern "C" {

kpt_status kpt_generic_parker_step(kpt_ctx* ctx, const kpt_generic_parker_params* p, int32_t slot, uint32_t step,
                                   double* dt_used, uint64_t* failures) {
    return generic_advance(ctx, p, slot, step, false, dt_used, failures);
}

kpt_status kpt_generic_parker_advance_deterministic(kpt_ctx* ctx, const kpt_generic_parker_params* p, int32_t slot,
                                                    double* dt_used, uint64_t* failures) {
    return generic_advance(ctx, p, slot, 0, true, dt_used, failures);
}

}  // extern "C"

// ---- ParticleSystem lifecycle -------------------------------------------------

// ParticleProperties::validate (include/ParticleSystem.hpp:155-165).
kpt_status kpt_capi::check_properties(kpt_ctx* ctx, const kpt_particle_properties* pr) {
    if (!pr) return fail(ctx, KPT_ERR_ARG, "particle properties pointer is null.");
    if (!std::isfinite(pr->rest_mass) || pr->rest_mass <= 0.0 || !std::isfinite(pr->charge) ||
        !std::isfinite(pr->speed_of_light) || pr->speed_of_light <= 0.0 || !std::isfinite(pr->energy_scale_erg) ||
        pr->energy_scale_erg <= 0.0)
        return fail(ctx, KPT_ERR_ARG,
                    "ParticleProperties: rest mass and speed of light must be positive finite dimensionless values, "
                    "charge must be finite, and the energy scale must be positive and finite.");
    return KPT_OK;
}

namespace {


}  // namespace

extern "C" {

kpt_status kpt_default_particle_properties(kpt_particle_properties* out) {
    if (!out) return KPT_ERR_ARG;
    out->rest_mass = 1.0;
    out->charge = 1.0;
    out->speed_of_light = 1.0;
    out->energy_scale_erg = 1.0;
    return KPT_OK;
}

kpt_status kpt_default_energy_split_policy(kpt_energy_split_policy* out) {
    if (!out) return KPT_ERR_ARG;
    out->energy_split_ratio = 2.0;
    out->minimum_child_weight = 0.0;
    return KPT_OK;
}

kpt_status kpt_split_energy_growth(kpt_ctx* ctx, const kpt_particle_properties* props,
                                   const kpt_energy_split_policy* policy, kpt_split_statistics* stats) {
    if (!ctx) return KPT_ERR_ARG;
    kpt_status st = check_properties(ctx, props);
    if (st != KPT_OK) return st;
    if (!policy) return fail(ctx, KPT_ERR_ARG, "split policy pointer is null.");
    // ParticleSplittingPolicy::validate (include/ParticleSystem.hpp:287-294)
    if (!std::isfinite(policy->energy_split_ratio) || policy->energy_split_ratio <= 1.0 ||
        !std::isfinite(policy->minimum_child_weight) || policy->minimum_child_weight < 0.0)
        return fail(ctx, KPT_ERR_ARG,
                    "ParticleSplittingPolicy: energy split ratio must be greater than one and minimum child weight "
                    "must be finite and non-negative.");
    kpt_split_statistics out{};
    if (ctx->count == 0 || ctx->count >= ctx->capacity) {
        if (stats) *stats = out;
        return KPT_OK;
    }
    (void)hipSetDevice(ctx->device);
    KPT_HIP(ctx, hipMemsetAsync(ctx->counters, 0, 5 * sizeof(unsigned long long), ctx->stream));
    KPT_HIP(ctx, kpt::launch_energy_split(ctx->P, ctx->count, ctx->capacity, ctx->next_root, *props,
                                          policy->energy_split_ratio, std::log(policy->energy_split_ratio),
                                          policy->minimum_child_weight, ctx->split_scratch, ctx->counters,
                                          ctx->stream));
    unsigned long long c[5];
    KPT_HIP(ctx, hipMemcpyAsync(c, ctx->counters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    out.active_count = c[0];
    out.split_count = c[1];
    out.skipped_capacity_count = c[2];
    out.skipped_invalid_baseline_count = c[3];
    out.skipped_minimum_weight_count = c[4];
    ctx->count += c[1];
    ctx->next_root += c[1];
    if (stats) *stats = out;
    return KPT_OK;
}

kpt_status kpt_reset_split_baselines(kpt_ctx* ctx, const kpt_particle_properties* props) {
    if (!ctx) return KPT_ERR_ARG;
    kpt_status st = check_properties(ctx, props);
    if (st != KPT_OK) return st;
    (void)hipSetDevice(ctx->device);
    KPT_HIP(ctx, kpt::launch_reset_baselines(ctx->P, ctx->count, *props, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KPT_OK;
}

kpt_status kpt_apply_boundaries(kpt_ctx* ctx, int32_t slot, int32_t periodic_mask, uint64_t* escaped) {
    if (!ctx) return KPT_ERR_ARG;
    if (slot != KPT_SLOT_LOWER && slot != KPT_SLOT_UPPER) return fail(ctx, KPT_ERR_ARG, "invalid field slot.");
    if (periodic_mask < 0 || periodic_mask > 3) return fail(ctx, KPT_ERR_ARG, "periodic mask must be in [0, 3].");
    const FieldSlot& f = ctx->slot[slot];
    if (!f.valid) return fail(ctx, KPT_ERR_STATE, "field slot is empty.");
    (void)hipSetDevice(ctx->device);
    KPT_HIP(ctx, hipMemsetAsync(ctx->counters, 0, sizeof(unsigned long long), ctx->stream));
    KPT_HIP(ctx, kpt::launch_boundaries(ctx->P, ctx->count, f.g, periodic_mask, ctx->counters, ctx->stream));
    unsigned long long e = 0;
    KPT_HIP(ctx, hipMemcpyAsync(&e, ctx->counters, sizeof(e), hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (escaped) *escaped = e;
    return KPT_OK;
}

kpt_status kpt_particle_kinematics(kpt_ctx* ctx, const kpt_particle_properties* props, double* gamma,
                                   double* speed, double* kinetic_energy, uint64_t count) {
    if (!ctx) return KPT_ERR_ARG;
    kpt_status st = check_properties(ctx, props);
    if (st != KPT_OK) return st;
    if (count > ctx->count) return fail(ctx, KPT_ERR_ARG, "count exceeds stored particles.");
    if (count == 0) return KPT_OK;
    (void)hipSetDevice(ctx->device);
    double* out[3] = {gamma, speed, kinetic_energy};
    double* dev[3] = {nullptr, nullptr, nullptr};
    kpt_status rc = KPT_OK;
    for (int k = 0; k < 3 && rc == KPT_OK; ++k) {
        if (!out[k]) continue;
        const hipError_t e = dev_alloc(&dev[k], count);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "hipMalloc");
    }
    if (rc == KPT_OK) {
        hipError_t e = kpt::launch_kinematics(ctx->P.p, count, *props, dev[0], dev[1], dev[2], ctx->stream);
        for (int k = 0; k < 3 && e == hipSuccess; ++k)
            if (out[k]) e = hipMemcpyAsync(out[k], dev[k], count * 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "particle kinematics");
    }
    for (double* d : dev) dev_free(d);
    return rc;
}

}  // extern "C"

// ---- generic focused-transport solver ------------------------------------------

namespace {

kpt_status check_focus(kpt_ctx* ctx, const kpt_generic_focus_params* p) {
    if (!p) return fail(ctx, KPT_ERR_ARG, "generic focused parameters pointer is null.");
    kpt_status st = check_generic(ctx, &p->model);
    if (st != KPT_OK) return st;
    if (!std::isfinite(p->minimum_magnetic_field_magnitude) || p->minimum_magnetic_field_magnitude < 0.0)
        return fail(ctx, KPT_ERR_ARG,
                    "FocusCoefficient: magnetic-field magnitude floor must be finite and non-negative.");
    // PitchAngleScatteringParameters::validate (include/LegencyModel.hpp:206-222)
    if (!std::isfinite(p->spectral_index) || p->spectral_index <= 1.0 ||
        !std::isfinite(p->resonance_regularization_h0) || p->resonance_regularization_h0 <= 0.0 ||
        !std::isfinite(p->scattering_minimum_magnetic_field_magnitude) ||
        p->scattering_minimum_magnetic_field_magnitude < 0.0 || !std::isfinite(p->scattering_minimum_sigma2) ||
        p->scattering_minimum_sigma2 < 0.0 || !std::isfinite(p->scattering_minimum_correlation_length) ||
        p->scattering_minimum_correlation_length < 0.0 || !std::isfinite(p->scattering_minimum_speed) ||
        p->scattering_minimum_speed < 0.0)
        return fail(ctx, KPT_ERR_ARG,
                    "LegencyModel::PitchAngleScatteringParameters: coefficients and floors must be finite, with "
                    "spectral_index > 1 and resonance_regularization_h0 > 0.");
    if (!std::isfinite(p->maximum_pitch_angle_cosine) || p->maximum_pitch_angle_cosine <= 0.0 ||
        p->maximum_pitch_angle_cosine > 1.0)
        return fail(ctx, KPT_ERR_ARG, "FocusTransportSolver: maximum |mu| must satisfy 0 < value <= 1.");
    return KPT_OK;
}

kpt_status focus_slot(kpt_ctx* ctx, int32_t slot, const FieldSlot** out) {
    if (slot != KPT_SLOT_LOWER && slot != KPT_SLOT_UPPER) return fail(ctx, KPT_ERR_ARG, "invalid field slot.");
    const FieldSlot& f = ctx->slot[slot];
    if (!f.valid || !f.foc_valid)
        return fail(ctx, KPT_ERR_STATE, "FocusTransportSolver: coefficient fields are not configured.");
    *out = &f;
    return KPT_OK;
}

kpt_status focus_dt(kpt_ctx* ctx, const kpt_generic_focus_params* p, const FieldSlot& f, double* dt) {
    const double large = 1.0e300;
    unsigned long long bits;
    std::memcpy(&bits, &large, 8);
    KPT_HIP(ctx, hipMemcpyAsync(ctx->counters + 6, &bits, 8, hipMemcpyHostToDevice, ctx->stream));
    KPT_HIP(ctx, kpt::launch_focus_min_dt(ctx->P, ctx->count, f.foc, f.g, *p, ctx->counters + 6, ctx->stream));
    KPT_HIP(ctx, hipMemcpyAsync(&bits, ctx->counters + 6, 8, hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double m;
    std::memcpy(&m, &bits, 8);
    *dt = (!std::isfinite(m) || m >= 0.5 * large) ? 0.0 : p->model.courant * m;
    return KPT_OK;
}

}  // namespace

extern "C" {

kpt_status kpt_generic_focus_default_params(kpt_generic_focus_params* p) {
    if (!p) return KPT_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    kpt_generic_parker_default_params(&p->model);
    p->spectral_index = 5.0 / 3.0;             // PitchAngleScatteringParameters (LegencyModel.hpp:196-203)
    p->resonance_regularization_h0 = 1.0e-6;
    p->maximum_pitch_angle_cosine = 1.0;       // FocusTransportSolver (FocusSolver.hpp:145)
    return KPT_OK;
}

kpt_status kpt_generic_focus_build(kpt_ctx* ctx, const kpt_generic_focus_params* p, int32_t slot,
                                   const double* dvdt) {
    if (!ctx) return KPT_ERR_ARG;
    kpt_status st = check_focus(ctx, p);
    if (st != KPT_OK) return st;
    if (slot != KPT_SLOT_LOWER && slot != KPT_SLOT_UPPER) return fail(ctx, KPT_ERR_ARG, "invalid field slot.");
    FieldSlot& f = ctx->slot[slot];
    if (!f.valid) return fail(ctx, KPT_ERR_STATE, "field slot is empty.");
    (void)hipSetDevice(ctx->device);
    const size_t elems = kpt::coefficient_plane(f.g.nx, f.g.ny) * KPT_GENERIC_FOCUS_COMPONENTS;
    if (elems > f.foc_cap) {
        dev_free(f.foc);
        f.foc = nullptr;
        f.foc_cap = 0;
        KPT_HIP(ctx, dev_alloc(&f.foc, elems));
        f.foc_cap = elems;
    }
    f.foc_valid = false;
    double* dv = nullptr;
    const size_t cells = static_cast<size_t>(f.g.nx) * f.g.ny;
    if (dvdt) {
        KPT_HIP(ctx, dev_alloc(&dv, 3 * cells));
        const hipError_t e = hipMemcpy(dv, dvdt, 3 * cells * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            dev_free(dv);
            return hip_fail(ctx, e, "hipMemcpy");
        }
    }
    int* bad = reinterpret_cast<int*>(ctx->counters + 7);
    int h = 0;
    hipError_t e = hipMemsetAsync(bad, 0, sizeof(int), ctx->stream);
    if (e == hipSuccess) e = kpt::launch_focus_pack(f.raw, dv, f.g, *p, f.foc, bad, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dev_free(dv);
    if (e != hipSuccess) return hip_fail(ctx, e, "generic focused basis");
    if (h)
        return fail(ctx, KPT_ERR_ARG,
                    "FocusCoefficient / LegencyModel: |B| must exceed the floor and V, Lc, Omega and sigma2 must be "
                    "positive wherever the coefficients are evaluated.");
    f.foc_valid = true;
    return KPT_OK;
}

kpt_status kpt_generic_focus_download(kpt_ctx* ctx, int32_t slot, double* out, size_t out_len) {
    if (!ctx || !out) return KPT_ERR_ARG;
    const FieldSlot* f;
    kpt_status st = focus_slot(ctx, slot, &f);
    if (st != KPT_OK) return st;
    const size_t plane = kpt::coefficient_plane(f->g.nx, f->g.ny);
    const int C = KPT_GENERIC_FOCUS_COMPONENTS;
    if (out_len < plane * C) return fail(ctx, KPT_ERR_ARG, "output buffer too small.");
    std::vector<double> tmp(plane * C);
    (void)hipSetDevice(ctx->device);
    KPT_HIP(ctx, hipMemcpy(tmp.data(), f->foc, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t n = 0; n < plane; ++n)
        for (int k = 0; k < C; ++k) out[n * C + k] = tmp[k * plane + n];
    return KPT_OK;
}

kpt_status kpt_generic_focus_time_step(kpt_ctx* ctx, const kpt_generic_focus_params* p, int32_t slot,
                                       double* dt) {
    if (!ctx || !dt) return KPT_ERR_ARG;
    kpt_status st = check_focus(ctx, p);
    if (st != KPT_OK) return st;
    const FieldSlot* f;
    st = focus_slot(ctx, slot, &f);
    if (st != KPT_OK) return st;
    (void)hipSetDevice(ctx->device);
    return focus_dt(ctx, p, *f, dt);
}

}  // extern "C"

namespace {

// FocusTransportSolver::advance_stochastic / advance_deterministic (include/
// FocusSolver.hpp:251-290).
kpt_status focus_advance(kpt_ctx* ctx, const kpt_generic_focus_params* p, int32_t slot, uint32_t step,
                         bool deterministic, double* dt_used, uint64_t* failures) {
    if (!ctx) return KPT_ERR_ARG;
    kpt_status st = check_focus(ctx, p);
    if (st != KPT_OK) return st;
    const FieldSlot* f;
    st = focus_slot(ctx, slot, &f);
    if (st != KPT_OK) return st;
    if (!deterministic && ctx->noise && (ctx->noise_particles < ctx->count || step >= ctx->noise_steps))
        return fail(ctx, KPT_ERR_STATE, "noise table does not cover this step of every stored particle.");
    (void)hipSetDevice(ctx->device);
    double dt = 0.0;
    st = focus_dt(ctx, p, *f, &dt);
    if (st != KPT_OK) return st;
    if (!std::isfinite(dt) || dt <= 0.0)
        return fail(ctx, KPT_ERR_STATE,
                    deterministic ? "FocusTransportSolver: adaptive deterministic time step is not positive and finite."
                                  : "FocusTransportSolver: adaptive stochastic time step is not positive and finite.");
    const uint32_t key[2] = {static_cast<uint32_t>(ctx->seed), static_cast<uint32_t>(ctx->seed >> 32)};
    KPT_HIP(ctx, hipMemsetAsync(ctx->counters + 5, 0, sizeof(unsigned long long), ctx->stream));
    KPT_HIP(ctx, kpt::launch_focus_step(ctx->P, ctx->count, f->foc, f->g, *p, dt, step, key, ctx->noise,
                                        ctx->noise_steps, deterministic ? 1 : 0, ctx->counters + 5, ctx->stream));
    unsigned long long nf = 0;
    KPT_HIP(ctx, hipMemcpyAsync(&nf, ctx->counters + 5, 8, hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (dt_used) *dt_used = dt;
    if (failures) *failures = nf;
    return KPT_OK;
}

}  // namespace

extern "C" {

kpt_status kpt_generic_focus_step(kpt_ctx* ctx, const kpt_generic_focus_params* p, int32_t slot, uint32_t step,
                                  double* dt_used, uint64_t* failures) {
    return focus_advance(ctx, p, slot, step, false, dt_used, failures);
}

kpt_status kpt_generic_focus_advance_deterministic(kpt_ctx* ctx, const kpt_generic_focus_params* p, int32_t slot,
                                                   double* dt_used, uint64_t* failures) {
    return focus_advance(ctx, p, slot, 0, true, dt_used, failures);
}

}  // extern "C"

