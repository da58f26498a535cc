// kpt_noalloc.hip — kpt_push_interval_noalloc: the interval push of
// kpt_push_interval with its per-chunk launch arguments and counters in a
// file-scope device array, so that a call allocates no device memory.
//
// The array `scratch` (one copy per GPU, held until the process ends):
//   [0, kMaxChunks * kArgsStride)  one PushArgs record per push chunk (the
//                                  device copy the FP32 push's near-null
//                                  sub-steps read through PushArgs::self)
//   then kMaxChunks x 4 counters   particle steps, failures, active at entry
// Host code never uses the array's name as an address: the set-up below looks
// up each device's copy with hipGetSymbolAddress and keeps that pointer.
// Every call on a GPU shares that GPU's array and overwrites what the last call
// left there; calls that overlap on one GPU (other contexts, other host threads)
// race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kpt_capi_internal.h"

using namespace kpt_capi;

namespace {

constexpr int kMaxChunks = KPT_NOALLOC_MAX_CHUNKS;
constexpr size_t kArgsStride = (sizeof(kpt::PushArgs) + 255) / 256 * 256;
constexpr int kCounterWords = 4;
constexpr size_t kCountersOffset = kMaxChunks * kArgsStride;
constexpr size_t kScratchBytes = kCountersOffset + kMaxChunks * kCounterWords * sizeof(unsigned long long);
static_assert(kScratchBytes <= KPT_NOALLOC_SCRATCH_BYTES, "the scratch layout outgrew its published size");

static __device__ __align__(256) unsigned char scratch[KPT_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};
hipError_t g_setup_error[kMaxDevices] = {};

// One chunk's launch arguments, passed by value, into their record; clears the
// chunk's counters (a kernel rather than a memset, so that the clear also holds
// when a caller captures the call into a graph and replays it).
__global__ void noalloc_stage_kernel(kpt::PushArgs a, kpt::PushArgs* record, unsigned long long* counters) {
    if (threadIdx.x == 0) *record = a;
    if (threadIdx.x < kCounterWords) counters[threadIdx.x] = 0ull;
}

}  // namespace

hipError_t kpt_capi::noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[device] != nullptr) return hipSuccess;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p == nullptr) e = hipErrorInvalidSymbol;
    g_setup_error[device] = e;
    if (e == hipSuccess) g_scratch[device] = static_cast<unsigned char*>(p);
    return e;
}

kpt_status kpt_push_interval_noalloc(kpt_ctx* ctx, const kpt_device_settings* s, double t0, uint32_t frame,
                                     kpt_push_stats* stats) {
    if (!ctx) return KPT_ERR_ARG;
    if (!s) return fail(ctx, KPT_ERR_ARG, "settings pointer is null.");
    const int chunks = s->push_chunks > 1 ? s->push_chunks : 1;
    if (chunks > kMaxChunks)
        return fail(ctx, KPT_ERR_ARG, "kpt_push_interval_noalloc: push_chunks " + std::to_string(chunks) +
                                          " exceeds the " + std::to_string(kMaxChunks) +
                                          " chunk records of its device scratch array.");
    unsigned char* base = nullptr;
    hipError_t setup = hipErrorInvalidDevice;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices) {
            base = g_scratch[ctx->device];
            setup = g_setup_error[ctx->device];
        }
    }
    if (base == nullptr)
        return fail(ctx, KPT_ERR_DEVICE, std::string("kpt_push_interval_noalloc: the device scratch array has no "
                                                     "address on this device (hipGetSymbolAddress: ") +
                                             hipGetErrorString(setup) + ").");
    kpt::PushArgs a;
    bool focused = false;
    kpt_status st = push_args(ctx, s, t0, frame, false, a, focused);
    if (st != KPT_OK) return st;
    const int model = ctx->slot[KPT_SLOT_LOWER].model;
    const bool fp32 = s->precision == KPT_PREC_FP32;
    (void)hipSetDevice(ctx->device);
    auto* counters = reinterpret_cast<unsigned long long*>(base + kCountersOffset);
    if (stats) KPT_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int k = 0; k < chunks; ++k) {
        auto* record = reinterpret_cast<kpt::PushArgs*>(base + static_cast<size_t>(k) * kArgsStride);
        a.P = ctx->P;
        a.first_chunk = k == 0 ? 1 : 0;
        a.t_stop = k == chunks - 1 ? a.t_end : t0 + s->mhd_dt * static_cast<double>(k + 1) / chunks;
        a.counters = counters + static_cast<size_t>(k) * kCounterWords;
        a.self = record;
        hipLaunchKernelGGL(noalloc_stage_kernel, dim3(1), dim3(64), 0, ctx->stream, a, record, a.counters);
        KPT_HIP(ctx, hipGetLastError());
        KPT_HIP(ctx, kpt::launch_push(a, model, focused, fp32, s->closure, ctx->stream));
    }
    if (!stats) {
        KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return KPT_OK;
    }
    KPT_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    unsigned long long c[kMaxChunks * kCounterWords];
    KPT_HIP(ctx, hipMemcpyAsync(c, counters, static_cast<size_t>(chunks) * kCounterWords * sizeof(c[0]),
                                hipMemcpyDeviceToHost, ctx->stream));
    KPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    KPT_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    stats->particle_steps = 0;
    stats->failures = 0;
    for (int k = 0; k < chunks; ++k) {
        stats->particle_steps += c[k * kCounterWords + 0];
        stats->failures += c[k * kCounterWords + 1];
    }
    stats->active_in = c[2];
    stats->kernel_ms = ms;
    stats->launches = ctx->count > 0 ? static_cast<uint32_t>(chunks) : 0u;
    stats->reserved = 0;
    return KPT_OK;
}
