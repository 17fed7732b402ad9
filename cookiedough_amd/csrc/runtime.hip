// runtime.hip -- context, errors, device memory, events.
#include "ckd_internal.h"

#include <atomic>

#include <cmath>
#include <cstdio>
#include <cstring>

namespace {
thread_local char g_lastError[512] = "";
}

namespace ckd {

ckd_status fail(ckd_status s, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_lastError, sizeof(g_lastError), fmt, ap);
    va_end(ap);
    return s;
}

// the allocation that fails next: -1 none, 0 the next one, n the one after n
std::atomic<int> g_failAlloc{-1};

hipError_t dev_alloc_raw(void **p, size_t bytes)
{
    int n = g_failAlloc.load();
    while (n >= 0 && !g_failAlloc.compare_exchange_weak(n, n - 1)) {
    }
    if (n == 0) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

ckd_status check_launch(const char *what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(CKD_EDEVICE, "%s: launch failed: %s", what, hipGetErrorString(e));
    return CKD_OK;
}

ckd_status stage(ckd_ctx *ctx, const void *host, size_t bytes, void **dev)
{
    if (bytes > ctx->stagingCap) {
        // the previous staging buffer may still be read by queued work
        if (ctx->staging) {
            CKD_HIP(hipStreamSynchronize(ctx->stream));
            CKD_HIP(hipFree(ctx->staging));
            ctx->staging = nullptr;
            ctx->stagingCap = 0;
        }
        size_t cap = 4096;
        while (cap < bytes) cap *= 2;
        if (ckd::dev_alloc(&ctx->staging, cap) != hipSuccess) {
            ctx->staging = nullptr;
            return fail(CKD_ENOMEM, "stage: ckd::dev_alloc(%zu) failed", cap);
        }
        ctx->stagingCap = cap;
    }
    // Reuse needs no wait: the copy is ordered on the context stream after
    // every kernel that read the previous contents.
    CKD_HIP(hipMemcpyAsync(ctx->staging, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    *dev = ctx->staging;
    return CKD_OK;
}

KernelTimer::KernelTimer(ckd_ctx *ctx, int kernel) : ctx_(ctx), kernel_(kernel)
{
    if (!ctx->timing || kernel < 0 || kernel >= CKD_KT_COUNT) return;
    hipEvent_t e = nullptr;
    if (!ctx->eventPool.empty()) {
        e = ctx->eventPool.back();
        ctx->eventPool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        return;
    }
    if (hipEventRecord(e, ctx->stream) != hipSuccess) {
        ctx->eventPool.push_back(e);
        return;
    }
    begin_ = e;
}

KernelTimer::~KernelTimer()
{
    if (!begin_) return;
    hipEvent_t e = nullptr;
    if (!ctx_->eventPool.empty()) {
        e = ctx_->eventPool.back();
        ctx_->eventPool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        ctx_->eventPool.push_back(begin_);
        return;
    }
    (void)hipEventRecord(e, ctx_->stream);
    ctx_->ktPending[kernel_].push_back({begin_, e, true});
}

ckd_status drain_timing(ckd_ctx *ctx)
{
    for (int k = 0; k < CKD_KT_COUNT; ++k) {
        for (auto &pr : ctx->ktPending[k]) {
            CKD_HIP(hipEventSynchronize(pr.end));
            float ms = 0.f;
            CKD_HIP(hipEventElapsedTime(&ms, pr.begin, pr.end));
            ctx->ktMs[k] += ms;
            ctx->ktLaunches[k] += 1;
            if (pr.pooled) {
                ctx->eventPool.push_back(pr.begin);
                ctx->eventPool.push_back(pr.end);
            }
        }
        ctx->ktPending[k].clear();
    }
    return CKD_OK;
}

} // namespace ckd

namespace {
using ckd::drain_timing;
} // namespace

extern "C" {

ckd_status ckd_debug_fail_alloc(int n)
{
    CKD_REQUIRE(n >= -1, "ckd_debug_fail_alloc: n %d below -1", n);
    ckd::g_failAlloc.store(n);
    return CKD_OK;
}

ckd_status ckd_kernel_timing(ckd_ctx *ctx, int enable)
{
    CKD_REQUIRE_CTX(ctx);
    ckd_status s = drain_timing(ctx);
    if (s != CKD_OK) return s;
    if (enable) {
        for (int k = 0; k < CKD_KT_COUNT; ++k) {
            ctx->ktMs[k] = 0.0;
            ctx->ktLaunches[k] = 0;
        }
    }
    ctx->timing = enable != 0;
    return CKD_OK;
}

ckd_status ckd_kernel_timing_read(ckd_ctx *ctx, int kernel, double *totalMs,
                                  unsigned long long *launches)
{
    CKD_REQUIRE(kernel >= 0 && kernel < CKD_KT_COUNT && totalMs && launches,
                "ckd_kernel_timing_read: bad kernel id %d or null output", kernel);
    CKD_REQUIRE_CTX(ctx);
    ckd_status s = drain_timing(ctx);
    if (s != CKD_OK) return s;
    *totalMs = ctx->ktMs[kernel];
    *launches = ctx->ktLaunches[kernel];
    return CKD_OK;
}


const char *ckd_last_error(void) { return g_lastError; }

ckd_status ckd_create(int device, ckd_ctx **out)
{
    CKD_REQUIRE(out != nullptr, "ckd_create: null output pointer");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return ckd::fail(CKD_EDEVICE, "ckd_create: no HIP device available");
    CKD_REQUIRE(device >= 0 && device < n, "ckd_create: device %d out of range (%d devices)",
                device, n);
    CKD_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    CKD_HIP(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return ckd::fail(CKD_EDEVICE, "ckd_create: device %d is %s, this library is built for gfx950",
                         device, prop.gcnArchName);
    {
        // the scratch array of ckd_tunnel_frame_noalloc, once per device
        const ckd_status s = ckd::noalloc_scratch_setup(device);
        if (s != CKD_OK) return s;
    }
    ckd_ctx *ctx = new ckd_ctx();
    ctx->device = device;
    ctx->cuCount = prop.multiProcessorCount;
    for (int i = 0; i < CKD_NUM_EVENTS; ++i) {
        if (hipEventCreate(&ctx->events[i]) != hipSuccess) {
            ckd_destroy(ctx);
            return ckd::fail(CKD_EDEVICE, "ckd_create: hipEventCreate failed");
        }
    }
    {
        // CalculateCosLUT (code/sincos-lut.cpp:9-16) with the host libm
        const float kPI = 3.1415926535897932384626433832795f;
        const float k2PI = 2.f * kPI;
        float lut[2049];
        for (unsigned i = 0; i < 2048; ++i) lut[i] = cosf((float)i * (k2PI / 2048.f));
        lut[2048] = lut[0];
        for (unsigned i = 0; i < 2049; ++i) ctx->hostLut[i] = lut[i];
        if (ckd::dev_alloc(&ctx->cosLut, sizeof(lut)) != hipSuccess ||
            hipMemcpy(ctx->cosLut, lut, sizeof(lut), hipMemcpyHostToDevice) != hipSuccess) {
            ckd_destroy(ctx);
            return ckd::fail(CKD_ENOMEM, "ckd_create: cosine LUT upload failed");
        }
    }
    if (hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->aux2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->forkEv, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->joinEv, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->tunEv, hipEventDisableTiming) != hipSuccess) {
        ckd_destroy(ctx);
        return ckd::fail(CKD_EDEVICE, "ckd_create: stream/event creation failed");
    }
    *out = ctx;
    return CKD_OK;
}

void ckd_destroy(ckd_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (int i = 0; i < CKD_NUM_EVENTS; ++i)
        if (ctx->events[i]) (void)hipEventDestroy(ctx->events[i]);
    for (int k = 0; k < CKD_KT_COUNT; ++k)
        for (auto &pr : ctx->ktPending[k])
            if (pr.pooled) {
                (void)hipEventDestroy(pr.begin);
                (void)hipEventDestroy(pr.end);
            }
    for (hipEvent_t e : ctx->eventPool) (void)hipEventDestroy(e);
    if (ctx->forkEv) (void)hipEventDestroy(ctx->forkEv);
    if (ctx->joinEv) (void)hipEventDestroy(ctx->joinEv);
    if (ctx->tunEv) (void)hipEventDestroy(ctx->tunEv);
    if (ctx->aux) (void)hipStreamDestroy(ctx->aux);
    if (ctx->aux2) (void)hipStreamDestroy(ctx->aux2);
    for (auto &pm : ctx->polar) {
        if (pm.map) (void)hipFree(pm.map);
        if (pm.inv) (void)hipFree(pm.inv);
    }
    void *bufs[] = {ctx->texQuads, ctx->cosLut, ctx->lsHeight, ctx->lsColor, ctx->lsFog, ctx->lsQuadH, ctx->lsQuadC, ctx->lsQueue, ctx->lsState,
                    ctx->tsHeight, ctx->tsColor,
                    ctx->tsFog, ctx->b26Scratch[0],
                    ctx->b26Scratch[1], ctx->staging, ctx->fdTex, ctx->fdTexHi,
                    ctx->stScratch, ctx->spikeBlurMaps, ctx->twHeight, ctx->twColor,
                    ctx->twTables, ctx->twBackground, ctx->ballMaps, ctx->ballImages, ctx->v2Scratch};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete ctx;
}

ckd_status ckd_set_stream(ckd_ctx *ctx, void *hipStream)
{
    CKD_REQUIRE_CTX(ctx);
    ctx->stream = (hipStream_t)hipStream;
    return CKD_OK;
}

ckd_status ckd_synchronize(ckd_ctx *ctx)
{
    CKD_REQUIRE_CTX(ctx);
    CKD_HIP(hipStreamSynchronize(ctx->stream));
    return CKD_OK;
}

ckd_status ckd_malloc(ckd_ctx *ctx, void **dptr, size_t bytes)
{
    CKD_REQUIRE(dptr != nullptr, "ckd_malloc: null output pointer");
    CKD_REQUIRE_CTX(ctx);
    *dptr = nullptr;
    if (bytes == 0) return CKD_OK;
    if (ckd::dev_alloc(dptr, bytes) != hipSuccess) {
        *dptr = nullptr;
        return ckd::fail(CKD_ENOMEM, "ckd_malloc: ckd::dev_alloc(%zu) failed", bytes);
    }
    return CKD_OK;
}

ckd_status ckd_free(ckd_ctx *ctx, void *dptr)
{
    CKD_REQUIRE_CTX(ctx);
    if (dptr) CKD_HIP(hipFree(dptr));
    return CKD_OK;
}

ckd_status ckd_memcpy_h2d(ckd_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    CKD_REQUIRE(bytes == 0 || (dst && src), "ckd_memcpy_h2d: null pointer");
    CKD_REQUIRE_CTX(ctx);
    if (bytes) CKD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return CKD_OK;
}

ckd_status ckd_memcpy_d2h(ckd_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    CKD_REQUIRE(bytes == 0 || (dst && src), "ckd_memcpy_d2h: null pointer");
    CKD_REQUIRE_CTX(ctx);
    if (bytes) {
        CKD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        CKD_HIP(hipStreamSynchronize(ctx->stream));
    }
    return CKD_OK;
}

ckd_status ckd_memcpy_d2d(ckd_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    CKD_REQUIRE(bytes == 0 || (dst && src), "ckd_memcpy_d2d: null pointer");
    CKD_REQUIRE_CTX(ctx);
    if (bytes) CKD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return CKD_OK;
}

ckd_status ckd_event_record(ckd_ctx *ctx, int slot)
{
    CKD_REQUIRE(slot >= 0 && slot < CKD_NUM_EVENTS, "ckd_event_record: slot %d out of range", slot);
    CKD_REQUIRE_CTX(ctx);
    CKD_HIP(hipEventRecord(ctx->events[slot], ctx->stream));
    return CKD_OK;
}

ckd_status ckd_event_elapsed_ms(ckd_ctx *ctx, int a, int b, float *ms)
{
    CKD_REQUIRE(a >= 0 && a < CKD_NUM_EVENTS && b >= 0 && b < CKD_NUM_EVENTS && ms,
                "ckd_event_elapsed_ms: bad slot or null output");
    CKD_REQUIRE_CTX(ctx);
    CKD_HIP(hipEventSynchronize(ctx->events[b]));
    CKD_HIP(hipEventElapsedTime(ms, ctx->events[a], ctx->events[b]));
    return CKD_OK;
}

} // extern "C"
