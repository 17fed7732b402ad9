// ckd_internal.h -- context state and error plumbing shared by the HIP
// translation units of libckd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "../../include/ckd.h"

#define CKD_NUM_EVENTS 16

// launch shapes of a new context
constexpr ckd_tuning kCkdDefaultTuning = {16, 32, 8, 32, 0, 2, 2};

struct ckd_ctx {
    int device = 0;
    int cuCount = 0;
    hipStream_t stream = nullptr;
    hipEvent_t events[CKD_NUM_EVENTS] = {};
    // second stream of the batched renderer (voxel and tunnel work runs
    // beside the persistent SDF kernel), joined through two events
    hipStream_t aux = nullptr;
    hipEvent_t forkEv = nullptr, joinEv = nullptr;
    // third stream: the batch's tunnel work beside its voxel work, joined
    // into the second through tunEv
    hipStream_t aux2 = nullptr;
    hipEvent_t tunEv = nullptr;
    // texel quads of the batched tunnel's texture (16 bytes per texel),
    // grown to the largest texture a batch has used
    uint4 *texQuads = nullptr;
    size_t texQuadBytes = 0;

    // interpolated cosine LUT, 2049 floats (CalculateCosLUT), on the
    // device and the host copy the per-frame set-up reads
    float *cosLut = nullptr;
    float hostLut[2049] = {};

    // the free-directional tunnel's two 1024x1024 textures
    // (Shadertoy_Create) and the FX-map scratch of the effects' Draw calls
    uint32_t *fdTex = nullptr, *fdTexHi = nullptr;
    uint32_t *stScratch = nullptr;
    size_t stScratchCap = 0;
    // march-step counter of the FX-map kernels (ckd_shadertoy_step_counter)
    unsigned long long *stSteps = nullptr;
    // the batched SDF kernel's two work counters (ckd_sdf_work_counter)
    unsigned long long *sdfWork = nullptr;
    // the close-up spike's two blur maps (Shadertoy_Create), for the output
    // size they were given for
    uint32_t *spikeBlurMaps = nullptr;
    unsigned spikeBlurW = 0, spikeBlurH = 0;

    // torus twister (Twister_Create): height map, colour map, the two
    // 512-entry ray tables and the W x H background
    uint8_t *twHeight = nullptr;
    uint32_t *twColor = nullptr, *twTables = nullptr, *twBackground = nullptr;
    unsigned twW = 0, twH = 0;
    // voxel ball (Ball_Create): 5 height maps + their mix (bytes), then 2
    // colour maps, 3 beam maps, the beam mix and the environment map; the
    // two backgrounds and the halo (W x H each); the largest map height
    uint8_t *ballMaps = nullptr;
    uint32_t *ballImages = nullptr;
    unsigned ballW = 0, ballH = 0;
    uint32_t ballTopHeight = 0;
    // batched ball / twister Draw: per-frame render targets and map mixes
    uint8_t *v2Scratch = nullptr;
    size_t v2Cap = 0;

    // landscape maps (Landscape_Create), and their 2x2 texel quads: texel
    // (u,v)'s four bilinear neighbours (wrapped) in one word of heights and
    // four words of colours, so a march step fetches them with two loads
    uint8_t *lsHeight = nullptr;
    uint32_t *lsColor = nullptr;
    uint32_t *lsFog = nullptr;
    uint32_t lsFog0 = 0;
    uint32_t *lsQuadH = nullptr;
    uint4 *lsQuadC = nullptr;
    // ckd_landscape_draw's work queue of march ranges: four control words
    // and one 64-bit state word per column of a launch's frames
    uint32_t *lsQueue = nullptr;
    unsigned long long *lsState = nullptr;

    // tunnelscape maps (Tunnelscape_Create)
    uint8_t *tsHeight = nullptr;
    uint32_t *tsColor = nullptr;
    uint32_t *tsFog = nullptr;
    uint32_t tsFog0 = 0;

    // polar maps (Polar_Create), interleaved (U,V) 24:8 per pixel, one set
    // per image size (the reference keeps 1280x720 and the FX map's 644x364)
    struct PolarMaps {
        int32_t *map = nullptr, *inv = nullptr;
        unsigned W = 0, H = 0;
    };
    static constexpr int kPolarSlots = 4;
    PolarMaps polar[kPolarSlots];
    int polarNext = 0; // slot the next new size takes

    // 2026 blur scratch (BoxBlur_Create): two images plus guard pixels
    uint32_t *b26Scratch[2] = {nullptr, nullptr};
    size_t b26Cap = 0;

    // per-call parameter staging (device), grown on demand
    void *staging = nullptr;
    size_t stagingCap = 0;

    // launch-shape tuning (ckd_autotune_batch / ckd_set_tuning)
    ckd_tuning tuning = kCkdDefaultTuning;

    // per-kernel HIP-event timing (ckd_kernel_timing)
    bool timing = false;
    struct KtPending {
        hipEvent_t begin, end;
        bool pooled; // false: the events belong to a graph
    };
    std::vector<KtPending> ktPending[CKD_KT_COUNT];
    double ktMs[CKD_KT_COUNT] = {};
    unsigned long long ktLaunches[CKD_KT_COUNT] = {};
    std::vector<hipEvent_t> eventPool;
};

namespace ckd {
// Fx_Blit_2x2 of nFrames FX maps ((W/2+4) x (H/2+4) each, back to back)
// into nFrames W x H frames (back to back), one launch; sizes as
// ckd_fx_blit_2x2 checks them
ckd_status fx_blit_frames(ckd_ctx *ctx, uint32_t *dst, const uint32_t *maps, unsigned dstW,
                          unsigned dstH, unsigned nFrames);
// dst = sat(dst + src) per byte over `rows` rows of `width` pixels (Add32 /
// BlitAdd32), rows destStride / srcStride pixels apart
ckd_status launch_blit_add(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned destStride,
                           unsigned srcStride, unsigned width, unsigned rows);

// Records a message for ckd_last_error() on this thread and returns s.
// the device address of ckd_tunnel_frame_noalloc's scratch array on the
// current device, looked up once per device (k_remap.hip)
ckd_status noalloc_scratch_setup(int device);

// hipMalloc for every device allocation the library makes, so that a test
// can make the n-th next one fail (ckd_debug_fail_alloc)
hipError_t dev_alloc_raw(void **p, size_t bytes);
template <typename T> hipError_t dev_alloc(T **p, size_t bytes)
{
    return dev_alloc_raw(reinterpret_cast<void **>(p), bytes);
}

ckd_status fail(ckd_status s, const char *fmt, ...)
    __attribute__((format(printf, 2, 3)));

// hipGetLastError after a launch: CKD_EDEVICE with the kernel name.
ckd_status check_launch(const char *what);

// The polar maps for W x H (ckd_polar_create), or nullptr.
const ckd_ctx::PolarMaps *polar_maps(const ckd_ctx *ctx, unsigned W, unsigned H);

// Device buffer of at least `bytes` owned by the context for per-call
// parameters; enqueues the host->device copy on the context stream.
ckd_status stage(ckd_ctx *ctx, const void *host, size_t bytes, void **dev);

// Waits for every pending kernel-timing event pair and adds it up.
ckd_status drain_timing(ckd_ctx *ctx);

// Redirects ctx->stream for the lifetime of the guard.
class StreamSwap {
public:
    StreamSwap(ckd_ctx *ctx, hipStream_t s) : ctx_(ctx), saved_(ctx->stream) { ctx->stream = s; }
    ~StreamSwap() { ctx_->stream = saved_; }
    StreamSwap(const StreamSwap &) = delete;
    StreamSwap &operator=(const StreamSwap &) = delete;

private:
    ckd_ctx *ctx_;
    hipStream_t saved_;
};

// Brackets one kernel launch with events when kernel timing is on.
class KernelTimer {
public:
    KernelTimer(ckd_ctx *ctx, int kernel);
    ~KernelTimer();
    KernelTimer(const KernelTimer &) = delete;
    KernelTimer &operator=(const KernelTimer &) = delete;

private:
    ckd_ctx *ctx_;
    int kernel_;
    hipEvent_t begin_ = nullptr;
};

} // namespace ckd

#define CKD_REQUIRE(cond, ...)                                                 \
    do {                                                                       \
        if (!(cond)) return ckd::fail(CKD_EINVAL, __VA_ARGS__);               \
    } while (0)

#define CKD_REQUIRE_CTX(ctx)                                                   \
    do {                                                                       \
        if (!(ctx)) return ckd::fail(CKD_EINVAL, "%s: null context", __func__); \
    } while (0)

#define CKD_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            return ckd::fail(CKD_EDEVICE, "%s: %s: %s", __func__, #call,      \
                             hipGetErrorString(e_));                           \
    } while (0)

static inline unsigned ckd_div_up(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }
