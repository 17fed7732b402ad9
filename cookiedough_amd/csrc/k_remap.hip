// k_remap.hip -- polar UV-map remap (code/polar.cpp) and the per-8x8-block
// UV grid remap of the tunnel benchmark.
//
// Both are one lane per output pixel: read the 24:8 (U,V), fetch four
// texels, blend them with ckd_bilerp32 (exactly bsamp32_32) and store.
#include "ckd_internal.h"
#include "ckd_pixel.h"
#include "ckd_grid.h"
#include "ckd_jobs.h"

#include <cfloat>
#include <cmath>
#include <mutex>
#include <vector>

namespace {

// CalculateMaps (code/polar.cpp:18-59), host FP32 with libm, no contraction
#pragma clang fp contract(off)
void calculate_maps(int32_t *pDest, int32_t *pInvDest, unsigned srcResX, unsigned srcResY,
                    unsigned destResX, unsigned destResY)
{
    const float kPI = 3.1415926535897932384626433832795f;
    const float halfResX = destResX / 2.f;
    const float halfResY = destResY / 2.f;
    unsigned iPixel = 0;
    const float maxDist = sqrtf(halfResX * halfResX + halfResY * halfResY);
    for (float Y = -halfResY; Y < halfResY; Y += 1.f) {
        for (float X = -halfResX + FLT_EPSILON; X < halfResX; X += 1.f) {
            const float distance = sqrtf(X * X + Y * Y) / maxDist;
            float theta = atan2f(Y, X);
            theta += kPI;
            theta /= kPI * 2.f;
            const float U = distance * (srcResX - 1.f);
            const float invU = (1.f - distance) * (srcResX - 1.f);
            const float V = theta * (srcResY - 1.f);
            pDest[iPixel] = U >= srcResX - 1.f ? (int32_t)(((srcResX - 2) << 8) | 0xff)
                                               : (int32_t)(U * 256.f);
            pInvDest[iPixel] = invU >= srcResX - 1.f ? (int32_t)(((srcResX - 2) << 8) | 0xff)
                                                     : (int32_t)(invU * 256.f);
            const int32_t v = V >= srcResY - 1.f ? (int32_t)(((srcResY - 2) << 8) | 0xff)
                                                 : (int32_t)(V * 256.f);
            pInvDest[iPixel + 1] = pDest[iPixel + 1] = v;
            iPixel += 2;
        }
    }
}
#pragma clang fp contract(on)

// Fetch32 + Polar_Blit_Tile (code/polar.cpp:82-133): no wrap, the map is
// clamped so that (U0+1, V0+1) stays inside the W x H source.
__global__ __launch_bounds__(256) void k_polar_blit(uint32_t *__restrict__ dst,
                                                    const uint32_t *__restrict__ src,
                                                    const int2 *__restrict__ map,
                                                    unsigned W, unsigned H)
{
    const unsigned x = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned y = blockIdx.y;
    if (x >= W) return;
    const size_t i = (size_t)y * W + x;
    const int2 uv = map[i];
    const unsigned U0 = (unsigned)(uv.x >> 8);
    const unsigned V0 = (unsigned)(uv.y >> 8) * W;
    const uint32_t fu = (uint32_t)uv.x & 0xff, fv = (uint32_t)uv.y & 0xff;
    const uint32_t p0 = src[U0 + V0], p1 = src[U0 + 1 + V0];
    const uint32_t p2 = src[U0 + V0 + W], p3 = src[U0 + 1 + V0 + W];
    dst[i] = ckd_bilerp32(p0, p1, p2, p3, fu, fv);
}

} // namespace


namespace {

__global__ __launch_bounds__(256) void k_grid_remap(uint32_t *__restrict__ dst, unsigned W,
                                                    unsigned H, const int2 *__restrict__ grid,
                                                    const uint32_t *__restrict__ tex,
                                                    unsigned texLog2)
{
    const unsigned x = blockIdx.x * 64 + (threadIdx.x & 63);
    const unsigned y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    const unsigned GW = (W + 7) / 8 + 1;
    dst[(size_t)y * W + x] =
        ckd::grid_pixel(reinterpret_cast<const int32_t *>(grid), GW, tex, texLog2, x, y);
}

} // namespace

namespace ckd {
const ckd_ctx::PolarMaps *polar_maps(const ckd_ctx *ctx, unsigned W, unsigned H)
{
    for (const ckd_ctx::PolarMaps &pm : ctx->polar)
        if (pm.map && pm.W == W && pm.H == H) return &pm;
    return nullptr;
}

void tunnel_job_geometry(TunJob &J, uint32_t *dst, uint32_t *tmp, const int32_t *grid, unsigned W,
                         unsigned H, int radius);
uint32_t tunnel_vblocks(unsigned W, unsigned H, int seg);
ckd_status launch_tunnel_jobs(ckd_ctx *ctx, const TunJob *jobs, int nJobs, const TunJob &single,
                              uint32_t totalRows, uint32_t totalVBlocks, unsigned maxW,
                              unsigned maxEdge, const uint32_t *tex, uint4 *quads, unsigned texLog2,
                              bool timed);
size_t tex_quad_bytes(unsigned texLog2);
ckd_status blur07_hv_oop(ckd_ctx *ctx, uint32_t *tmp, uint32_t *dst, const uint32_t *src,
                         unsigned W, unsigned H, float strength);

ckd_status launch_grid_remap(ckd_ctx *ctx, uint32_t *dst, unsigned W, unsigned H,
                             const int32_t *grid, const uint32_t *tex, unsigned texLog2)
{
    dim3 g(ckd_div_up(W, 64), ckd_div_up(H, 4));
    k_grid_remap<<<g, 256, 0, ctx->stream>>>(dst, W, H, reinterpret_cast<const int2 *>(grid), tex,
                                             texLog2);
    return check_launch("grid_remap");
}
} // namespace ckd

extern "C" {

ckd_status ckd_polar_create(ckd_ctx *ctx, unsigned W, unsigned H)
{
    CKD_REQUIRE(W >= 4 && H >= 4 && W <= 8192 && H <= 8192 && (W % 2) == 0 && (H % 2) == 0,
                "ckd_polar_create: bad size %ux%u (even, 4..8192)", W, H);
    CKD_REQUIRE_CTX(ctx);
    const size_t n = (size_t)W * H * 2;
    std::vector<int32_t> map(n), inv(n);
    calculate_maps(map.data(), inv.data(), W, H, W, H);
    // the size's own slot, else the oldest one
    int slot = -1;
    for (int i = 0; i < ckd_ctx::kPolarSlots; ++i)
        if (ctx->polar[i].map && ctx->polar[i].W == W && ctx->polar[i].H == H) slot = i;
    if (slot < 0) {
        slot = ctx->polarNext;
        ctx->polarNext = (ctx->polarNext + 1) % ckd_ctx::kPolarSlots;
    }
    ckd_ctx::PolarMaps &pm = ctx->polar[slot];
    if (pm.map) {
        // queued blits may still read the maps being replaced
        CKD_HIP(hipStreamSynchronize(ctx->stream));
        (void)hipFree(pm.map);
        (void)hipFree(pm.inv);
        pm = ckd_ctx::PolarMaps();
    }
    if (ckd::dev_alloc(&pm.map, n * 4) != hipSuccess || ckd::dev_alloc(&pm.inv, n * 4) != hipSuccess) {
        if (pm.map) (void)hipFree(pm.map);
        pm = ckd_ctx::PolarMaps();
        return ckd::fail(CKD_ENOMEM, "ckd_polar_create: allocation failed");
    }
    CKD_HIP(hipMemcpy(pm.map, map.data(), n * 4, hipMemcpyHostToDevice));
    CKD_HIP(hipMemcpy(pm.inv, inv.data(), n * 4, hipMemcpyHostToDevice));
    pm.W = W;
    pm.H = H;
    return CKD_OK;
}

ckd_status ckd_polar_blit(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned W,
                          unsigned H, int inverse)
{
    CKD_REQUIRE(dst && src, "ckd_polar_blit: null buffer");
    CKD_REQUIRE(dst != src, "ckd_polar_blit: dst must differ from src");
    CKD_REQUIRE_CTX(ctx);
    const ckd_ctx::PolarMaps *pm = ckd::polar_maps(ctx, W, H);
    if (!pm)
        return ckd::fail(CKD_ESTATE, "ckd_polar_blit: no polar maps for %ux%u (ckd_polar_create)",
                         W, H);
    dim3 g(ckd_div_up(W, 256), H);
    k_polar_blit<<<g, 256, 0, ctx->stream>>>(
        dst, src, reinterpret_cast<const int2 *>(inverse ? pm->inv : pm->map), W, H);
    return ckd::check_launch("polar_blit");
}

ckd_status ckd_grid_remap(ckd_ctx *ctx, uint32_t *dst, unsigned W, unsigned H,
                          const int32_t *grid, const uint32_t *tex, unsigned texLog2)
{
    CKD_REQUIRE(dst && grid && tex, "ckd_grid_remap: null buffer");
    CKD_REQUIRE(W > 0 && H > 0 && W <= 16384 && H <= 16384, "ckd_grid_remap: bad size %ux%u", W, H);
    CKD_REQUIRE(texLog2 >= 1 && texLog2 <= 14, "ckd_grid_remap: texLog2 %u out of 1..14", texLog2);
    CKD_REQUIRE_CTX(ctx);
    return ckd::launch_grid_remap(ctx, dst, W, H, grid, tex, texLog2);
}

ckd_status ckd_tunnel_frame(ckd_ctx *ctx, uint32_t *dst, uint32_t *tmp, unsigned W, unsigned H,
                            const int32_t *grid, const uint32_t *tex, unsigned texLog2, int radius)
{
    CKD_REQUIRE(dst && tmp && grid && tex, "ckd_tunnel_frame: null buffer");
    CKD_REQUIRE(W > 0 && H > 0 && W <= 4096 && H <= 4096, "ckd_tunnel_frame: bad size %ux%u", W, H);
    CKD_REQUIRE(texLog2 >= 1 && texLog2 <= 14, "ckd_tunnel_frame: texLog2 %u out of 1..14", texLog2);
    CKD_REQUIRE(radius >= 1 && radius <= 127 && (unsigned)(2 * radius + 1) <= (W < H ? W : H),
                "ckd_tunnel_frame: radius %d out of range", radius);
    CKD_REQUIRE_CTX(ctx);
    TunJob J;
    ckd::tunnel_job_geometry(J, dst, tmp, grid, W, H, radius);
    J.rowBase = 0;
    J.vBase = 0;
    return ckd::launch_tunnel_jobs(ctx, nullptr, 1, J, H, ckd::tunnel_vblocks(W, H, ctx->tuning.tunnelVSeg),
                                   W, (unsigned)radius, tex, nullptr, texLog2, true);
}

// ---- the tunnel frame without allocation ---------------------------------
//
// ckd_tunnel_frame_noalloc keeps the frame's horizontal-blur image in
// `scratch`, a fixed array of device memory: every GPU the library runs on
// holds one, 14.7 MB, until the process ends.  On the host the array's name
// is only a stand-in for the symbol; the device address of each GPU's copy
// comes from hipGetSymbolAddress, called once per device with that device
// current (ckd_create), and is the only thing host code passes to kernels.

} // extern "C"

namespace {

constexpr size_t kNoAllocPixels = 2560u * 1440u;
static __device__ __align__(256) uint32_t scratch[kNoAllocPixels];

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
uint32_t *g_scratchPtr[kMaxDevices] = {};

uint32_t *scratch_of(int device)
{
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    return device >= 0 && device < kMaxDevices ? g_scratchPtr[device] : nullptr;
}

} // namespace

namespace ckd {

ckd_status noalloc_scratch_setup(int device)
{
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (device < 0 || device >= kMaxDevices)
        return fail(CKD_EDEVICE, "device %d beyond the %d the library keeps scratch for", device,
                    kMaxDevices);
    if (g_scratchPtr[device]) return CKD_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(CKD_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                    hipGetErrorString(e));
    g_scratchPtr[device] = static_cast<uint32_t *>(p);
    return CKD_OK;
}

} // namespace ckd

extern "C" {

ckd_status ckd_tunnel_frame_noalloc(ckd_ctx *ctx, uint32_t *dst, unsigned W, unsigned H,
                                    const int32_t *grid, const uint32_t *tex, unsigned texLog2,
                                    int radius)
{
    CKD_REQUIRE(dst && grid && tex, "ckd_tunnel_frame_noalloc: null buffer");
    CKD_REQUIRE(W > 0 && H > 0 && W <= 4096 && H <= 4096, "ckd_tunnel_frame_noalloc: bad size %ux%u",
                W, H);
    CKD_REQUIRE((size_t)W * H <= kNoAllocPixels,
                "ckd_tunnel_frame_noalloc: %ux%u exceeds its scratch of %zu pixels (2560x1440)", W, H,
                kNoAllocPixels);
    CKD_REQUIRE(texLog2 >= 1 && texLog2 <= 14, "ckd_tunnel_frame_noalloc: texLog2 %u out of 1..14",
                texLog2);
    CKD_REQUIRE(radius >= 1 && radius <= 127 && (unsigned)(2 * radius + 1) <= (W < H ? W : H),
                "ckd_tunnel_frame_noalloc: radius %d out of range", radius);
    CKD_REQUIRE_CTX(ctx);
    uint32_t *tmp = scratch_of(ctx->device);
    if (!tmp) return ckd::fail(CKD_ESTATE, "ckd_tunnel_frame_noalloc: no scratch on device %d", ctx->device);
    TunJob J;
    ckd::tunnel_job_geometry(J, dst, tmp, grid, W, H, radius);
    J.rowBase = 0;
    J.vBase = 0;
    return ckd::launch_tunnel_jobs(ctx, nullptr, 1, J, H, ckd::tunnel_vblocks(W, H, ctx->tuning.tunnelVSeg),
                                   W, (unsigned)radius, tex, nullptr, texLog2, false);
}

} // extern "C"
