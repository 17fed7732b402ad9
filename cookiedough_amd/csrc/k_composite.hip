// k_composite.hip -- the compositing blends of code/util.cpp (Mix32,
// MixOver32, Sub32, Excl32, SoftLight32/A/AA, Overlay32/A, Darken32_50,
// MulSrc32/A, MixSrc32/S, BlitSrc32/A, Fade32) and Zoom32.
//
// Streaming, HBM-bound kernels: one template per blend over a per-pixel
// function; contiguous aligned buffers take four pixels per lane (dwordx4),
// strided ones one pixel per lane.  The integer results are the
// reference's bit for bit, including the unsigned wrap-around of the
// scalar blends when a blended channel falls below the destination (the
// wrapped bits that reach the packed pixel are kept).
#include "ckd_internal.h"
#include "ckd_pixel.h"
#include "ckd_blend.h"

#include <cmath>

namespace {


template <int kOp>
__global__ __launch_bounds__(256) void k_comp_flat(uint4 *__restrict__ dst,
                                                   const uint4 *__restrict__ src, size_t n4,
                                                   uint32_t p0, uint32_t p1)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        uint4 d = dst[i];
        const uint4 s = kOp == OP_FADE ? make_uint4(0, 0, 0, 0) : src[i];
        d.x = blend_px<kOp>(d.x, s.x, p0, p1);
        d.y = blend_px<kOp>(d.y, s.y, p0, p1);
        d.z = blend_px<kOp>(d.z, s.z, p0, p1);
        d.w = blend_px<kOp>(d.w, s.w, p0, p1);
        dst[i] = d;
    }
}

template <int kOp>
__global__ __launch_bounds__(256) void k_comp_rows(uint32_t *__restrict__ dst,
                                                   const uint32_t *__restrict__ src,
                                                   unsigned width, unsigned rows,
                                                   unsigned dstStride, unsigned srcStride,
                                                   uint32_t p0, uint32_t p1)
{
    const unsigned x = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned y = blockIdx.y;
    if (x >= width || y >= rows) return;
    uint32_t *d = dst + (size_t)y * dstStride + x;
    const uint32_t s = kOp == OP_FADE ? 0u : src[(size_t)y * srcStride + x];
    *d = blend_px<kOp>(*d, s, p0, p1);
}

// Zoom32 (code/util.cpp:58-81): nearest neighbour, FP32 without contraction
__global__ __launch_bounds__(256) void k_zoom32(uint32_t *__restrict__ dst,
                                                const uint32_t *__restrict__ src, unsigned xRes,
                                                unsigned yRes, float scale, float sOffsX,
                                                float sOffsY)
{
#pragma clang fp contract(off)
    const unsigned iX = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned iY = blockIdx.y;
    if (iX >= xRes || iY >= yRes) return;
    const float sY = sOffsY + (float)(int)iY * scale;
    const float sX = sOffsX + (float)iX * scale;
    const float f = sY * (float)xRes + sX;
    const size_t n = (size_t)xRes * yRes;
    // the reference reads src[unsigned(f)]; indices past the image (only
    // reachable through rounding at the last row) read its last pixel
    size_t idx = (size_t)(unsigned)f;
    if (idx >= n) idx = n - 1;
    dst[(size_t)iY * xRes + iX] = src[idx];
}

unsigned grid_for(size_t items, unsigned block, unsigned cap)
{
    const size_t g = (items + block - 1) / block;
    return (unsigned)(g < cap ? (g ? g : 1) : cap);
}

template <int kOp>
ckd_status launch(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned width, unsigned rows,
                  unsigned dstStride, unsigned srcStride, uint32_t p0, uint32_t p1, const char *name)
{
    if (width == 0 || rows == 0) return CKD_OK;
    const bool flat = dstStride == width && (kOp == OP_FADE || srcStride == width) &&
                      (((uintptr_t)dst | (uintptr_t)src) & 15) == 0 &&
                      (((size_t)width * rows) & 3) == 0;
    if (flat) {
        const size_t n4 = (size_t)width * rows / 4;
        k_comp_flat<kOp><<<grid_for(n4, 256, 16384), 256, 0, ctx->stream>>>(
            reinterpret_cast<uint4 *>(dst), reinterpret_cast<const uint4 *>(src), n4, p0, p1);
    } else {
        dim3 grid(ckd_div_up(width, 256), rows);
        k_comp_rows<kOp><<<grid, 256, 0, ctx->stream>>>(dst, src, width, rows, dstStride, srcStride,
                                                         p0, p1);
    }
    return ckd::check_launch(name);
}

// a flat blend over numPixels: rows of at most 2^30 pixels
template <int kOp>
ckd_status flat_op(ckd_ctx *ctx, const char *name, uint32_t *dst, const uint32_t *src,
                   unsigned numPixels, uint32_t p0 = 0, uint32_t p1 = 0)
{
    CKD_REQUIRE(numPixels == 0 || (dst && (src || kOp == OP_FADE)), "%s: null buffer", name);
    CKD_REQUIRE_CTX(ctx);
    return launch<kOp>(ctx, dst, src, numPixels, 1, numPixels, numPixels, p0, p1, name);
}

} // namespace

namespace ckd {

// Mix32 / SoftLight32AA / Overlay32A on device buffers, for the effects'
// Draw functions
ckd_status comp_mix32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned n, uint32_t alpha)
{
    return launch<OP_MIX>(ctx, dst, src, n, 1, n, n, alpha & 0xffu, 0, "mix32");
}

ckd_status comp_softlight32aa(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned n,
                              uint32_t iA)
{
    return launch<OP_SOFTLIGHT_AA>(ctx, dst, src, n, 1, n, n, iA, 0, "softlight32aa");
}

ckd_status comp_overlay32a(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned n)
{
    return launch<OP_OVERLAY_A>(ctx, dst, src, n, 1, n, n, 0, 0, "overlay32a");
}

// saturatef(alpha) * 255, truncated (code/util.cpp:311-312)
uint32_t softlight_alpha(float alpha)
{
#pragma clang fp contract(off)
    float a = (alpha < 1.f) ? alpha : 1.f; // std::min<float>(1, alpha)
    a = (0.f < a) ? a : 0.f;               // std::max<float>(0, .)
    return (uint32_t)(a * 255.f);
}

} // namespace ckd

extern "C" {

ckd_status ckd_mix32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned numPixels,
                     uint8_t alpha)
{
    return flat_op<OP_MIX>(ctx, "ckd_mix32", dst, src, numPixels, alpha);
}

ckd_status ckd_mixover32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned numPixels)
{
    return flat_op<OP_MIXOVER>(ctx, "ckd_mixover32", dst, src, numPixels);
}

ckd_status ckd_sub32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned numPixels)
{
    return flat_op<OP_SUB>(ctx, "ckd_sub32", dst, src, numPixels);
}

ckd_status ckd_excl32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned numPixels)
{
    return flat_op<OP_EXCL>(ctx, "ckd_excl32", dst, src, numPixels);
}

ckd_status ckd_softlight32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned numPixels)
{
    return flat_op<OP_SOFTLIGHT>(ctx, "ckd_softlight32", dst, src, numPixels);
}

ckd_status ckd_softlight32a(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned numPixels)
{
    return flat_op<OP_SOFTLIGHT_A>(ctx, "ckd_softlight32a", dst, src, numPixels);
}

ckd_status ckd_softlight32aa(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned numPixels,
                             float alpha)
{
    return flat_op<OP_SOFTLIGHT_AA>(ctx, "ckd_softlight32aa", dst, src, numPixels,
                                    ckd::softlight_alpha(alpha));
}

ckd_status ckd_overlay32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned numPixels)
{
    return flat_op<OP_OVERLAY>(ctx, "ckd_overlay32", dst, src, numPixels);
}

ckd_status ckd_overlay32a(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned numPixels)
{
    return flat_op<OP_OVERLAY_A>(ctx, "ckd_overlay32a", dst, src, numPixels);
}

ckd_status ckd_darken32_50(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned numPixels)
{
    return flat_op<OP_DARKEN50>(ctx, "ckd_darken32_50", dst, src, numPixels);
}

ckd_status ckd_mulsrc32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned numPixels)
{
    return flat_op<OP_MULSRC>(ctx, "ckd_mulsrc32", dst, src, numPixels);
}

ckd_status ckd_mulsrc32a(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned numPixels)
{
    return flat_op<OP_MULSRC_A>(ctx, "ckd_mulsrc32a", dst, src, numPixels);
}

ckd_status ckd_mixsrc32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned numPixels)
{
    return flat_op<OP_MIXSRC>(ctx, "ckd_mixsrc32", dst, src, numPixels);
}

ckd_status ckd_fade32(ckd_ctx *ctx, uint32_t *dst, unsigned numPixels, uint32_t RGB, uint8_t alpha)
{
    return flat_op<OP_FADE>(ctx, "ckd_fade32", dst, nullptr, numPixels, RGB, alpha);
}

ckd_status ckd_mixsrc32s(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned resX,
                         unsigned resY, unsigned srcStride)
{
    CKD_REQUIRE((resX == 0 || resY == 0) || (dst && src), "ckd_mixsrc32s: null buffer");
    CKD_REQUIRE(resX <= (1u << 20) && resY <= (1u << 20) && srcStride <= (1u << 24),
                "ckd_mixsrc32s: size %ux%u / stride %u out of range", resX, resY, srcStride);
    CKD_REQUIRE_CTX(ctx);
    return launch<OP_MIXSRC>(ctx, dst, src, resX, resY, resX, srcStride, 0, 0, "mixsrc32s");
}

ckd_status ckd_blitsrc32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned destResX,
                         unsigned srcResX, unsigned yRes)
{
    CKD_REQUIRE((srcResX == 0 || yRes == 0) || (dst && src), "ckd_blitsrc32: null buffer");
    CKD_REQUIRE(srcResX <= destResX, "ckd_blitsrc32: srcResX %u exceeds destResX %u", srcResX,
                destResX);
    CKD_REQUIRE(destResX <= (1u << 24) && yRes <= (1u << 20), "ckd_blitsrc32: size out of range");
    CKD_REQUIRE_CTX(ctx);
    return launch<OP_MIXSRC>(ctx, dst, src, srcResX, yRes, destResX, srcResX, 0, 0, "blitsrc32");
}

ckd_status ckd_blitsrc32a(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned destResX,
                          unsigned srcResX, unsigned yRes, float alpha)
{
    CKD_REQUIRE((srcResX == 0 || yRes == 0) || (dst && src), "ckd_blitsrc32a: null buffer");
    CKD_REQUIRE(srcResX <= destResX, "ckd_blitsrc32a: srcResX %u exceeds destResX %u", srcResX,
                destResX);
    CKD_REQUIRE(destResX <= (1u << 24) && yRes <= (1u << 20), "ckd_blitsrc32a: size out of range");
    // the reference asserts alpha in [0, 255] (code/util.cpp:731)
    CKD_REQUIRE(alpha >= 0.f && alpha <= 255.f, "ckd_blitsrc32a: alpha %g outside [0, 255]",
                (double)alpha);
    CKD_REQUIRE_CTX(ctx);
    uint32_t packed;
    {
#pragma clang fp contract(off)
        packed = 0x01010101u * (uint32_t)(alpha * 255.f);
    }
    return launch<OP_MIXSRC_A>(ctx, dst, src, srcResX, yRes, destResX, srcResX, packed, 0,
                               "blitsrc32a");
}

ckd_status ckd_zoom32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned xRes,
                      unsigned yRes, float scale)
{
    CKD_REQUIRE(dst && src, "ckd_zoom32: null buffer");
    CKD_REQUIRE(dst != src, "ckd_zoom32: dst must differ from src");
    CKD_REQUIRE(xRes >= 1 && yRes >= 1 && xRes <= 16384 && yRes <= 16384,
                "ckd_zoom32: bad size %ux%u", xRes, yRes);
    // the reference only zooms in (asserts scale <= 1, code/util.cpp:60)
    CKD_REQUIRE(scale > 0.f && scale <= 1.f, "ckd_zoom32: scale %g outside (0, 1]", (double)scale);
    CKD_REQUIRE_CTX(ctx);
    float sOffsX, sOffsY;
    {
#pragma clang fp contract(off)
        sOffsY = ((float)yRes - ((float)yRes * scale)) / 2.f;
        sOffsX = ((float)xRes - ((float)xRes * scale)) / 2.f;
    }
    dim3 grid(ckd_div_up(xRes, 256), yRes);
    k_zoom32<<<grid, 256, 0, ctx->stream>>>(dst, src, xRes, yRes, scale, sOffsX, sOffsY);
    return ckd::check_launch("zoom32");
}

} // extern "C"
