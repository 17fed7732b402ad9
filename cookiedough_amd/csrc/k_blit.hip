// k_blit.hip -- memset32, Fx_Blit_2x2 and the additive blits.
//
// All of these are streaming, HBM-bound kernels: each thread moves 16 bytes
// per access where the layout allows it (dwordx4 loads/stores).
#include "ckd_internal.h"
#include "ckd_pixel.h"

namespace {

// memset32 (code/util.h:57-67): fill with a 32-bit value.  A kernel rather
// than hipMemsetD32Async so that the fill can sit in a captured graph.
__global__ __launch_bounds__(256) void k_memset32(uint4 *__restrict__ dst4, uint32_t *__restrict__ tail,
                                                  size_t n4, size_t nTail, uint32_t value)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const uint4 v = make_uint4(value, value, value, value);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
        dst4[i] = v;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nTail) tail[t] = value;
}

// Fx_Blit_2x2 (code/fx-blitter.cpp:27-75).  Thread (x, y) expands FX-map
// pixel (x, y) into the 2x2 output block at (2x, 2y); it reads its right
// and lower neighbours (guard column dstW/2 and guard row dstH/2 included).
__global__ __launch_bounds__(256) void k_fx_blit_2x2(uint32_t *__restrict__ dst,
                                                     const uint32_t *__restrict__ src,
                                                     unsigned dstW, unsigned mapW,
                                                     unsigned halfW, size_t dstFrame,
                                                     size_t mapFrame)
{
    const unsigned x = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned y = blockIdx.y;
    if (x >= halfW) return;
    // blockIdx.z: the frame of a batch, frames dstFrame / mapFrame pixels apart
    dst += blockIdx.z * dstFrame;
    src += blockIdx.z * mapFrame;
    const uint32_t *r0 = src + (size_t)y * mapW + x;
    const uint32_t *r1 = r0 + mapW;
    const uint32_t s00 = r0[0], s10 = r0[1], s01 = r1[0], s11 = r1[1];
    const uint32_t avgH0 = ckd_avg_u8x4(s00, s10);
    const uint32_t avgH1 = ckd_avg_u8x4(s01, s11);
    uint2 *top = reinterpret_cast<uint2 *>(dst + (size_t)(2 * y) * dstW) + x;
    uint2 *bot = reinterpret_cast<uint2 *>(dst + (size_t)(2 * y + 1) * dstW) + x;
    *top = make_uint2(s00, avgH0);
    *bot = make_uint2(ckd_avg_u8x4(s00, s01), ckd_avg_u8x4(avgH0, avgH1));
}

// Add32 / BlitAdd32 / BlitAdd32A (code/util.cpp:132-145, 754-796):
// dst = sat(dst + ((src*alphaLane) >> 8)) per byte; alphaLanes == 0 selects
// the plain add.  Thread per pixel over a (width x rows) region.
template <bool kAlpha>
__global__ __launch_bounds__(256) void k_blit_add(uint32_t *__restrict__ dst,
                                                  const uint32_t *__restrict__ src,
                                                  unsigned destStride, unsigned srcStride,
                                                  unsigned width, unsigned rows,
                                                  uint32_t alphaLanes)
{
    const unsigned x = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned y = blockIdx.y;
    if (x >= width || y >= rows) return;
    uint32_t s = src[(size_t)y * srcStride + x];
    if (kAlpha) {
        uint32_t m = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            m |= ((((s >> (8 * k)) & 0xff) * ((alphaLanes >> (8 * k)) & 0xff)) >> 8) << (8 * k);
        s = m;
    }
    uint32_t *d = dst + (size_t)y * destStride + x;
    *d = ckd_adds_u8x4(*d, s);
}

// flat variant for contiguous buffers: 4 pixels per thread
__global__ __launch_bounds__(256) void k_add32_flat(uint4 *__restrict__ dst,
                                                    const uint4 *__restrict__ src, size_t n4)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        uint4 d = dst[i];
        const uint4 s = src[i];
        d.x = ckd_adds_u8x4(d.x, s.x);
        d.y = ckd_adds_u8x4(d.y, s.y);
        d.z = ckd_adds_u8x4(d.z, s.z);
        d.w = ckd_adds_u8x4(d.w, s.w);
        dst[i] = d;
    }
}

// TapeWarp32 (code/util.cpp:552-603) with lutsinf/lutcosf
// (code/sincos-lut.h:13-26): FP32 without contraction, so the 24:8 sample
// positions equal the host's bit for bit.
#pragma clang fp contract(off)
CKD_DEV float lutcos(const float *__restrict__ lut, float angle, float scale)
{
    angle = fabsf(angle);
    angle *= scale;
    const int index = (int)angle & 2047;
    const float frac = angle - truncf(angle);
    return lut[index] + (lut[index + 1] - lut[index]) * frac;
}

__global__ __launch_bounds__(256) void k_tape_warp(uint32_t *__restrict__ dst,
                                                   const uint32_t *__restrict__ src,
                                                   const float *__restrict__ lut, unsigned xRes,
                                                   unsigned yRes, float strength, float speed,
                                                   float scale, float halfPi)
{
    const unsigned iX = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned iY = blockIdx.y;
    if (iX >= xRes) return;
    // ARRESTED_DEV_LEGACY lutsinf(a) = lutcosf(a + pi/2)
    const float dX = lutcos(lut, (float)(int)iY * speed + halfPi, scale) * strength * 1.f;
    const float dY = lutcos(lut, (float)iX * speed, scale) * strength * 1.f;
    float tX = (float)iX + dX;
    float tY = (float)iY + dY;
    if (tX < 0.f) tX = 0.f;
    else if (tX >= (float)xRes - 1.f) tX = (float)xRes - 2.f;
    if (tY < 0.f) tY = 0.f;
    else if (tY >= (float)yRes - 1.f) tY = (float)yRes - 2.f;
    const int U = (int)(tX * 256.f), V = (int)(tY * 256.f);
    const unsigned U0 = (unsigned)(U >> 8);
    const unsigned V0 = (unsigned)(V >> 8) * xRes;
    const uint32_t fu = (uint32_t)U & 0xff, fv = (uint32_t)V & 0xff;
    dst[(size_t)iY * xRes + iX] = ckd_bilerp32(src[U0 + V0], src[U0 + 1 + V0], src[U0 + V0 + xRes],
                                               src[U0 + 1 + V0 + xRes], fu, fv);
}
#pragma clang fp contract(on)

unsigned grid_for(size_t items, unsigned block, unsigned cap)
{
    size_t g = (items + block - 1) / block;
    if (g > cap) g = cap;
    if (g == 0) g = 1;
    return (unsigned)g;
}

} // namespace

namespace ckd {

// shared with the pipeline kernels of other translation units
ckd_status launch_memset32(ckd_ctx *ctx, uint32_t *dst, uint32_t value, size_t n)
{
    // align the vector part to 16 bytes
    size_t head = 0;
    while (head < n && ((uintptr_t)(dst + head) & 15)) ++head;
    if (head) {
        k_memset32<<<1, 64, 0, ctx->stream>>>(nullptr, dst, 0, head, value);
    }
    const size_t rest = n - head;
    const size_t n4 = rest / 4, nTail = rest % 4;
    if (n4 || nTail)
        k_memset32<<<grid_for(n4 ? n4 : 1, 256, 8192), 256, 0, ctx->stream>>>(
            reinterpret_cast<uint4 *>(dst + head), dst + head + 4 * n4, n4, nTail, value);
    return check_launch("memset32");
}

ckd_status launch_blit_add(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned destStride,
                           unsigned srcStride, unsigned width, unsigned rows)
{
    if (width == 0 || rows == 0) return CKD_OK;
    const bool flat = destStride == width && srcStride == width &&
                      (((uintptr_t)dst | (uintptr_t)src) & 15) == 0 &&
                      (((size_t)width * rows) & 3) == 0;
    if (flat) {
        const size_t n4 = (size_t)width * rows / 4;
        k_add32_flat<<<grid_for(n4, 256, 16384), 256, 0, ctx->stream>>>(
            reinterpret_cast<uint4 *>(dst), reinterpret_cast<const uint4 *>(src), n4);
    } else {
        dim3 grid(ckd_div_up(width, 256), rows);
        k_blit_add<false><<<grid, 256, 0, ctx->stream>>>(dst, src, destStride, srcStride, width,
                                                          rows, 0);
    }
    return check_launch("blit_add32");
}

} // namespace ckd

namespace ckd {

ckd_status fx_blit_frames(ckd_ctx *ctx, uint32_t *dst, const uint32_t *maps, unsigned dstW,
                          unsigned dstH, unsigned nFrames)
{
    if (nFrames == 0) return CKD_OK;
    const unsigned halfW = dstW / 2, mapW = halfW + 4;
    const size_t mapFrame = (size_t)mapW * (dstH / 2 + 4), dstFrame = (size_t)dstW * dstH;
    for (unsigned f0 = 0; f0 < nFrames; f0 += 65535) {
        const unsigned n = nFrames - f0 < 65535 ? nFrames - f0 : 65535;
        dim3 grid(ckd_div_up(halfW, 256), dstH / 2, n);
        k_fx_blit_2x2<<<grid, 256, 0, ctx->stream>>>(dst + f0 * dstFrame, maps + f0 * mapFrame, dstW,
                                                     mapW, halfW, dstFrame, mapFrame);
    }
    return check_launch("fx_blit_2x2");
}

} // namespace ckd

extern "C" {

ckd_status ckd_tape_warp32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned xRes,
                           unsigned yRes, float strength, float speed)
{
    CKD_REQUIRE(dst && src, "ckd_tape_warp32: null buffer");
    CKD_REQUIRE(dst != src, "ckd_tape_warp32: dst must differ from src");
    CKD_REQUIRE(xRes >= 2 && yRes >= 2 && xRes <= 16384 && yRes <= 16384,
                "ckd_tape_warp32: bad size %ux%u", xRes, yRes);
    CKD_REQUIRE(strength == strength && speed == speed && strength < 1e30f && strength > -1e30f &&
                    speed < 1e30f && speed > -1e30f,
                "ckd_tape_warp32: strength and speed must be finite");
    CKD_REQUIRE_CTX(ctx);
    float scale, halfPi;
    {
#pragma clang fp contract(off)
        const float kPI = 3.1415926535897932384626433832795f;
        scale = (1.f / (2.f * kPI)) * 2048.f;
        halfPi = kPI * 0.5f;
    }
    dim3 grid(ckd_div_up(xRes, 256), yRes);
    k_tape_warp<<<grid, 256, 0, ctx->stream>>>(dst, src, ctx->cosLut, xRes, yRes, strength, speed,
                                                scale, halfPi);
    return ckd::check_launch("tape_warp32");
}

ckd_status ckd_memset32(ckd_ctx *ctx, uint32_t *dst, uint32_t value, size_t numInts)
{
    CKD_REQUIRE(numInts == 0 || dst, "ckd_memset32: null destination");
    CKD_REQUIRE_CTX(ctx);
    if (numInts == 0) return CKD_OK;
    return ckd::launch_memset32(ctx, dst, value, numInts);
}

ckd_status ckd_fx_blit_2x2(ckd_ctx *ctx, uint32_t *dst, const uint32_t *fxMap, unsigned dstW,
                           unsigned dstH)
{
    CKD_REQUIRE(dst && fxMap, "ckd_fx_blit_2x2: null buffer");
    CKD_REQUIRE(dstW >= 8 && dstH >= 2 && (dstW % 8) == 0 && (dstH % 2) == 0 &&
                    dstW <= 16384 && dstH <= 16384,
                "ckd_fx_blit_2x2: %ux%u: need even sizes with (W/2) %% 4 == 0", dstW, dstH);
    CKD_REQUIRE(((uintptr_t)dst & 7) == 0, "ckd_fx_blit_2x2: dst must be 8-byte aligned");
    CKD_REQUIRE_CTX(ctx);
    return ckd::fx_blit_frames(ctx, dst, fxMap, dstW, dstH, 1);
}

ckd_status ckd_add32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned numPixels)
{
    CKD_REQUIRE(numPixels == 0 || (dst && src), "ckd_add32: null buffer");
    CKD_REQUIRE_CTX(ctx);
    if (numPixels == 0) return CKD_OK;
    // one "row" of numPixels (rows of at most 2^31 pixels)
    return ckd::launch_blit_add(ctx, dst, src, numPixels, numPixels, numPixels, 1);
}

ckd_status ckd_blit_add32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned destResX,
                          unsigned srcResX, unsigned yRes)
{
    CKD_REQUIRE((srcResX == 0 || yRes == 0) || (dst && src), "ckd_blit_add32: null buffer");
    CKD_REQUIRE(srcResX <= destResX, "ckd_blit_add32: srcResX %u > destResX %u", srcResX,
                destResX);
    CKD_REQUIRE_CTX(ctx);
    return ckd::launch_blit_add(ctx, dst, src, destResX, srcResX, srcResX, yRes);
}

ckd_status ckd_blit_add32a(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned destResX,
                           unsigned srcResX, unsigned yRes, float alpha)
{
    CKD_REQUIRE((srcResX == 0 || yRes == 0) || (dst && src), "ckd_blit_add32a: null buffer");
    CKD_REQUIRE(srcResX <= destResX, "ckd_blit_add32a: srcResX %u > destResX %u", srcResX,
                destResX);
    // unsigned(alpha*255.f) is undefined for negative / huge values in C++
    CKD_REQUIRE(alpha >= 0.f && alpha * 255.f < 4294967040.f, "ckd_blit_add32a: alpha %g out of range",
                (double)alpha);
    CKD_REQUIRE_CTX(ctx);
    if (srcResX == 0 || yRes == 0) return CKD_OK;
    // c2vISSE16(0x01010101 * unsigned(alpha*255.f)): byte k of the product
    const uint32_t lanes = 0x01010101u * (uint32_t)(alpha * 255.f);
    dim3 grid(ckd_div_up(srcResX, 256), yRes);
    k_blit_add<true><<<grid, 256, 0, ctx->stream>>>(dst, src, destResX, srcResX, srcResX, yRes,
                                                     lanes);
    return ckd::check_launch("blit_add32a");
}

} // extern "C"
