/*
 * orc_blend.c -- the compositing blends of code/util.cpp (Mix32, MixOver32,
 * Sub32, Excl32, SoftLight32/A/AA, Overlay32/A, Darken32_50, MulSrc32/A,
 * MixSrc32/S, BlitSrc32/A, Fade32, Zoom32), restated per pixel.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * The SSE forms widen each byte to a 16-bit lane; their products and sums
 * are taken modulo 2^16 here exactly as _mm_mullo_epi16 / _mm_add_epi16 do,
 * and the final _mm_packus_epi16 saturates the signed 16-bit lane to a
 * byte.  The scalar forms keep the reference's unsigned 32-bit arithmetic,
 * including its wrap-around when a blended channel falls below the
 * destination (SoftLight32A/AA, Overlay32A): the wrapped bits that reach
 * the packed pixel are part of the result.
 */
#include "ckd_oracle.h"

#include <stdint.h>

/* _mm_packus_epi16 of one signed 16-bit lane */
static uint32_t packus16(uint32_t lane16)
{
    const int16_t v = (int16_t)(uint16_t)lane16;
    return v < 0 ? 0u : (v > 255 ? 255u : (uint32_t)v);
}

static uint32_t ch(uint32_t px, int k) { return (px >> (8 * k)) & 0xffu; }

/* ((d << 8) + mullo(a, s - d)) >> 8 of 16-bit lanes, then packus */
static uint32_t mix_lane(uint32_t d, uint32_t s, uint32_t a)
{
    const uint32_t delta = (a * (uint32_t)(uint16_t)(s - d)) & 0xffffu;
    const uint32_t sum = ((d << 8) + delta) & 0xffffu;
    return packus16(sum >> 8);
}

/* per-lane alphas a[4] */
static uint32_t mix_px(uint32_t d, uint32_t s, const uint32_t a[4])
{
    uint32_t r = 0;
    for (int k = 0; k < 4; ++k) r |= mix_lane(ch(d, k), ch(s, k), a[k]) << (8 * k);
    return r;
}

/* Mix32, code/util.cpp:83-96 */
void orc_mix32(uint32_t *dst, const uint32_t *src, unsigned n, uint8_t alpha)
{
    const uint32_t a[4] = {alpha, alpha, alpha, alpha};
    for (unsigned i = 0; i < n; ++i) dst[i] = mix_px(dst[i], src[i], a);
}

/* MixOver32, code/util.cpp:148-177 */
void orc_mixover32(uint32_t *dst, const uint32_t *src, unsigned n)
{
    for (unsigned i = 0; i < n; ++i) {
        const uint32_t d = dst[i], s = src[i];
        const unsigned A1 = 0xffu - (s >> 24);
        unsigned c[3];
        for (int k = 0; k < 3; ++k) {
            c[k] = ((ch(s, k) * (0xffu - A1)) >> 8) + ((ch(d, k) * A1) >> 8);
            if (c[k] > 255) c[k] = 255;
        }
        dst[i] = (c[2] << 16) | (c[1] << 8) | c[0];
    }
}

/* Sub32, code/util.cpp:179-192 */
void orc_sub32(uint32_t *dst, const uint32_t *src, unsigned n)
{
    for (unsigned i = 0; i < n; ++i) {
        uint32_t r = 0;
        for (int k = 0; k < 4; ++k) r |= packus16(ch(dst[i], k) - ch(src[i], k)) << (8 * k);
        dst[i] = r;
    }
}

/* Excl32, code/util.cpp:194-224 */
void orc_excl32(uint32_t *dst, const uint32_t *src, unsigned n)
{
    for (unsigned i = 0; i < n; ++i) {
        const uint32_t d = dst[i], s = src[i];
        unsigned c[3];
        for (int k = 0; k < 3; ++k) {
            const unsigned a = ch(s, k), b = ch(d, k);
            c[k] = a + b - ((2u * a * b) >> 8);
        }
        dst[i] = ((d >> 24) << 24) | (c[2] << 16) | (c[1] << 8) | c[0];
    }
}

/* SoftLightBlend, code/util.cpp:227-240 */
static unsigned soft_light(unsigned A, unsigned B)
{
    const unsigned dA = (A / 2) + 64;
    if (B < 128) return (2 * dA * B) / 256;
    return 255 - (2 * (255 - dA) * (255 - B)) / 256;
}

/* SoftLight32, code/util.cpp:242-271 */
void orc_softlight32(uint32_t *dst, const uint32_t *src, unsigned n)
{
    for (unsigned i = 0; i < n; ++i) {
        const uint32_t d = dst[i], s = src[i];
        const unsigned R = soft_light(ch(s, 2), ch(d, 2));
        const unsigned G = soft_light(ch(s, 1), ch(d, 1));
        const unsigned B = soft_light(ch(s, 0), ch(d, 0));
        dst[i] = ((d >> 24) << 24) | (R << 16) | (G << 8) | B;
    }
}

/* the blend towards (R,G,B) by a of SoftLight32A/AA and Overlay32A, in
 * unsigned 32-bit arithmetic: c2 + (((c - c2) * a) >> 8) */
static uint32_t blend_wrap(unsigned c2, unsigned c, unsigned a)
{
    return c2 + (((c - c2) * a) >> 8);
}

/* SoftLight32A, code/util.cpp:274-306 */
void orc_softlight32a(uint32_t *dst, const uint32_t *src, unsigned n)
{
    for (unsigned i = 0; i < n; ++i) {
        const uint32_t d = dst[i], s = src[i];
        const unsigned A1 = s >> 24;
        const uint32_t R = blend_wrap(ch(d, 2), soft_light(ch(s, 2), ch(d, 2)), A1);
        const uint32_t G = blend_wrap(ch(d, 1), soft_light(ch(s, 1), ch(d, 1)), A1);
        const uint32_t B = blend_wrap(ch(d, 0), soft_light(ch(s, 0), ch(d, 0)), A1);
        dst[i] = (R << 16) | (G << 8) | B;
    }
}

/* SoftLight32AA, code/util.cpp:309-346; saturatef(alpha)*255 truncated */
void orc_softlight32aa(uint32_t *dst, const uint32_t *src, unsigned n, float alpha)
{
    /* std::max<float>(0, std::min<float>(1, alpha)) */
    float a = (alpha < 1.f) ? alpha : 1.f;
    a = (0.f < a) ? a : 0.f;
    a = a * 255.f;
    const unsigned iA = (unsigned)a;
    for (unsigned i = 0; i < n; ++i) {
        const uint32_t d = dst[i], s = src[i];
        const uint32_t R = blend_wrap(ch(d, 2), soft_light(ch(s, 2), ch(d, 2)), iA);
        const uint32_t G = blend_wrap(ch(d, 1), soft_light(ch(s, 1), ch(d, 1)), iA);
        const uint32_t B = blend_wrap(ch(d, 0), soft_light(ch(s, 0), ch(d, 0)), iA);
        dst[i] = (iA << 24) | (R << 16) | (G << 8) | B;
    }
}

/* the integer overlay of one channel, code/util.cpp:449-451 */
static unsigned overlay(unsigned bottom, unsigned top)
{
    return bottom < 128 ? (2 * bottom * top / 255) : (255 - 2 * (255 - bottom) * (255 - top) / 255);
}

/* Overlay32 (the integer form), code/util.cpp:434-456 */
void orc_overlay32(uint32_t *dst, const uint32_t *src, unsigned n)
{
    for (unsigned i = 0; i < n; ++i) {
        const uint32_t b = dst[i], t = src[i];
        dst[i] = (overlay(ch(b, 2), ch(t, 2)) << 16) | (overlay(ch(b, 1), ch(t, 1)) << 8) |
                 overlay(ch(b, 0), ch(t, 0));
    }
}

/* Overlay32A (the integer form), code/util.cpp:485-515 */
void orc_overlay32a(uint32_t *dst, const uint32_t *src, unsigned n)
{
    for (unsigned i = 0; i < n; ++i) {
        const uint32_t b = dst[i], t = src[i];
        const unsigned A = t >> 24;
        const uint32_t R = blend_wrap(ch(b, 2), overlay(ch(b, 2), ch(t, 2)), A);
        const uint32_t G = blend_wrap(ch(b, 1), overlay(ch(b, 1), ch(t, 1)), A);
        const uint32_t B = blend_wrap(ch(b, 0), overlay(ch(b, 0), ch(t, 0)), A);
        dst[i] = (R << 16) | (G << 8) | B;
    }
}

/* Darken32_50, code/util.cpp:518-549 */
void orc_darken32_50(uint32_t *dst, const uint32_t *src, unsigned n)
{
    for (unsigned i = 0; i < n; ++i) {
        const uint32_t d = dst[i], s = src[i];
        unsigned c[3];
        for (int k = 0; k < 3; ++k) {
            const unsigned a = ch(s, k), b = ch(d, k);
            c[k] = (b + (a < b ? a : b)) >> 1;
        }
        dst[i] = ((d >> 24) << 24) | (c[2] << 16) | (c[1] << 8) | c[0];
    }
}

/* MulSrc32, code/util.cpp:605-618: (s*d) >> 8 of 16-bit lanes */
void orc_mulsrc32(uint32_t *dst, const uint32_t *src, unsigned n)
{
    for (unsigned i = 0; i < n; ++i) {
        uint32_t r = 0;
        for (int k = 0; k < 4; ++k)
            r |= packus16(((ch(src[i], k) * ch(dst[i], k)) & 0xffffu) >> 8) << (8 * k);
        dst[i] = r;
    }
}

/* MulSrc32A, code/util.cpp:620-634: every lane times the source alpha */
void orc_mulsrc32a(uint32_t *dst, const uint32_t *src, unsigned n)
{
    for (unsigned i = 0; i < n; ++i) {
        const uint32_t a = src[i] >> 24;
        uint32_t r = 0;
        for (int k = 0; k < 4; ++k) r |= packus16(((a * ch(dst[i], k)) & 0xffffu) >> 8) << (8 * k);
        dst[i] = r;
    }
}

static uint32_t mix_src_px(uint32_t d, uint32_t s)
{
    const uint32_t a = s >> 24;
    const uint32_t al[4] = {a, a, a, a};
    return mix_px(d, s, al);
}

/* MixSrc32, code/util.cpp:661-676 */
void orc_mixsrc32(uint32_t *dst, const uint32_t *src, unsigned n)
{
    for (unsigned i = 0; i < n; ++i) dst[i] = mix_src_px(dst[i], src[i]);
}

/* MixSrc32S, code/util.cpp:636-659 */
void orc_mixsrc32s(uint32_t *dst, const uint32_t *src, unsigned resX, unsigned resY,
                   unsigned srcStride)
{
    for (unsigned y = 0; y < resY; ++y)
        for (unsigned x = 0; x < resX; ++x)
            dst[(size_t)y * resX + x] = mix_src_px(dst[(size_t)y * resX + x], src[(size_t)y * srcStride + x]);
}

/* BlitSrc32, code/util.cpp:707-727 */
void orc_blitsrc32(uint32_t *dst, const uint32_t *src, unsigned destResX, unsigned srcResX,
                   unsigned yRes)
{
    for (unsigned y = 0; y < yRes; ++y)
        for (unsigned x = 0; x < srcResX; ++x)
            dst[(size_t)y * destResX + x] = mix_src_px(dst[(size_t)y * destResX + x],
                                                       src[(size_t)y * srcResX + x]);
}

/* BlitSrc32A, code/util.cpp:729-752: per lane a = (srcA * f_k) >> 8 with
 * f_k the bytes of 0x01010101 * unsigned(alpha * 255.f) (mod 2^32) */
void orc_blitsrc32a(uint32_t *dst, const uint32_t *src, unsigned destResX, unsigned srcResX,
                    unsigned yRes, float alpha)
{
    const uint32_t packed = 0x01010101u * (uint32_t)(int64_t)(alpha * 255.f);
    for (unsigned y = 0; y < yRes; ++y)
        for (unsigned x = 0; x < srcResX; ++x) {
            uint32_t *d = dst + (size_t)y * destResX + x;
            const uint32_t s = src[(size_t)y * srcResX + x];
            uint32_t a[4];
            for (int k = 0; k < 4; ++k) a[k] = (((s >> 24) * ch(packed, k)) & 0xffffu) >> 8;
            *d = mix_px(*d, s, a);
        }
}

/* Fade32, code/util.cpp:798-811 */
void orc_fade32(uint32_t *dst, unsigned n, uint32_t RGB, uint8_t alpha)
{
    const uint32_t a[4] = {alpha, alpha, alpha, alpha};
    for (unsigned i = 0; i < n; ++i) dst[i] = mix_px(dst[i], RGB, a);
}

/* Zoom32, code/util.cpp:58-81 (nearest neighbour, scale <= 1) */
void orc_zoom32(uint32_t *dst, const uint32_t *src, unsigned xRes, unsigned yRes, float scale)
{
    const float sOffsY = ((float)yRes - ((float)yRes * scale)) / 2.f;
    const float sOffsX = ((float)xRes - ((float)xRes * scale)) / 2.f;
    for (int iY = 0; iY < (int)yRes; ++iY) {
        const float sY = sOffsY + (float)iY * scale;
        for (unsigned iX = 0; iX < xRes; ++iX) {
            const float sX = sOffsX + (float)iX * scale;
            const unsigned sIndex = (unsigned)(int64_t)(sY * (float)xRes + sX);
            dst[(size_t)iY * xRes + iX] = src[sIndex];
        }
    }
}
