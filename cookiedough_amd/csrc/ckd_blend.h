// ckd_blend.h -- per-pixel compositing blends of code/util.cpp, shared by
// the blend kernels and the effects whose Draw functions end with one.
//
// The integer results are the reference's bit for bit, including the
// unsigned wrap-around of the scalar blends when a blended channel falls
// below the destination (the wrapped bits that reach the packed pixel are
// kept).
#pragma once

#include "ckd_pixel.h"

namespace {

enum CompOp {
    OP_MIX,          // p0 = alpha (0..255)
    OP_MIXOVER,
    OP_SUB,
    OP_EXCL,
    OP_SOFTLIGHT,
    OP_SOFTLIGHT_A,
    OP_SOFTLIGHT_AA, // p0 = alpha (0..255)
    OP_OVERLAY,
    OP_OVERLAY_A,
    OP_DARKEN50,
    OP_MULSRC,
    OP_MULSRC_A,
    OP_MIXSRC,
    OP_MIXSRC_A,     // p0 = the four byte lanes of 0x01010101 * fixed alpha
    OP_FADE,         // p0 = colour, p1 = alpha (0..255); no source
};

CKD_DEV uint32_t chan(uint32_t px, int k) { return (px >> (8 * k)) & 0xffu; }

// ((d<<8) + a*(s-d)) >> 8 of the reference's 16-bit lanes: with a <= 255 the
// exact value d*(256-a) + s*a lies in [0, 65280], so the mod-2^16 sum is
// exact and two channels share a register without carries
CKD_DEV uint32_t mix_uniform(uint32_t d, uint32_t s, uint32_t a)
{
    const uint32_t M = 0x00ff00ffu;
    const ckd_u16x2 f = ckd_splat16(a), nf = (unsigned short)256 - f;
    return ckd_u32_of(ckd_lerp2x(ckd_u16x2_of(d & M), ckd_u16x2_of(s & M), nf, f)) |
           (ckd_u32_of(ckd_lerp2x(ckd_u16x2_of(ckd_odd_bytes(d)), ckd_u16x2_of(ckd_odd_bytes(s)), nf, f))
            << 8);
}

// SoftLightBlend (code/util.cpp:227-240)
CKD_DEV uint32_t soft_light(uint32_t A, uint32_t B)
{
    const uint32_t dA = (A >> 1) + 64;
    return B < 128 ? (2 * dA * B) >> 8 : 255 - ((2 * (255 - dA) * (255 - B)) >> 8);
}

// the integer overlay of one channel (code/util.cpp:449-451); the divisor
// 255 is exact integer division
CKD_DEV uint32_t overlay(uint32_t b, uint32_t t)
{
    return b < 128 ? (2 * b * t) / 255 : 255 - (2 * (255 - b) * (255 - t)) / 255;
}

// c2 + (((c - c2) * a) >> 8) in unsigned 32-bit arithmetic
CKD_DEV uint32_t blend_wrap(uint32_t c2, uint32_t c, uint32_t a) { return c2 + (((c - c2) * a) >> 8); }

template <int kOp> CKD_DEV uint32_t blend_px(uint32_t d, uint32_t s, uint32_t p0, uint32_t p1)
{
    if constexpr (kOp == OP_MIX) return mix_uniform(d, s, p0);
    if constexpr (kOp == OP_FADE) return mix_uniform(d, p0, p1);
    if constexpr (kOp == OP_MIXSRC) return mix_uniform(d, s, s >> 24);
    if constexpr (kOp == OP_MIXSRC_A) {
        // per lane a_k = (srcA * f_k) >> 8 (code/util.cpp:741-746)
        uint32_t r = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t a = ((s >> 24) * chan(p0, k)) >> 8;
            r |= ((chan(d, k) * (256 - a) + chan(s, k) * a) >> 8) << (8 * k);
        }
        return r;
    }
    if constexpr (kOp == OP_SUB) {
        uint32_t r = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int v = (int)chan(d, k) - (int)chan(s, k);
            r |= (uint32_t)(v > 0 ? v : 0) << (8 * k);
        }
        return r;
    }
    if constexpr (kOp == OP_MULSRC || kOp == OP_MULSRC_A) {
        uint32_t r = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            r |= ((chan(d, k) * (kOp == OP_MULSRC ? chan(s, k) : (s >> 24))) >> 8) << (8 * k);
        return r;
    }
    if constexpr (kOp == OP_MIXOVER) {
        // code/util.cpp:148-177: weight srcA on the source, 255-srcA on dst
        const uint32_t sa = s >> 24, da = 0xffu - sa;
        uint32_t r = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = ((chan(s, k) * sa) >> 8) + ((chan(d, k) * da) >> 8);
            r |= (v > 255 ? 255 : v) << (8 * k);
        }
        return r;
    }
    if constexpr (kOp == OP_EXCL || kOp == OP_DARKEN50 || kOp == OP_SOFTLIGHT) {
        uint32_t r = d & 0xff000000u;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t a = chan(s, k), b = chan(d, k);
            uint32_t v;
            if constexpr (kOp == OP_EXCL) v = a + b - ((2 * a * b) >> 8);
            if constexpr (kOp == OP_DARKEN50) v = (b + (a < b ? a : b)) >> 1;
            if constexpr (kOp == OP_SOFTLIGHT) v = soft_light(a, b);
            r |= v << (8 * k);
        }
        return r;
    }
    if constexpr (kOp == OP_OVERLAY)
        return (overlay(chan(d, 2), chan(s, 2)) << 16) | (overlay(chan(d, 1), chan(s, 1)) << 8) |
               overlay(chan(d, 0), chan(s, 0));
    if constexpr (kOp == OP_SOFTLIGHT_A || kOp == OP_SOFTLIGHT_AA || kOp == OP_OVERLAY_A) {
        const uint32_t a = kOp == OP_SOFTLIGHT_AA ? p0 : (s >> 24);
        uint32_t c[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t b = chan(d, k);
            const uint32_t t = kOp == OP_OVERLAY_A ? overlay(b, chan(s, k)) : soft_light(chan(s, k), b);
            c[k] = blend_wrap(b, t, a);
        }
        const uint32_t alpha = kOp == OP_SOFTLIGHT_AA ? (p0 << 24) : 0u;
        return alpha | (c[2] << 16) | (c[1] << 8) | c[0];
    }
    return d;
}

} // namespace
