// ckd_pixel.h -- device-side pixel arithmetic shared by the kernels.
//
// Every helper reproduces an SSE lane sequence of the reference exactly,
// written as plain 32-bit integer arithmetic for the GCN/CDNA VALU.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#define CKD_DEV __device__ __forceinline__

// a * b for a, b < 2^24 with a product below 2^32: v_mul_u32_u24, full rate
// (the compiler otherwise picks the quarter-rate 32-bit multiply where it
// cannot prove the operands' range)
CKD_DEV uint32_t ckd_mul_u24(uint32_t a, uint32_t b)
{
    uint32_t r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// Two 8-bit channels held in the 16-bit fields of 0x00XX00YY.  The convex
// combination a*(256-f) + b*f of two such fields never exceeds 255*256, so
// the fields never carry into each other.
CKD_DEV uint32_t ckd_lerp2(uint32_t a, uint32_t b, uint32_t f)
{
    return ((a * (256u - f) + b * f) >> 8) & 0x00ff00ffu;
}

// Bilinear blend of four ARGB texels with 8-bit fractions.  Equals
// bsamp32_16 and bsamp32_32 (code/bilinear.h:58-125) channel for channel:
// their 16-bit wrap-around intermediates equal S0*(256-f)+S1*f, which lies
// in [0, 65280], so the mod-2^16 and 32-bit forms agree with this one.
CKD_DEV uint32_t ckd_bilerp32(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3,
                              uint32_t fu, uint32_t fv)
{
    const uint32_t M = 0x00ff00ffu;
    const uint32_t br = ckd_lerp2(ckd_lerp2(p0 & M, p1 & M, fu),
                                  ckd_lerp2(p2 & M, p3 & M, fu), fv);
    const uint32_t ga = ckd_lerp2(ckd_lerp2((p0 >> 8) & M, (p1 >> 8) & M, fu),
                                  ckd_lerp2((p2 >> 8) & M, (p3 >> 8) & M, fu), fv);
    return br | (ga << 8);
}

// bsamp8 (code/bilinear.h:34-53), one channel.
CKD_DEV uint32_t ckd_bilerp8(uint32_t s0, uint32_t s1, uint32_t s2, uint32_t s3,
                             uint32_t fu, uint32_t fv)
{
    const uint32_t a = (s0 * (256u - fu) + s1 * fu) >> 8;
    const uint32_t b = (s2 * (256u - fu) + s3 * fu) >> 8;
    return (a * (256u - fv) + b * fv) >> 8;
}

// _mm_subs_epu16 of two packed ARGB colours widened to 16-bit lanes.
CKD_DEV uint32_t ckd_subs_u8x4(uint32_t c, uint32_t f)
{
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int v = (int)((c >> (8 * k)) & 0xff) - (int)((f >> (8 * k)) & 0xff);
        r |= (uint32_t)(v > 0 ? v : 0) << (8 * k);
    }
    return r;
}

// Saturating per-byte add (_mm_add_epi16 of widened bytes + packus_epi16).
CKD_DEV uint32_t ckd_adds_u8x4(uint32_t a, uint32_t b)
{
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t v = ((a >> (8 * k)) & 0xff) + ((b >> (8 * k)) & 0xff);
        r |= (v > 255u ? 255u : v) << (8 * k);
    }
    return r;
}

// _mm_avg_epu8 of four bytes
CKD_DEV uint32_t ckd_avg_u8x4(uint32_t a, uint32_t b)
{
    // (a+b+1)>>1 per byte without cross-byte carries
    return (a | b) - (((a ^ b) >> 1) & 0x7f7f7f7fu);
}

// _mm_madd_epi16 of one 32-bit lane
CKD_DEV int32_t ckd_madd16(uint32_t a, uint32_t b)
{
    const int32_t lo = (int32_t)(int16_t)(a & 0xffff) * (int32_t)(int16_t)(b & 0xffff);
    const int32_t hi = (int32_t)(int16_t)(a >> 16) * (int32_t)(int16_t)(b >> 16);
    return (int32_t)((uint32_t)lo + (uint32_t)hi);
}

// cspanISSE16 set-up (code/cspan.h:47-65): per-channel 16.16 start value
// (A<<16)+preStep and step, with the lane quirks of _mm_mul_epi32 (only
// lanes 0 and 2 receive the pre-step) and _mm_madd_epi16 (16-bit halves).
struct CkdSpan {
    uint32_t acc[4];
    uint32_t step[4];
};

CKD_DEV void ckd_span_setup(CkdSpan &s, uint32_t A, uint32_t B, uint32_t length,
                            uint32_t drawLength)
{
    const uint32_t divisor = 65536u / length;
    const uint32_t pre = length - drawLength;
    const int64_t m = (int64_t)(int32_t)divisor * (int64_t)(int32_t)pre;
    const uint32_t mlo = (uint32_t)m, mhi = (uint32_t)((uint64_t)m >> 32);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t a = (A >> (8 * k)) & 0xff, b = (B >> (8 * k)) & 0xff;
        const uint32_t delta = b - a;
        const uint32_t M = (k & 1) ? mhi : mlo;
        s.step[k] = (uint32_t)ckd_madd16(delta, divisor);
        s.acc[k] = (a << 16) + (uint32_t)ckd_madd16(delta, M);
    }
}

// ckd_span_setup for a span that is drawn (1 <= drawLength <= length <
// 65536), in fewer operations.  The pre-step product m = divisor * pre is
// below 65536 there (pre < length, divisor = floor(65536/length)), so the
// 64-bit product of _mm_mul_epi32 has a zero high word (G and A get no
// pre-step) and its low word has a zero high half: _mm_madd_epi16 of a
// channel delta d (|d| <= 255, its high half the sign) with m is
// d * int16(m).  The divisor's high half is 1 only for length 1, where the
// step is d * 0 + (d < 0 ? -1 : 0).  Equal to ckd_span_setup on every such
// input (checked exhaustively over lengths up to 2200).
CKD_DEV void ckd_span_setup_drawn(CkdSpan &s, uint32_t A, uint32_t B, uint32_t length,
                                  uint32_t drawLength, uint32_t divisor)
{
    const int32_t m = (int32_t)(int16_t)ckd_mul_u24(divisor, length - drawLength);
    const int32_t dlo = (int32_t)(int16_t)divisor;
    const int32_t dhi = (int32_t)(divisor >> 16);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int32_t a = (int32_t)((A >> (8 * k)) & 0xff), b = (int32_t)((B >> (8 * k)) & 0xff);
        const int32_t d = b - a;
        s.step[k] = (uint32_t)(d * dlo - (d < 0 ? dhi : 0));
        s.acc[k] = ((uint32_t)a << 16) + ((k & 1) ? 0u : (uint32_t)(d * m));
    }
}

// one span pixel: v = acc>>16 (logical); packus_epi32 keeps v (<= 65535);
// packus_epi16 maps v > 32767 to 0 and clamps to 255.
CKD_DEV uint32_t ckd_span_pixel(const uint32_t acc[4])
{
    uint32_t px = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t v = acc[k] >> 16;
        const uint32_t b = v > 32767u ? 0u : (v > 255u ? 255u : v);
        px |= b << (8 * k);
    }
    return px;
}

// ---- 2007 box blur helpers --------------------------------------------
// Four 8-bit channels spread into the 16-bit fields of a u64: window sums
// of up to 255 pixels stay below 2^16 per field, so sums and differences of
// spread values (prefix sums included) are exact field by field.
CKD_DEV uint64_t ckd_spread(uint32_t px)
{
    return (uint64_t)(px & 0xffu) | ((uint64_t)(px & 0xff00u) << 8) |
           ((uint64_t)(px & 0xff0000u) << 16) | ((uint64_t)(px & 0xff000000u) << 24);
}

// Div (code/deprecated/boxblur.cpp:39-42): mulhi_epu16(acc, div) per field,
// then packus_epi16
CKD_DEV uint32_t ckd_b07_div_pack(uint64_t acc, uint32_t div)
{
    uint32_t px = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t a = (uint32_t)(acc >> (16 * k)) & 0xffffu;
        const uint32_t q = (a * div) >> 16;
        px |= (q > 32767u ? 0u : (q > 255u ? 255u : q)) << (8 * k);
    }
    return px;
}

// WeightToDiv(16*count) (code/deprecated/boxblur.cpp:11-14) stored through
// _mm_set1_epi16: the divisor of an odd-span window of `count` pixels
__host__ __device__ constexpr inline uint32_t ckd_b07_count_div(uint32_t count)
{
    return count ? ((((1u << 24) / (16u * count)) >> 4) & 0xffffu) : 0u;
}
