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
// cannot prove the operands' range); a uniform operand stays in an SGPR
CKD_DEV uint32_t ckd_mul_u24(uint32_t a, uint32_t b)
{
    return __umul24(a, b);
}

// the same product where only its low bits are used: the compiler turns
// __umul24 there into v_mul_lo_u32 (quarter rate), so it is written out
CKD_DEV uint32_t ckd_mul_u24_v(uint32_t a, uint32_t b)
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

// Two 16-bit lanes, for the packed 16-bit integer VALU operations
// (v_pk_mul_lo_u16, v_pk_mad_u16, v_pk_lshrrev_b16, v_pk_add_u16, ...): one
// instruction works on both channels of a 0x00XX00YY register, and the
// lanes never carry into each other.
typedef unsigned short ckd_u16x2 __attribute__((ext_vector_type(2)));

CKD_DEV ckd_u16x2 ckd_u16x2_of(uint32_t v) { return __builtin_bit_cast(ckd_u16x2, v); }
CKD_DEV uint32_t ckd_u32_of(ckd_u16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// v_perm_b32: byte k of the result is byte sel.k of the 8-byte value
// hi:lo (0-3 select lo's bytes, 4-7 hi's, 0x0c gives 0)
CKD_DEV uint32_t ckd_perm(uint32_t hi, uint32_t lo, uint32_t sel)
{
    return __builtin_amdgcn_perm(hi, lo, sel);
}

// (v >> 8) & 0x00ff00ff: bytes 1 and 3 into the low bytes of the 16-bit fields
CKD_DEV uint32_t ckd_odd_bytes(uint32_t v) { return ckd_perm(0u, v, 0x0c030c01u); }

// a*(256-f) + b*f >> 8 in each 16-bit lane: f and nf = 256 - f are splat
// into both lanes; every lane stays below 2^16 (ckd_lerp2), so the packed
// mod-2^16 arithmetic is exact
CKD_DEV ckd_u16x2 ckd_lerp2x(ckd_u16x2 a, ckd_u16x2 b, ckd_u16x2 nf, ckd_u16x2 f)
{
    return (a * nf + b * f) >> (unsigned short)8;
}

CKD_DEV ckd_u16x2 ckd_splat16(uint32_t f) { return ckd_u16x2{(unsigned short)f, (unsigned short)f}; }

// Bilinear blend of four ARGB texels with 8-bit fractions.  Equals
// bsamp32_16 and bsamp32_32 (code/bilinear.h:58-125) channel for channel:
// their 16-bit wrap-around intermediates equal S0*(256-f)+S1*f, which lies
// in [0, 65280], so the mod-2^16 and 32-bit forms agree with this one.
// B,R and G,A go through the packed 16-bit lerps in pairs.
CKD_DEV uint32_t ckd_bilerp32(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3,
                              uint32_t fu, uint32_t fv)
{
    const uint32_t M = 0x00ff00ffu;
    const ckd_u16x2 f = ckd_splat16(fu), g = ckd_splat16(fv);
    const ckd_u16x2 nf = (unsigned short)256 - f, ng = (unsigned short)256 - g;
    // the vertical blend before its >> 8: each 16-bit lane's quotient by 256
    // is its high byte, so one byte permute packs B, G, R, A from bytes 1
    // and 3 of br and ga
    const ckd_u16x2 br = ckd_lerp2x(ckd_u16x2_of(p0 & M), ckd_u16x2_of(p1 & M), nf, f) * ng +
                         ckd_lerp2x(ckd_u16x2_of(p2 & M), ckd_u16x2_of(p3 & M), nf, f) * g;
    const ckd_u16x2 ga =
        ckd_lerp2x(ckd_u16x2_of(ckd_odd_bytes(p0)), ckd_u16x2_of(ckd_odd_bytes(p1)), nf, f) * ng +
        ckd_lerp2x(ckd_u16x2_of(ckd_odd_bytes(p2)), ckd_u16x2_of(ckd_odd_bytes(p3)), nf, f) * g;
    return ckd_perm(ckd_u32_of(ga), ckd_u32_of(br), 0x07030501u);
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
    const uint32_t M = 0x00ff00ffu;
    const ckd_u16x2 br = __builtin_elementwise_sub_sat(ckd_u16x2_of(c & M), ckd_u16x2_of(f & M));
    const ckd_u16x2 ga =
        __builtin_elementwise_sub_sat(ckd_u16x2_of(ckd_odd_bytes(c)), ckd_u16x2_of(ckd_odd_bytes(f)));
    return ckd_u32_of(br) | (ckd_u32_of(ga) << 8);
}

// Saturating per-byte add (_mm_add_epi16 of widened bytes + packus_epi16),
// two channels per packed 16-bit add and min.
CKD_DEV uint32_t ckd_adds_u8x4(uint32_t a, uint32_t b)
{
    const uint32_t M = 0x00ff00ffu;
    const ckd_u16x2 top = ckd_splat16(255u);
    const ckd_u16x2 br = __builtin_elementwise_min(ckd_u16x2_of(a & M) + ckd_u16x2_of(b & M), top);
    const ckd_u16x2 ga =
        __builtin_elementwise_min(ckd_u16x2_of(ckd_odd_bytes(a)) + ckd_u16x2_of(ckd_odd_bytes(b)), top);
    return ckd_u32_of(br) | (ckd_u32_of(ga) << 8);
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
// step of ckd_span_setup is d * 0 + (d < 0 ? -1 : 0); such a span draws one
// pixel, from acc, and never adds its step, so the step here is d * int16
// of the divisor for every length.  acc, and step for length >= 2, equal
// ckd_span_setup's.
CKD_DEV void ckd_span_setup_drawn(CkdSpan &s, uint32_t A, uint32_t B, uint32_t length,
                                  uint32_t drawLength, uint32_t divisor)
{
    const int32_t m = (int32_t)(int16_t)ckd_mul_u24_v(divisor, length - drawLength);
    const int32_t dlo = (int32_t)(int16_t)divisor;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int32_t a = (int32_t)((A >> (8 * k)) & 0xff), b = (int32_t)((B >> (8 * k)) & 0xff);
        const int32_t d = b - a;
        s.step[k] = (uint32_t)(d * dlo);
        s.acc[k] = ((uint32_t)a << 16) + ((k & 1) ? 0u : (uint32_t)(d * m));
    }
}

// one span pixel: v = acc>>16 (logical); packus_epi32 keeps v (<= 65535);
// packus_epi16 maps v > 32767 to 0 and clamps to 255.  v > 32767 is a
// negative acc, so the channel is byte 2 of acc clamped (signed) to
// [0, 0xffffff]: one v_med3_i32 per channel, two byte permutes per pixel.
CKD_DEV uint32_t ckd_span_pixel(const uint32_t acc[4])
{
    uint32_t m[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) m[k] = (uint32_t)min(max((int32_t)acc[k], 0), 0xffffff);
    return ckd_perm(m[1], m[0], 0x0c0c0602u) | ckd_perm(m[3], m[2], 0x06020c0cu);
}

// ---- 2007 box blur helpers --------------------------------------------
// Four 8-bit channels spread into the 16-bit fields of a u64: window sums
// of up to 255 pixels stay below 2^16 per field, so sums and differences of
// spread values (prefix sums included) are exact field by field.
CKD_DEV uint64_t ckd_spread(uint32_t px)
{
    const uint32_t lo = ckd_perm(0u, px, 0x0c010c00u);  // 0x00b1_00b0
    const uint32_t hi = ckd_perm(0u, px, 0x0c030c02u);  // 0x00b3_00b2
    return ((uint64_t)hi << 32) | lo;
}

// Div (code/deprecated/boxblur.cpp:39-42): mulhi_epu16(acc, div) per field,
// then packus_epi16, for a window sum of `count` pixels and its divisor
// div = ckd_b07_count_div(count).  Each field of acc is at most 255*count
// and div at most 2^16/count, so every quotient is at most 255: packus
// never saturates, and the quotient is byte 2 of the 32-bit product (both
// factors below 2^16, so v_mul_u32_u24 is exact).
CKD_DEV uint32_t ckd_b07_div_pack(uint64_t acc, uint32_t div)
{
    const uint32_t lo = (uint32_t)acc, hi = (uint32_t)(acc >> 32);
    const uint32_t q0 = __umul24(lo & 0xffffu, div), q1 = __umul24(lo >> 16, div);
    const uint32_t q2 = __umul24(hi & 0xffffu, div), q3 = __umul24(hi >> 16, div);
    return ckd_perm(q1, q0, 0x0c0c0602u) | ckd_perm(q3, q2, 0x06020c0cu);
}

// WeightToDiv(16*count) (code/deprecated/boxblur.cpp:11-14) stored through
// _mm_set1_epi16: the divisor of an odd-span window of `count` pixels
__host__ __device__ constexpr inline uint32_t ckd_b07_count_div(uint32_t count)
{
    return count ? ((((1u << 24) / (16u * count)) >> 4) & 0xffffu) : 0u;
}
