// ckd_grid.h -- per-pixel fetch of the per-8x8-block UV grid remap.
#pragma once

#include "ckd_pixel.h"

namespace ckd {

// The texture half of grid_pixel, from (U, V) << 6: 18:14 fixed point, as
// the per-block interpolation (8 * corner + i * step) produces them before
// its >> 6.  Texel column (U >> 8) & mask is bits [14, 14 + texLog2) of
// X, the next column ((U >> 8) + 1) & mask those bits of X + 2^14, the
// fraction U & 0xff bits [6, 14) (texLog2 <= 14).
CKD_DEV uint32_t grid_sample_q14(const uint32_t *__restrict__ tex, unsigned texLog2, uint32_t X,
                                 uint32_t Y)
{
    const uint32_t u0 = __builtin_amdgcn_ubfe(X, 14, texLog2);
    const uint32_t u1 = __builtin_amdgcn_ubfe(X + (1u << 14), 14, texLog2);
    const uint32_t v0 = __builtin_amdgcn_ubfe(Y, 14, texLog2) << texLog2;
    const uint32_t v1 = __builtin_amdgcn_ubfe(Y + (1u << 14), 14, texLog2) << texLog2;
    // 32-bit byte offsets from the texture's base (below 2^30), so each
    // fetch adds one offset register to a uniform base
    auto at = [tex](uint32_t k) { return *(const uint32_t *)((const char *)tex + (k << 2)); };
    return ckd_bilerp32(at(u0 + v0), at(u1 + v0), at(u0 + v1), at(u1 + v1),
                        __builtin_amdgcn_ubfe(X, 6, 8), __builtin_amdgcn_ubfe(Y, 6, 8));
}

// grid_sample_q14 from the texture's texel quads: quads[v * 2^texLog2 + u]
// holds texels (u,v), (u+1,v), (u,v+1), (u+1,v+1), indices wrapped, so the
// four texels of a sample are one 16-byte load.
CKD_DEV uint32_t grid_sample_quad(const uint4 *__restrict__ quads, unsigned texLog2, uint32_t X,
                                  uint32_t Y)
{
    const uint32_t u0 = __builtin_amdgcn_ubfe(X, 14, texLog2);
    const uint32_t v0 = __builtin_amdgcn_ubfe(Y, 14, texLog2);
    // byte offset below 2^32 (texLog2 <= 13)
    const uint32_t k = ((v0 << texLog2) + u0) << 4;
    const uint4 q = *(const uint4 *)((const char *)quads + k);
    // the fractions as one v_bfe_u32 each (the compiler otherwise splits
    // them into a 16-bit shift and a mask)
    uint32_t fu = __builtin_amdgcn_ubfe(X, 6, 8), fv = __builtin_amdgcn_ubfe(Y, 6, 8);
    asm("" : "+v"(fu), "+v"(fv));
    return ckd_bilerp32(q.x, q.y, q.z, q.w, fu, fv);
}

// The bilinear sample at the interpolated (U, V), both 24:8 fixed point
// (the bits grid_sample_q14 reads are U's bits [0, 8 + texLog2)).
CKD_DEV uint32_t grid_sample(const uint32_t *__restrict__ tex, unsigned texLog2, int U, int V)
{
    return grid_sample_q14(tex, texLog2, (uint32_t)U << 6, (uint32_t)V << 6);
}

// Grid remap of one pixel (build-defined law):
//   corners C00,C10,C01,C11 of block (x>>3, y>>3), local (i,j) = (x&7, y&7)
//   L = C00*(8-j) + C01*j, R = C10*(8-j) + C11*j, UV = (L*(8-i) + R*i) >> 6
// then bsamp_prepUVs (code/bilinear.h:10-31) with wrap mask 2^texLog2-1 and
// bsamp32_32 (code/bilinear.h:86-125).
// The grid is read as (U, V) int pairs through IntPtr: `const int32_t *`, or
// a pointer into the global address space when the caller's grid comes from
// memory (global loads rather than flat ones).
template <typename IntPtr>
CKD_DEV uint32_t grid_pixel(IntPtr grid, unsigned GW, const uint32_t *__restrict__ tex,
                            unsigned texLog2, unsigned x, unsigned y)
{
    const unsigned bx = x >> 3, by = y >> 3;
    const int i = (int)(x & 7), j = (int)(y & 7);
    const auto g0 = grid + 2 * ((size_t)by * GW + bx), g1 = g0 + 2 * (size_t)GW;
    const int c00x = g0[0], c00y = g0[1], c10x = g0[2], c10y = g0[3];
    const int c01x = g1[0], c01y = g1[1], c11x = g1[2], c11y = g1[3];
    const int Lu = c00x * (8 - j) + c01x * j, Ru = c10x * (8 - j) + c11x * j;
    const int Lv = c00y * (8 - j) + c01y * j, Rv = c10y * (8 - j) + c11y * j;
    return grid_sample(tex, texLog2, (Lu * (8 - i) + Ru * i) >> 6, (Lv * (8 - i) + Rv * i) >> 6);
}

} // namespace ckd
