// ckd_grid.h -- per-pixel fetch of the per-8x8-block UV grid remap.
#pragma once

#include "ckd_pixel.h"

namespace ckd {

// Grid remap of one pixel (build-defined law):
//   corners C00,C10,C01,C11 of block (x>>3, y>>3), local (i,j) = (x&7, y&7)
//   L = C00*(8-j) + C01*j, R = C10*(8-j) + C11*j, UV = (L*(8-i) + R*i) >> 6
// then bsamp_prepUVs (code/bilinear.h:10-31) with wrap mask 2^texLog2-1 and
// bsamp32_32 (code/bilinear.h:86-125).
CKD_DEV uint32_t grid_pixel(const int2 *__restrict__ grid, unsigned GW,
                            const uint32_t *__restrict__ tex, unsigned texLog2, unsigned x,
                            unsigned y)
{
    const unsigned bx = x >> 3, by = y >> 3;
    const int i = (int)(x & 7), j = (int)(y & 7);
    const int2 *g = grid + (size_t)by * GW + bx;
    const int2 c00 = g[0], c10 = g[1], c01 = g[GW], c11 = g[GW + 1];
    const int Lu = c00.x * (8 - j) + c01.x * j, Ru = c10.x * (8 - j) + c11.x * j;
    const int Lv = c00.y * (8 - j) + c01.y * j, Rv = c10.y * (8 - j) + c11.y * j;
    const int U = (Lu * (8 - i) + Ru * i) >> 6;
    const int V = (Lv * (8 - i) + Rv * i) >> 6;
    const unsigned mask = (1u << texLog2) - 1;
    const unsigned u0 = ((unsigned)(U >> 8)) & mask, u1 = ((unsigned)(U >> 8) + 1) & mask;
    const unsigned v0 = (((unsigned)(V >> 8)) & mask) << texLog2;
    const unsigned v1 = (((unsigned)(V >> 8) + 1) & mask) << texLog2;
    return ckd_bilerp32(tex[u0 + v0], tex[u1 + v0], tex[u0 + v1], tex[u1 + v1],
                        (uint32_t)U & 0xff, (uint32_t)V & 0xff);
}

} // namespace ckd
