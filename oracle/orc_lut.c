/*
 * orc_lut.c -- interpolated cosine LUT (code/sincos-lut.h, sincos-lut.cpp).
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 */
#include "ckd_oracle.h"

#include <math.h>

#define ORC_COS_TAB 2048

/* code/sincos-lut.cpp:9-16 */
void orc_cos_lut(float lut[2049])
{
    const float kPI = 3.1415926535897932384626433832795f;
    const float k2PI = 2.f * kPI;
    for (unsigned i = 0; i < ORC_COS_TAB; ++i)
        lut[i] = cosf((float)i * (k2PI / (float)ORC_COS_TAB));
    lut[ORC_COS_TAB] = lut[0];
}

/* code/sincos-lut.h:13-19 */
float orc_lutcosf(const float *lut, float angle)
{
    const float kPI = 3.1415926535897932384626433832795f;
    const float k2PI = 2.f * kPI;
    angle = fabsf(angle);
    angle *= (1.f / k2PI) * (float)ORC_COS_TAB;
    const int index = (int)angle & (ORC_COS_TAB - 1);
    const float frac = angle - truncf(angle);
    return lut[index] + (lut[index + 1] - lut[index]) * frac;
}

/* code/sincos-lut.h:21-26 (ARRESTED_DEV_LEGACY: cos(angle + pi/2) = -sin) */
float orc_lutsinf(const float *lut, float angle)
{
    const float kPI = 3.1415926535897932384626433832795f;
    return orc_lutcosf(lut, angle + kPI * 0.5f);
}

/* code/util.cpp:552-603 (TapeWarp32).  The fetch's row stride is xRes;
 * the reference uses kResX, equal to xRes for its 1280-wide calls. */
void orc_tape_warp32(uint32_t *dst, const uint32_t *src, unsigned xRes, unsigned yRes,
                     float strength, float speed, const float *lut)
{
    for (int iY = 0; iY < (int)yRes; ++iY) {
        for (unsigned iX = 0; iX < xRes; ++iX) {
            const float distance = 1.f;
            const float dX = orc_lutsinf(lut, iY * speed) * strength * distance;
            const float dY = orc_lutcosf(lut, iX * speed) * strength * distance;
            float tX = iX + dX;
            float tY = iY + dY;
            if (tX < 0.f) tX = 0.f;
            else if (tX >= xRes - 1.f) tX = xRes - 2.f;
            if (tY < 0.f) tY = 0.f;
            else if (tY >= yRes - 1.f) tY = yRes - 2.f;
            const int U = orc_ftofp24(tX), V = orc_ftofp24(tY);
            const unsigned U0 = (unsigned)(U >> 8);
            const unsigned V0 = (unsigned)(V >> 8) * xRes;
            const unsigned fu = ((unsigned)U & 0xff) * 0x01010101u;
            const unsigned fv = ((unsigned)V & 0xff) * 0x01010101u;
            uint16_t v[4];
            orc_bsamp32_16(src, U0, V0, U0 + 1, V0 + xRes, fu, fv, v);
            uint32_t px = 0;
            for (int c = 0; c < 4; ++c) px |= (uint32_t)(v[c] > 255 ? 255 : v[c]) << (8 * c);
            dst[(size_t)iY * xRes + iX] = px;
        }
    }
}
