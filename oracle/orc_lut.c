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
