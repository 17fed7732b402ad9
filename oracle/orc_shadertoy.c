/*
 * orc_shadertoy.c -- the reference's own raymarched effects
 * (code/shadertoy.cpp) and their helpers (code/shadertoy-util.h),
 * restated in scalar C, one FX-map pixel at a time.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * The float operations follow the reference's order one for one (FP32,
 * no contraction: this file is built with -ffp-contract=off).  Helpers
 * that the reference runs on four SSE lanes are restated per lane,
 * including the padding lane where it reaches the output (Desaturate,
 * GammaAdj, ToPixel4).  Two choices are explicit:
 *
 *  - rsqrtMode 0: the reciprocal square roots of vNorm4 and Specular are
 *    1/sqrtf (correctly rounded, portable);
 *    rsqrtMode 1: the x86 rsqrtss estimate, as the reference's
 *    _mm_rsqrt_ps (its table is CPU-model specific, so this mode only
 *    agrees with a reference built and run on the same machine).
 *  - GammaAdj uses exp_ps/log_ps of 3rdparty/sse_mathfun.h, which are the
 *    Cephes expf/logf polynomials (S. L. Moshier) in J. Pommier's 4-lane
 *    form; they are restated below step for step so the result is exact.
 *
 * Map geometry: an output of resX x resY has an FX map of
 * (resX/2+4) x (resY/2+4) (code/fx-blitter.h:13-17) and aspect resY/resX
 * (code/main.h:43-44).
 */
#include "ckd_oracle.h"

#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <immintrin.h>

#define K_PI 3.1415926535897932384626433832795f
#define K_2PI (2.f * K_PI)
#define K_GOLDEN_RATIO 1.61803398875f
#define K_GOLDEN_ANGLE 2.39996f
#define K_EPSILON 1.1920928955078125e-07f

typedef struct st_ctx {
    const float *lut;
    int rsq;
    unsigned mapW, mapH;
    float aspect, oneOverAspect;
} st_ctx;

static float lcos(const st_ctx *c, float a) { return orc_lutcosf(c->lut, a); }
static float lsin(const st_ctx *c, float a) { return orc_lutsinf(c->lut, a); }

static uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* _mm_dp_ps(a, b, 0xff): (a0*b0 + a1*b1) + (a2*b2 + a3*b3) */
static float dp4(const float a[4], const float b[4])
{
    return (a[0] * b[0] + a[1] * b[1]) + (a[2] * b[2] + a[3] * b[3]);
}

/* Vector3::Dot, code/Std3DMath-stripped/Vector3.h:19-22 */
static float dot3(const float a[4], const float b[4])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* the x86 rsqrtss estimate of this CPU (probe for fixtures written with it) */
float orc_rsqrtss(float x) { return _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(x))); }

static float rsqrt_of(const st_ctx *c, float x)
{
    if (c->rsq) return _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(x)));
    return 1.f / sqrtf(x);
}

/* vNorm4 / vFastNorm3, code/shadertoy-util.h:78-101: all four lanes */
static void vnorm(const st_ctx *c, float v[4])
{
    const float r = rsqrt_of(c, dp4(v, v));
    for (int k = 0; k < 4; ++k) v[k] *= r;
}

/* vFastLen3, code/shadertoy-util.h:69-76 */
static float vlen(const float p[4]) { return sqrtf(dp4(p, p)); }

/* vLerp4, code/shadertoy-util.h:62-67 */
static void vlerp4(float out[4], const float A[4], const float B[4], float f)
{
    for (int k = 0; k < 4; ++k) out[k] = A[k] + f * (B[k] - A[k]);
}

/* rotX/rotY/rotZ, code/shadertoy-util.h:31-59 */
static void rot_x(const st_ctx *c, float angle, float *Y, float *Z)
{
    const float cs = lcos(c, angle), sn = lsin(c, angle);
    const float ry = cs * *Y + -sn * *Z;
    const float rz = sn * *Y + cs * *Z;
    *Y = ry;
    *Z = rz;
}

static void rot_y(const st_ctx *c, float angle, float *X, float *Z)
{
    const float cs = lcos(c, angle), sn = lsin(c, angle);
    const float rx = cs * *X + sn * *Z;
    const float rz = -sn * *X + cs * *Z;
    *X = rx;
    *Z = rz;
}

static void rot_z(const st_ctx *c, float angle, float *X, float *Y)
{
    const float cs = lcos(c, angle), sn = lsin(c, angle);
    const float rx = cs * *X + sn * *Y;
    const float ry = -sn * *X + cs * *Y;
    *X = rx;
    *Y = ry;
}

/* ToUV_FxMap, code/shadertoy-util.h:123-130 */
static void to_uv(const st_ctx *c, unsigned iX, unsigned iY, float scale, float *u, float *v)
{
    float fX = (float)iX, fY = (float)iY;
    fX *= 1.f / (float)c->mapW;
    fY *= 1.f / (float)c->mapH;
    *u = (fX - 0.5f) * scale * c->oneOverAspect;
    *v = (fY - 0.5f) * scale;
}

/* MichielPal, code/shadertoy-util.h:193-198 */
static void michiel_pal(const st_ctx *c, float phase, float out[4])
{
    out[0] = .1f - lcos(c, phase / 3.f) / (19.f * 0.5f);
    out[1] = .1f;
    out[2] = .1f + lcos(c, phase / 14.f) / 4.f;
    out[3] = 0.f;
}

/* Desaturate, code/shadertoy-util.h:244-249 (the padding lane becomes
 * amount*luma) */
static void desaturate(const float color[4], float amount, float out[4])
{
    const float w[4] = {0.0722f, 0.7152f, 0.2126f, 0.f};
    const float luma = dp4(w, color);
    const float l4[4] = {luma, luma, luma, luma};
    vlerp4(out, color, l4, amount);
}

/* ExpFog, code/shadertoy-util.h:237-241 */
static float exp_fog(float distance, float scale)
{
    return 1.f - expf(-scale * distance * distance * distance);
}

/* SSE minps/maxps: the second operand when either is NaN */
static float sse_min(float a, float b) { return a < b ? a : b; }
static float sse_max(float a, float b) { return a > b ? a : b; }

/* log_ps, 3rdparty/sse_mathfun.h (Cephes logf), one lane */
float orc_log_ps(float x)
{
    const int invalid = x <= 0.f;
    x = sse_max(x, bitsf(0x00800000u));
    int32_t emm0 = (int32_t)(fbits(x) >> 23);
    x = bitsf((fbits(x) & ~0x7f800000u) | fbits(0.5f));
    emm0 -= 0x7f;
    float e = (float)emm0;
    e = e + 1.f;
    const int mask = x < (float)0.707106781186547524;
    const float tmp0 = mask ? x : 0.f;
    x = x - 1.f;
    e = e - (mask ? 1.f : 0.f);
    x = x + tmp0;
    const float z = x * x;
    float y = (float)7.0376836292E-2;
    y = y * x; y = y + (float)-1.1514610310E-1;
    y = y * x; y = y + (float)1.1676998740E-1;
    y = y * x; y = y + (float)-1.2420140846E-1;
    y = y * x; y = y + (float)1.4249322787E-1;
    y = y * x; y = y + (float)-1.6668057665E-1;
    y = y * x; y = y + (float)2.0000714765E-1;
    y = y * x; y = y + (float)-2.4999993993E-1;
    y = y * x; y = y + (float)3.3333331174E-1;
    y = y * x;
    y = y * z;
    float tmp = e * (float)-2.12194440e-4;
    y = y + tmp;
    tmp = z * 0.5f;
    y = y - tmp;
    tmp = e * (float)0.693359375;
    x = x + y;
    x = x + tmp;
    if (invalid) x = bitsf(0xffffffffu);
    return x;
}

/* exp_ps, 3rdparty/sse_mathfun.h (Cephes expf), one lane */
float orc_exp_ps(float x)
{
    x = sse_min(x, 88.3762626647949f);
    x = sse_max(x, -88.3762626647949f);
    float fx = x * (float)1.44269504088896341;
    fx = fx + 0.5f;
    float tmp = (float)(int32_t)fx;
    const float mask = tmp > fx ? 1.f : 0.f;
    fx = tmp - mask;
    tmp = fx * (float)0.693359375;
    float z = fx * (float)-2.12194440e-4;
    x = x - tmp;
    x = x - z;
    z = x * x;
    float y = (float)1.9875691500E-4;
    y = y * x; y = y + (float)1.3981999507E-3;
    y = y * x; y = y + (float)8.3334519073E-3;
    y = y * x; y = y + (float)4.1665795894E-2;
    y = y * x; y = y + (float)1.6666665459E-1;
    y = y * x; y = y + (float)5.0000001201E-1;
    y = y * z;
    y = y + x;
    y = y + 1.f;
    const int32_t n = (int32_t)fx + 0x7f;
    const float pow2n = bitsf((uint32_t)n << 23);
    return y * pow2n;
}

/* GammaAdj, code/shadertoy-util.h:186-190 */
static void gamma_adj(float c[4], float gamma)
{
    for (int k = 0; k < 4; ++k) c[k] = orc_exp_ps(gamma * orc_log_ps(c[k]));
}

/* cvtps_epi32 (nearest even; NaN and out of range give INT_MIN), then
 * packus_epi32 and packus_epi16 of one lane */
static uint32_t pack_lane(float v)
{
    int32_t i;
    if (v == v && v < 2147483648.f && v >= -2147483648.f) i = (int32_t)nearbyintf(v);
    else i = INT32_MIN;
    if (i < 0) i = 0;
    const uint32_t u16 = i > 65535 ? 65535u : (uint32_t)i;
    const int16_t s16 = (int16_t)(uint16_t)u16;
    return s16 < 0 ? 0u : (s16 > 255 ? 255u : (uint32_t)s16);
}

/* ToPixel4, code/shadertoy-util.h:136-146 (one of its four pixels) */
uint32_t orc_topixel(const float c[4])
{
    uint32_t px = 0;
    for (int k = 0; k < 4; ++k) px |= pack_lane(255.f * c[k]) << (8 * k);
    return px;
}

/* ToPixel4_NoConv, code/shadertoy-util.h:148-157 */
static uint32_t topixel_noconv(const float c[4])
{
    uint32_t px = 0;
    for (int k = 0; k < 4; ++k) px |= pack_lane(c[k]) << (8 * k);
    return px;
}

/* Q3_rsqrtf<2>, code/q3-rsqrt.h:23-41 */
static float q3_rsqrt2(float x)
{
    const float half = 0.5f * x;
    int32_t iX = (int32_t)fbits(x);
    iX = 0x5f3759df - (iX >> 1);
    x = bitsf((uint32_t)iX);
    for (int i = 0; i < 2; ++i) x = x * (1.5f - half * x * x);
    return x;
}

/* Specular, code/shadertoy-util.h:251-279 */
static float specular(const st_ctx *c, const float origin[4], const float pos[4],
                      const float normal[4], const float lightDir[4], float power)
{
    float V[4], H[4];
    for (int k = 0; k < 4; ++k) V[k] = origin[k] - pos[k];
    const float rV = rsqrt_of(c, dp4(V, V));
    for (int k = 0; k < 4; ++k) V[k] *= rV;
    for (int k = 0; k < 4; ++k) H[k] = lightDir[k] + V[k];
    const float rH = rsqrt_of(c, dp4(H, H));
    for (int k = 0; k < 4; ++k) H[k] *= rH;
    const float cosAng = dp4(normal, H);
    return (fbits(cosAng) >> 31) == 0 ? powf(cosAng, power) : 0.f;
}

/* std::max<float>(a, b) and std::min<float>(a, b) */
static float std_max(float a, float b) { return (a < b) ? b : a; }
static float std_min(float a, float b) { return (b < a) ? b : a; }

/* fracf, code/Std3DMath-stripped/Math.h:49 */
static float fracf_(float v) { return v - truncf(v); }

static void set3(float v[4], float x, float y, float z) { v[0] = x; v[1] = y; v[2] = z; v[3] = 0.f; }

/* ---- plasma: code/shadertoy.cpp:200-274 -------------------------------- */

/* fPlasma, code/shadertoy.cpp:200-208 */
static float f_plasma(const st_ctx *c, const float p[4], float time)
{
    const float sine = 0.2f * lsin(c, p[0] - p[1]);
    const float fX = sine + lcos(c, p[0] * 0.33f);
    const float fY = sine + lcos(c, p[1] * 0.43f);
    const float fZ = sine + lcos(c, (5.f * time + p[2]) * 0.53f);
    return sqrtf(fX * fX + fY * fY + fZ * fZ) - 0.8f;
}

typedef struct plasma_frame {
    float colMulA[4], colMulB[4], time, dirCos, dirSin, gamma;
} plasma_frame;

/* per-frame part of RenderPlasmaMap, code/shadertoy.cpp:213-228;
 * p = time, speed, hue, gamma, desaturation */
static void plasma_setup(const st_ctx *c, const float *p, plasma_frame *f)
{
    float pal[4];
    michiel_pal(c, p[2], pal);
    desaturate(pal, p[4], f->colMulA);
    /* Vector3(__m128) keeps the padding lane; Desaturate(colMulA, .8f) */
    desaturate(f->colMulA, 0.8f, f->colMulB);
    f->time = p[0] * p[1];
    const float angle = f->time * 0.314f * 0.5f;
    f->dirCos = lcos(c, angle);
    f->dirSin = lsin(c, angle);
    f->gamma = p[3];
}

/* per-pixel body, code/shadertoy.cpp:236-268 */
static uint32_t plasma_pixel(const st_ctx *c, const plasma_frame *f, unsigned iX, unsigned iY)
{
    float u, v;
    to_uv(c, iX, iY, 4.f, &u, &v);
    float d[4];
    set3(d, f->dirCos * u * c->aspect - f->dirSin * 0.75f, v, f->dirSin * u + f->dirCos * 0.75f);
    float total = 0.f, march = 0.f;
    float hit[4] = {0.f, 0.f, 0.f, 0.f};
    for (int iStep = 0; iStep < 24; ++iStep) {
        march = f_plasma(c, hit, f->time);
        total += march * (0.5f * K_GOLDEN_RATIO);
        hit[0] = d[0] * total;
        hit[1] = d[1] * total;
        hit[2] = d[2] * total;
    }
    float half[4];
    set3(half, hit[0] * 0.5f, hit[1] * 0.5f, hit[2] * 0.5f);
    const float f2 = f_plasma(c, half, f->time);
    float col[4];
    /* Vector3 arithmetic: three lanes, padding 0 */
    set3(col, f->colMulA[0] * march + f->colMulB[0] * f2, f->colMulA[1] * march + f->colMulB[1] * f2,
         f->colMulA[2] * march + f->colMulB[2] * f2);
    const float mul = 8.f - d[0] * 0.5f;
    set3(col, col[0] * mul, col[1] * mul, col[2] * mul);
    gamma_adj(col, f->gamma);
    return orc_topixel(col);
}

/* ---- nautilus: code/shadertoy.cpp:288-383 ------------------------------ */

typedef struct nautilus_frame {
    float g[3], diffColor[4], time, roll, cosHitOffs, funkCos;
} nautilus_frame;

/* fNautilus, code/shadertoy.cpp:289-298 */
static float f_nautilus(const st_ctx *c, const nautilus_frame *f, const float p[4], float time)
{
    const float cosX = lcos(c, lcos(c, p[0] + f->g[0]) * p[0] - lcos(c, p[1] + f->g[1]) * p[1]);
    const float cosY = lcos(c, p[2] * 0.33f * p[0] - f->g[2] * p[1]);
    const float cosZ = lcos(c, p[0] + p[1] + p[2] * 0.8f + time);
    const float dotted = cosX * cosX + cosY * cosY + cosZ * cosZ;
    return dotted * 0.5f - .7f;
}

/* code/shadertoy.cpp:303-327; p = time, roll, hue, speed, desaturation, blur */
static void nautilus_setup(const st_ctx *c, const float *p, nautilus_frame *f)
{
    const float hue = p[2];
    f->roll = p[1];
    f->time = p[0] * p[3];
    f->g[0] = f->time * 0.125f;
    f->g[1] = f->time / 9.f;
    f->g[2] = lcos(c, f->time * 0.1428f);
    float col[4];
    set3(col, .1f - lcos(c, hue / 3.f) / 19.f, .1f, .1f + lcos(c, hue / 14.f) / 8.f);
    desaturate(col, p[4], f->diffColor);
    f->cosHitOffs = lcos(c, f->time * 0.314f * 0.5f);
    f->funkCos = lcos(c, f->time * K_GOLDEN_RATIO * 0.1f);
}

static void nautilus_grad(const st_ctx *c, const nautilus_frame *f, const float p[4], float march,
                          float nOffs, float n[4])
{
    float q[4];
    set3(q, p[0] + nOffs, p[1], p[2]);
    const float nx = march - f_nautilus(c, f, q, f->time);
    set3(q, p[0], p[1] + nOffs, p[2]);
    const float ny = march - f_nautilus(c, f, q, f->time);
    set3(q, p[0], p[1], p[2] + nOffs);
    const float nz = march - f_nautilus(c, f, q, f->time);
    set3(n, nx, ny, nz);
    vnorm(c, n);
}

/* code/shadertoy.cpp:338-378 */
static uint32_t nautilus_pixel(const st_ctx *c, const nautilus_frame *f, unsigned iX, unsigned iY)
{
    float u, v;
    to_uv(c, iX, iY, 2.f, &u, &v);
    float d[4];
    set3(d, u * c->aspect, v, 1.f);
    rot_z(c, f->roll * f->time, &d[0], &d[1]);
    vnorm(c, d);
    float hit[4] = {0.f, 0.f, 0.f, 0.f};
    float total = 0.01f, march = 1.f;
    for (int iStep = 0; march > 0.01f && iStep < 48; ++iStep) {
        hit[0] = d[0] * total;
        hit[1] = d[1] * total;
        hit[2] = d[2] * total;
        march = f_nautilus(c, f, hit, f->time);
        total += march * 0.628f;
    }
    float n[4];
    nautilus_grad(c, f, hit, march, 0.15f, n);
    float diffuse = n[2] * 0.1f;
    const float spec = powf(std_max(0.f, dot3(n, d)), 16.f);
    float hitOffs[4], funk[4];
    set3(hitOffs, hit[0] + f->cosHitOffs, hit[1] + f->cosHitOffs, hit[2] + f->cosHitOffs);
    nautilus_grad(c, f, hitOffs, march, 0.15f, funk);
    const float yMod = fracf_(hit[1] * 0.3f + funk[0] * 0.628f + funk[1] * f->funkCos);
    diffuse *= yMod * yMod * yMod;
    const float s = 1.56f * total + spec;
    const float b = spec * K_GOLDEN_RATIO * 0.2f;
    float col[4];
    set3(col, diffuse + f->diffColor[0] * s, diffuse + f->diffColor[1] * s,
         diffuse + f->diffColor[2] * s);
    set3(col, col[0] + b, col[1] + b, col[2] + b);
    gamma_adj(col, 1.44f);
    return orc_topixel(col);
}

/* ---- spikey: code/shadertoy.cpp:412-659 -------------------------------- */

/* fSpikey1 / fSpikey2, code/shadertoy.cpp:418-430 */
static float f_spikey(const st_ctx *c, const float g[4], const float p[4], float scale)
{
    const float radius = 1.35f + scale * lcos(c, g[1] * p[1] - g[0]) + scale * lcos(c, g[2] * p[0] + g[0]);
    return vlen(p) - radius;
}

#define SPIKE1_SCALE (K_GOLDEN_ANGLE * 0.1f)
#define SPIKE2_SCALE (K_GOLDEN_RATIO * 0.1f)

static void spikey_grad(const st_ctx *c, const float g[4], const float p[4], float march,
                        float nOffs, float scale, float n[4])
{
    float q[4];
    set3(q, p[0] + nOffs, p[1], p[2]);
    const float nx = march - f_spikey(c, g, q, scale);
    set3(q, p[0], p[1] + nOffs, p[2]);
    const float ny = march - f_spikey(c, g, q, scale);
    set3(q, p[0], p[1], p[2] + nOffs);
    const float nz = march - f_spikey(c, g, q, scale);
    set3(n, nx, ny, nz);
    vnorm(c, n);
}

typedef struct spikey_frame {
    int mode; /* 0 distant, 1 distant specular-only, 2 close */
    float g[4], diffColor[4], origin[4];
    float roll, specPow, gamma, xOffs, yOffs, zOffsFinal, warmup, normalGrain;
} spikey_frame;

/* easeInOutElasticf, code/synth-math-easings.h:190-202 */
static float ease_in_out_elastic(float x)
{
    const float c5 = (2.f * K_PI) / 4.5f;
    return (0.f == x) ? 0.f
         : (1.f == x) ? 1.f
         : (x < 0.5f) ? -(powf(2.f, 20.f * x - 10.f) * sinf((20.f * x - 11.125f) * c5)) * 0.5f
                      : (powf(2.f, -20.f * x + 10.f) * sinf((20.f * x - 11.125f) * c5)) * 0.5f + 1.f;
}

/* p = time, speed, roll, specular, desaturation, hue, gamma,
 *     distXOffs, distYOffs, distZOffs, distWarmup,
 *     closeXOffs, closeYOffs, closeZOffs, closeZOffsScale, closeNormalGrain,
 *     closeScale, closeRim, closeAspectMul,
 *     closeMixBlurMap, closeMixBlur, closeMixMapBlur, closeMixBlurOpacity */
static void spikey_setup(const st_ctx *c, const float *p, int close, spikey_frame *f)
{
    const float time = p[0], speed = p[1];
    f->roll = p[2];
    f->specPow = p[3];
    f->gamma = p[6];
    float pal[4];
    michiel_pal(c, p[5], pal);
    desaturate(pal, p[4], f->diffColor);
    if (close) {
        /* code/shadertoy.cpp:436-459 */
        f->mode = 2;
        f->xOffs = p[11];
        f->yOffs = p[12];
        f->zOffsFinal = ease_in_out_elastic(p[13]) * p[14];
        f->normalGrain = p[15];
        const float scale = p[16];
        const int aspectMul = 0 != (int)roundf(p[18]);
        f->g[0] = speed * time;
        f->g[1] = 16.f * scale;
        f->g[2] = aspectMul ? c->aspect * 22.f * scale : 22.f * scale;
        f->g[3] = 0.f;
        set3(f->origin, 0.2f, 0.f, -2.23f);
    } else if (0.f == p[10]) {
        /* code/shadertoy.cpp:527-547 */
        f->mode = 0;
        f->xOffs = p[7];
        f->yOffs = p[8];
        f->g[0] = speed * time; f->g[1] = 16.f; f->g[2] = 16.f; f->g[3] = K_EPSILON;
        set3(f->origin, 0.f, 0.f, -2.614f + p[9]);
    } else {
        /* code/shadertoy.cpp:600-610, warmup = 1 + track (code/shadertoy.cpp:720) */
        f->mode = 1;
        f->warmup = 1.f + p[10];
        f->g[0] = speed * time; f->g[1] = 8.f; f->g[2] = 16.f; f->g[3] = 0.f;
        set3(f->origin, 0.f, 0.f, -3.314f);
    }
}

static uint32_t spikey_pixel(const st_ctx *c, const spikey_frame *f, unsigned iX, unsigned iY)
{
    float u, v, d[4], hit[4] = {0.f, 0.f, 0.f, 0.f}, n[4], col[4];
    const float *o = f->origin;
    float march = 1.f, total = 0.f;
    if (f->mode == 0) {
        /* code/shadertoy.cpp:558-593 */
        to_uv(c, iX, iY, 2.f, &u, &v);
        set3(d, u + f->xOffs, v + f->yOffs, 1.f);
        rot_z(c, f->roll, &d[0], &d[1]);
        vnorm(c, d);
        for (int iStep = 0; march > 0.001f && iStep < 48; ++iStep) {
            hit[0] = o[0] + d[0] * total;
            hit[1] = o[1] + d[1] * total;
            hit[2] = o[2] + d[2] * total;
            march = f_spikey(c, f->g, hit, SPIKE2_SCALE);
            march *= 0.314f;
            total += march;
        }
        spikey_grad(c, f->g, hit, march, K_PI * 0.02f, SPIKE2_SCALE, n);
        const float diffuse = std_max(0.f, n[2] * 0.8f + n[1] * 0.2f);
        const float spec = powf(dot3(n, d), f->specPow);
        const float distance = hit[2] - o[2];
        float a[4];
        for (int k = 0; k < 4; ++k) a[k] = (f->diffColor[k] + spec) * diffuse;
        const float one[4] = {1.f, 1.f, 1.f, 1.f};
        vlerp4(col, a, one, exp_fog(distance, 0.133f));
        gamma_adj(col, f->gamma);
    } else if (f->mode == 1) {
        /* code/shadertoy.cpp:620-653 */
        to_uv(c, iX, iY, K_GOLDEN_RATIO, &u, &v);
        set3(d, u * c->aspect, v, 1.f);
        rot_z(c, f->roll, &d[0], &d[1]);
        vnorm(c, d);
        for (int iStep = 0; iStep < 36; ++iStep) {
            hit[0] = o[0] + d[0] * total;
            hit[1] = o[1] + d[1] * total;
            hit[2] = o[2] + d[2] * total;
            march = f_spikey(c, f->g, hit, SPIKE2_SCALE);
            total += march * 0.075f * K_GOLDEN_RATIO;
        }
        spikey_grad(c, f->g, hit, march, 0.01f, SPIKE2_SCALE, n);
        const float distance = hit[2] - o[2];
        const float spec = f->warmup * powf(std_max(0.f, dot3(n, d)), f->specPow);
        const float s4[4] = {spec, spec, spec, spec}, z4[4] = {0.f, 0.f, 0.f, 0.f};
        vlerp4(col, s4, z4, exp_fog(distance, 0.0133f));
    } else {
        /* code/shadertoy.cpp:468-517 */
        to_uv(c, iX, iY, 2.f, &u, &v);
        set3(d, (u + f->xOffs) * c->aspect, v + f->yOffs, 1.f + f->zOffsFinal);
        rot_z(c, f->roll, &d[0], &d[1]);
        vnorm(c, d);
        for (int iStep = 0; march > 0.0001f && iStep < 32; ++iStep) {
            hit[0] = o[0] + d[0] * total;
            hit[1] = o[1] + d[1] * total;
            hit[2] = o[2] + d[2] * total;
            march = f_spikey(c, f->g, hit, SPIKE1_SCALE);
            total += march * (0.05f * K_PI);
        }
        spikey_grad(c, f->g, hit, march, f->normalGrain, SPIKE1_SCALE, n);
        /* the rim factor max(1, min(0, x)) is 1 for every x: diffuse is kept */
        const float diffuse = n[2];
        const float spec = powf(std_max(0.f, dot3(n, d)), f->specPow);
        const float distance = hit[2] - o[2];
        float a[4];
        for (int k = 0; k < 4; ++k) a[k] = (f->diffColor[k] + spec) * diffuse;
        const float one[4] = {1.f, 1.f, 1.f, 1.f};
        vlerp4(col, a, one, exp_fog(distance, K_GOLDEN_RATIO * 0.1f));
        gamma_adj(col, f->gamma);
    }
    return orc_topixel(col);
}

/* ---- sinuses: code/shadertoy.cpp:870-977 ------------------------------- */

/* fSinMap, code/shadertoy.cpp:878-898 */
static float f_sinmap(const st_ctx *c, const float p[4])
{
    const float pZ = p[2];
    const float zMod = pZ * 0.314f;
    const float pathCos = lcos(c, zMod);
    const float pathCos2 = lcos(c, zMod + (K_2PI / 4.f)) * K_GOLDEN_RATIO;
    const float pX = p[0] - (pathCos2 * 2.f - pathCos * 1.5f);
    const float pY = p[1] - (pathCos * 3.14f + pathCos2);
    const float aX = pX * 0.315f * 1.25f + lsin(c, pZ * (0.814f * 1.25f));
    const float aY = pY * 0.315f * 1.25f + lsin(c, pX * (0.814f * 1.25f));
    const float aZ = pZ * 0.315f * 1.25f + lsin(c, pY * (0.814f * 1.25f));
    const float cosX = lcos(c, aX), cosY = lcos(c, aY), cosZ = lcos(c, aZ);
    const float length = sqrtf(cosX * cosX + cosY * cosY + cosZ * cosZ);
    return (length - 1.025f) * 1.33f;
}

typedef struct sinuses_frame {
    float origin[4], diffColor[4], specPow, roll, offsX, gamma;
} sinuses_frame;

/* code/shadertoy.cpp:904-924; p = time, specular, roll, speed, offsX,
 * gamma, hue, desaturation */
static void sinuses_setup(const st_ctx *c, const float *p, sinuses_frame *f)
{
    f->specPow = 1.f + p[1];
    f->roll = p[2];
    f->offsX = p[4];
    f->gamma = p[5];
    float pal[4];
    michiel_pal(c, p[6], pal);
    desaturate(pal, p[7], f->diffColor);
    /* fSinPath, code/shadertoy.cpp:869-875 */
    const float time = p[0] * p[3];
    const float timeMod = time * 0.314f;
    const float sine = lsin(c, timeMod), cosine = lcos(c, timeMod);
    set3(f->origin, sine * 2.f * K_GOLDEN_RATIO - cosine * 1.5f, cosine * 3.14f + sine * K_GOLDEN_RATIO,
         time);
}

static uint32_t sinuses_pixel(const st_ctx *c, const sinuses_frame *f, unsigned iX, unsigned iY)
{
    /* code/shadertoy.cpp:935-970 */
    float u, v, d[4], hit[4] = {0.f, 0.f, 0.f, 0.f}, n[4], col[4], q[4];
    const float *o = f->origin;
    to_uv(c, iX, iY, 2.f, &u, &v);
    set3(d, (u + f->offsX) * c->aspect, v, 0.314f);
    rot_z(c, f->roll, &d[0], &d[1]);
    vnorm(c, d);
    float march = 1.f, total = 0.f;
    for (int iStep = 0; march > 0.01f && iStep < 32; ++iStep) {
        hit[0] = o[0] + d[0] * total;
        hit[1] = o[1] + d[1] * total;
        hit[2] = o[2] + d[2] * total;
        march = f_sinmap(c, hit);
        total += march * 0.814f;
    }
    set3(q, hit[0] + 0.2f, hit[1], hit[2]);
    const float nx = march - f_sinmap(c, q);
    set3(q, hit[0], hit[1] + 0.2f, hit[2]);
    const float ny = march - f_sinmap(c, q);
    set3(q, hit[0], hit[1], hit[2] + 0.2f);
    const float nz = march - f_sinmap(c, q);
    set3(n, nx, ny, nz);
    vnorm(c, n);
    float diffuse = n[2] * 0.7f + 0.3f * n[1];
    diffuse = 0.2f + 0.8f * diffuse;
    const float spec = powf(dot3(n, d), f->specPow);
    const float distance = hit[2] - o[2];
    float a[4];
    for (int k = 0; k < 4; ++k) a[k] = (f->diffColor[k] + spec) * diffuse;
    const float one[4] = {1.f, 1.f, 1.f, 1.f};
    vlerp4(col, a, one, exp_fog(distance, 0.03f));
    gamma_adj(col, f->gamma);
    return orc_topixel(col);
}

/* ---- laura: code/shadertoy.cpp:993-1103 -------------------------------- */

/* fLaura, code/shadertoy.cpp:993-996 */
static float f_laura(const st_ctx *c, const float p[4])
{
    return lcos(c, p[0]) + lcos(c, p[1]) + lcos(c, p[2]) + 1.f;
}

typedef struct laura_frame {
    float origin[4], diffColor[4], yaw, pitch, rollT;
} laura_frame;

/* code/shadertoy.cpp:1015-1029; p = time, speed, yaw, pitch, roll, hue,
 * saturate */
static void laura_setup(const st_ctx *c, const float *p, laura_frame *f)
{
    float pal[4];
    michiel_pal(c, p[5], pal);
    desaturate(pal, p[6], f->diffColor);
    set3(f->origin, 0.f, 0.f, p[1] * p[0]);
    f->yaw = p[2];
    f->pitch = p[3];
    f->rollT = p[4] * p[0];
}

static uint32_t laura_pixel(const st_ctx *c, const laura_frame *f, unsigned iX, unsigned iY)
{
    /* code/shadertoy.cpp:1040-1095 */
    float u, v, d[4], hit[4] = {0.f, 0.f, 0.f, 0.f}, n[4], col[4], q[4];
    const float *o = f->origin;
    to_uv(c, iX, iY, 2.f, &u, &v);
    set3(d, u * c->aspect, v, K_PI);
    rot_y(c, f->yaw, &d[0], &d[2]);
    rot_x(c, f->pitch, &d[1], &d[2]);
    rot_z(c, f->rollT, &d[0], &d[1]);
    vnorm(c, d);
    float march = 0.f, total = 0.f;
    for (int iStep = 0; iStep < 32; ++iStep) {
        hit[0] = o[0] + d[0] * total;
        hit[1] = o[1] + d[1] * total;
        hit[2] = o[2] + d[2] * total;
        march = f_laura(c, hit);
        total += march * 0.5f;
    }
    /* LauraNormal, code/shadertoy.cpp:998-1010 */
    set3(q, hit[0] + 0.1628f, hit[1], hit[2]);
    const float nx = f_laura(c, q) - march;
    set3(q, hit[0], hit[1] + 0.1628f, hit[2]);
    const float ny = f_laura(c, q) - march;
    set3(q, hit[0], hit[1], hit[2] + 0.1628f);
    const float nz = f_laura(c, q) - march;
    set3(n, nx, ny, nz);
    vnorm(c, n);
    float lightPos[4], lightDir[4];
    set3(lightPos, o[0] - d[0], o[1] - d[1], o[2] - d[2]);
    set3(lightDir, lightPos[0] - hit[0], lightPos[1] - hit[1], lightPos[2] - hit[2]);
    vnorm(c, lightDir);
    const float diffuse = std_max(0.3f, dot3(n, lightDir));
    const float distance = hit[2] - o[2];
    const float spec = specular(c, o, hit, n, lightDir, 4.f);
    /* the rim factor max(1, min(0, x)) is 1 for every x: diffuse is kept */
    const float ds = diffuse + spec;
    const float q3 = q3_rsqrt2(spec + diffuse);
    float a[4];
    for (int k = 0; k < 4; ++k) a[k] = f->diffColor[k] * ds;
    const float b[4] = {q3, q3, q3, q3};
    vlerp4(col, a, b, exp_fog(distance, 0.001f));
    gamma_adj(col, 1.44f);
    return orc_topixel(col);
}

/* ---- free-directional tunnel: code/shadertoy.cpp:746-838 --------------- */

typedef struct fdtunnel_frame {
    float boxy, flowerScale, flowerFreq, flowerPhase, speed, roll, pitch, radius, uMul, vMul;
    float fog[2], time;
} fdtunnel_frame;

/* code/shadertoy.cpp:752-768; p = time, boxy, flowerScale, flowerFreq,
 * flowerPhase, speed, roll, pitch, radius, mulU, mulV, litTiles, litBlur,
 * fog1, fog2 */
static void fdtunnel_setup(const float *p, fdtunnel_frame *f)
{
    const float time = p[0];
    f->boxy = p[1];
    f->flowerScale = p[2];
    f->flowerFreq = p[3];
    f->flowerPhase = p[4] * time;
    f->speed = p[5];
    f->roll = p[6] * time;
    f->pitch = p[7] * time;
    f->radius = p[8];
    f->uMul = p[9];
    f->vMul = p[10];
    f->fog[0] = p[13];
    f->fog[1] = p[14];
    f->time = time * f->speed;
}

/* (int) of a float as x86 cvttss2si: NaN and out of range give INT_MIN */
static int32_t x86_cvtt(float v)
{
    if (v == v && v < 2147483648.f && v >= -2147483648.f) return (int32_t)v;
    return INT32_MIN;
}

/* ftofp24 (code/util.h:195-197), bsamp_prepUVs with 1023/10 and
 * bsamp32_32f (code/bilinear.h:128-134) */
static void fd_sample(const uint32_t *tex, float U, float V, float out[4])
{
    unsigned U0, V0, U1, V1, fu, fv;
    orc_bsamp_prep(x86_cvtt(U * 256.f), x86_cvtt(V * 256.f), 1023, 10, &U0, &V0, &U1, &V1, &fu, &fv);
    const uint32_t px = orc_bsamp32_32_packed(tex, U0, V0, U1, V1, fu, fv);
    for (int k = 0; k < 4; ++k) out[k] = (float)((px >> (8 * k)) & 0xff);
}

static void fdtunnel_pixel(const st_ctx *c, const fdtunnel_frame *f, const uint32_t *tex,
                           const uint32_t *texHi, unsigned iX, unsigned iY, uint32_t *color,
                           uint32_t *glow)
{
    /* code/shadertoy.cpp:780-825 */
    float u, v, d[4];
    to_uv(c, iX, iY, 2.f, &u, &v);
    set3(d, u, v, 1.f);
    rot_x(c, f->pitch, &d[1], &d[2]);
    rot_z(c, f->roll, &d[0], &d[1]);
    vnorm(c, d);
    float A = d[0] * d[0] + d[1] * d[1];
    A += f->flowerScale * lcos(c, atan2f(d[1], d[0]) * f->flowerFreq + f->flowerPhase);
    const float absX = fabsf(d[0]), absY = fabsf(d[1]);
    const float box = absX > absY ? absX : absY;
    /* smoothstepf, code/Std3DMath-stripped/Math.h:59-63 */
    float t = f->boxy;
    t = t * t * (3.f - 2.f * t);
    A = A + (box - A) * t;
    A += K_EPSILON;
    A = 1.f / A;
    const float T = f->radius * A;
    const float T2 = T * 0.912f;
    const float i1[3] = {d[0] * T, d[1] * T, d[2] * T};
    const float i2[3] = {d[0] * T2, d[1] * T2, d[2] * T2};
    const float U = atan2f(i1[1], i1[0]) / K_PI;
    const float V = i1[2] + f->time * f->speed;
    const float U2 = atan2f(i2[1], i2[0]) / K_PI;
    const float V2 = i2[2] + f->time * f->speed;
    /* clampf(0, 1, x) = std::max(0, std::min(1, x)) */
    const float shade = std_max(0.f, std_min(1.f, 1.f - expf(-0.006f * T * T)));
    float col[4], gcol[4], out[4];
    fd_sample(tex, U * f->uMul, V * f->vMul, col);
    fd_sample(texHi, U2 * f->uMul, V2 * f->vMul, gcol);
    const float f0[4] = {f->fog[0], f->fog[0], f->fog[0], f->fog[0]};
    const float f1[4] = {f->fog[1], f->fog[1], f->fog[1], f->fog[1]};
    vlerp4(out, col, f0, shade);
    *color = topixel_noconv(out);
    vlerp4(out, gcol, f1, shade);
    *glow = topixel_noconv(out);
}

/* ---- frame entry points ------------------------------------------------- */

static void st_init(st_ctx *c, unsigned resX, unsigned resY, const float *lut, int rsq)
{
    c->lut = lut;
    c->rsq = rsq;
    c->mapW = resX / 2 + 4;
    c->mapH = resY / 2 + 4;
    c->aspect = (float)resY / (float)resX;
    c->oneOverAspect = 1.f / c->aspect;
}

void orc_st_map(int effect, uint32_t *dst, uint32_t *glow, unsigned resX, unsigned resY,
                const float *params, const float *lut, const uint32_t *tex,
                const uint32_t *texHi, int rsqrtMode, int nthreads)
{
    st_ctx c;
    st_init(&c, resX, resY, lut, rsqrtMode);
    plasma_frame pf;
    nautilus_frame nf;
    spikey_frame sf;
    sinuses_frame sif;
    laura_frame lf;
    fdtunnel_frame tf;
    switch (effect) {
    case ORC_ST_PLASMA: plasma_setup(&c, params, &pf); break;
    case ORC_ST_NAUTILUS: nautilus_setup(&c, params, &nf); break;
    case ORC_ST_SPIKEY_DISTANT: spikey_setup(&c, params, 0, &sf); break;
    case ORC_ST_SPIKEY_CLOSE: spikey_setup(&c, params, 1, &sf); break;
    case ORC_ST_SINUSES: sinuses_setup(&c, params, &sif); break;
    case ORC_ST_LAURA: laura_setup(&c, params, &lf); break;
    case ORC_ST_FDTUNNEL: fdtunnel_setup(params, &tf); break;
    default: return;
    }
    #pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
    for (unsigned iY = 0; iY < c.mapH; ++iY) {
        for (unsigned iX = 0; iX < c.mapW; ++iX) {
            const size_t i = (size_t)iY * c.mapW + iX;
            switch (effect) {
            case ORC_ST_PLASMA: dst[i] = plasma_pixel(&c, &pf, iX, iY); break;
            case ORC_ST_NAUTILUS: dst[i] = nautilus_pixel(&c, &nf, iX, iY); break;
            case ORC_ST_SPIKEY_DISTANT:
            case ORC_ST_SPIKEY_CLOSE: dst[i] = spikey_pixel(&c, &sf, iX, iY); break;
            case ORC_ST_SINUSES: dst[i] = sinuses_pixel(&c, &sif, iX, iY); break;
            case ORC_ST_LAURA: dst[i] = laura_pixel(&c, &lf, iX, iY); break;
            case ORC_ST_FDTUNNEL: fdtunnel_pixel(&c, &tf, tex, texHi, iX, iY, dst + i, glow + i); break;
            }
        }
    }
}

/* BoxBlurScale, code/deprecated/boxblur.h:13-19 */
static float box_blur_scale(float strength)
{
    if (strength != 0.f) strength = std_max(1.f, std_min(100.f, strength)) * 0.01f;
    return strength;
}

/* clampf(lo, hi, v), code/Std3DMath-stripped/Math.h:37-40 */
static float clampf_(float lo, float hi, float v) { return std_max(lo, std_min(hi, v)); }

/* the close-up's mix-blur layer, code/shadertoy.cpp:665-701: a blend of the
 * two blur maps (Mix32), blurs (BoxBlur32 in place), SoftLight32AA of the
 * blurred copy of the map by the blur map, Overlay32A back onto the map */
static int spikey_mix_blur(uint32_t *map, unsigned mapW, unsigned mapH, const float *params,
                           const uint32_t *blurMaps, int nthreads)
{
    const float mbOpacity = std_max(0.f, std_min(1.f, params[22]));
    if (!(mbOpacity > 0.f)) return 0;
    if (!blurMaps) return -1;
    const size_t n = (size_t)mapW * mapH;
    const float mbMap = clampf_(0.f, 1.f, params[19]);
    const float mbBlur = clampf_(0.f, 100.f, params[20]);
    const float mbMapBlur = clampf_(0.f, 100.f, params[21]);
    uint32_t *blurMap = (uint32_t *)malloc(n * 4), *fx1 = (uint32_t *)malloc(n * 4);
    if (!blurMap || !fx1) { free(blurMap); free(fx1); return -2; }
    if (0.f == mbMap) memcpy(blurMap, blurMaps, n * 4);
    else if (1.f == mbMap) memcpy(blurMap, blurMaps + n, n * 4);
    else {
        memcpy(blurMap, blurMaps, n * 4);
        orc_mix32(blurMap, blurMaps + n, (unsigned)n, (uint8_t)(unsigned)(mbMap * 255.f));
    }
    memcpy(fx1, map, n * 4);
    if (mbMapBlur >= 1.f) orc_blur07(blurMap, blurMap, mapW, mapH, box_blur_scale(mbMapBlur), nthreads);
    if (mbBlur >= 1.f) orc_blur07(fx1, fx1, mapW, mapH, box_blur_scale(mbBlur), nthreads);
    orc_softlight32aa(fx1, blurMap, (unsigned)n, tanhf(mbBlur + mbOpacity));
    orc_overlay32a(map, fx1, (unsigned)n);
    free(blurMap);
    free(fx1);
    return 0;
}

/* the effects' Draw functions: Plasma_Draw (code/shadertoy.cpp:276-280),
 * Nautilus_Draw (:385-396), Spikey_Draw (:661-727), Sinuses_Draw
 * (:984-988), Laura_Draw (:1105-1109), Tunnel_Draw (:840-861).  blurMaps:
 * the close-up's two blur maps, back to back, each the FX map's size (may
 * be NULL when the mix-blur layer is off; -1 when it is on without them). */
int orc_st_draw(int effect, uint32_t *dst, unsigned resX, unsigned resY, const float *params,
                const float *lut, const uint32_t *tex, const uint32_t *texHi,
                const uint32_t *blurMaps, int rsqrtMode, int nthreads)
{
    const unsigned mapW = resX / 2 + 4, mapH = resY / 2 + 4;
    const size_t n = (size_t)mapW * mapH;
    uint32_t *map = (uint32_t *)calloc(n, 4), *glow = (uint32_t *)calloc(n, 4);
    if (!map || !glow) { free(map); free(glow); return -2; }
    orc_st_map(effect, map, glow, resX, resY, params, lut, tex, texHi, rsqrtMode, nthreads);
    int rc = 0;
    if (effect == ORC_ST_SPIKEY_CLOSE)
        rc = spikey_mix_blur(map, mapW, mapH, params, blurMaps, nthreads);
    if (effect == ORC_ST_SPIKEY_DISTANT && params[10] != 0.f)
        orc_blur07_h(map, map, mapW, mapH, box_blur_scale(1.f + params[10]), nthreads);
    if (effect == ORC_ST_FDTUNNEL && 0 != (int)roundf(params[11])) {
        const float litBlur = clampf_(0.f, 100.f, params[12]);
        if (litBlur >= 1.f) orc_blur07(glow, glow, mapW, mapH, box_blur_scale(litBlur), nthreads);
        orc_add32(map, glow, (unsigned)n);
    }
    if (rc == 0) {
        orc_fx_blit_2x2(dst, map, resX, resY);
        if (effect == ORC_ST_NAUTILUS) {
            const float blur = box_blur_scale(params[5]);
            if (0.f != blur) orc_blur07(dst, dst, resX, resY, blur, nthreads);
        }
    }
    free(map);
    free(glow);
    return rc;
}
