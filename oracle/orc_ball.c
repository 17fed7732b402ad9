/*
 * orc_ball.c -- the voxel ball (code/ball.cpp), the torus twister
 * (code/torus-twister.cpp) and Polar_BlitA (code/polar.cpp:157-198), with
 * their Draw compositions.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * The rays are the reference's integer arithmetic; the per-frame tables
 * (height projections, lighting norms, ray directions) are the reference's
 * FP32 host code, here with the same libm.  Rays write only their own row:
 * the reference writes past the end of a 1280-pixel row (into the next row,
 * or past the image for the last one) when a projected height exceeds the
 * row, which happens only for map heights above 1279*256/radius; inputs
 * are kept below that.
 */
#include "ckd_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define K_PI 3.1415926535897932384626433832795f
#define K_2PI (2.f * K_PI)
#define K_GOLDEN_RATIO 1.61803398875f
#define K_GOLDEN_ANGLE 2.39996f
#define K_EPSILON 1.1920928955078125e-07f
#define MAP_AND 1023u
#define MAP_SHIFT 10u

static float std_max(float a, float b) { return (a < b) ? b : a; }
static float std_min(float a, float b) { return (b < a) ? b : a; }
static float clampf_(float lo, float hi, float v) { return std_max(lo, std_min(hi, v)); }
static int clampi_(int lo, int hi, int v) { return v < lo ? lo : (v > hi ? hi : v); }
static int geti(float v) { return (int)roundf(v); }

/* _mm_adds_epu16 */
static uint32_t adds16(uint32_t a, uint32_t b) { const uint32_t s = a + b; return s > 65535u ? 65535u : s; }

/* bsamp32_16 with its four 16-bit lanes (code/bilinear.h:58-83) */
static void sample16(const uint32_t *tex, int U, int V, uint16_t out[4])
{
    unsigned U0, V0, U1, V1, fu, fv;
    orc_bsamp_prep(U, V, MAP_AND, MAP_SHIFT, &U0, &V0, &U1, &V1, &fu, &fv);
    orc_bsamp32_16(tex, U0, V0, U1, V1, fu, fv, out);
}

static unsigned sample8(const uint8_t *tex, int U, int V)
{
    unsigned U0, V0, U1, V1, fu, fv;
    orc_bsamp_prep(U, V, MAP_AND, MAP_SHIFT, &U0, &V0, &U1, &V1, &fu, &fv);
    return orc_bsamp8(tex, U0, V0, U1, V1, fu, fv);
}

/* (int)(x*256.f) as x86 cvttss2si: NaN and out of range give INT_MIN */
static int fp24(float v)
{
    const float s = v * 256.f;
    return (s == s && s < 2147483648.f && s >= -2147483648.f) ? (int)s : INT32_MIN;
}

/* ---- Polar_BlitA, code/polar.cpp:157-198 -------------------------------- */

/* dst = dst + srcA*(s - dst)/256 per channel, s = Fetch16 of src through
 * the map (code/polar.cpp:95-106); srcA is the fetched colour's alpha */
void orc_polar_blit_a(uint32_t *dst, const uint32_t *src, const int *map, unsigned W,
                      unsigned H, unsigned srcStride, int nthreads)
{
    #pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int y = 0; y < (int)H; ++y) {
        for (unsigned x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            const int U = map[2 * i], V = map[2 * i + 1];
            const unsigned U0 = (unsigned)(U >> 8);
            const unsigned V0 = (unsigned)(V >> 8) * srcStride;
            uint16_t s[4];
            orc_bsamp32_16(src, U0, V0, U0 + 1, V0 + srcStride, ((unsigned)U & 0xff) * 0x01010101u,
                           ((unsigned)V & 0xff) * 0x01010101u, s);
            const uint32_t a = s[3];
            uint32_t px = 0;
            for (int k = 0; k < 4; ++k) {
                const uint32_t d = (dst[i] >> (8 * k)) & 0xff;
                const uint32_t delta = (a * (uint32_t)(uint16_t)(s[k] - d)) & 0xffffu;
                const uint32_t sum = ((d << 8) + delta) & 0xffffu;
                const int16_t v = (int16_t)(uint16_t)(sum >> 8);
                px |= (uint32_t)(v < 0 ? 0 : (v > 255 ? 255 : v)) << (8 * k);
            }
            dst[i] = px;
        }
    }
}

/* ---- torus twister, code/torus-twister.cpp -------------------------------- */

#define TW_RAY_LENGTH 512u

/* vtwister_precalc, code/torus-twister.cpp:118-134 */
void orc_twister_tables(uint32_t heightProj[512], uint32_t heightProjNorm[512])
{
    for (unsigned iAngle = 0; iAngle < TW_RAY_LENGTH; ++iAngle) {
        const float angle = K_PI / (float)(TW_RAY_LENGTH - 1) * (float)iAngle;
        const float scale = 600.f * sinf(angle);
        heightProj[iAngle] = (uint32_t)(int64_t)scale;
        const float cosine = cosf(angle * 0.99f);
        heightProjNorm[iAngle] = cosine > 0.f ? (uint32_t)(int64_t)(255.f * powf(cosine, 2.f)) : 0u;
    }
}

/* vtwister_ray, code/torus-twister.cpp:39-85 */
static void twister_ray(uint32_t *pDest, int curX, int curY, int dX, const uint8_t *hmap,
                        const uint32_t *cmap, const uint32_t *hp, const uint32_t *hn)
{
    unsigned lastHeight = 0, lastDrawn = 0;
    uint16_t lastColor[4];
    sample16(cmap, curX, curY, lastColor);
    const int direction = dX < 0 ? -1 : 1;
    for (unsigned iStep = 0; iStep < TW_RAY_LENGTH; ++iStep) {
        curX = (int)((uint32_t)curX - (uint32_t)dX);
        const unsigned mapHeight = sample8(hmap, curX, curY);
        uint16_t color[4];
        sample16(cmap, curX, curY, color);
        const unsigned diffuse = mapHeight * hn[iStep] >> 8;
        for (int k = 0; k < 4; ++k) color[k] = (uint16_t)adds16(color[k], diffuse & 0xffff);
        const unsigned height = mapHeight * hp[iStep] >> 8;
        if (height > lastDrawn) {
            const unsigned drawLength = height - lastDrawn;
            orc_cspan16(pDest, direction, height - lastHeight, drawLength, lastColor, color);
            pDest += (ptrdiff_t)drawLength * direction;
            lastDrawn = height;
        }
        lastHeight = height;
        memcpy(lastColor, color, sizeof(color));
    }
}

/* vtwister, code/torus-twister.cpp:89-116, into a W x H target (the
 * reference's 1280 x 720); the two rays of a row start at its centre */
void orc_twister_frame(uint32_t *target, unsigned W, unsigned H, const uint8_t *hmap,
                       const uint32_t *cmap, float time, float speed, float shearSpeed,
                       int nthreads)
{
    uint32_t hp[TW_RAY_LENGTH], hn[TW_RAY_LENGTH];
    orc_twister_tables(hp, hn);
    const float fMapSize = 1024.f;
    const float fMapSizeHH = (fMapSize * 0.5f) - 0.5f;
    const float fMapSizeHHH = (fMapSize * 0.25f) - 0.5f;
    const float mapStepY = fMapSize / (float)(H - 1);
    const float shearStep = K_2PI / (float)(H - 1);
    #pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int iRay = 0; iRay < (int)H; ++iRay) {
        const float shearAngle = (float)iRay * shearStep;
        const float mapY = (float)(unsigned)iRay * mapStepY;
        const int fromX = fp24(fMapSizeHH + fMapSizeHHH * sinf(time * shearSpeed + shearAngle));
        const int fromY = fp24(mapY + time * speed);
        uint32_t *row = target + (size_t)iRay * W + (W >> 1);
        twister_ray(row, fromX, fromY, 512, hmap, cmap, hp, hn);
        twister_ray(row - 1, (int)((uint32_t)fromX - 512u), fromY, -512, hmap, cmap, hp, hn);
    }
}

/* Twister_Draw, code/torus-twister.cpp:164-187: memset32 the target, the
 * twister, the optional in-place horizontal box blur, then the background
 * with Polar_BlitA (forward map) of the target on top.  p = time, speed,
 * shearSpeed, blur. */
int orc_twister_draw(uint32_t *dst, uint32_t *target, unsigned W, unsigned H,
                     const uint8_t *hmap, const uint32_t *cmap, const uint32_t *background,
                     const int *polarMap, const float *p, int nthreads)
{
    memset(target, 0, (size_t)W * H * 4);
    orc_twister_frame(target, W, H, hmap, cmap, p[0], p[1], p[2], nthreads);
    if (0.f != p[3]) {
        const float blur = std_max(1.f, std_min(100.f, p[3])) * 0.01f;
        orc_blur07_h(target, target, W, H, blur, nthreads);
    }
    memcpy(dst, background, (size_t)W * H * 4);
    orc_polar_blit_a(dst, target, polarMap, W, H, W, nthreads);
    return 0;
}

/* ---- voxel ball, code/ball.cpp ------------------------------------------- */

#define BALL_MAX_RAY 1024

typedef struct ball_tables {
    unsigned rayLength;
    uint32_t heightProj[BALL_MAX_RAY];
    int32_t norm[BALL_MAX_RAY][3];
} ball_tables;

/* vball_precalc, code/ball.cpp:263-286 */
static void ball_precalc(float radiusTrack, float rayLengthTrack, ball_tables *t)
{
    const float radius = clampf_(1.f, 1920.f, radiusTrack);
    t->rayLength = (unsigned)clampi_(1, BALL_MAX_RAY, geti(rayLengthTrack));
    const float angStepSin = K_PI / (float)(t->rayLength - 1);
    const float angStepCos = angStepSin * 0.99f;
    for (unsigned i = 0; i < t->rayLength; ++i) {
        t->heightProj[i] = (uint32_t)(int64_t)(radius * sinf(angStepSin * (float)i));
        const float cosine = cosf(angStepCos * (float)i);
        if (cosine >= 0.f) {
            t->norm[i][0] = (int)(255.f * powf(cosine, K_GOLDEN_RATIO));
            t->norm[i][1] = (int)(255.f * powf(cosine, K_GOLDEN_ANGLE));
            t->norm[i][2] = (int)(255.f * powf(cosine, K_PI));
        } else {
            t->norm[i][0] = t->norm[i][1] = t->norm[i][2] = 0;
        }
    }
}

/* vball_ray_beams, code/ball.cpp:77-205 (row of W pixels) */
static void ball_ray_beams(uint32_t *pDest, unsigned W, int curX, int curY, int dX, int dY,
                           const ball_tables *t, const uint8_t *hmapMix, const uint32_t *cmap,
                           const uint32_t *beamMapMix, unsigned beamAtten, unsigned lowLight,
                           float beamAlphaMin)
{
    unsigned lastHeight = 0, lastDrawn = 0;
    uint16_t lastColor[4];
    sample16(cmap, curX, curY, lastColor);
    uint32_t beamAccum[4] = {0, 0, 0, 0};
    for (unsigned iStep = 0; iStep < t->rayLength; ++iStep) {
        curX = (int)((uint32_t)curX - (uint32_t)dX);
        curY = (int)((uint32_t)curY - (uint32_t)dY);
        const unsigned mapHeight = sample8(hmapMix, curX, curY);
        uint16_t color[4], beam[4];
        sample16(cmap, curX, curY, color);
        sample16(beamMapMix, curX, curY, beam);
        const uint32_t hi = mapHeight * (uint32_t)t->norm[iStep][0] >> 8;
        const uint32_t lo = mapHeight * (uint32_t)t->norm[iStep][1] >> 8;
        const uint32_t diffuse = hi + (((uint32_t)(int32_t)(lo - hi) * lowLight) >> 8);
        const uint32_t lit = diffuse & 0xffffu; /* _mm_set1_epi16 */
        for (int k = 0; k < 4; ++k) {
            const uint32_t b = ((uint32_t)beam[k] * beamAtten & 0xffffu) >> 8;
            const uint32_t bl = (b * lit & 0xffffu) >> 8;
            beamAccum[k] = adds16(beamAccum[k], bl);
            color[k] = (uint16_t)adds16(adds16(color[k], beamAccum[k]), lit);
        }
        const unsigned height = mapHeight * t->heightProj[iStep] >> 8;
        if (height > lastDrawn) {
            orc_cspan16(pDest + lastDrawn, 1, height - lastHeight, height - lastDrawn, lastColor, color);
            lastDrawn = height;
        }
        lastHeight = height;
        memcpy(lastColor, color, sizeof(color));
    }
    /* beam extrusion, code/ball.cpp:176-204 */
    uint32_t beamCol = 0;
    for (int k = 0; k < 4; ++k) {
        const int16_t v = (int16_t)(uint16_t)beamAccum[k];
        beamCol |= (uint32_t)(v < 0 ? 0 : (v > 255 ? 255 : v)) << (8 * k);
    }
    const unsigned remainder = (W - 1) - lastDrawn;
    beamCol &= 0xffffff;
    const unsigned beamR = beamCol >> 16, beamG = (beamCol >> 8) & 0xff, beamB = beamCol & 0xff;
    const unsigned mulR = (unsigned)(0.0722f * 65536.f);
    const unsigned mulG = (unsigned)(0.7152f * 65536.f);
    const unsigned mulB = (unsigned)(0.2126f * 65536.f);
    const unsigned luminosity = ((beamR * mulR) >> 16) + ((beamG * mulG) >> 16) + ((beamB * mulB) >> 16);
    const float fLuminosity = (float)luminosity;
    const float alphaStep = 1.f / (float)(remainder - 1);
    float curStep = 0.f;
    for (unsigned i = 0; i < remainder; ++i) {
        /* smoothstepf(a, b, t), code/Std3DMath-stripped/Math.h:59-63 */
        float s = curStep;
        s = s * s * (3.f - 2.f * s);
        const float fBeamAlpha = beamAlphaMin + (fLuminosity - beamAlphaMin) * s;
        const unsigned beamAlpha = (unsigned)(int64_t)fBeamAlpha;
        pDest[lastDrawn++] = beamCol | (beamAlpha << 24);
        curStep += alphaStep;
    }
}

/* vball_ray_no_beams, code/ball.cpp:207-261 */
static void ball_ray_no_beams(uint32_t *pDest, int curX, int curY, int dX, int dY,
                              const ball_tables *t, const uint8_t *hmapMix, const uint32_t *cmap,
                              const uint32_t *envMap)
{
    uint16_t lastColor[4];
    sample16(cmap, curX, curY, lastColor);
    int envU = (1024 >> 1) << 8, envV = envU;
    unsigned lastHeight = 0, lastDrawn = 0;
    for (unsigned iStep = 0; iStep < t->rayLength; ++iStep) {
        curX = (int)((uint32_t)curX - (uint32_t)dX);
        curY = (int)((uint32_t)curY - (uint32_t)dY);
        envU = (int)((uint32_t)envU - ((uint32_t)dX << 1));
        envV = (int)((uint32_t)envV - ((uint32_t)dY << 1));
        const unsigned mapHeight = sample8(hmapMix, curX, curY);
        uint16_t color[4], env[4];
        sample16(cmap, curX, curY, color);
        sample16(envMap, (int)((uint32_t)envU + mapHeight), (int)((uint32_t)envV + mapHeight), env);
        const unsigned height = mapHeight * t->heightProj[iStep] >> 8;
        const uint32_t diffuse = mapHeight * (uint32_t)t->norm[iStep][2] >> 8;
        const uint32_t full = (32u + diffuse) < 255u ? 32u + diffuse : 255u;
        for (int k = 0; k < 4; ++k) {
            const uint32_t e = ((uint32_t)env[k] * full & 0xffffu) >> 8;
            color[k] = (uint16_t)adds16(adds16(color[k], e), diffuse & 0xffffu);
        }
        if (height > lastDrawn) {
            orc_cspan16(pDest + lastDrawn, 1, height - lastHeight, height - lastDrawn, lastColor, color);
            lastDrawn = height;
        }
        lastHeight = height;
        memcpy(lastColor, color, sizeof(color));
    }
}

/* the per-ray set-up of vball (code/ball.cpp:290-333): the fan deltas of
 * voxel::calc_fandeltas (code/voxel-shared.h:18-31) as 24:8, and the
 * rotated origin */
void orc_ball_rays(unsigned H, float time, float rayLengthTrack, float rotateOffsX,
                   float rotateOffsY, int *fromXY, int *dXY)
{
    const unsigned rayLength = (unsigned)clampi_(1, BALL_MAX_RAY, geti(rayLengthTrack));
    const float timeScale = (float)rayLength * (0.25f / 1024.f);
    const float fMapDim = 1024.f, fMapHalf = fMapDim * 0.5f;
    fromXY[0] = fp24(fMapDim * sinf(time * timeScale) + fMapHalf + rotateOffsX);
    fromXY[1] = fp24(fMapDim * cosf(time * timeScale) + fMapHalf + rotateOffsY);
    const float delta = K_2PI / (float)(H - 1);
    for (unsigned iRay = 0; iRay < H; ++iRay) {
        const float curAngle = (float)iRay * delta;
        float X = cosf(curAngle), Y = sinf(curAngle);
        if (fabsf(X + Y) > K_EPSILON) {
            const float length = 1.f / sqrtf(X * X + Y * Y);
            X *= length;
            Y *= length;
        }
        dXY[2 * iRay] = fp24(X);
        dXY[2 * iRay + 1] = fp24(Y);
    }
}

/* Ball_Draw, code/ball.cpp:442-511.  maps: heightMaps (5 x 1024^2 bytes),
 * colorMaps (2 x 1024^2), beamMaps (3 x 1024^2), envMap (1024^2),
 * backgrounds (2 x W*H), halo (W*H).  p = time, blur, radius, rayLength,
 * spikes, hasBeams, baseShapeIndex, speed, beamAttenuation, beamAlphaMin,
 * rotateOffsX, rotateOffsY, beams1, beams2, beams3, lowBeams.  target is
 * the W x H render target (kept between calls, as the reference's
 * g_renderTarget[0]: the beam rays do not write its last column). */
int orc_ball_draw(uint32_t *dst, uint32_t *target, unsigned W, unsigned H,
                  const uint8_t *heightMaps, const uint32_t *colorMaps, const uint32_t *beamMaps,
                  const uint32_t *envMap, const uint32_t *backgrounds, const uint32_t *halo,
                  const int *polarMap, const float *p, int nthreads)
{
    const size_t mapPx = 1024u * 1024u;
    uint8_t *hmapMix = (uint8_t *)malloc(mapPx);
    uint32_t *beamMix = (uint32_t *)malloc(mapPx * 4);
    ball_tables *t = (ball_tables *)malloc(sizeof(ball_tables));
    int *dXY = (int *)malloc(sizeof(int) * 2 * H);
    if (!hmapMix || !beamMix || !t || !dXY) { free(hmapMix); free(beamMix); free(t); free(dXY); return -2; }
    const int hasBeams = geti(p[5]) != 0;
    const unsigned iBase = (unsigned)clampi_(1, 4, geti(p[6]));
    memcpy(hmapMix, heightMaps + iBase * mapPx, mapPx);
    const uint8_t spikes = (uint8_t)geti(p[4]);
    if (0 != spikes)
        orc_mix32((uint32_t *)hmapMix, (const uint32_t *)heightMaps, (unsigned)(mapPx / 4), spikes);
    if (hasBeams) {
        memset(beamMix, 0, mapPx * 4);
        for (int k = 0; k < 3; ++k) {
            const float a = std_max(0.f, std_min(1.f, p[12 + k]));
            if (a > 0.f) orc_blit_add32a(beamMix, beamMaps + k * mapPx, 1024, 1024, 1024, a);
        }
    }
    /* vball(g_renderTarget[0], time * speed) */
    const float time = p[0] * p[7];
    ball_precalc(p[2], p[3], t);
    const unsigned beamAtten = (unsigned)clampi_(0, 255, geti(p[8]));
    const float beamAlphaMin = clampf_(0.f, 255.f, p[9]);
    const unsigned lowLight = (unsigned)clampi_(0, 255, geti(p[15]));
    int from[2];
    orc_ball_rays(H, time, p[3], p[10], p[11], from, dXY);
    if (!hasBeams) memset(target, 0, (size_t)W * H * 4);
    #pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
    for (int iRay = 0; iRay < (int)H; ++iRay) {
        uint32_t *row = target + (size_t)iRay * W;
        if (hasBeams)
            ball_ray_beams(row, W, from[0], from[1], dXY[2 * iRay], dXY[2 * iRay + 1], t, hmapMix,
                           colorMaps, beamMix, beamAtten, lowLight, beamAlphaMin);
        else
            ball_ray_no_beams(row, from[0], from[1], dXY[2 * iRay], dXY[2 * iRay + 1], t, hmapMix,
                              colorMaps + mapPx, envMap);
    }
    if (0.f != p[1]) {
        const float blur = std_max(1.f, std_min(100.f, p[1])) * 0.01f;
        orc_blur07_h(target, target, W, H, blur, nthreads);
    }
    memcpy(dst, backgrounds + (hasBeams ? 0 : (size_t)W * H), (size_t)W * H * 4);
    orc_polar_blit_a(dst, target, polarMap, W, H, W, nthreads);
    if (hasBeams) orc_softlight32a(dst, halo, W * H);
    free(hmapMix);
    free(beamMix);
    free(t);
    free(dXY);
    return 0;
}
