/*
 * orc_int.c -- integer / fixed-point part of the oracle.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Each function
 * restates the cited reference function with plain scalar C, lane by lane,
 * keeping the reference's wrap-around, saturation and truncation behaviour.
 */
#include "ckd_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* helpers: SSE lane semantics written out                                 */
/* ---------------------------------------------------------------------- */

static inline uint16_t adds_u16(uint16_t a, uint16_t b)
{
    unsigned s = (unsigned)a + b;
    return (uint16_t)(s > 65535u ? 65535u : s);
}

static inline uint16_t subs_u16(uint16_t a, uint16_t b)
{
    return (uint16_t)(a > b ? a - b : 0);
}

/* _mm_packus_epi16 of one lane: signed int16 -> u8 with saturation */
static inline uint32_t packus16(uint16_t v)
{
    int16_t s = (int16_t)v;
    return s < 0 ? 0u : (s > 255 ? 255u : (uint32_t)s);
}

/* _mm_packus_epi32 of one lane: signed int32 -> u16 with saturation */
static inline uint16_t packus32(int32_t v)
{
    return (uint16_t)(v < 0 ? 0 : (v > 65535 ? 65535 : v));
}

static inline void unpack16(uint32_t px, uint16_t out[4])
{
    for (int c = 0; c < 4; ++c) out[c] = (uint16_t)((px >> (8 * c)) & 0xff);
}

static inline uint32_t pack16(const uint16_t v[4])
{
    uint32_t px = 0;
    for (int c = 0; c < 4; ++c) px |= packus16(v[c]) << (8 * c);
    return px;
}

int orc_ftofp24(float v)
{
    /* code/util.h:195-197: (int)(value*256.f), truncation toward zero */
    return (int)(v * 256.f);
}

/* ---------------------------------------------------------------------- */
/* a1..a3 bilinear samplers                                                */
/* ---------------------------------------------------------------------- */

/* code/bilinear.h:10-31 */
void orc_bsamp_prep(int U, int V, unsigned mapAnd, unsigned mapShift,
                    unsigned *U0, unsigned *V0, unsigned *U1, unsigned *V1,
                    unsigned *fracU, unsigned *fracV)
{
    unsigned u0 = (unsigned)(U >> 8); /* arithmetic shift of signed 24:8 */
    unsigned v0 = (unsigned)(V >> 8);
    unsigned u1 = u0 + 1, v1 = v0 + 1;
    *U0 = u0 & mapAnd;
    *V0 = (v0 & mapAnd) << mapShift;
    *U1 = u1 & mapAnd;
    *V1 = (v1 & mapAnd) << mapShift;
    *fracU = ((unsigned)U & 0xff) * 0x01010101u;
    *fracV = ((unsigned)V & 0xff) * 0x01010101u;
}

/* code/bilinear.h:34-53 (signed int arithmetic, arithmetic shifts) */
unsigned orc_bsamp8(const uint8_t *tex, unsigned U0, unsigned V0, unsigned U1,
                    unsigned V1, unsigned fracU, unsigned fracV)
{
    const int S0 = tex[U0 + V0];
    const int S1 = tex[U1 + V0];
    const int S2 = tex[U0 + V1];
    const int S3 = tex[U1 + V1];
    const int fu = (int)(fracU & 0xff), fv = (int)(fracV & 0xff);
    const int dX1 = (S1 - S0) * fu;
    const int dX2 = (S3 - S2) * fu;
    const int S01 = ((S0 << 8) + dX1) >> 8;
    const int S23 = ((S2 << 8) + dX2) >> 8;
    const int dY = (S23 - S01) * fv;
    return (unsigned)(((S01 << 8) + dY) >> 8);
}

/* code/bilinear.h:58-83: 16-bit lanes, mullo/add wrap mod 2^16, logical >> */
void orc_bsamp32_16(const uint32_t *tex, unsigned U0, unsigned V0, unsigned U1,
                    unsigned V1, unsigned fracU, unsigned fracV, uint16_t out[4])
{
    uint16_t S0[4], S1[4], S2[4], S3[4];
    unpack16(tex[U0 + V0], S0);
    unpack16(tex[U1 + V0], S1);
    unpack16(tex[U0 + V1], S2);
    unpack16(tex[U1 + V1], S3);
    const uint16_t fu = (uint16_t)(fracU & 0xff), fv = (uint16_t)(fracV & 0xff);
    for (int c = 0; c < 4; ++c) {
        uint16_t dX1 = (uint16_t)((uint16_t)(S1[c] - S0[c]) * fu);
        uint16_t dX2 = (uint16_t)((uint16_t)(S3[c] - S2[c]) * fu);
        uint16_t S01 = (uint16_t)((uint16_t)((uint16_t)(S0[c] << 8) + dX1) >> 8);
        uint16_t S23 = (uint16_t)((uint16_t)((uint16_t)(S2[c] << 8) + dX2) >> 8);
        uint16_t dY = (uint16_t)((uint16_t)(S23 - S01) * fv);
        out[c] = (uint16_t)((uint16_t)((uint16_t)(S01 << 8) + dY) >> 8);
    }
}

/* code/bilinear.h:86-125, followed by the packus_epi32/packus_epi16 pair of
 * Polar_Blit_Tile (code/polar.cpp:125-128) */
uint32_t orc_bsamp32_32_packed(const uint32_t *tex, unsigned U0, unsigned V0,
                               unsigned U1, unsigned V1, unsigned fracU,
                               unsigned fracV)
{
    const uint32_t p0 = tex[U0 + V0], p1 = tex[U1 + V0];
    const uint32_t p2 = tex[U0 + V1], p3 = tex[U1 + V1];
    const uint32_t fu = fracU & 0xff, fv = fracV & 0xff;
    uint32_t px = 0;
    for (int c = 0; c < 4; ++c) {
        const int32_t S0 = (p0 >> (8 * c)) & 0xff, S1 = (p1 >> (8 * c)) & 0xff;
        const int32_t S2 = (p2 >> (8 * c)) & 0xff, S3 = (p3 >> (8 * c)) & 0xff;
        uint32_t S01 = (uint32_t)((S0 << 8) + (S1 - S0) * (int32_t)fu) >> 8;
        uint32_t S23 = (uint32_t)((S2 << 8) + (S3 - S2) * (int32_t)fu) >> 8;
        uint32_t r = ((S01 << 8) + (S23 - S01) * fv) >> 8;
        px |= packus16(packus32((int32_t)r)) << (8 * c);
    }
    return px;
}

/* ---------------------------------------------------------------------- */
/* a4 cspanISSE16                                                           */
/* ---------------------------------------------------------------------- */

/* _mm_madd_epi16 of one 32-bit lane */
static inline int32_t madd16(uint32_t a, uint32_t b)
{
    int64_t lo = (int64_t)(int16_t)(a & 0xffff) * (int16_t)(b & 0xffff);
    int64_t hi = (int64_t)(int16_t)(a >> 16) * (int16_t)(b >> 16);
    return (int32_t)(uint32_t)(lo + hi);
}

/* code/cspan.h:47-78 (line numbers of the inline function cspanISSE16) */
void orc_cspan16(uint32_t *dst, ptrdiff_t destIncr, unsigned length,
                 unsigned drawLength, const uint16_t A[4], const uint16_t B[4])
{
    /* _mm_set1_epi32(65536/length) */
    const uint32_t divisor = 65536u / length;
    const unsigned preSteps = length - drawLength;
    /* _mm_mul_epi32(divisor, set1(preSteps)): signed 32x32->64 of lanes 0,2 */
    const int64_t m = (int64_t)(int32_t)divisor * (int64_t)(int32_t)preSteps;
    const uint32_t M[4] = {(uint32_t)m, (uint32_t)((uint64_t)m >> 32),
                           (uint32_t)m, (uint32_t)((uint64_t)m >> 32)};
    uint32_t acc[4];
    int32_t step[4];
    for (int c = 0; c < 4; ++c) {
        const uint32_t delta = (uint32_t)((int32_t)B[c] - (int32_t)A[c]);
        const int32_t preStep = madd16(delta, M[c]);
        step[c] = madd16(delta, divisor);
        acc[c] = ((uint32_t)A[c] << 16) + (uint32_t)preStep;
    }
    while (drawLength--) {
        uint32_t px = 0;
        for (int c = 0; c < 4; ++c) {
            const uint32_t v = acc[c] >> 16;          /* _mm_srli_epi32   */
            px |= packus16(packus32((int32_t)v)) << (8 * c);
        }
        *dst = px;
        dst += destIncr;
        for (int c = 0; c < 4; ++c) acc[c] += (uint32_t)step[c];
    }
}

/* ---------------------------------------------------------------------- */
/* a5..a7 voxel landscape                                                   */
/* ---------------------------------------------------------------------- */

#define LS_MAP_AND 1023u
#define LS_MAP_SHIFT 10u
#define LS_RAY_LENGTH 512u

/* code/landscape.cpp:56-107 */
void orc_vscape_ray(uint32_t *pColumn, unsigned W, unsigned H,
                    const uint8_t *heightMap, const uint32_t *colorMap,
                    const uint32_t *fog256, int mapScale, int mapTilt,
                    int curX, int curY, int dX, int dY, float fishMul)
{
    int lastHeight = (int)H;
    int lastDrawnHeight = (int)H;

    const unsigned U = ((unsigned)(curX >> 8)) & LS_MAP_AND;
    const unsigned V = (((unsigned)(curY >> 8)) & LS_MAP_AND) << LS_MAP_SHIFT;
    uint16_t lastColor[4];
    unpack16(colorMap[U | V], lastColor);

    const int fpFishMul = orc_ftofp24(fabsf(fishMul));

    for (unsigned iStep = 0; iStep < LS_RAY_LENGTH; ++iStep) {
        curX += dX;
        curY += dY;

        unsigned U0, V0, U1, V1, fracU, fracV;
        orc_bsamp_prep(curX, curY, LS_MAP_AND, LS_MAP_SHIFT, &U0, &V0, &U1, &V1,
                       &fracU, &fracV);
        const unsigned mapHeight =
            orc_bsamp8(heightMap, U0, V0, U1, V1, fracU, fracV);
        uint16_t color[4], fog[4];
        orc_bsamp32_16(colorMap, U0, V0, U1, V1, fracU, fracV, color);
        unpack16(fog256[iStep >> 1], fog);
        for (int c = 0; c < 4; ++c) color[c] = subs_u16(color[c], fog[c]);

        /* int / (int*unsigned) is an unsigned division; the numerator is
         * never negative here */
        int height = 255 - (int)mapHeight;
        height <<= 16;
        height = (int)((unsigned)height / ((unsigned)fpFishMul * (iStep + 1)));
        height *= mapScale;
        height >>= 8;
        height += mapTilt;

        if (height < lastDrawnHeight) {
            const unsigned drawLength = (unsigned)(lastDrawnHeight - height);
            orc_cspan16(pColumn + (ptrdiff_t)height * W, (ptrdiff_t)W,
                        (unsigned)(lastHeight - height), drawLength, color,
                        lastColor);
            lastDrawnHeight = height;
        }
        lastHeight = height;
        memcpy(lastColor, color, sizeof(color));
    }
}

/* code/landscape.cpp:32 and :169, at W x H.  At 1280x720 this is exactly the
 * reference constexpr: 1024 * ((720/1280) * (kPI*0.1f)). */
static float ls_rayY(unsigned W, unsigned H)
{
    const float kPI = 3.1415926535897932384626433832795f;
    const float aspect = (float)H / (float)W;
    const float viewLenScale = aspect * (kPI * 0.1f);
    return 1024.f * viewLenScale;
}

/* code/landscape.cpp:164-190 (per-column FP32 ray set-up).  The column
 * spacing 0.25 is rescaled by 1280/W so every resolution keeps the 1280-px
 * field of view; at W=1280 the factor is exactly 1. */
void orc_vscape_setup(unsigned W, unsigned H, const orc_landscape_cam *cam,
                      int *fpX1, int *fpY1, int *dXo, int *dYo, float *fishMulO)
{
    const float X1 = cam->X1, Y1 = cam->Y1;
    *fpX1 = orc_ftofp24(X1);
    *fpY1 = orc_ftofp24(Y1);
    const float rayY = ls_rayY(W, H);
    const float colScale = 0.25f * (1280.f / (float)W);
    const float halfW = (float)W * 0.5f;
    for (unsigned iRay = 0; iRay < W; ++iRay) {
        const float rayX = colScale * ((float)iRay - halfW);
        float rotRayX = rayX, rotRayY = rayY;
        /* voxel::vrot2D, code/voxel-shared.h:8-14 */
        {
            const float rx = cam->viewCos * rotRayX - cam->viewSin * rotRayY;
            const float ry = cam->viewSin * rotRayX + cam->viewCos * rotRayY;
            rotRayX = rx;
            rotRayY = ry;
        }
        const float X2 = X1 + rotRayX;
        const float Y2 = Y1 + rotRayY;
        float dX = X2 - X1;
        float dY = Y2 - Y1;
        /* voxel::vnorm2D, code/voxel-shared.h:16-24 */
        if (fabsf(dX + dY) > FLT_EPSILON) {
            const float length = 1.f / sqrtf(dX * dX + dY * dY);
            dX *= length;
            dY *= length;
        }
        fishMulO[iRay] = rayY / sqrtf(rotRayX * rotRayX + rotRayY * rotRayY);
        dXo[iRay] = orc_ftofp24(dX);
        dYo[iRay] = orc_ftofp24(dY);
    }
}

void orc_memset32(uint32_t *dst, uint32_t value, size_t n)
{
    /* code/util.h:57-67 */
    for (size_t i = 0; i < n; ++i) dst[i] = value;
}


/* Per-resolution scale of the reference's 720-row constants kMapScale (512)
 * and s_mapTilt (code/landscape.cpp:34-36,124): multiplied by H/720 with
 * integer arithmetic, exact at H=720. */
static int ls_scale_rows(int v720, unsigned H)
{
    return (int)(((long long)v720 * (long long)H) / 720);
}

/* code/landscape.cpp:228-243 (memset32 + vscape) without the optional
 * TapeWarp32, and code/landscape.cpp:109-191 with the camera passed in. */
void orc_landscape_frame(uint32_t *dst, unsigned W, unsigned H,
                         const uint8_t *heightMap, const uint32_t *colorMap,
                         const uint32_t *fog256, const orc_landscape_cam *cam,
                         int nthreads)
{
    int fpX1, fpY1;
    int *dX = (int *)malloc(sizeof(int) * W);
    int *dY = (int *)malloc(sizeof(int) * W);
    float *fish = (float *)malloc(sizeof(float) * W);
    orc_vscape_setup(W, H, cam, &fpX1, &fpY1, dX, dY, fish);
    orc_memset32(dst, fog256[0], (size_t)W * H);
    const int mapScale = ls_scale_rows(512, H);
    const int mapTilt = ls_scale_rows(cam->mapTilt, H);
    #pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int iRay = 0; iRay < (int)W; ++iRay)
        orc_vscape_ray(dst + iRay, W, H, heightMap, colorMap, fog256, mapScale,
                       mapTilt, fpX1, fpY1, dX[iRay], dY[iRay], fish[iRay]);
    free(dX);
    free(dY);
    free(fish);
}

/* ---------------------------------------------------------------------- */
/* a8 tunnelscape rows                                                      */
/* ---------------------------------------------------------------------- */

#define TS_MAP_AND 2047u
#define TS_MAP_SHIFT 11u

/* code/tunnelscape.cpp:44-100.  Heights of the map must stay <= 159: the
 * reference divides a negative height by an unsigned step count otherwise
 * (code/tunnelscape.cpp:79) and then writes outside the row. */
void orc_tscape_ray(uint32_t *pRow, unsigned W, const uint8_t *heightMap,
                    const uint32_t *colorMap, const uint32_t *fog256,
                    int curX, int curY, int dX, int dY)
{
    int lastHeight = (int)W;
    int lastDrawnHeight = (int)W;
    const unsigned U = ((unsigned)(curX >> 8)) & TS_MAP_AND;
    const unsigned V = (((unsigned)(curY >> 8)) & TS_MAP_AND) << TS_MAP_SHIFT;
    uint16_t lastColor[4];
    unpack16(colorMap[U | V], lastColor);
    const float kMapViewLenScale = (720.f / 1280.f) * 0.5f;

    for (unsigned iStep = 0; iStep < LS_RAY_LENGTH; ++iStep) {
        curX += dX;
        curY += dY;
        unsigned U0, V0, U1, V1, fracU, fracV;
        orc_bsamp_prep(curX, curY, TS_MAP_AND, TS_MAP_SHIFT, &U0, &V0, &U1, &V1,
                       &fracU, &fracV);
        const unsigned mapHeight =
            orc_bsamp8(heightMap, U0, V0, U1, V1, fracU, fracV);
        uint16_t color[4], fog[4];
        orc_bsamp32_16(colorMap, U0, V0, U1, V1, fracU, fracV, color);
        unpack16(fog256[iStep >> 1], fog);
        for (int c = 0; c < 4; ++c) color[c] = subs_u16(color[c], fog[c]);

        int height = 255 - (int)mapHeight;
        height -= 96;
        height = (int)((unsigned)height << 8);
        height = (int)((float)height / kMapViewLenScale);
        height = (int)((unsigned)height / (iStep + 1));
        height *= 160;
        height >>= 8;
        height += 120;

        if (height < lastDrawnHeight) {
            const unsigned drawLength = (unsigned)(lastDrawnHeight - height);
            orc_cspan16(pRow, 1, (unsigned)(lastHeight - height), drawLength,
                        color, lastColor);
            lastDrawnHeight = height;
            pRow += drawLength;
        }
        lastHeight = height;
        memcpy(lastColor, color, sizeof(color));
    }
}

/* code/tunnelscape.cpp:103-134 (tscape) and :168-171, parameters explicit */
void orc_tscape_frame(uint32_t *dst, unsigned W, unsigned H,
                      const uint8_t *heightMap2048, const uint32_t *colorMap2048,
                      const uint32_t *fog256, float stepU, float stepV,
                      float speed, float time, int nthreads)
{
    const float kGoldenRatio = 1.61803398875f;
    const float kOneOverAspect = 1.f / ((float)H / (float)W);
    const float mapStepX = 2048.f / (float)(H - 1);
    const float speedMul = sqrtf(stepU * stepU + stepV * stepV) * speed;
    const float fromY = 1024.f + speedMul * time;
    const int dX = orc_ftofp24(stepV);
    const int dY = orc_ftofp24(kOneOverAspect * stepU);
    const int fpFromY = orc_ftofp24(fromY);
    orc_memset32(dst, fog256[0], (size_t)W * H);
    #pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int iRay = 0; iRay < (int)H; ++iRay) {
        const float mapX = (float)(unsigned)iRay * mapStepX;
        const float fromX = mapX + stepU * time * kGoldenRatio;
        orc_tscape_ray(dst + (size_t)iRay * W, W, heightMap2048, colorMap2048,
                       fog256, orc_ftofp24(fromX), fpFromY, dX, dY);
    }
}

/* ---------------------------------------------------------------------- */
/* a9/a10 polar                                                             */
/* ---------------------------------------------------------------------- */

/* code/polar.cpp:18-59 (FP32 libm, same loop structure) */
void orc_polar_maps(int *pDest, int *pInvDest, unsigned srcResX, unsigned srcResY,
                    unsigned destResX, unsigned destResY)
{
    const float kPI = 3.1415926535897932384626433832795f;
    const float halfResX = destResX / 2.f;
    const float halfResY = destResY / 2.f;
    unsigned iPixel = 0;
    const float maxDist = sqrtf(halfResX * halfResX + halfResY * halfResY);
    for (float Y = -halfResY; Y < halfResY; Y += 1.f) {
        for (float X = -halfResX + FLT_EPSILON; X < halfResX; X += 1.f) {
            const float distance = sqrtf(X * X + Y * Y) / maxDist;
            float theta = atan2f(Y, X);
            theta += kPI;
            theta /= kPI * 2.f;
            const float U = distance * (srcResX - 1.f);
            const float invU = (1.f - distance) * (srcResX - 1.f);
            const float V = theta * (srcResY - 1.f);
            if (U >= srcResX - 1.f)
                pDest[iPixel] = (int)(((srcResX - 2) << 8) | 0xff);
            else
                pDest[iPixel] = orc_ftofp24(U);
            if (invU >= srcResX - 1.f)
                pInvDest[iPixel] = (int)(((srcResX - 2) << 8) | 0xff);
            else
                pInvDest[iPixel] = orc_ftofp24(invU);
            if (V >= srcResY - 1.f)
                pInvDest[iPixel + 1] = pDest[iPixel + 1] = (int)(((srcResY - 2) << 8) | 0xff);
            else
                pInvDest[iPixel + 1] = pDest[iPixel + 1] = orc_ftofp24(V);
            iPixel += 2;
        }
    }
}

/* code/polar.cpp:82-93 (Fetch32) + :108-133 (Polar_Blit_Tile) */
void orc_polar_blit(uint32_t *dst, const uint32_t *src, const int *map,
                    unsigned W, unsigned H, unsigned srcStride, int nthreads)
{
    #pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int y = 0; y < (int)H; ++y) {
        for (unsigned x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            const int U = map[2 * i], V = map[2 * i + 1];
            const unsigned U0 = (unsigned)(U >> 8);
            const unsigned V0 = (unsigned)(V >> 8) * srcStride;
            const unsigned fracU = ((unsigned)U & 0xff) * 0x01010101u;
            const unsigned fracV = ((unsigned)V & 0xff) * 0x01010101u;
            dst[i] = orc_bsamp32_32_packed(src, U0, V0, U0 + 1, V0 + srcStride,
                                           fracU, fracV);
        }
    }
}

/* ---------------------------------------------------------------------- */
/* a11 grid remap (build-defined; no reference counterpart)                 */
/* ---------------------------------------------------------------------- */

/* Grid: (ceil(W/8)+1) x (ceil(H/8)+1) corners, each a 24:8 (U,V) pair, row
 * major.  Pixel (x,y) lies in block (x>>3, y>>3) at local (i,j) = (x&7, y&7).
 * Integer bilinear interpolation of the four corner UVs:
 *   L = C00*(8-j) + C01*j,  R = C10*(8-j) + C11*j,  UV = (L*(8-i) + R*i) >> 6
 * (arithmetic shift, i.e. floor division by 64), then the reference's
 * wrapped bilinear fetch bsamp_prepUVs + bsamp32_32 (code/bilinear.h:10-125)
 * on a 2^texLog2 square ARGB texture, packed like Polar_Blit_Tile. */
void orc_grid_remap(uint32_t *dst, unsigned W, unsigned H, const int *grid,
                    const uint32_t *tex, unsigned texLog2, int nthreads)
{
    const unsigned GW = (W + 7) / 8 + 1;
    const unsigned mask = (1u << texLog2) - 1;
    #pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int y = 0; y < (int)H; ++y) {
        const unsigned by = (unsigned)y >> 3, j = (unsigned)y & 7;
        for (unsigned x = 0; x < W; ++x) {
            const unsigned bx = x >> 3, i = x & 7;
            const int *c00 = grid + 2 * ((size_t)by * GW + bx);
            const int *c10 = c00 + 2;
            const int *c01 = c00 + 2 * GW;
            const int *c11 = c01 + 2;
            int uv[2];
            for (int k = 0; k < 2; ++k) {
                const int L = c00[k] * (int)(8 - j) + c01[k] * (int)j;
                const int R = c10[k] * (int)(8 - j) + c11[k] * (int)j;
                uv[k] = (L * (int)(8 - i) + R * (int)i) >> 6;
            }
            unsigned U0, V0, U1, V1, fu, fv;
            orc_bsamp_prep(uv[0], uv[1], mask, texLog2, &U0, &V0, &U1, &V1, &fu, &fv);
            dst[(size_t)y * W + x] = orc_bsamp32_32_packed(tex, U0, V0, U1, V1, fu, fv);
        }
    }
}

/* ---------------------------------------------------------------------- */
/* a12 2007 box blur                                                        */
/* ---------------------------------------------------------------------- */

/* code/deprecated/boxblur.cpp:11-14; the caller stores it with
 * _mm_set1_epi16, i.e. truncated to 16 bits (65536 -> 0). */
static uint16_t b07_weight_to_div(unsigned weight)
{
    return (uint16_t)(((65536u * 256u) / weight) >> 4);
}

/* code/deprecated/boxblur.cpp:17-25 */
static inline void b07_add(uint16_t acc[4], uint16_t rem[4], uint32_t px,
                           unsigned shift)
{
    for (int c = 0; c < 4; ++c) {
        const uint16_t p = (uint16_t)((px >> (8 * c)) & 0xff);
        acc[c] = adds_u16(acc[c], rem[c]);
        rem[c] = (uint16_t)(p >> shift);
        acc[c] = adds_u16(acc[c], subs_u16(p, rem[c]));
    }
}

/* code/deprecated/boxblur.cpp:28-36 */
static inline void b07_sub(uint16_t acc[4], uint16_t rem[4], uint32_t px,
                           unsigned shift)
{
    for (int c = 0; c < 4; ++c) {
        const uint16_t p = (uint16_t)((px >> (8 * c)) & 0xff);
        acc[c] = subs_u16(acc[c], rem[c]);
        rem[c] = (uint16_t)(p >> shift);
        acc[c] = subs_u16(acc[c], subs_u16(p, rem[c]));
    }
}

/* code/deprecated/boxblur.cpp:39-42: mulhi_epu16 then packus_epi16 */
static inline uint32_t b07_div(const uint16_t acc[4], uint16_t div)
{
    uint32_t px = 0;
    for (int c = 0; c < 4; ++c) {
        const uint16_t q = (uint16_t)(((uint32_t)acc[c] * div) >> 16);
        px |= packus16(q) << (8 * c);
    }
    return px;
}

/* code/deprecated/boxblur.cpp:53-56: span = clampi(1,255,unsigned(s*255.f)).
 * Negative or NaN strengths are not defined by the reference (float to
 * unsigned conversion); they map to 1 here. */
int orc_blur07_span(float strength)
{
    const float f = strength * 255.f;
    if (!(f >= 0.f)) return 1;
    unsigned k = f >= 4294967040.f ? 0xffffffffu : (unsigned)f;
    int ki = (int)k;
    if (ki > 255) ki = 255;
    if (ki < 1) ki = 1;
    return ki;
}

typedef struct b07_setup {
    unsigned span, edgeSpan, shift, median;
    int subEdges;
    uint16_t edgeDivs[256];
    uint16_t fullDiv;
} b07_setup;

/* code/deprecated/boxblur.cpp:52-76 */
static void b07_make_setup(float strength, b07_setup *s)
{
    s->span = (unsigned)orc_blur07_span(strength);
    s->subEdges = (s->span & 1) == 0;
    s->edgeSpan = s->span >> 1;
    s->shift = 1 + (unsigned)(!s->subEdges) * 7;
    s->median = s->edgeSpan + (unsigned)(!s->subEdges);
    const unsigned startWeight = (s->median << 4) + ((unsigned)s->subEdges << 3);
    unsigned cw = startWeight;
    for (unsigned i = 0; i < s->median; ++i) {
        s->edgeDivs[i] = b07_weight_to_div(cw);
        cw += 16;
    }
    s->fullDiv = b07_weight_to_div(s->span << 4);
}

/* one line of HorizontalBoxBlur32 / VerticalBoxBlur32
 * (code/deprecated/boxblur.cpp:81-125 and :165-219); `stride` is 1 for rows
 * and xRes for columns.  Reads and writes interleave exactly as in the
 * reference, so dst == src reproduces the in-place recurrence. */
static void b07_line(uint32_t *dst, const uint32_t *src, size_t stride,
                     unsigned len, const b07_setup *s)
{
    size_t destIndex = 0, addPos = 0, subPos = 0;
    uint16_t acc[4] = {0, 0, 0, 0}, addRem[4] = {0, 0, 0, 0}, subRem[4] = {0, 0, 0, 0};
    const unsigned fullPassLen = len - (s->median + s->edgeSpan);

    for (unsigned i = 0; i < s->edgeSpan; ++i) {
        b07_add(acc, addRem, src[addPos], s->shift);
        addPos += stride;
    }
    for (unsigned i = 0; i < s->median; ++i) {
        b07_add(acc, addRem, src[addPos], s->shift);
        addPos += stride;
        dst[destIndex] = b07_div(acc, s->edgeDivs[i]);
        destIndex += stride;
    }
    for (unsigned i = 0; i < fullPassLen; ++i) {
        b07_add(acc, addRem, src[addPos], s->shift);
        addPos += stride;
        b07_sub(acc, subRem, src[subPos], s->shift);
        subPos += stride;
        dst[destIndex] = b07_div(acc, s->fullDiv);
        destIndex += stride;
    }
    if (s->subEdges)
        for (int c = 0; c < 4; ++c) acc[c] = adds_u16(acc[c], addRem[c]);
    for (unsigned i = s->edgeSpan; i > 0; --i) {
        b07_sub(acc, subRem, src[subPos], s->shift);
        subPos += stride;
        dst[destIndex] = b07_div(acc, s->edgeDivs[i - 1]);
        destIndex += stride;
    }
}

/* code/deprecated/boxblur.cpp:44-126.  Requires span <= xRes (the reference
 * walks out of the row otherwise); nothing is written when it does not hold. */
void orc_blur07_h(uint32_t *dst, const uint32_t *src, unsigned xRes,
                  unsigned yRes, float strength, int nthreads)
{
    b07_setup s;
    b07_make_setup(strength, &s);
    if (s.span > xRes) return;
    #pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int y = 0; y < (int)yRes; ++y)
        b07_line(dst + (size_t)y * xRes, src + (size_t)y * xRes, 1, xRes, &s);
}

/* code/deprecated/boxblur.cpp:128-220 (requires span <= yRes) */
void orc_blur07_v(uint32_t *dst, const uint32_t *src, unsigned xRes,
                  unsigned yRes, float strength, int nthreads)
{
    b07_setup s;
    b07_make_setup(strength, &s);
    if (s.span > yRes) return;
    #pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int x = 0; x < (int)xRes; ++x)
        b07_line(dst + x, src + x, xRes, yRes, &s);
}

/* code/deprecated/boxblur.cpp:222-231: H (dst<-src) then V in place on dst */
void orc_blur07(uint32_t *dst, const uint32_t *src, unsigned xRes,
                unsigned yRes, float strength, int nthreads)
{
    orc_blur07_h(dst, src, xRes, yRes, strength, nthreads);
    orc_blur07_v(dst, dst, xRes, yRes, strength, nthreads);
}

/* ---------------------------------------------------------------------- */
/* a14..a16 blitters                                                        */
/* ---------------------------------------------------------------------- */

static inline uint32_t avg_u8x4(uint32_t a, uint32_t b)
{
    uint32_t r = 0;
    for (int c = 0; c < 4; ++c) {
        const uint32_t x = (a >> (8 * c)) & 0xff, y = (b >> (8 * c)) & 0xff;
        r |= ((x + y + 1) >> 1) << (8 * c);              /* _mm_avg_epu8 */
    }
    return r;
}

/* code/fx-blitter.cpp:27-75.  The FX map is (dstW/2+4) x (dstH/2+4); the
 * guard column dstW/2 and guard row dstH/2 are read. */
void orc_fx_blit_2x2(uint32_t *dst, const uint32_t *src, unsigned dstW,
                     unsigned dstH)
{
    const unsigned mapW = dstW / 2 + 4;
    for (unsigned iY = 0; iY < dstH / 2; ++iY) {
        const uint32_t *r0 = src + (size_t)iY * mapW;
        const uint32_t *r1 = r0 + mapW;
        uint32_t *top = dst + (size_t)(2 * iY) * dstW;
        uint32_t *bot = top + dstW;
        for (unsigned x = 0; x < dstW / 2; ++x) {
            const uint32_t avgH0 = avg_u8x4(r0[x], r0[x + 1]);
            const uint32_t avgH1 = avg_u8x4(r1[x], r1[x + 1]);
            top[2 * x] = r0[x];
            top[2 * x + 1] = avgH0;
            bot[2 * x] = avg_u8x4(r0[x], r1[x]);
            bot[2 * x + 1] = avg_u8x4(avgH0, avgH1);
        }
    }
}

static inline uint32_t add_sat_u8x4(uint32_t d, uint32_t s)
{
    uint32_t r = 0;
    for (int c = 0; c < 4; ++c) {
        uint32_t v = ((d >> (8 * c)) & 0xff) + ((s >> (8 * c)) & 0xff);
        r |= (v > 255 ? 255 : v) << (8 * c);
    }
    return r;
}

/* code/util.cpp:132-145 */
void orc_add32(uint32_t *dst, const uint32_t *src, unsigned n)
{
    for (unsigned i = 0; i < n; ++i) dst[i] = add_sat_u8x4(dst[i], src[i]);
}

/* code/util.cpp:754-773 */
void orc_blit_add32(uint32_t *dst, const uint32_t *src, unsigned destResX,
                    unsigned srcResX, unsigned yRes)
{
    for (unsigned y = 0; y < yRes; ++y)
        for (unsigned x = 0; x < srcResX; ++x)
            dst[(size_t)y * destResX + x] =
                add_sat_u8x4(dst[(size_t)y * destResX + x], src[(size_t)y * srcResX + x]);
}

/* code/util.cpp:775-796: s' = (s * a8) >> 8 per 16-bit lane, then add */
void orc_blit_add32a(uint32_t *dst, const uint32_t *src, unsigned destResX,
                     unsigned srcResX, unsigned yRes, float alpha)
{
    /* c2vISSE16(0x01010101 * unsigned(alpha*255.f)): lane c is byte c of
     * the wrapped product (all equal to a8 when a8 <= 255) */
    const uint32_t m = 0x01010101u * (uint32_t)(alpha * 255.f);
    for (unsigned y = 0; y < yRes; ++y)
        for (unsigned x = 0; x < srcResX; ++x) {
            const uint32_t s = src[(size_t)y * srcResX + x];
            uint32_t sm = 0;
            for (int c = 0; c < 4; ++c)
                sm |= ((((s >> (8 * c)) & 0xff) * ((m >> (8 * c)) & 0xff)) >> 8) << (8 * c);
            dst[(size_t)y * destResX + x] = add_sat_u8x4(dst[(size_t)y * destResX + x], sm);
        }
}

/* ---------------------------------------------------------------------- */
/* c2 frame                                                                 */
/* ---------------------------------------------------------------------- */

/* remap into dst, blur07 H (dst->tmp0) and V (tmp0->tmp1) out of place with
 * span 2r+1 (strength (2r+1.5)/255), then BlitAdd32(dst, tmp1). */
void orc_tunnel_frame(uint32_t *dst, uint32_t *tmp0, uint32_t *tmp1,
                      unsigned W, unsigned H, const int *gridUV,
                      const uint32_t *tex, unsigned texLog2, int radius,
                      int nthreads)
{
    const float strength = (2.f * (float)radius + 1.5f) / 255.f;
    orc_grid_remap(dst, W, H, gridUV, tex, texLog2, nthreads);
    orc_blur07_h(tmp0, dst, W, H, strength, nthreads);
    orc_blur07_v(tmp1, tmp0, W, H, strength, nthreads);
    orc_blit_add32(dst, tmp1, W, W, H);
}
