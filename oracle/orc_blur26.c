/*
 * orc_blur26.c -- the 2026 multi-pass box blur (code/boxblur.cpp).
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Restates HorzBlur32 (code/boxblur.cpp:87-212) lane by lane, including its
 * quirks: the head lerp reads pLine[head+2] and the tail lerp pLine[tail+1],
 * so every row reads two pixels past its end (the next row's first two
 * pixels, or whatever follows the buffer for the last row).  Callers of this
 * restatement pass buffers that hold two readable guard pixels after the
 * image; the product defines those two pixels as zero past the last row.
 */
#include "ckd_oracle.h"

#include <stdint.h>
#include <string.h>

#define B26_MAX_RADIUS 500

/* code/boxblur.cpp:45-56 (iAdd / iSub share the lerp term) */
static inline int32_t b26_lerp(int32_t A, int32_t B, int32_t iAlpha)
{
    /* _mm_srai_epi32(_mm_mullo_epi32(B-A, alpha), 16) */
    const int32_t prod = (int32_t)((uint32_t)(B - A) * (uint32_t)iAlpha);
    return A + (prod >> 16);
}

/* code/boxblur.cpp:69-79 then v2cISSE32 (code/util.h:152-154) */
static inline uint32_t b26_div_pack(const int32_t sum[4], int64_t iScale)
{
    const uint32_t k = (uint32_t)(uint64_t)iScale; /* mul_epu32 uses the low dword */
    uint32_t px = 0;
    for (int c = 0; c < 4; ++c) {
        const uint64_t prod = (uint64_t)(uint32_t)sum[c] * (uint64_t)k;
        const uint64_t r = prod >> 22;
        /* _mm_hadd_epi32 adds the two dwords of each 64-bit lane */
        const int32_t lane = (int32_t)((uint32_t)r + (uint32_t)(r >> 32));
        const int32_t u16 = lane < 0 ? 0 : (lane > 65535 ? 65535 : lane);
        const int16_t s16 = (int16_t)(uint16_t)u16;
        const uint32_t b = s16 < 0 ? 0u : (s16 > 255 ? 255u : (uint32_t)s16);
        px |= b << (8 * c);
    }
    return px;
}

static inline void b26_unpack(uint32_t px, int32_t v[4])
{
    for (int c = 0; c < 4; ++c) v[c] = (int32_t)((px >> (8 * c)) & 0xff);
}

/* code/boxblur.cpp:87-212 */
static void horz_blur32(uint32_t *pDest, uint32_t *pScratch, const uint32_t *pSrc,
                        unsigned xRes, unsigned yRes, unsigned writeStrideCol,
                        unsigned writeStrideRow, float strength, float gain,
                        unsigned numPasses)
{
    strength *= 0.01f;
    const float r0 = strength * (float)((xRes - 2) / 2);
    const float radius = r0 < (float)B26_MAX_RADIUS ? r0 : (float)B26_MAX_RADIUS;
    const unsigned iSpan = (unsigned)radius;

    const float scale = 1.f / ((2.f - gain) * radius + 1.f);
    const float alpha = radius - (float)iSpan;
    const int64_t iScale = (int64_t)(scale * (float)(1 << 22));
    const int32_t iAlpha = (int32_t)(65536.f * alpha);

    const float halfScale = scale * 0.5f;
    const float dScale = halfScale / (float)iSpan;
    int64_t spanScales[B26_MAX_RADIUS];
    for (unsigned i = 0; i < iSpan; ++i)
        spanScales[i] = (int64_t)((halfScale + (float)i * dScale) * (float)(1 << 22));

    const uint32_t *pRead = pSrc;
    if ((numPasses & 1) == 0) {
        uint32_t *t = pDest;
        pDest = pScratch;
        pScratch = t;
    }

    for (unsigned iPass = 0; iPass < numPasses; ++iPass) {
        unsigned colStride = xRes, rowStride = 1;
        if (iPass == numPasses - 1) {
            colStride = writeStrideCol;
            rowStride = writeStrideRow;
        }
        for (unsigned iY = 0; iY < yRes; ++iY) {
            size_t writeIdx = (size_t)iY * colStride;
            const uint32_t *pLine = pRead + (size_t)iY * xRes;
            int32_t iSum[4] = {0, 0, 0, 0}, v[4], hA[4], hB[4], tA[4], tB[4];
            unsigned tail = 0, head = 0;

            for (unsigned i = 0; i < iSpan; ++i) {
                b26_unpack(pLine[head++], v);
                for (int c = 0; c < 4; ++c) iSum[c] += v[c];
            }
            b26_unpack(pLine[head], v);
            for (int c = 0; c < 4; ++c)
                iSum[c] += (int32_t)((uint32_t)v[c] * (uint32_t)iAlpha) >> 16;

            b26_unpack(pLine[head + 1], hA);
            for (unsigned i = 0; i < iSpan; ++i) {
                pDest[writeIdx] = b26_div_pack(iSum, spanScales[i]);
                writeIdx += rowStride;
                b26_unpack(pLine[head + 2], hB);
                for (int c = 0; c < 4; ++c) iSum[c] += b26_lerp(hA[c], hB[c], iAlpha);
                memcpy(hA, hB, sizeof(hA));
                ++head;
            }

            b26_unpack(pLine[tail], tA);
            for (unsigned i = 0; i < xRes - iSpan * 2; ++i) {
                pDest[writeIdx] = b26_div_pack(iSum, iScale);
                writeIdx += rowStride;
                b26_unpack(pLine[head + 2], hB);
                for (int c = 0; c < 4; ++c) iSum[c] += b26_lerp(hA[c], hB[c], iAlpha);
                memcpy(hA, hB, sizeof(hA));
                ++head;
                b26_unpack(pLine[tail + 1], tB);
                for (int c = 0; c < 4; ++c) iSum[c] -= b26_lerp(tA[c], tB[c], iAlpha);
                memcpy(tA, tB, sizeof(tA));
                ++tail;
            }

            for (unsigned i = iSpan; i > 0; --i) {
                pDest[writeIdx] = b26_div_pack(iSum, spanScales[i - 1]);
                writeIdx += rowStride;
                b26_unpack(pLine[tail + 1], tB);
                for (int c = 0; c < 4; ++c) iSum[c] -= b26_lerp(tA[c], tB[c], iAlpha);
                memcpy(tA, tB, sizeof(tA));
                ++tail;
            }
        }
        pRead = pDest;
        uint32_t *t = pDest;
        pDest = pScratch;
        pScratch = t;
    }
}

/* code/boxblur.cpp:215-268 (a plain transpose) */
static void transpose32(uint32_t *pDest, const uint32_t *pSrc, unsigned xRes,
                        unsigned yRes)
{
    for (unsigned y = 0; y < yRes; ++y)
        for (unsigned x = 0; x < xRes; ++x)
            pDest[(size_t)x * yRes + y] = pSrc[(size_t)y * xRes + x];
}

/* code/boxblur.cpp:270-279 */
void orc_blur26_h(uint32_t *dst, const uint32_t *src, unsigned xRes,
                  unsigned yRes, float strength, float gain, unsigned passes,
                  uint32_t *scratch0, uint32_t *scratch1)
{
    (void)scratch1;
    horz_blur32(dst, scratch0, src, xRes, yRes, xRes, 1, strength, gain, passes);
}

/* code/boxblur.cpp:281-297 */
void orc_blur26_v(uint32_t *dst, const uint32_t *src, unsigned xRes,
                  unsigned yRes, float strength, float gain, unsigned passes,
                  uint32_t *scratch0, uint32_t *scratch1)
{
    transpose32(scratch1, src, xRes, yRes);
    horz_blur32(dst, scratch0, scratch1, yRes, xRes, 1, xRes, strength, gain, passes);
}

/* code/boxblur.cpp:299-318 */
void orc_blur26(uint32_t *dst, const uint32_t *src, unsigned xRes,
                unsigned yRes, float strength, float gain, unsigned passes,
                uint32_t *scratch0, uint32_t *scratch1)
{
    horz_blur32(scratch1, scratch0, src, xRes, yRes, 1, yRes, strength, gain, passes);
    horz_blur32(dst, scratch0, scratch1, yRes, xRes, 1, xRes, strength, gain, passes);
}
