/* harness over code/landscape.cpp (+ bilinear.h, cspan.h it includes) */
#include "code/landscape.cpp"
#include "ref_common.h"

/* code/bilinear.h:10-31 */
REF_API void ref_bsamp_prep(int U, int V, unsigned mapAnd, unsigned mapShift,
                            unsigned *out6)
{
    bsamp_prepUVs(U, V, mapAnd, mapShift, out6[0], out6[1], out6[2], out6[3],
                  out6[4], out6[5]);
}

/* code/bilinear.h:34-53 */
REF_API unsigned ref_bsamp8(const uint8_t *tex, unsigned U0, unsigned V0,
                            unsigned U1, unsigned V1, unsigned fu, unsigned fv)
{
    return bsamp8(tex, U0, V0, U1, V1, fu, fv);
}

/* code/bilinear.h:58-83, result narrowed with v2cISSE16 */
REF_API uint32_t ref_bsamp32_16(const uint32_t *tex, unsigned U0, unsigned V0,
                                unsigned U1, unsigned V1, unsigned fu, unsigned fv)
{
    return v2cISSE16(bsamp32_16(tex, U0, V0, U1, V1, fu, fv));
}

/* code/bilinear.h:86-125, result narrowed with v2cISSE32 */
REF_API uint32_t ref_bsamp32_32(const uint32_t *tex, unsigned U0, unsigned V0,
                                unsigned U1, unsigned V1, unsigned fu, unsigned fv)
{
    return v2cISSE32(bsamp32_32(tex, U0, V0, U1, V1, fu, fv));
}

/* code/cspan.h:47-78; A and B are packed ARGB colours widened by c2vISSE16 */
REF_API void ref_cspan16(uint32_t *dst, int destIncr, unsigned length,
                         unsigned drawLength, uint32_t A, uint32_t B)
{
    cspanISSE16(dst, destIncr, length, drawLength, c2vISSE16(A), c2vISSE16(B));
}

/* code/landscape.cpp:56-107 over caller-supplied maps / fog LUT / tilt */
REF_API void ref_vscape_ray(uint32_t *pColumn, const uint8_t *heightMap,
                            const uint32_t *colorMap, const uint32_t *fog256,
                            int mapTilt, int curX, int curY, int dX, int dY,
                            float fishMul)
{
    s_pHeightMap = const_cast<uint8_t *>(heightMap);
    s_pColorMap = const_cast<uint32_t *>(colorMap);
    for (int i = 0; i < 256; ++i) s_fogGradientUnp[i] = c2vISSE16(fog256[i]);
    s_mapTilt = mapTilt;
    vscape_ray(pColumn, curX, curY, dX, dY, fishMul);
}

/* code/util.h:57-67 */
REF_API void ref_memset32(uint32_t *dst, uint32_t value, size_t n)
{
    memset32(dst, (int)value, n);
}

/* Landscape_Draw's fill and vscape's column loop (code/landscape.cpp:170-191,
 * :228-233) over the reference's vscape_ray, OpenMP over columns as there,
 * for a 1280x720 frame; the per-column ray parameters (24:8 origin,
 * direction and fishMul) come from the caller */
REF_API void ref_vscape_columns(uint32_t *pDest, const uint8_t *heightMap,
                                const uint32_t *colorMap, const uint32_t *fog256,
                                int mapTilt, int fpX1, int fpY1, const int *dX,
                                const int *dY, const float *fishMul)
{
    s_pHeightMap = const_cast<uint8_t *>(heightMap);
    s_pColorMap = const_cast<uint32_t *>(colorMap);
    for (int i = 0; i < 256; ++i) s_fogGradientUnp[i] = c2vISSE16(fog256[i]);
    s_mapTilt = mapTilt;
    memset32(pDest, fog256[0], kResX * kResY);
    #pragma omp parallel for schedule(static)
    for (unsigned iRay = 0; iRay < kResX; ++iRay)
        vscape_ray(pDest + iRay, fpX1, fpY1, dX[iRay], dY[iRay], fishMul[iRay]);
}
