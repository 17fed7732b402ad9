/* harness over code/polar.cpp */
#include "code/polar.cpp"
#include "ref_common.h"

/* code/polar.cpp:18-59 at any size */
REF_API void ref_polar_maps(int *map, int *invMap, unsigned srcResX,
                            unsigned srcResY, unsigned destResX, unsigned destResY)
{
    CalculateMaps(map, invMap, srcResX, srcResY, destResX, destResY);
}

/* code/polar.cpp:61-72 */
REF_API int ref_polar_create(void) { return Polar_Create() ? 1 : 0; }
REF_API void ref_polar_destroy(void) { Polar_Destroy(); }

/* code/polar.cpp:135-154 (1280x720, 16-byte aligned buffers) */
REF_API void ref_polar_blit(uint32_t *dst, const uint32_t *src, int inverse)
{
    Polar_Blit(dst, src, inverse != 0);
}

/* code/polar.cpp:157-198 (1280x720; the forward blit's 32-row tiles reach
 * 16 rows past the image: dst needs 16 guard rows, and the forward map
 * must be padded with ref_polar_pad_forward first) */
REF_API void ref_polar_blit_a(uint32_t *dst, const uint32_t *src, int inverse)
{
    Polar_BlitA(dst, src, inverse != 0);
}

/* Replaces the forward map by a copy with 16 zero rows appended, so the
 * forward blits' last tile row reads inside its allocation (the map values
 * for the image itself are unchanged). */
REF_API void ref_polar_pad_forward(void)
{
    const size_t rows = kTargetResY + 16;
    int *padded = static_cast<int *>(mallocAligned(kTargetResX * rows * 2 * sizeof(int), kAlignTo));
    memset(padded, 0, kTargetResX * rows * 2 * sizeof(int));
    memcpy(padded, s_pMap, kTargetResX * kTargetResY * 2 * sizeof(int));
    freeAligned(s_pMap);
    s_pMap = padded;
}

/* code/polar.cpp:200-218 (644x364).  Its 64-row (inverse 32-row) tiles of
 * 64 (32) pixels run past the image's last column into the next row and
 * past its last row: dst needs 24 guard rows, and the maps must be padded
 * with ref_polar_pad_2x2 first.  The pixels inside the image are those of a
 * per-pixel blit with the 644x364 maps. */
REF_API void ref_polar_blit_2x2(uint32_t *dst, const uint32_t *src, int inverse)
{
    Polar_Blit_2x2(dst, src, inverse != 0);
}

/* Replaces both FX-map maps by copies with 24 zero rows appended. */
REF_API void ref_polar_pad_2x2(void)
{
    const size_t rows = kFxMapResY + 24;
    int **maps[2] = {&s_pMap2x2, &s_pInvMap2x2};
    for (int **m : maps) {
        int *padded = static_cast<int *>(mallocAligned(kFxMapResX * rows * 2 * sizeof(int), kAlignTo));
        memset(padded, 0, kFxMapResX * rows * 2 * sizeof(int));
        memcpy(padded, *m, kFxMapSize * 2 * sizeof(int));
        freeAligned(*m);
        *m = padded;
    }
}
