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
