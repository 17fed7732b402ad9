/* harness over code/deprecated/boxblur.cpp */
#include "code/deprecated/boxblur.cpp"
#include "ref_common.h"

REF_API void ref_blur07_h(uint32_t *dst, const uint32_t *src, unsigned xRes,
                          unsigned yRes, float strength)
{
    HorizontalBoxBlur32(dst, src, xRes, yRes, strength);
}

REF_API void ref_blur07_v(uint32_t *dst, const uint32_t *src, unsigned xRes,
                          unsigned yRes, float strength)
{
    VerticalBoxBlur32(dst, src, xRes, yRes, strength);
}

REF_API void ref_blur07(uint32_t *dst, const uint32_t *src, unsigned xRes,
                        unsigned yRes, float strength)
{
    BoxBlur32(dst, src, xRes, yRes, strength);
}
