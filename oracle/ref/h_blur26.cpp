/* harness over code/boxblur.cpp.  The scratch buffers are zeroed before
 * every call so that the two pixels the reference reads past each image
 * (code/boxblur.cpp:185,195) are defined. */
#include "code/boxblur.cpp"
#include "ref_common.h"
#include <string.h>

static void zero_scratch()
{
    if (!s_pScratch[0]) BoxBlur_Create();
    memset(s_pScratch[0], 0, kScratchSize * sizeof(uint32_t));
    memset(s_pScratch[1], 0, kScratchSize * sizeof(uint32_t));
}

REF_API void ref_blur26_h(uint32_t *dst, const uint32_t *src, unsigned xRes,
                          unsigned yRes, float strength, float gain, unsigned passes)
{
    zero_scratch();
    BoxBlur_Horz32(dst, src, xRes, yRes, strength, gain, passes);
}

REF_API void ref_blur26_v(uint32_t *dst, const uint32_t *src, unsigned xRes,
                          unsigned yRes, float strength, float gain, unsigned passes)
{
    zero_scratch();
    BoxBlur_Vert32(dst, src, xRes, yRes, strength, gain, passes);
}

REF_API void ref_blur26(uint32_t *dst, const uint32_t *src, unsigned xRes,
                        unsigned yRes, float strength, float gain, unsigned passes)
{
    zero_scratch();
    BoxBlur_32(dst, src, xRes, yRes, strength, gain, passes);
}
