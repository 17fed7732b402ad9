/* harness over code/util.cpp and code/sincos-lut.cpp */
#include "code/util.cpp"
#include "code/sincos-lut.cpp"
#include "ref_common.h"

REF_API void ref_add32(uint32_t *dst, const uint32_t *src, unsigned n)
{
    Add32(dst, src, n);
}

REF_API void ref_blit_add32(uint32_t *dst, const uint32_t *src, unsigned destResX,
                            unsigned srcResX, unsigned yRes)
{
    BlitAdd32(dst, src, destResX, srcResX, yRes);
}

REF_API void ref_blit_add32a(uint32_t *dst, const uint32_t *src, unsigned destResX,
                             unsigned srcResX, unsigned yRes, float alpha)
{
    BlitAdd32A(dst, src, destResX, srcResX, yRes, alpha);
}

/* code/sincos-lut.cpp:9-16, code/sincos-lut.h:13-26 */
REF_API void ref_cos_lut(float *out2049)
{
    CalculateCosLUT();
    for (unsigned i = 0; i <= kCosTabSize; ++i) out2049[i] = g_cosLUT[i];
}
REF_API float ref_lutcosf(float a) { return lutcosf(a); }
REF_API float ref_lutsinf(float a) { return lutsinf(a); }

/* code/util.cpp:552-603 (needs the LUT: call ref_cos_lut first) */
REF_API void ref_tape_warp32(uint32_t *dst, const uint32_t *src, unsigned xRes,
                             unsigned yRes, float strength, float speed)
{
    TapeWarp32(dst, src, xRes, yRes, strength, speed);
}
