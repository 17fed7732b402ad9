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

/* the compositing blends, code/util.cpp:58-811 */
REF_API void ref_mix32(uint32_t *d, const uint32_t *s, unsigned n, uint8_t a) { Mix32(d, s, n, a); }
REF_API void ref_mixover32(uint32_t *d, const uint32_t *s, unsigned n) { MixOver32(d, s, n); }
REF_API void ref_sub32(uint32_t *d, const uint32_t *s, unsigned n) { Sub32(d, s, n); }
REF_API void ref_excl32(uint32_t *d, const uint32_t *s, unsigned n) { Excl32(d, s, n); }
REF_API void ref_softlight32(uint32_t *d, const uint32_t *s, unsigned n) { SoftLight32(d, s, n); }
REF_API void ref_softlight32a(uint32_t *d, const uint32_t *s, unsigned n) { SoftLight32A(d, s, n); }
REF_API void ref_softlight32aa(uint32_t *d, const uint32_t *s, unsigned n, float a)
{
    SoftLight32AA(d, s, n, a);
}
REF_API void ref_overlay32(uint32_t *d, const uint32_t *s, unsigned n) { Overlay32(d, s, n); }
REF_API void ref_overlay32a(uint32_t *d, const uint32_t *s, unsigned n) { Overlay32A(d, s, n); }
REF_API void ref_darken32_50(uint32_t *d, const uint32_t *s, unsigned n) { Darken32_50(d, s, n); }
REF_API void ref_mulsrc32(uint32_t *d, const uint32_t *s, unsigned n) { MulSrc32(d, s, n); }
REF_API void ref_mulsrc32a(uint32_t *d, const uint32_t *s, unsigned n) { MulSrc32A(d, s, n); }
REF_API void ref_mixsrc32(uint32_t *d, const uint32_t *s, unsigned n) { MixSrc32(d, s, n); }
REF_API void ref_mixsrc32s(uint32_t *d, const uint32_t *s, unsigned x, unsigned y, unsigned stride)
{
    MixSrc32S(d, s, x, y, stride);
}
REF_API void ref_blitsrc32(uint32_t *d, const uint32_t *s, unsigned dx, unsigned sx, unsigned y)
{
    BlitSrc32(d, s, dx, sx, y);
}
REF_API void ref_blitsrc32a(uint32_t *d, const uint32_t *s, unsigned dx, unsigned sx, unsigned y,
                            float a)
{
    BlitSrc32A(d, s, dx, sx, y, a);
}
REF_API void ref_fade32(uint32_t *d, unsigned n, uint32_t rgb, uint8_t a) { Fade32(d, n, rgb, a); }
REF_API void ref_zoom32(uint32_t *d, const uint32_t *s, unsigned x, unsigned y, float scale)
{
    Zoom32(d, s, x, y, scale);
}

/* OpenMP team size of the reference's parallel loops (their pragmas take
 * the runtime default) */
REF_API void ref_omp_threads(int n) { omp_set_num_threads(n); }
