/* harness over code/fx-blitter.cpp (1280x720 output, 644x364 FX map) */
#include "code/fx-blitter.cpp"
#include "ref_common.h"

REF_API void ref_fx_blit_2x2(uint32_t *dst, const uint32_t *src)
{
    Fx_Blit_2x2(dst, src);
}
