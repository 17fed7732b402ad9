/* harness over code/fx-blitter.cpp (1280x720 output, 644x364 FX map) */
#include "code/fx-blitter.cpp"
#include "ref_common.h"

REF_API void ref_fx_blit_2x2(uint32_t *dst, const uint32_t *src)
{
    Fx_Blit_2x2(dst, src);
}

/* FxBlitter_DrawTestPattern (code/fx-blitter.cpp:77-100), which fills
 * g_pFxMap[0]: the maps are created first when they do not exist */
REF_API int ref_fx_test_pattern(uint32_t *dst)
{
    if (nullptr == g_pFxMap[0] && !FxBlitter_Create()) return -1;
    FxBlitter_DrawTestPattern(dst);
    return 0;
}
