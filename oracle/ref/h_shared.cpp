/* harness over code/shared-resources.cpp: the shared render targets and
 * the grey gradient the voxel ball reads (Shared_Create also loads logos
 * through DevIL, absent here, so the harness sets up the two it needs) */
#include "code/shared-resources.cpp"
#include "ref_common.h"

REF_API int ref_shared_init(void)
{
    /* code/shared-resources.cpp:15-23 */
    for (unsigned i = 0; i < kNumGradients; ++i) g_gradientUnp16[i] = c2vISSE16(i * 0x01010101);
    for (unsigned t = 0; t < kNumRenderTargets; ++t)
        if (!g_renderTarget[t])
            g_renderTarget[t] = static_cast<uint32_t *>(mallocAligned(kTargetBytes, kAlignTo));
    return 0;
}

/* the render target the ball and twister draw into (1280x720) */
REF_API uint32_t *ref_shared_target(unsigned index) { return g_renderTarget[index]; }

/* zero a render target (the voxel ball's beam rays leave its last column as
 * it was, so a reproducible run starts from a known target) */
REF_API void ref_shared_clear(unsigned index) { memset(g_renderTarget[index], 0, kTargetBytes); }
