/* harness over code/tunnelscape.cpp */
#include <unistd.h>
#include "code/tunnelscape.cpp"
#include "ref_common.h"

/* code/tunnelscape.cpp:44-100 over caller-supplied 2048^2 maps / fog LUT */
REF_API void ref_tscape_ray(uint32_t *pRow, const uint8_t *heightMap,
                            const uint32_t *colorMap, const uint32_t *fog256,
                            int curX, int curY, int dX, int dY)
{
    s_pHeightMap = const_cast<uint8_t *>(heightMap);
    s_pColorMap = const_cast<uint32_t *>(colorMap);
    for (int i = 0; i < 256; ++i) s_fogGradientUnp[i] = c2vISSE16(fog256[i]);
    tscape_ray(pRow, curX, curY, dX, dY);
}

/* Tunnelscape_Create (code/tunnelscape.cpp:138-162) on the caller's maps and
 * fog LUT in place of the PNG/JPEG assets (their loader, DevIL, is absent);
 * the four tracks are read from trackDir/sync/ through the reference's
 * Rocket player, which ref_st_rocket has started */
REF_API int ref_tunnelscape_setup(const char *trackDir, const uint8_t *heightMap,
                                  const uint32_t *colorMap, const uint32_t *fog256)
{
    s_pHeightMap = const_cast<uint8_t *>(heightMap);
    s_pColorMap = const_cast<uint32_t *>(colorMap);
    s_pFogGradient = const_cast<uint32_t *>(fog256);
    for (int i = 0; i < 256; ++i) s_fogGradientUnp[i] = c2vISSE16(fog256[i]);
    char cwd[4096];
    if (!getcwd(cwd, sizeof(cwd)) || chdir(trackDir) != 0) return -1;
    trackStarsStepU = Rocket::AddTrack("starsTunnel:stepU");
    trackStarsStepV = Rocket::AddTrack("starsTunnel:stepV");
    trackStarsSpeed = Rocket::AddTrack("starsTunnel:Speed");
    trackStarsBlur = Rocket::AddTrack("starsTunnel:Blur");
    return chdir(cwd) == 0 ? 0 : -1;
}

/* the first two statements of Tunnelscape_Draw (code/tunnelscape.cpp:170-171):
 * the fog fill and tscape's 720 rows into a 1280x720 target */
REF_API void ref_tscape(uint32_t *pTarget, float time)
{
    memset32(pTarget, s_pFogGradient[0], kTargetResX * kTargetResY);
    tscape(pTarget, time);
}

/* Tunnelscape_Draw (code/tunnelscape.cpp:168-186): rows into
 * g_renderTarget[0], inverse Polar_Blit into pDest, two in-place BoxBlur32 */
REF_API void ref_tunnelscape_draw(uint32_t *pDest, float time, float delta)
{
    Tunnelscape_Draw(pDest, time, delta);
}
