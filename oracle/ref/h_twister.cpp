/* harness over code/torus-twister.cpp (1280x720): the harness installs
 * caller maps in the statics Twister_Create loads through DevIL, runs its
 * table set-up and binds its Rocket tracks */
#include <unistd.h>
#include "code/torus-twister.cpp"
#include "ref_common.h"

REF_API int ref_twister_setup(const char *trackDir, const uint8_t *heightMap,
                              const uint32_t *colorMap, const uint32_t *background)
{
    vtwister_precalc();
    s_pHeightMap = const_cast<uint8_t *>(heightMap);
    s_pColorMap = const_cast<uint32_t *>(colorMap);
    s_pBackground = const_cast<uint32_t *>(background);
    /* code/torus-twister.cpp:153-155, the track files read from trackDir/sync/ */
    char cwd[4096];
    if (!getcwd(cwd, sizeof(cwd)) || chdir(trackDir) != 0) return -1;
    trackTwisterSpeed = Rocket::AddTrack("twister:Speed");
    trackTwisterShearSpeed = Rocket::AddTrack("twister::ShearSpeed");
    trackTwisterBlur = Rocket::AddTrack("twister:Blur");
    return chdir(cwd) == 0 ? 0 : -1;
}

REF_API void ref_twister_draw(uint32_t *pDest, float time, float delta)
{
    Twister_Draw(pDest, time, delta);
}

/* the projection tables vtwister_precalc fills (code/torus-twister.cpp:118-134) */
REF_API void ref_twister_tables(unsigned *heightProj, unsigned *heightProjNorm)
{
    for (unsigned i = 0; i < kRayLength; ++i) {
        heightProj[i] = s_heightProj[i];
        heightProjNorm[i] = s_heightProjNorm[i];
    }
}
