/* harness over code/ball.cpp (1280x720).  Ball_Create loads its maps
 * through DevIL (absent here): the harness installs caller maps in the
 * same statics, allocates the two mix maps as Ball_Create does, and binds
 * the effect's Rocket tracks (the player is started by ref_st_rocket). */
#include <unistd.h>
#include "code/ball.cpp"
#include "ref_common.h"

REF_API int ref_ball_setup(const char *trackDir, const uint8_t *heightMaps5,
                           const uint32_t *colorMaps2, const uint32_t *beamMaps3,
                           const uint32_t *envMap, const uint32_t *backgrounds2,
                           const uint32_t *halo)
{
    const size_t mapPx = kMapSize * kMapSize;
    for (int i = 0; i < 5; ++i) s_pHeightMap[i] = const_cast<uint8_t *>(heightMaps5 + i * mapPx);
    for (int i = 0; i < 2; ++i) s_pColorMap[i] = const_cast<uint32_t *>(colorMaps2 + i * mapPx);
    for (int i = 0; i < 3; ++i) s_pBeamMaps[i] = const_cast<uint32_t *>(beamMaps3 + i * mapPx);
    s_pEnvMap = const_cast<uint32_t *>(envMap);
    for (int i = 0; i < 2; ++i) s_pBackgrounds[i] = const_cast<uint32_t *>(backgrounds2 + i * kOutputSize);
    s_pHalo = const_cast<uint32_t *>(halo);
    if (!s_heightMapMix) {
        s_heightMapMix = static_cast<uint8_t *>(mallocAligned(mapPx * sizeof(uint8_t), kAlignTo));
        s_pBeamMapMix = static_cast<uint32_t *>(mallocAligned(mapPx * sizeof(uint32_t), kAlignTo));
    }
    /* code/ball.cpp:413-428, the track files read from trackDir/sync/ */
    char cwd[4096];
    if (!getcwd(cwd, sizeof(cwd)) || chdir(trackDir) != 0) return -1;
    trackBallBlur = Rocket::AddTrack("ball:Blur");
    trackBallRadius = Rocket::AddTrack("ball:Radius");
    trackBallRayLength = Rocket::AddTrack("ball:RayLength");
    trackBallSpikes = Rocket::AddTrack("ball:Spikes");
    trackBallHasBeams = Rocket::AddTrack("ball:HasBeams");
    trackBallBaseShapeIndex = Rocket::AddTrack("ball:BaseShapeIndex");
    trackBallSpeed = Rocket::AddTrack("ball:Speed");
    trackBallBeamAtten = Rocket::AddTrack("ball:BeamAttenuation");
    trackBallBeamAlphaMin = Rocket::AddTrack("ball:BeamAlphaMin");
    trackBallRotateOffsX = Rocket::AddTrack("ball:RotateOffsX");
    trackBallRotateOffsY = Rocket::AddTrack("ball:RotateOffsY");
    trackBallBeams1 = Rocket::AddTrack("ball:Beams1");
    trackBallBeams2 = Rocket::AddTrack("ball:Beams2");
    trackBallBeams3 = Rocket::AddTrack("ball:Beams3");
    trackBallLowBeams = Rocket::AddTrack("ball:BallLowBeams");
    return chdir(cwd) == 0 ? 0 : -1;
}

/* Ball_Draw into pDest (1280x720 plus 16 guard rows: the forward
 * Polar_BlitA writes past the image) */
REF_API void ref_ball_draw(uint32_t *pDest, float time, float delta) { Ball_Draw(pDest, time, delta); }
