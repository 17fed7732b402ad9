/* harness over code/shadertoy.cpp (1280x720 output, 644x364 FX map).
 *
 * The effects read their parameters through Rocket::getf; the harness
 * starts the reference's own Rocket player (code/rocket.cpp over
 * 3rdparty/rocket-stripped/lib, SYNC_PLAYER) from track files the caller
 * writes, binds the effects' track handles as Shadertoy_Create does
 * (code/shadertoy.cpp:106-172; Shadertoy_Create itself also loads PNG
 * assets through DevIL, absent here), and selects the row. */
#include <unistd.h>
#include "code/shadertoy.cpp"
#include "ref_common.h"

namespace Rocket { extern double s_rocketRow; }

/* effect ids, shared with the oracle */
enum { ST_PLASMA, ST_NAUTILUS, ST_SPIKEY_DISTANT, ST_SPIKEY_CLOSE, ST_SINUSES, ST_LAURA,
       ST_FDTUNNEL };

REF_API int ref_st_init(void)
{
    CalculateCosLUT();
    if (nullptr == g_pFxMap[0] && !FxBlitter_Create()) return -1;
    return 0;
}

/* (re)start the player with the track files under dir/sync/ */
REF_API int ref_st_rocket(const char *dir)
{
    char cwd[4096];
    if (!getcwd(cwd, sizeof(cwd)) || chdir(dir) != 0) return -1;
    const bool ok = Rocket::Launch();
    struct { SyncTrack *track; const char *name; } tracks[] = {
        {&trackLauraSpeed, "laura:Speed"}, {&trackLauraYaw, "laura:Yaw"},
        {&trackLauraPitch, "laura:Pitch"}, {&trackLauraRoll, "laura:Roll"},
        {&trackLauraHue, "laura:Hue"}, {&trackLauraSaturate, "laura:Saturate"},
        {&trackNautilusRoll, "nautilus:Roll"}, {&trackNautilusBlur, "nautilus:Blur"},
        {&trackNautilusHue, "nautilus:Hue"}, {&trackNautilusSpeed, "nautilus:Speed"},
        {&trackNautilusDesaturation, "nautilus:Desat"},
        {&trackSpikeSpeed, "spike:Speed"}, {&trackSpikeRoll, "spike:Roll"},
        {&trackSpikeSpecular, "spike:Specular"}, {&trackSpikeDesaturation, "spike:Desaturation"},
        {&trackDistSpikeWarmup, "distSpike:Warmup"}, {&trackSpikeHue, "spike:Hue"},
        {&trackSpikeGamma, "spike:Gamma"}, {&trackDistSpikeX, "distSpike:xOffs"},
        {&trackDistSpikeY, "distSpike:yOffs"}, {&trackDistSpikeZ, "distSpike:zOffs"},
        {&trackCloseSpikeX, "closeSpike:xOffs"}, {&trackCloseSpikeY, "closeSpike:yOffs"},
        {&trackCloseSpikeZ, "closeSpike:zOffs"}, {&trackCloseSpikeZScale, "closeSpike:zOffsScale"},
        {&trackCloseSpikeNormalGrain, "closeSpike:NormalGrain"},
        {&trackCloseSpikeScale, "closeSpike:Scale"}, {&trackCloseSpikeRim, "closeSpike:Rim"},
        {&trackCloseSpikeAspectMul, "closeSpike:AspectMul"},
        {&trackCloseMixBlurMap, "closeSpike:MixBlurMap"}, {&trackCloseMixBlur, "closeSpike:MixBlur"},
        {&trackCloseMixMapBlur, "closeSpike:MixMapBlur"},
        {&trackCloseMixBlurOpacity, "closeSpike:MixBlurOpacity"},
        {&trackSinusesSpecular, "sinusesTunnel:Specular"}, {&trackSinusesRoll, "sinusesTunnel:Roll"},
        {&trackSinusesSpeed, "sinusesTunnel:Speed"}, {&trackSinusesOffsX, "sinusesTunnel:OffsX"},
        {&trackSinusesGamma, "sinusesTunnel:Gamma"}, {&trackSinusesHue, "sinusesTunnel:Hue"},
        {&trackSinusesDesat, "sinusesTunnel:Desaturation"},
        {&trackPlasmaSpeed, "plasma:Speed"}, {&trackPlasmaHue, "plasma:Hue"},
        {&trackPlasmaGamma, "plasma:Gamma"}, {&trackPlasmaDesat, "plasma:Desaturation"},
        {&trackTunnelBoxy, "tunnel:Boxy"}, {&trackTunnelFlowerScale, "tunnel:FlowerScale"},
        {&trackTunnelFlowerFreq, "tunnel:FlowerFreq"}, {&trackTunnelFlowerPhase, "tunnel:FlowerPhase"},
        {&trackTunnelSpeed, "tunnel:Speed"}, {&trackTunnelRoll, "tunnel:Roll"},
        {&trackTunnelPitch, "tunnel:Pitch"}, {&trackTunnelRadius, "tunnel:Radius"},
        {&trackTunnelMulU, "tunnel:MulU"}, {&trackTunnelMulV, "tunnel:MulV"},
        {&trackTunnelLitTiles, "tunnel:LitTiles"}, {&trackTunnelLitBlur, "tunnel:LitBlur"},
        {&trackTunnelFog1, "tunnel:Fog1"}, {&trackTunnelFog2, "tunnel:Fog2"},
    };
    for (auto &t : tracks) *t.track = Rocket::AddTrack(t.name);
    const int rc = chdir(cwd);
    return (ok && rc == 0) ? 0 : -1;
}

REF_API void ref_st_row(double row) { Rocket::s_rocketRow = row; }

REF_API double ref_st_get(const char *name) { return Rocket::get(Rocket::AddTrack(name)); }

/* the free-directional tunnel's two 1024x1024 textures */
REF_API void ref_st_textures(uint32_t *tex, uint32_t *texHighlights)
{
    s_pFDTunnelTex = tex;
    s_pFDTunnelTexHighlights = texHighlights;
}

/* the close-up's two blur maps (644x364 each) and the blend target
 * Shadertoy_Create allocates (code/shadertoy.cpp:180-185) */
REF_API void ref_st_blurmaps(uint32_t *map1, uint32_t *map2)
{
    s_pSpikeBlurMaps[0] = map1;
    s_pSpikeBlurMaps[1] = map2;
    if (nullptr == s_pSpikeBlurMap)
        s_pSpikeBlurMap = static_cast<uint32_t*>(mallocAligned((1280*720)/2 * sizeof(uint32_t), kAlignTo));
}

/* one FX map (644x364) of an effect; the distant spike picks its renderer
 * as Spikey_Draw does (code/shadertoy.cpp:712-725) */
REF_API void ref_st_map(int effect, uint32_t *dst, uint32_t *glow, float time)
{
    switch (effect) {
    case ST_PLASMA: RenderPlasmaMap(dst, time); break;
    case ST_NAUTILUS: RenderNautilusMap_2x2(dst, time); break;
    case ST_SPIKEY_DISTANT: {
        const float warmup = Rocket::getf(trackDistSpikeWarmup);
        if (0.f == warmup) RenderSpikeyMap_2x2_Distant(dst, time);
        else RenderSpikeyMap_2x2_Distant_SpecularOnly(dst, time, 1.f + warmup);
        break;
    }
    case ST_SPIKEY_CLOSE: RenderSpikeyMap_2x2_Close(dst, time); break;
    case ST_SINUSES: RenderSinMap_2x2(dst, time); break;
    case ST_LAURA: RenderLaura_2x2(dst, time); break;
    case ST_FDTUNNEL: RenderTunnelMap_2x2(dst, glow, time); break;
    }
}

/* the effect's Draw function into a 1280x720 frame */
REF_API void ref_st_draw(int effect, uint32_t *dst, float time, float delta)
{
    switch (effect) {
    case ST_PLASMA: Plasma_Draw(dst, time, delta); break;
    case ST_NAUTILUS: Nautilus_Draw(dst, time, delta); break;
    case ST_SPIKEY_DISTANT: Spikey_Draw(dst, time, delta, false); break;
    case ST_SPIKEY_CLOSE: Spikey_Draw(dst, time, delta, true); break;
    case ST_SINUSES: Sinuses_Draw(dst, time, delta); break;
    case ST_LAURA: Laura_Draw(dst, time, delta); break;
    case ST_FDTUNNEL: Tunnel_Draw(dst, time, delta); break;
    }
}

/* helpers of code/shadertoy-util.h, one lane / pixel at a time */
REF_API void ref_st_gamma(const float *in4, float gamma, float *out4)
{
    __m128 v = _mm_loadu_ps(in4);
    _mm_storeu_ps(out4, Shadertoy::GammaAdj(v, gamma));
}

REF_API float ref_st_expfog(float distance, float scale)
{
    return Shadertoy::ExpFog(distance, scale);
}

REF_API void ref_st_vlerp4(const float *a4, const float *b4, float factor, float *out4)
{
    _mm_storeu_ps(out4, Shadertoy::vLerp4(_mm_loadu_ps(a4), _mm_loadu_ps(b4), factor));
}

REF_API void ref_st_topixel4(const float *in16, uint32_t *out4)
{
    __m128 c[4];
    for (int i = 0; i < 4; ++i) c[i] = _mm_loadu_ps(in16 + 4 * i);
    _mm_storeu_si128(reinterpret_cast<__m128i *>(out4), Shadertoy::ToPixel4(c));
}
