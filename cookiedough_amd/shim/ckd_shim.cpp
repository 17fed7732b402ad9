// ckd_shim.cpp -- the reference's effect / filter / blit entry points over
// the libckd C ABI.  See ckd_shim.h.
//
// One process-wide context is created by the first *_Create of any part and
// destroyed by the last *_Destroy.  Host buffers are uploaded into device
// staging images the shim owns, processed, and downloaded before the call
// returns; errors go to SetLastError, as the reference's Create functions
// report theirs (code/main.h:49).
#include "ckd_shim.h"

#include <cmath>
#include <cstring>
#include <string>

#include "../../include/ckd.h"
#include "ckd_host_api.h"

namespace {

constexpr unsigned kResX = 1280, kResY = 720;                  // code/main.h:37-38
constexpr unsigned kFxMapResX = kResX / 2 + 4, kFxMapResY = kResY / 2 + 4; // code/fx-blitter.h:16-17
constexpr float kPI = 3.1415926535897932384626433832795f;

ckd_ctx *g_ctx = nullptr;
int g_users = 0;

void report(const char *what)
{
    SetLastError(std::string(what) + ": " + ckd_last_error());
}

bool acquire()
{
    if (!g_ctx && ckd_create(0, &g_ctx) != CKD_OK) {
        report("libckd context");
        return false;
    }
    ++g_users;
    return true;
}

void release()
{
    if (g_users > 0 && --g_users == 0) {
        ckd_destroy(g_ctx);
        g_ctx = nullptr;
    }
}

// The reference's free functions (the blits and blends of util.cpp, the
// blurs, Fx_Blit_2x2) need no Create call.  The first of them made while no
// part holds the context takes one reference for them, held until the
// process ends; later calls reuse it.
bool g_lazyHeld = false;

bool ensure_ctx()
{
    if (g_ctx) return true;
    if (!acquire()) return false;
    g_lazyHeld = true;
    return true;
}

// The context reference of one part (landscape, polar, ...), taken by its
// Create call.  A Create that fails after taking it gives it back when the
// guard goes out of scope; a Create that succeeds keeps it (done()) until
// the part's Destroy.  Creating a part that exists destroys it first.
struct PartRef {
    bool &ready;
    bool took = false;
    explicit PartRef(bool &r) : ready(r) {}
    bool take()
    {
        if (ready) {
            ready = false;
            release();
        }
        return took = acquire();
    }
    bool done()
    {
        ready = true;
        took = false;
        return true;
    }
    ~PartRef()
    {
        if (took) release();
    }
};

void drop(bool &ready)
{
    if (ready) release();
    ready = false;
}

// device image that grows on demand
struct DevImage {
    uint32_t *p = nullptr;
    size_t cap = 0;
    uint32_t *get(size_t pixels)
    {
        if (pixels <= cap) return p;
        if (p) ckd_free(g_ctx, p);
        p = nullptr;
        cap = 0;
        void *q = nullptr;
        if (ckd_malloc(g_ctx, &q, pixels * 4) != CKD_OK) return nullptr;
        p = static_cast<uint32_t *>(q);
        cap = pixels;
        return p;
    }
};

DevImage g_a, g_b;

bool up(uint32_t *dev, const uint32_t *host, size_t pixels)
{
    return ckd_memcpy_h2d(g_ctx, dev, host, pixels * 4) == CKD_OK;
}

bool down(uint32_t *host, const uint32_t *dev, size_t pixels)
{
    return ckd_memcpy_d2h(g_ctx, host, dev, pixels * 4) == CKD_OK;
}

// unary filter: upload src, f(devDst, devSrc), download; dst == src keeps
// the reference's in-place semantics where f supports them
template <typename F>
void filter(const char *name, uint32_t *pDest, const uint32_t *pSrc, size_t pixels, bool inPlace,
            F f)
{
    if (!ensure_ctx()) return;
    uint32_t *a = g_a.get(pixels);
    uint32_t *b = inPlace ? a : g_b.get(pixels);
    if (!a || !b || !up(a, pSrc, pixels) || f(b, a) != CKD_OK || !down(pDest, b, pixels))
        report(name);
}

// binary blit: dst = f(dst, src)
template <typename F>
void blit2(const char *name, uint32_t *pDest, size_t dstPixels, const uint32_t *pSrc,
           size_t srcPixels, F f)
{
    if (!ensure_ctx()) return;
    uint32_t *a = g_a.get(dstPixels), *b = g_b.get(srcPixels);
    if (!a || !b || !up(a, pDest, dstPixels) || !up(b, pSrc, srcPixels) || f(a, b) != CKD_OK ||
        !down(pDest, a, dstPixels))
        report(name);
}

// ---- landscape state (code/landscape.cpp:15-47, 109-168) -----------------
bool s_lsReady = false;
SyncTrack trackVoxelScapeForward, trackVoxelScapeTilt, trackWarpSpeed, trackWarpStrength;
float s_padTilt = 0.f, s_viewAngle = 0.f, s_strafeX = 0.f, s_strafeY = 0.f;
float s_padMoveX = 0.f, s_padMoveY = 0.f;

float getf(SyncTrack t) { return (float)Rocket::get(t); }

float clampf(float lo, float hi, float v) { return std::fmax(lo, std::fmin(hi, v)); }

// ---- tunnelscape state (code/tunnelscape.cpp:19-20) ----------------------
bool s_tsReady = false;

// ---- filters: whether each part's Create holds a context reference -------
bool s_polarReady = false, s_fxReady = false, s_blurReady = false;
SyncTrack trackStarsStepU, trackStarsStepV, trackStarsSpeed, trackStarsBlur;

// ---- raymarched effects, ball, twister ------------------------------------
// Each effect's Rocket tracks, in the order of its ckd_*_params fields after
// `time`, come from the library (ckd_effect_tracks); a Draw call reads them
// into the struct.
template <typename P> struct FxTracks {
    static constexpr size_t N = sizeof(P) / sizeof(float) - 1;
    SyncTrack t[N] = {};
    bool add(const char *effect)
    {
        const char *const *names = nullptr;
        unsigned n = 0;
        if (ckd_effect_tracks(effect, &names, &n) != CKD_OK || n != N) return false;
        for (size_t i = 0; i < N; ++i) t[i] = Rocket::AddTrack(names[i]);
        return true;
    }
    // time, then the track values (Rocket::getf), into the struct of floats
    P read(float time) const
    {
        float v[N + 1];
        v[0] = time;
        for (size_t i = 0; i < N; ++i) v[i + 1] = (float)Rocket::get(t[i]);
        P p;
        memcpy(&p, v, sizeof(p));
        return p;
    }
};

bool s_stReady = false;
FxTracks<ckd_plasma_params> s_plasmaTracks;
FxTracks<ckd_nautilus_params> s_nautilusTracks;
FxTracks<ckd_spikey_params> s_spikeyTracks;
FxTracks<ckd_sinuses_params> s_sinusesTracks;
FxTracks<ckd_laura_params> s_lauraTracks;
FxTracks<ckd_fdtunnel_params> s_tunnelTracks;
uint32_t *s_spikeBlurMaps[2] = {nullptr, nullptr}; // the close-up's mix-blur layer

// ---- voxel ball and torus twister (code/ball.cpp:337-441,
// code/torus-twister.cpp:136-158) ------------------------------------------
bool s_ballReady = false, s_twReady = false;
FxTracks<ckd_ball_params> s_ballTracks;
FxTracks<ckd_twister_params> s_twisterTracks;
uint32_t *s_ballBackground = nullptr; // Ball_GetBackground
// the render target the ball and the twister draw into (the reference's
// g_renderTarget[0]), cleared when allocated; the beam rays leave its last
// column as it was
DevImage g_target;

uint32_t *render_target(size_t pixels)
{
    const bool fresh = g_target.cap < pixels;
    uint32_t *t = g_target.get(pixels);
    if (t && fresh && ckd_memset32(g_ctx, t, 0, pixels) != CKD_OK) return nullptr;
    return t;
}

// renders one frame on the GPU into the staging image and downloads it
template <typename F> void st_draw(const char *name, uint32_t *pDest, F render)
{
    if (!s_stReady) return;
    const size_t n = (size_t)kResX * kResY;
    uint32_t *a = g_a.get(n);
    if (!a || render(a) != CKD_OK || !down(pDest, a, n)) report(name);
}

} // namespace

// ===================================================================== parts

int ckd_shim_context_refs() { return g_users; }

uint32_t *g_pFxMap[kShimNumFxMaps] = {nullptr};

bool Landscape_Create()
{
    PartRef part(s_lsReady);
    if (!part.take()) return false;
    const uint8_t *hm = Image_Load8("assets/scape/D17.png");
    const uint32_t *cm = Image_Load32("assets/scape/C17W-edit.png");
    if (!hm || !cm) return false;
    const uint32_t *fog = Image_Load32("assets/scape/foggradient.jpg");
    if (!fog) return false;
    if (ckd_landscape_create(g_ctx, hm, cm, fog) != CKD_OK) {
        report("Landscape_Create");
        return false;
    }
    trackVoxelScapeForward = Rocket::AddTrack("voxelScape:Forward");
    trackVoxelScapeTilt = Rocket::AddTrack("voxelScape:Tilt");
    trackWarpSpeed = Rocket::AddTrack("voxelScape:WarpSpeed");
    trackWarpStrength = Rocket::AddTrack("voxelScape:WarpStrength");
    return part.done();
}

void Landscape_Destroy() { drop(s_lsReady); }

void Landscape_Draw(uint32_t *pDest, float time, float delta)
{
    (void)time;
    if (!s_lsReady) return;
    // the camera of vscape (code/landscape.cpp:111-168)
    PadState pad;
    Gamepad_Update(pad);
    const float kMaxTiltDiff = 90.f;
    const float tiltStep = std::fmin(delta, 1.f);
    if (s_padTilt < kMaxTiltDiff && pad.rightY > 0.f) s_padTilt += tiltStep;
    if (s_padTilt > -kMaxTiltDiff && pad.rightY < 0.f) s_padTilt -= tiltStep;
    const float tilt = clampf(-kMaxTiltDiff, kMaxTiltDiff, getf(trackVoxelScapeTilt) + s_padTilt);
    const int mapTilt = 90 + int(tilt);
    const float kAspect = (float)kResY / kResX;
    float viewMul = 1.f / kAspect;
    viewMul *= delta * 0.01f;
    s_viewAngle += viewMul * pad.lShoulder;
    s_viewAngle -= viewMul * pad.rShoulder;
    const float viewCos = cosf(s_viewAngle);
    const float viewSin = sinf(s_viewAngle);
    if (pad.rightX != 0.f) {
        const float strafe = pad.rightX * delta;
        s_strafeX += viewCos * strafe;
        s_strafeY += viewSin * strafe;
    }
    const float padMove = (pad.leftY != 0.f) ? -pad.leftY * delta : 0.f;
    s_padMoveX += -viewSin * padMove;
    s_padMoveY += viewCos * padMove;
    const float fwd = getf(trackVoxelScapeForward);
    float moveX = -viewSin * fwd;
    float moveY = viewCos * fwd;
    moveX += s_padMoveX;
    moveY += s_padMoveY;
    ckd_landscape_cam cam{moveX + s_strafeX, moveY + s_strafeY, viewCos, viewSin, mapTilt};

    // Landscape_Draw (code/landscape.cpp:228-243)
    const size_t n = (size_t)kResX * kResY;
    const float warpStrength = getf(trackWarpStrength);
    uint32_t *a = g_a.get(n), *b = g_b.get(n);
    if (!a || !b || ckd_landscape_draw(g_ctx, a, kResX, kResY, &cam, 1) != CKD_OK) {
        report("Landscape_Draw");
        return;
    }
    uint32_t *out = a;
    if (warpStrength != 0.f) {
        if (ckd_tape_warp32(g_ctx, b, a, kResX, kResY, getf(trackWarpSpeed), warpStrength) != CKD_OK) {
            report("Landscape_Draw (TapeWarp32)");
            return;
        }
        out = b;
    }
    if (!down(pDest, out, n)) report("Landscape_Draw");
}

bool Tunnelscape_Create()
{
    PartRef part(s_tsReady);
    if (!part.take()) return false;
    const uint8_t *hm = Image_Load8("assets/scape/tscape-D7-edit.png");
    const uint32_t *cm = Image_Load32("assets/scape/tscape-C7W-edit.png");
    if (!hm || !cm) return false;
    const uint32_t *fog = Image_Load32("assets/scape/foggradient.jpg");
    if (!fog) return false;
    if (ckd_tunnelscape_create(g_ctx, hm, cm, fog) != CKD_OK ||
        ckd_polar_create(g_ctx, kResX, kResY) != CKD_OK) {
        report("Tunnelscape_Create");
        return false;
    }
    trackStarsStepU = Rocket::AddTrack("starsTunnel:stepU");
    trackStarsStepV = Rocket::AddTrack("starsTunnel:stepV");
    trackStarsSpeed = Rocket::AddTrack("starsTunnel:Speed");
    trackStarsBlur = Rocket::AddTrack("starsTunnel:Blur");
    return part.done();
}

void Tunnelscape_Destroy() { drop(s_tsReady); }

void Tunnelscape_Draw(uint32_t *pDest, float time, float delta)
{
    (void)delta;
    if (!s_tsReady) return;
    const size_t n = (size_t)kResX * kResY;
    uint32_t *a = g_a.get(n), *b = g_b.get(n);
    if (!a || !b ||
        ckd_tunnelscape_rows(g_ctx, a, kResX, kResY, getf(trackStarsStepU), getf(trackStarsStepV),
                             getf(trackStarsSpeed), time) != CKD_OK ||
        ckd_polar_blit(g_ctx, b, a, kResX, kResY, 1) != CKD_OK) {
        report("Tunnelscape_Draw");
        return;
    }
    float blur = getf(trackStarsBlur);
    if (blur != 0.f) {
        // BoxBlurScale (code/deprecated/boxblur.h:13-19), then BoxBlur32 in
        // place twice (code/tunnelscape.cpp:176-185)
        blur = clampf(1.f, 100.f, blur) * 0.01f;
        if (ckd_blur07(g_ctx, b, b, kResX, kResY, blur) != CKD_OK ||
            ckd_blur07(g_ctx, b, b, kResX, kResY, blur) != CKD_OK) {
            report("Tunnelscape_Draw (blur)");
            return;
        }
    }
    if (!down(pDest, b, n)) report("Tunnelscape_Draw");
}

bool Shadertoy_Create()
{
    PartRef part(s_stReady);
    if (!part.take()) return false;
    if (!s_plasmaTracks.add("plasma") || !s_nautilusTracks.add("nautilus") ||
        !s_spikeyTracks.add("spikey") || !s_sinusesTracks.add("sinuses") ||
        !s_lauraTracks.add("laura") || !s_tunnelTracks.add("fdtunnel")) {
        report("Shadertoy_Create (tracks)");
        return false;
    }
    // the free-directional tunnel's 1024x1024 textures and the close-up's
    // blur maps (code/shadertoy.cpp:174-185)
    const uint32_t *tex = Image_Load32("assets/shadertoy/nytrik-hextexture.png");
    const uint32_t *texHi = Image_Load32("assets/shadertoy/nytrik-hextexture-fx.png");
    if (!tex || !texHi) return false;
    s_spikeBlurMaps[0] = Image_Load32("assets/shadertoy/close-up-blur-map-1.png");
    s_spikeBlurMaps[1] = Image_Load32("assets/shadertoy/close-up-blur-map-2.png");
    if (!s_spikeBlurMaps[0] || !s_spikeBlurMaps[1]) return false;
    if (ckd_fdtunnel_create(g_ctx, tex, texHi) != CKD_OK ||
        ckd_spikey_blurmaps(g_ctx, s_spikeBlurMaps[0], s_spikeBlurMaps[1], kResX, kResY) != CKD_OK) {
        report("Shadertoy_Create");
        return false;
    }
    return part.done();
}

void Shadertoy_Destroy() { drop(s_stReady); }

void Plasma_Draw(uint32_t *pDest, float time, float delta)
{
    (void)delta;
    const ckd_plasma_params p = s_plasmaTracks.read(time);
    st_draw("Plasma_Draw", pDest,
            [&](uint32_t *d) { return ckd_plasma_draw(g_ctx, d, kResX, kResY, &p, 1); });
}

void Nautilus_Draw(uint32_t *pDest, float time, float delta)
{
    (void)delta;
    const ckd_nautilus_params p = s_nautilusTracks.read(time);
    st_draw("Nautilus_Draw", pDest,
            [&](uint32_t *d) { return ckd_nautilus_draw(g_ctx, d, kResX, kResY, &p, 1); });
}

void Spikey_Draw(uint32_t *pDest, float time, float delta, bool close)
{
    (void)delta;
    const ckd_spikey_params p = s_spikeyTracks.read(time);
    st_draw("Spikey_Draw", pDest, [&](uint32_t *d) {
        return ckd_spikey_draw(g_ctx, d, kResX, kResY, &p, 1, close ? 1 : 0);
    });
}

void Sinuses_Draw(uint32_t *pDest, float time, float delta)
{
    (void)delta;
    const ckd_sinuses_params p = s_sinusesTracks.read(time);
    st_draw("Sinuses_Draw", pDest,
            [&](uint32_t *d) { return ckd_sinuses_draw(g_ctx, d, kResX, kResY, &p, 1); });
}

void Laura_Draw(uint32_t *pDest, float time, float delta)
{
    (void)delta;
    const ckd_laura_params p = s_lauraTracks.read(time);
    st_draw("Laura_Draw", pDest,
            [&](uint32_t *d) { return ckd_laura_draw(g_ctx, d, kResX, kResY, &p, 1); });
}

void Tunnel_Draw(uint32_t *pDest, float time, float delta)
{
    (void)delta;
    const ckd_fdtunnel_params p = s_tunnelTracks.read(time);
    st_draw("Tunnel_Draw", pDest,
            [&](uint32_t *d) { return ckd_fdtunnel_draw(g_ctx, d, kResX, kResY, &p, 1); });
}

bool Polar_Create()
{
    PartRef part(s_polarReady);
    if (!part.take()) return false;
    // the output's maps and the FX map's (code/polar.cpp:61-70)
    if (ckd_polar_create(g_ctx, kResX, kResY) != CKD_OK ||
        ckd_polar_create(g_ctx, kFxMapResX, kFxMapResY) != CKD_OK) {
        report("Polar_Create");
        return false;
    }
    return part.done();
}

void Polar_Destroy() { drop(s_polarReady); }

void Polar_Blit(uint32_t *pDest, const uint32_t *pSrc, bool inverse)
{
    filter("Polar_Blit", pDest, pSrc, (size_t)kResX * kResY, false, [&](uint32_t *d, uint32_t *s) {
        return ckd_polar_blit(g_ctx, d, s, kResX, kResY, inverse ? 1 : 0);
    });
}

bool FxBlitter_Create()
{
    PartRef part(s_fxReady);
    if (!part.take()) return false;
    // one allocation for the four maps, 16-byte aligned (code/fx-blitter.cpp:10-20)
    free(g_pFxMap[0]);
    for (auto &m : g_pFxMap) m = nullptr;
    const size_t mapPixels = (size_t)kFxMapResX * kFxMapResY;
    void *base = nullptr;
    if (posix_memalign(&base, 16, mapPixels * 4 * kShimNumFxMaps) != 0) {
        SetLastError("FxBlitter_Create: out of host memory");
        return false;
    }
    memset(base, 0, mapPixels * 4 * kShimNumFxMaps);
    for (size_t i = 0; i < kShimNumFxMaps; ++i)
        g_pFxMap[i] = static_cast<uint32_t *>(base) + i * mapPixels;
    return part.done();
}

void FxBlitter_Destroy()
{
    free(g_pFxMap[0]);
    for (auto &m : g_pFxMap) m = nullptr;
    drop(s_fxReady);
}

void Fx_Blit_2x2(uint32_t *pDest, const uint32_t *pSrc)
{
    if (!ensure_ctx()) return;
    const size_t mapPixels = (size_t)kFxMapResX * kFxMapResY, n = (size_t)kResX * kResY;
    uint32_t *a = g_a.get(mapPixels), *b = g_b.get(n);
    if (!a || !b || !up(a, pSrc, mapPixels) || ckd_fx_blit_2x2(g_ctx, b, a, kResX, kResY) != CKD_OK ||
        !down(pDest, b, n))
        report("Fx_Blit_2x2");
}

// code/fx-blitter.cpp:77-100: the test pattern (upper half rows, lower half
// columns of alternating black and white) written into g_pFxMap[0], then
// Fx_Blit_2x2 of it
void FxBlitter_DrawTestPattern(uint32_t *pDest)
{
    if (!g_pFxMap[0]) {
        SetLastError("FxBlitter_DrawTestPattern: FxBlitter_Create has not been called");
        return;
    }
    for (unsigned y = 0; y < kFxMapResY; ++y)
        for (unsigned x = 0; x < kFxMapResX; ++x) {
            const unsigned bit = y < kFxMapResY / 2 ? y : x;
            g_pFxMap[0][y * kFxMapResX + x] = (bit & 1) ? 0xffffffffu : 0u;
        }
    Fx_Blit_2x2(pDest, g_pFxMap[0]);
}

bool BoxBlur_Create()
{
    PartRef part(s_blurReady);
    return part.take() && part.done();
}

void BoxBlur_Destroy() { drop(s_blurReady); }

void BoxBlur_Horz32(uint32_t *pDest, const uint32_t *pSrc, unsigned xRes, unsigned yRes,
                    float strength, float gain, unsigned numPasses)
{
    filter("BoxBlur_Horz32", pDest, pSrc, (size_t)xRes * yRes, false, [&](uint32_t *d, uint32_t *s) {
        return ckd_blur26_h(g_ctx, d, s, xRes, yRes, strength, gain, numPasses);
    });
}

void BoxBlur_Vert32(uint32_t *pDest, const uint32_t *pSrc, unsigned xRes, unsigned yRes,
                    float strength, float gain, unsigned numPasses)
{
    filter("BoxBlur_Vert32", pDest, pSrc, (size_t)xRes * yRes, false, [&](uint32_t *d, uint32_t *s) {
        return ckd_blur26_v(g_ctx, d, s, xRes, yRes, strength, gain, numPasses);
    });
}

void BoxBlur_32(uint32_t *pDest, const uint32_t *pSrc, unsigned xRes, unsigned yRes, float strength,
                float gain, unsigned numPasses)
{
    filter("BoxBlur_32", pDest, pSrc, (size_t)xRes * yRes, false, [&](uint32_t *d, uint32_t *s) {
        return ckd_blur26(g_ctx, d, s, xRes, yRes, strength, gain, numPasses);
    });
}

void HorizontalBoxBlur32(uint32_t *pDest, const uint32_t *pSrc, unsigned int xRes,
                         unsigned int yRes, float strength)
{
    filter("HorizontalBoxBlur32", pDest, pSrc, (size_t)xRes * yRes, pDest == pSrc,
           [&](uint32_t *d, uint32_t *s) { return ckd_blur07_h(g_ctx, d, s, xRes, yRes, strength); });
}

void VerticalBoxBlur32(uint32_t *pDest, const uint32_t *pSrc, unsigned int xRes, unsigned int yRes,
                       float strength)
{
    filter("VerticalBoxBlur32", pDest, pSrc, (size_t)xRes * yRes, pDest == pSrc,
           [&](uint32_t *d, uint32_t *s) { return ckd_blur07_v(g_ctx, d, s, xRes, yRes, strength); });
}

void BoxBlur32(uint32_t *pDest, const uint32_t *pSrc, unsigned int xRes, unsigned int yRes,
               float strength)
{
    filter("BoxBlur32", pDest, pSrc, (size_t)xRes * yRes, pDest == pSrc,
           [&](uint32_t *d, uint32_t *s) { return ckd_blur07(g_ctx, d, s, xRes, yRes, strength); });
}

void Add32(uint32_t *pDest, const uint32_t *pSrc, unsigned numPixels)
{
    blit2("Add32", pDest, numPixels, pSrc, numPixels,
          [&](uint32_t *d, uint32_t *s) { return ckd_add32(g_ctx, d, s, numPixels); });
}

void BlitAdd32(uint32_t *pDest, const uint32_t *pSrc, unsigned destResX, unsigned srcResX,
               unsigned yRes)
{
    blit2("BlitAdd32", pDest, (size_t)destResX * yRes, pSrc, (size_t)srcResX * yRes,
          [&](uint32_t *d, uint32_t *s) { return ckd_blit_add32(g_ctx, d, s, destResX, srcResX, yRes); });
}

void BlitAdd32A(uint32_t *pDest, const uint32_t *pSrc, unsigned destResX, unsigned srcResX,
                unsigned yRes, float alpha)
{
    blit2("BlitAdd32A", pDest, (size_t)destResX * yRes, pSrc, (size_t)srcResX * yRes,
          [&](uint32_t *d, uint32_t *s) {
              return ckd_blit_add32a(g_ctx, d, s, destResX, srcResX, yRes, alpha);
          });
}

void TapeWarp32(uint32_t *pDest, const uint32_t *pSrc, unsigned xRes, unsigned yRes,
                float strength, float speed)
{
    filter("TapeWarp32", pDest, pSrc, (size_t)xRes * yRes, false, [&](uint32_t *d, uint32_t *s) {
        return ckd_tape_warp32(g_ctx, d, s, xRes, yRes, strength, speed);
    });
}

// ---- compositing blends (code/util.cpp:58-811) ---------------------------

#define CKD_SHIM_FLAT_BLEND(Name, fn)                                                           \
    void Name(uint32_t *pDest, const uint32_t *pSrc, unsigned numPixels)                        \
    {                                                                                           \
        blit2(#Name, pDest, numPixels, pSrc, numPixels,                                         \
              [&](uint32_t *d, uint32_t *s) { return fn(g_ctx, d, s, numPixels); });            \
    }

CKD_SHIM_FLAT_BLEND(MixOver32, ckd_mixover32)
CKD_SHIM_FLAT_BLEND(Sub32, ckd_sub32)
CKD_SHIM_FLAT_BLEND(Excl32, ckd_excl32)
CKD_SHIM_FLAT_BLEND(SoftLight32, ckd_softlight32)
CKD_SHIM_FLAT_BLEND(SoftLight32A, ckd_softlight32a)
CKD_SHIM_FLAT_BLEND(Overlay32, ckd_overlay32)
CKD_SHIM_FLAT_BLEND(Overlay32A, ckd_overlay32a)
CKD_SHIM_FLAT_BLEND(Darken32_50, ckd_darken32_50)
CKD_SHIM_FLAT_BLEND(MulSrc32, ckd_mulsrc32)
CKD_SHIM_FLAT_BLEND(MulSrc32A, ckd_mulsrc32a)
CKD_SHIM_FLAT_BLEND(MixSrc32, ckd_mixsrc32)

#undef CKD_SHIM_FLAT_BLEND

void Mix32(uint32_t *pDest, const uint32_t *pSrc, unsigned numPixels, uint8_t alpha)
{
    blit2("Mix32", pDest, numPixels, pSrc, numPixels,
          [&](uint32_t *d, uint32_t *s) { return ckd_mix32(g_ctx, d, s, numPixels, alpha); });
}

void SoftLight32AA(uint32_t *pDest, const uint32_t *pSrc, unsigned numPixels, float alpha)
{
    blit2("SoftLight32AA", pDest, numPixels, pSrc, numPixels, [&](uint32_t *d, uint32_t *s) {
        return ckd_softlight32aa(g_ctx, d, s, numPixels, alpha);
    });
}

void MixSrc32S(uint32_t *pDest, const uint32_t *pSrc, unsigned destResX, unsigned destResY,
               unsigned srcStride)
{
    // the source rows reach (destResY-1)*srcStride + destResX pixels
    const size_t srcPixels = destResY ? (size_t)(destResY - 1) * srcStride + destResX : 0;
    blit2("MixSrc32S", pDest, (size_t)destResX * destResY, pSrc, srcPixels,
          [&](uint32_t *d, uint32_t *s) {
              return ckd_mixsrc32s(g_ctx, d, s, destResX, destResY, srcStride);
          });
}

void BlitSrc32(uint32_t *pDest, const uint32_t *pSrc, unsigned destResX, unsigned srcResX,
               unsigned yRes)
{
    blit2("BlitSrc32", pDest, (size_t)destResX * yRes, pSrc, (size_t)srcResX * yRes,
          [&](uint32_t *d, uint32_t *s) { return ckd_blitsrc32(g_ctx, d, s, destResX, srcResX, yRes); });
}

void BlitSrc32A(uint32_t *pDest, const uint32_t *pSrc, unsigned destResX, unsigned srcResX,
                unsigned yRes, float alpha)
{
    blit2("BlitSrc32A", pDest, (size_t)destResX * yRes, pSrc, (size_t)srcResX * yRes,
          [&](uint32_t *d, uint32_t *s) {
              return ckd_blitsrc32a(g_ctx, d, s, destResX, srcResX, yRes, alpha);
          });
}

void Fade32(uint32_t *pDest, unsigned int numPixels, uint32_t RGB, uint8_t alpha)
{
    filter("Fade32", pDest, pDest, numPixels, true,
           [&](uint32_t *d, uint32_t *) { return ckd_fade32(g_ctx, d, numPixels, RGB, alpha); });
}

void Zoom32(uint32_t *pDest, const uint32_t *pSrc, unsigned xRes, unsigned yRes, float scale)
{
    filter("Zoom32", pDest, pSrc, (size_t)xRes * yRes, false, [&](uint32_t *d, uint32_t *s) {
        return ckd_zoom32(g_ctx, d, s, xRes, yRes, scale);
    });
}

// Polar_Blit at the FX map's size, 644x364 (code/polar.cpp:200-218)
void Polar_Blit_2x2(uint32_t *pDest, const uint32_t *pSrc, bool inverse)
{
    filter("Polar_Blit_2x2", pDest, pSrc, (size_t)kFxMapResX * kFxMapResY, false,
           [&](uint32_t *d, uint32_t *s) {
               return ckd_polar_blit(g_ctx, d, s, kFxMapResX, kFxMapResY, inverse ? 1 : 0);
           });
}

void Polar_BlitA(uint32_t *pDest, const uint32_t *pSrc, bool inverse)
{
    const size_t n = (size_t)kResX * kResY;
    blit2("Polar_BlitA", pDest, n, pSrc, n, [&](uint32_t *d, uint32_t *s) {
        return ckd_polar_blit_a(g_ctx, d, s, kResX, kResY, inverse ? 1 : 0);
    });
}

// ---- voxel ball ---------------------------------------------------------

bool Ball_Create()
{
    PartRef part(s_ballReady);
    if (!part.take()) return false;
    static const char *const kHeightMaps[5] = {"assets/ball/hmap_1_1k.jpg", "assets/ball/hmap_4_1k.jpg",
                                               "assets/ball/hmap_2_1k.jpg", "assets/ball/hmap_3_1k.jpg",
                                               "assets/ball/hmap_5_1k.jpg"};
    static const char *const kColorMaps[2] = {"assets/ball/colormap_1k.jpg",
                                              "assets/ball/colormap_2_1k.jpg"};
    static const char *const kBeamMaps[3] = {"assets/ball/beammap_1k_1.jpg",
                                             "assets/ball/beammap_1k_2.jpg",
                                             "assets/ball/beammap_1k_3-2.jpg"};
    static const char *const kBackgrounds[2] = {"assets/ball/nytrik-background_1280x720.png",
                                                "assets/ball/nytrik-background-2-1280x720.png"};
    constexpr size_t kMap = 1024 * 1024, kOut = (size_t)kResX * kResY;
    std::string err;
    uint8_t *heights = new uint8_t[5 * kMap];
    uint32_t *images = new uint32_t[6 * kMap + 3 * kOut];
    uint32_t *colors = images, *beams = images + 2 * kMap, *env = images + 5 * kMap;
    uint32_t *bgs = images + 6 * kMap, *halo = bgs + 2 * kOut;
    bool ok = true;
    for (int i = 0; ok && i < 5; ++i) {
        const uint8_t *m = Image_Load8(kHeightMaps[i]);
        if ((ok = m != nullptr)) memcpy(heights + i * kMap, m, kMap);
    }
    for (int i = 0; ok && i < 2; ++i) {
        const uint32_t *m = Image_Load32(kColorMaps[i]);
        if ((ok = m != nullptr)) memcpy(colors + i * kMap, m, kMap * 4);
    }
    for (int i = 0; ok && i < 3; ++i) {
        const uint32_t *m = Image_Load32(kBeamMaps[i]);
        if ((ok = m != nullptr)) memcpy(beams + i * kMap, m, kMap * 4);
    }
    const uint32_t *e = ok ? Image_Load32("assets/ball/envmap3_1k.jpg") : nullptr;
    if (ok && (ok = e != nullptr)) memcpy(env, e, kMap * 4);
    uint32_t *bg[2] = {nullptr, nullptr};
    for (int i = 0; ok && i < 2; ++i) {
        bg[i] = Image_Load32(kBackgrounds[i]);
        if ((ok = bg[i] != nullptr)) memcpy(bgs + i * kOut, bg[i], kOut * 4);
    }
    const uint32_t *h = ok ? Image_Load32("assets/ball/halo.png") : nullptr;
    if (ok && (ok = h != nullptr)) memcpy(halo, h, kOut * 4);
    if (ok && ckd_ball_create(g_ctx, heights, colors, beams, env, bgs, halo, kResX, kResY) != CKD_OK) {
        report("Ball_Create");
        ok = false;
    }
    delete[] heights;
    delete[] images;
    if (!ok) return false;
    if (!s_ballTracks.add("ball")) {
        report("Ball_Create (tracks)");
        return false;
    }
    s_ballBackground = bg[0];
    return part.done();
}

void Ball_Destroy() { drop(s_ballReady); }

void Ball_Draw(uint32_t *pDest, float time, float delta)
{
    (void)delta;
    if (!s_ballReady) return;
    const ckd_ball_params p = s_ballTracks.read(time);
    const size_t n = (size_t)kResX * kResY;
    uint32_t *a = g_a.get(n), *t = render_target(n);
    if (!a || !t || ckd_ball_draw(g_ctx, a, t, kResX, kResY, &p, 1) != CKD_OK || !down(pDest, a, n))
        report("Ball_Draw");
}

uint32_t *Ball_GetBackground() { return s_ballBackground; }

bool Ball_HasBeams()
{
    // Rocket::geti: int(roundf(getf(track)))
    return (int)roundf((float)Rocket::get(s_ballTracks.t[4])) != 0;
}

// ---- torus twister --------------------------------------------------------

bool Twister_Create()
{
    PartRef part(s_twReady);
    if (!part.take()) return false;
    const uint8_t *hm = Image_Load8("assets/twister/hmap_2_1k.jpg");
    const uint32_t *cm = Image_Load32("assets/twister/colormap_1k.jpg");
    const uint32_t *bg = (hm && cm) ? Image_Load32("assets/twister/nytrik-background_1280x720.png")
                                    : nullptr;
    if (!bg || ckd_twister_create(g_ctx, hm, cm, bg, kResX, kResY) != CKD_OK) {
        if (bg) report("Twister_Create");
        return false;
    }
    if (!s_twisterTracks.add("twister")) {
        report("Twister_Create (tracks)");
        return false;
    }
    return part.done();
}

void Twister_Destroy() { drop(s_twReady); }

void Twister_Draw(uint32_t *pDest, float time, float delta)
{
    (void)delta;
    if (!s_twReady) return;
    const ckd_twister_params p = s_twisterTracks.read(time);
    const size_t n = (size_t)kResX * kResY;
    uint32_t *a = g_a.get(n), *t = render_target(n);
    if (!a || !t || ckd_twister_draw(g_ctx, a, t, kResX, kResY, &p, 1) != CKD_OK || !down(pDest, a, n))
        report("Twister_Draw");
}
