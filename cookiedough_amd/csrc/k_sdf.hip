// k_sdf.hip -- sphere tracer of the synthetic SDF scene.
//
// Structure of the reference tracers (RenderSpikeyMap_2x2_Distant,
// code/shadertoy.cpp:525-598): `for (step = 0; march > eps && step < max;
// ++step) { hit = o + d*total; march = sdf(hit); total += march; }` (plus a
// far clip), shading wherever the march stopped, ExpFog, a vLerp4 towards
// the fog colour with alpha, GammaAdj and ToPixel4 packing
// (code/shadertoy-util.h:62,136-146,186-190,237-241).
//
// The march, the hit point and the central-difference normal are FP32.
// Shading (Lambert, albedo, fog blend, gamma) is FP32 for CKD_SHADE_FP32 and
// packed FP16 (two channels per register, FP32 for the fog exponent and
// the final 255 scale) for CKD_SHADE_FP16.
//
// A 256-lane block renders a 16x16 tile as four 8x8 wave tiles, so the
// lanes of a wave march neighbouring rays.
#include "ckd_internal.h"
#include "ckd_pixel.h"


namespace {

typedef _Float16 h2 __attribute__((ext_vector_type(2)));

struct SdfCam {
    float o[3], f[3], r[3], u[3];
};

CKD_DEV float scene_dist(const ckd_sdf_scene &sc, float px, float py, float pz)
{
    float d = py - sc.groundY;
    for (int i = 0; i < sc.nSpheres; ++i) {
        const float dx = px - sc.sphere[i][0];
        const float dy = py - sc.sphere[i][1];
        const float dz = pz - sc.sphere[i][2];
        const float l = sqrtf(dx * dx + dy * dy + dz * dz) - sc.sphere[i][3];
        d = fminf(d, l);
    }
    return d;
}

CKD_DEV int scene_id(const ckd_sdf_scene &sc, float px, float py, float pz)
{
    float d = py - sc.groundY;
    int best = sc.nSpheres;
    for (int i = 0; i < sc.nSpheres; ++i) {
        const float dx = px - sc.sphere[i][0];
        const float dy = py - sc.sphere[i][1];
        const float dz = pz - sc.sphere[i][2];
        const float l = sqrtf(dx * dx + dy * dy + dz * dz) - sc.sphere[i][3];
        if (l < d) {
            d = l;
            best = i;
        }
    }
    return best;
}

CKD_DEV uint32_t to_chan(float c)
{
    // _mm_cvtps_epi32 (round to nearest even) of 255*c, max 0, saturate
    const float v = rintf(255.f * c);
    return v > 0.f ? (v < 255.f ? (uint32_t)v : 255u) : 0u;
}

struct MarchResult {
    float hx, hy, hz, total;
    int steps;
};

CKD_DEV MarchResult march(const ckd_sdf_scene &sc, const SdfCam &cam, unsigned W, unsigned H,
                          unsigned x, unsigned y, float &dx, float &dy, float &dz)
{
    const float fH = (float)H;
    const float u = (2.f * (float)x - (float)W) / fH;
    const float v = (fH - 2.f * (float)y) / fH;
    dx = cam.f[0] + u * cam.r[0] + v * cam.u[0];
    dy = cam.f[1] + u * cam.r[1] + v * cam.u[1];
    dz = cam.f[2] + u * cam.r[2] + v * cam.u[2];
    const float il = 1.f / sqrtf(dx * dx + dy * dy + dz * dz);
    dx *= il;
    dy *= il;
    dz *= il;
    MarchResult m;
    m.hx = cam.o[0];
    m.hy = cam.o[1];
    m.hz = cam.o[2];
    float d = 1.f, total = 0.f;
    int step;
    for (step = 0; d > sc.eps && step < sc.maxSteps && total < sc.tMax; ++step) {
        m.hx = cam.o[0] + dx * total;
        m.hy = cam.o[1] + dy * total;
        m.hz = cam.o[2] + dz * total;
        d = scene_dist(sc, m.hx, m.hy, m.hz);
        total += d;
    }
    m.total = total;
    m.steps = step;
    return m;
}

template <bool kHalf>
CKD_DEV uint32_t shade(const ckd_sdf_scene &sc, const MarchResult &m)
{
    const float h = 1e-3f;
    float nx = scene_dist(sc, m.hx + h, m.hy, m.hz) - scene_dist(sc, m.hx - h, m.hy, m.hz);
    float ny = scene_dist(sc, m.hx, m.hy + h, m.hz) - scene_dist(sc, m.hx, m.hy - h, m.hz);
    float nz = scene_dist(sc, m.hx, m.hy, m.hz + h) - scene_dist(sc, m.hx, m.hy, m.hz - h);
    const int id = scene_id(sc, m.hx, m.hy, m.hz);
    const float t3 = m.total * m.total * m.total;
    if (!kHalf) {
        const float nl = sqrtf(nx * nx + ny * ny + nz * nz);
        const float inl = nl > 0.f ? 1.f / nl : 0.f;
        float diff = (nx * sc.light[0] + ny * sc.light[1] + nz * sc.light[2]) * inl;
        diff = diff > 0.f ? diff : 0.f;
        const float lit = sc.ambient + (1.f - sc.ambient) * diff;
        const float fog = 1.f - expf(-sc.fogDensity * t3);
        uint32_t px = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float base = sc.albedo[id][c] * lit;
            const float col = base + fog * (sc.fogColor[c] - base);
            const float g = col > 0.f ? expf(sc.gamma * logf(col)) : 0.f;
            px |= to_chan(g) << (8 * c);
        }
        const float ga = fog > 0.f ? expf(sc.gamma * logf(fog)) : 0.f;
        return px | (to_chan(ga) << 24);
    } else {
        // the normalisation scale and the dot product with the light stay
        // FP32 (rsqrt of a sum of squares of small differences); albedo,
        // lighting, fog blend and gamma run on packed halves
        const float nl2 = nx * nx + ny * ny + nz * nz;
        const float inl = nl2 > 0.f ? rsqrtf(nl2) : 0.f;
        const float diffF = (nx * sc.light[0] + ny * sc.light[1] + nz * sc.light[2]) * inl;
        const h2 diff = h2((_Float16)(diffF > 0.f ? diffF : 0.f));
        const h2 amb = h2((_Float16)sc.ambient);
        const h2 lit = (h2((_Float16)1.f) - amb) * diff + amb;
        const _Float16 fogE = __builtin_elementwise_exp2((_Float16)(-sc.fogDensity * t3 * 1.44269504f));
        const h2 fog = h2((_Float16)1.f - fogE);
        const h2 albBG = {(_Float16)sc.albedo[id][0], (_Float16)sc.albedo[id][1]};
        const h2 albRA = {(_Float16)sc.albedo[id][2], (_Float16)0.f};
        const h2 fogBG = {(_Float16)sc.fogColor[0], (_Float16)sc.fogColor[1]};
        const h2 fogRA = {(_Float16)sc.fogColor[2], (_Float16)1.f};
        const h2 baseBG = albBG * lit, baseRA = albRA * lit;
        const h2 cBG = fog * (fogBG - baseBG) + baseBG;
        const h2 cRA = fog * (fogRA - baseRA) + baseRA;
        const h2 g = h2((_Float16)sc.gamma);
        const h2 tiny = h2((_Float16)6.1e-5f);
        // c^gamma = exp2(gamma*log2(c)); c == 0 stays 0
        h2 pBG = __builtin_elementwise_exp2(g * __builtin_elementwise_log2(__builtin_elementwise_max(cBG, tiny)));
        h2 pRA = __builtin_elementwise_exp2(g * __builtin_elementwise_log2(__builtin_elementwise_max(cRA, tiny)));
        const float b = cBG.x > (_Float16)0.f ? (float)pBG.x : 0.f;
        const float gg = cBG.y > (_Float16)0.f ? (float)pBG.y : 0.f;
        const float r = cRA.x > (_Float16)0.f ? (float)pRA.x : 0.f;
        const float a = cRA.y > (_Float16)0.f ? (float)pRA.y : 0.f;
        return to_chan(b) | (to_chan(gg) << 8) | (to_chan(r) << 16) | (to_chan(a) << 24);
    }
}

CKD_DEV void count_steps(unsigned long long *stepCount, int steps)
{
    if (!stepCount) return;
    unsigned long long s = (unsigned long long)steps;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(stepCount, s);
}

template <bool kHalf>
__global__ __launch_bounds__(256) void k_sdf(uint32_t *__restrict__ dst, unsigned W, unsigned H,
                                             const ckd_sdf_scene sc, const SdfCam cam,
                                             unsigned long long *__restrict__ stepCount)
{
    const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned x = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const unsigned y = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    const bool inside = x < W && y < H;
    int steps = 0;
    if (inside) {
        float dx, dy, dz;
        const MarchResult m = march(sc, cam, W, H, x, y, dx, dy, dz);
        steps = m.steps;
        dst[(size_t)y * W + x] = shade<kHalf>(sc, m);
    }
    count_steps(stepCount, steps);
}

ckd_status validate_scene(const ckd_sdf_scene *sc)
{
    CKD_REQUIRE(sc != nullptr, "sdf: null scene");
    CKD_REQUIRE(sc->nSpheres >= 0 && sc->nSpheres <= 32, "sdf: nSpheres %d outside 0..32",
                sc->nSpheres);
    CKD_REQUIRE(sc->maxSteps >= 1 && sc->maxSteps <= 1024, "sdf: maxSteps %d outside 1..1024",
                sc->maxSteps);
    return CKD_OK;
}

} // namespace

namespace ckd {

ckd_status launch_sdf(ckd_ctx *ctx, uint32_t *dst, unsigned W, unsigned H,
                      const ckd_sdf_scene &sc, const float *cam12, int shading,
                      unsigned long long *stepCount)
{
    SdfCam cam;
    for (int i = 0; i < 3; ++i) {
        cam.o[i] = cam12[i];
        cam.f[i] = cam12[3 + i];
        cam.r[i] = cam12[6 + i];
        cam.u[i] = cam12[9 + i];
    }
    dim3 grid(ckd_div_up(W, 16), ckd_div_up(H, 16));
    if (shading == CKD_SHADE_FP16)
        k_sdf<true><<<grid, 256, 0, ctx->stream>>>(dst, W, H, sc, cam, stepCount);
    else
        k_sdf<false><<<grid, 256, 0, ctx->stream>>>(dst, W, H, sc, cam, stepCount);
    return check_launch("sdf");
}

} // namespace ckd

extern "C" {

ckd_status ckd_sdf_render(ckd_ctx *ctx, uint32_t *dst, unsigned W, unsigned H,
                          const ckd_sdf_scene *scene, const float *cam, int shading,
                          unsigned long long *stepCount)
{
    CKD_REQUIRE(dst && cam, "ckd_sdf_render: null buffer");
    CKD_REQUIRE(W > 0 && H > 0 && W <= 16384 && H <= 16384, "ckd_sdf_render: bad size %ux%u", W, H);
    CKD_REQUIRE(shading == CKD_SHADE_FP32 || shading == CKD_SHADE_FP16,
                "ckd_sdf_render: unknown shading mode %d", shading);
    ckd_status s = validate_scene(scene);
    if (s != CKD_OK) return s;
    CKD_REQUIRE_CTX(ctx);
    return ckd::launch_sdf(ctx, dst, W, H, *scene, cam, shading, stepCount);
}

} // extern "C"
