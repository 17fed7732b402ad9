/*
 * orc_sdf.c -- FP64 / FP32 CPU render of the synthetic SDF scene.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * The scene is build-defined (no counterpart in the reference) and mirrors
 * the structure of the reference's sphere tracers, RenderSpikeyMap_2x2_Distant
 * in particular (code/shadertoy.cpp:525-598):
 *   - the march loop `for (step=0; march > eps && step < max; ++step)
 *     { hit = o + d*total; march = sdf(hit); total += march; }`
 *     (code/shadertoy.cpp:566-576), plus a far clip `total < tMax`;
 *   - the pixel is shaded wherever the march stopped (no hit/miss branch);
 *   - ExpFog (code/shadertoy-util.h:237-241): fog = 1 - exp(-s*d^3);
 *   - vLerp4 of (B,G,R,0) towards (fogB,fogG,fogR,1) (code/shadertoy-util.h:62);
 *   - GammaAdj: exp(gamma*log(c)) per lane incl. alpha (code/shadertoy-util.h:186);
 *   - ToPixel4: round-to-nearest of 255*c, max 0, unsigned saturation
 *     (code/shadertoy-util.h:136-146).
 * Scene: N spheres + ground plane y = groundY; Lambert with ambient term;
 * normals by central differences (6 SDF evaluations, h = 1e-3).
 */
#include "ckd_oracle.h"

#include <math.h>
#include <stdint.h>

#define SDF_NORMAL_H 1e-3

static inline unsigned orc_max_chan_diff(unsigned cur, uint32_t a, uint32_t b)
{
    for (int c = 0; c < 4; ++c) {
        const int x = (int)((a >> (8 * c)) & 0xff), y = (int)((b >> (8 * c)) & 0xff);
        const unsigned e = (unsigned)(x > y ? x - y : y - x);
        if (e > cur) cur = e;
    }
    return cur;
}

/* round-to-nearest-even like _mm_cvtps_epi32 under the default MXCSR */
static inline uint32_t to_chan(double c)
{
    double v = 255.0 * c;
    if (!(v == v)) return 0;
    double r = nearbyint(v);
    if (r < 0) return 0;
    if (r > 255) return 255;
    return (uint32_t)r;
}

/* The march's stopping rule is perturbed to measure how much a pixel's
 * colour depends on it: the threshold scaled by 1 +- SDF_COND_EPS_REL, and
 * the step budget one step shorter or longer.  Each perturbed march is a
 * prefix or a continuation of the actual one, so it costs a few extra steps
 * and one more shading where the stop moves. */
#define SDF_COND_EPS_REL 0.01

#define DEFINE_SDF_RENDER(NAME, REAL, SQRT, EXP, LOG)                              \
static inline REAL NAME##_scene(const orc_sdf_scene *sc, REAL px, REAL py, REAL pz, \
                                int *id)                                           \
{                                                                                  \
    REAL d = py - (REAL)sc->groundY;                                               \
    int best = sc->nSpheres;                                                       \
    for (int i = 0; i < sc->nSpheres; ++i) {                                       \
        const REAL dx = px - (REAL)sc->sphere[i][0];                               \
        const REAL dy = py - (REAL)sc->sphere[i][1];                               \
        const REAL dz = pz - (REAL)sc->sphere[i][2];                               \
        const REAL l = SQRT(dx * dx + dy * dy + dz * dz) - (REAL)sc->sphere[i][3]; \
        if (l < d) {                                                               \
            d = l;                                                                 \
            best = i;                                                              \
        }                                                                          \
    }                                                                              \
    if (id) *id = best;                                                            \
    return d;                                                                      \
}                                                                                  \
                                                                                   \
/* march state: the last point sampled, the distance there, the distance  */     \
/* travelled (including that last distance) and the steps taken            */     \
typedef struct { REAL hx, hy, hz, march, total; int step; } NAME##_state;         \
                                                                                   \
/* one step of the reference tracers' loop (code/shadertoy.cpp:566-576)    */     \
static inline void NAME##_step(const orc_sdf_scene *sc, const REAL *cam,          \
                               const REAL *d, NAME##_state *s)                    \
{                                                                                  \
    s->hx = cam[0] + d[0] * s->total;                                              \
    s->hy = cam[1] + d[1] * s->total;                                              \
    s->hz = cam[2] + d[2] * s->total;                                              \
    s->march = NAME##_scene(sc, s->hx, s->hy, s->hz, 0);                           \
    s->total += s->march;                                                          \
    ++s->step;                                                                     \
}                                                                                  \
                                                                                   \
static uint32_t NAME##_shade(const orc_sdf_scene *sc, const NAME##_state *s)       \
{                                                                                  \
    const REAL hx = s->hx, hy = s->hy, hz = s->hz, total = s->total;               \
    const REAL h = (REAL)SDF_NORMAL_H;                                             \
    REAL nx = NAME##_scene(sc, hx + h, hy, hz, 0) - NAME##_scene(sc, hx - h, hy, hz, 0); \
    REAL ny = NAME##_scene(sc, hx, hy + h, hz, 0) - NAME##_scene(sc, hx, hy - h, hz, 0); \
    REAL nz = NAME##_scene(sc, hx, hy, hz + h, 0) - NAME##_scene(sc, hx, hy, hz - h, 0); \
    const REAL nl = SQRT(nx * nx + ny * ny + nz * nz);                             \
    const REAL inl = nl > (REAL)0 ? (REAL)1 / nl : (REAL)0;                         \
    nx *= inl; ny *= inl; nz *= inl;                                               \
    int id;                                                                        \
    (void)NAME##_scene(sc, hx, hy, hz, &id);                                       \
    REAL diff = nx * (REAL)sc->light[0] + ny * (REAL)sc->light[1] + nz * (REAL)sc->light[2]; \
    if (!(diff > (REAL)0)) diff = 0;                                               \
    const REAL lit = (REAL)sc->ambient + ((REAL)1 - (REAL)sc->ambient) * diff;     \
    const REAL fog = (REAL)1 - EXP(-(REAL)sc->fogDensity * total * total * total); \
    REAL c4[4];                                                                    \
    for (int c = 0; c < 3; ++c) {                                                  \
        const REAL base = (REAL)sc->albedo[id][c] * lit;                           \
        c4[c] = base + fog * ((REAL)sc->fogColor[c] - base);                       \
    }                                                                              \
    c4[3] = fog;                                                                   \
    uint32_t px = 0;                                                               \
    for (int c = 0; c < 4; ++c) {                                                  \
        const REAL g = c4[c] > (REAL)0 ? EXP((REAL)sc->gamma * LOG(c4[c])) : (REAL)0; \
        px |= to_chan((double)g) << (8 * c);                                       \
    }                                                                              \
    return px;                                                                     \
}                                                                                  \
                                                                                   \
static uint32_t NAME##_pixel(const orc_sdf_scene *sc, const REAL cam[12],           \
                             unsigned W, unsigned H, unsigned x, unsigned y,       \
                             uint16_t *stepsOut, uint8_t *condOut)                 \
{                                                                                  \
    const REAL u = ((REAL)2 * (REAL)x - (REAL)W) / (REAL)H;                        \
    const REAL v = ((REAL)H - (REAL)2 * (REAL)y) / (REAL)H;                        \
    REAL d[3];                                                                     \
    d[0] = cam[3] + u * cam[6] + v * cam[9];                                       \
    d[1] = cam[4] + u * cam[7] + v * cam[10];                                      \
    d[2] = cam[5] + u * cam[8] + v * cam[11];                                      \
    const REAL il = (REAL)1 / SQRT(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);        \
    d[0] *= il; d[1] *= il; d[2] *= il;                                            \
    const REAL eps = (REAL)sc->eps, tMax = (REAL)sc->tMax;                         \
    const REAL epsHi = eps * (REAL)(1 + SDF_COND_EPS_REL);                         \
    const REAL epsLo = eps * (REAL)(1 - SDF_COND_EPS_REL);                         \
    NAME##_state s = {cam[0], cam[1], cam[2], (REAL)1, (REAL)0, 0};                \
    NAME##_state sHi = s, sPrev = s;                                               \
    int haveHi = 0;                                                                \
    while (s.march > eps && s.step < sc->maxSteps && s.total < tMax) {             \
        sPrev = s;                                                                 \
        NAME##_step(sc, cam, d, &s);                                               \
        if (!haveHi && s.march <= epsHi) {                                         \
            sHi = s;                                                               \
            haveHi = 1;                                                            \
        }                                                                          \
    }                                                                              \
    if (stepsOut) *stepsOut = (uint16_t)s.step;                                    \
    const uint32_t px = NAME##_shade(sc, &s);                                      \
    if (condOut) {                                                                 \
        unsigned cond = 0;                                                         \
        /* a threshold 1% higher stops at the first distance <= epsHi */          \
        if (haveHi && sHi.step < s.step)                                           \
            cond = orc_max_chan_diff(cond, NAME##_shade(sc, &sHi), px);            \
        /* a threshold 1% lower carries on while the distance > epsLo */           \
        if (s.march <= eps && s.march > epsLo) {                                   \
            NAME##_state sLo = s;                                                  \
            while (sLo.march > epsLo && sLo.step < sc->maxSteps && sLo.total < tMax) \
                NAME##_step(sc, cam, d, &sLo);                                     \
            cond = orc_max_chan_diff(cond, NAME##_shade(sc, &sLo), px);            \
        }                                                                          \
        /* a march that used every step: one step fewer, one step more */          \
        if (s.march > eps && s.step >= sc->maxSteps && s.total < tMax) {           \
            if (sPrev.step > 0)                                                    \
                cond = orc_max_chan_diff(cond, NAME##_shade(sc, &sPrev), px);      \
            NAME##_state sMore = s;                                                \
            NAME##_step(sc, cam, d, &sMore);                                       \
            cond = orc_max_chan_diff(cond, NAME##_shade(sc, &sMore), px);          \
        }                                                                          \
        *condOut = (uint8_t)cond;                                                  \
    }                                                                              \
    return px;                                                                     \
}

DEFINE_SDF_RENDER(sdf64, double, sqrt, exp, log)
DEFINE_SDF_RENDER(sdf32, float, sqrtf, expf, logf)

/* camera basis for frame time t: origin, forward*focal, right, up
 * (an orbit around the scene centre looking at (0,-1,0)) */
static void sdf_camera(const orc_sdf_scene *sc, float t, double cam[12])
{
    const double a = (double)sc->camSpeed * (double)t;
    const double ox = (double)sc->camRadius * sin(a);
    const double oy = (double)sc->camHeight;
    const double oz = -(double)sc->camRadius * cos(a);
    double fx = 0.0 - ox, fy = -1.0 - oy, fz = 0.0 - oz;
    double l = sqrt(fx * fx + fy * fy + fz * fz);
    fx /= l; fy /= l; fz /= l;
    /* right = normalize(cross(up, f)) with up = (0,1,0) */
    double rx = fz, ry = 0.0, rz = -fx;
    l = sqrt(rx * rx + rz * rz);
    rx /= l; rz /= l;
    /* camUp = cross(f, right) */
    const double ux = fy * rz - fz * ry;
    const double uy = fz * rx - fx * rz;
    const double uz = fx * ry - fy * rx;
    cam[0] = ox; cam[1] = oy; cam[2] = oz;
    cam[3] = fx * sc->focal; cam[4] = fy * sc->focal; cam[5] = fz * sc->focal;
    cam[6] = rx; cam[7] = ry; cam[8] = rz;
    cam[9] = ux; cam[10] = uy; cam[11] = uz;
}

void orc_sdf_camera(const orc_sdf_scene *sc, float t, float camOut[12])
{
    double cam[12];
    sdf_camera(sc, t, cam);
    for (int i = 0; i < 12; ++i) camOut[i] = (float)cam[i];
}

/* rows rowStart, rowStart+rowStep, ... (compacted into dst).  steps and
 * cond (optional) receive per pixel the march's step count and its
 * conditioning: the largest per-channel change of the pixel when the
 * stopping threshold is scaled by 1 +- 1% or the step budget moves by one.
 * Where that exceeds the error bar, FP32 rounding of the distances (of
 * order 1e-6 against eps = 1e-3) decides the colour, not the scene: a ray
 * passing a surface at about eps, or still moving when its steps run out. */
void orc_sdf_rows_ex(uint32_t *dst, unsigned W, unsigned H, const orc_sdf_scene *sc,
                     float t, int prec, unsigned rowStart, unsigned rowStep,
                     uint16_t *steps, uint8_t *cond, int nthreads)
{
    /* both precisions start from the float camera the product receives */
    float camf[12];
    orc_sdf_camera(sc, t, camf);
    double cam64[12];
    float cam32[12];
    for (int i = 0; i < 12; ++i) {
        cam64[i] = camf[i];
        cam32[i] = camf[i];
    }
    if (rowStep == 0) rowStep = 1;
    const unsigned nRows = rowStart < H ? (H - rowStart + rowStep - 1) / rowStep : 0;
    #pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
    for (int r = 0; r < (int)nRows; ++r) {
        const unsigned y = rowStart + (unsigned)r * rowStep;
        for (unsigned x = 0; x < W; ++x) {
            const size_t o = (size_t)r * W + x;
            uint16_t *so = steps ? steps + o : 0;
            uint8_t *co = cond ? cond + o : 0;
            dst[o] = prec == 64 ? sdf64_pixel(sc, cam64, W, H, x, y, so, co)
                                : sdf32_pixel(sc, cam32, W, H, x, y, so, co);
        }
    }
}

void orc_sdf_rows(uint32_t *dst, unsigned W, unsigned H, const orc_sdf_scene *sc,
                  float t, int prec, unsigned rowStart, unsigned rowStep,
                  uint16_t *steps, int nthreads)
{
    orc_sdf_rows_ex(dst, W, H, sc, t, prec, rowStart, rowStep, steps, 0, nthreads);
}

void orc_sdf_frame(uint32_t *dst, unsigned W, unsigned H,
                   const orc_sdf_scene *scene, float t, int prec,
                   uint16_t *steps, int nthreads)
{
    orc_sdf_rows(dst, W, H, scene, t, prec, 0, 1, steps, nthreads);
}
