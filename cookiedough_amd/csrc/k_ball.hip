// k_ball.hip -- the voxel ball (code/ball.cpp), the torus twister
// (code/torus-twister.cpp) and Polar_BlitA (code/polar.cpp:157-198).
//
// Both casters march one ray per lane across a 1024^2 height map and draw
// colored spans along a row of the render target (the ball: one ray per row
// from its start; the twister: two per row, outwards from the centre).
// Per-frame tables (height projections, lighting norms, ray directions and
// origins) are the reference's FP32 host code, computed here on the host
// in the same order and staged once per frame.  Spans stay inside their row:
// the reference writes past the row end when a projected height exceeds it
// (map heights above 1279*256/radius), and such frames are rejected.
#include "ckd_internal.h"
#include "ckd_pixel.h"
#include "ckd_blend.h"
#include "ckd_b07.h"

#include <cmath>
#include <cstring>
#include <vector>

#pragma clang fp contract(off)

namespace {

constexpr float kPI = 3.1415926535897932384626433832795f;
constexpr float kGoldenRatio = 1.61803398875f;
constexpr float kGoldenAngle = 2.39996f;
constexpr float kEpsilon = 1.1920928955078125e-07f;
constexpr unsigned kTwRay = 512;     // code/torus-twister.cpp:29
constexpr unsigned kBallMaxRay = 1024; // code/ball.cpp:56
constexpr size_t kMapPx = 1024 * 1024;

// ---- device: samplers and spans with 16-bit lanes -------------------------

struct Uv {
    unsigned u0, u1, v0, v1, fu, fv;
};

// bsamp_prepUVs (code/bilinear.h:10-31) on a 1024^2 map
CKD_DEV Uv prep(int U, int V)
{
    Uv r;
    r.u0 = (unsigned)(U >> 8) & 1023u;
    r.u1 = ((unsigned)(U >> 8) + 1u) & 1023u;
    r.v0 = ((unsigned)(V >> 8) & 1023u) << 10;
    r.v1 = (((unsigned)(V >> 8) + 1u) & 1023u) << 10;
    r.fu = (unsigned)U & 0xffu;
    r.fv = (unsigned)V & 0xffu;
    return r;
}

CKD_DEV uint32_t samp8(const uint8_t *__restrict__ m, const Uv &t)
{
    return ckd_bilerp8(m[t.u0 + t.v0], m[t.u1 + t.v0], m[t.u0 + t.v1], m[t.u1 + t.v1], t.fu, t.fv);
}

// bsamp32_16 (code/bilinear.h:58-83): four channel lanes (each <= 255)
CKD_DEV void samp16(const uint32_t *__restrict__ m, const Uv &t, uint32_t c[4])
{
    const uint32_t p = ckd_bilerp32(m[t.u0 + t.v0], m[t.u1 + t.v0], m[t.u0 + t.v1], m[t.u1 + t.v1],
                                    t.fu, t.fv);
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = (p >> (8 * k)) & 0xffu;
}

// _mm_adds_epu16
CKD_DEV uint32_t adds16(uint32_t a, uint32_t b)
{
    const uint32_t s = a + b;
    return s > 65535u ? 65535u : s;
}

// cspanISSE16 (code/cspan.h:47-78) with 16-bit lane colours A and B: the
// lanes are zero-extended to 32 bits and delta = B - A taken modulo 2^32,
// then split into the two 16-bit halves _mm_madd_epi16 multiplies
CKD_DEV void span16(uint32_t *__restrict__ dst, int incr, uint32_t length, uint32_t drawLength,
                    const uint32_t A[4], const uint32_t B[4])
{
    const uint32_t divisor = 65536u / length;
    const uint32_t pre = length - drawLength;
    const int64_t m = (int64_t)(int32_t)divisor * (int64_t)(int32_t)pre;
    const uint32_t mlo = (uint32_t)m, mhi = (uint32_t)((uint64_t)m >> 32);
    uint32_t acc[4], step[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t delta = B[k] - A[k];
        step[k] = (uint32_t)ckd_madd16(delta, divisor);
        acc[k] = (A[k] << 16) + (uint32_t)ckd_madd16(delta, (k & 1) ? mhi : mlo);
    }
    for (uint32_t i = 0; i < drawLength; ++i) {
        *dst = ckd_span_pixel(acc);
        dst += incr;
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[k] += step[k];
    }
}

// ---- one wave per ray --------------------------------------------------------
//
// The reference marches a ray step by step: step i samples the maps at
// from - (i+1)*d, projects a height H_i, and when H_i exceeds the furthest
// pixel drawn so far (D_i = max(0, H_0..H_{i-1})) draws the span
// [D_i, H_i) from the previous step's colour to its own, with length
// H_i - H_{i-1}.  Only D_i (a running maximum) and, with beams, the
// saturating beam sum (min(65535, prefix sum) for non-negative terms) carry
// from step to step, so a 64-lane wave takes 64 steps at once: every lane
// samples its own step, wave prefix scans give D_i and the beam sums, the
// previous step's height and colour come from the neighbouring lane, and
// lane 63's values carry into the next 64 steps.  Every span is the one the
// sequential loop draws, pixel for pixel.

constexpr unsigned kWave = 64;

CKD_DEV uint32_t wave_incl_max(uint32_t v, unsigned lane)
{
#pragma unroll
    for (unsigned off = 1; off < kWave; off <<= 1) {
        const uint32_t o = __shfl_up(v, off, kWave);
        if (lane >= off) v = o > v ? o : v;
    }
    return v;
}

CKD_DEV uint32_t wave_incl_sum(uint32_t v, unsigned lane)
{
#pragma unroll
    for (unsigned off = 1; off < kWave; off <<= 1) {
        const uint32_t o = __shfl_up(v, off, kWave);
        if (lane >= off) v += o;
    }
    return v;
}

// the state one chunk of 64 steps hands to the next
struct RayCarry {
    uint32_t drawn = 0;      // D: furthest pixel drawn
    uint32_t height = 0;     // the last step's height
    uint32_t colLo, colHi;   // the last step's colour, two 16-bit lanes each
};

// draws this lane's span and advances the carry; `last` is the lane of the
// chunk's last step
CKD_DEV void chunk_spans(uint32_t *__restrict__ base, int dir, unsigned lane, unsigned last,
                         bool active, uint32_t height, const uint32_t color[4], RayCarry &c)
{
    const uint32_t lo = color[0] | (color[1] << 16), hi = color[2] | (color[3] << 16);
    uint32_t pLo = __shfl_up(lo, 1, kWave), pHi = __shfl_up(hi, 1, kWave);
    uint32_t pH = __shfl_up(height, 1, kWave);
    const uint32_t incl = wave_incl_max(height, lane);
    uint32_t D = __shfl_up(incl, 1, kWave);
    if (lane == 0) {
        pLo = c.colLo;
        pHi = c.colHi;
        pH = c.height;
        D = 0;
    }
    D = D > c.drawn ? D : c.drawn;
    if (active && height > D) {
        const uint32_t A[4] = {pLo & 0xffffu, pLo >> 16, pHi & 0xffffu, pHi >> 16};
        span16(base + (ptrdiff_t)dir * (ptrdiff_t)D, dir, height - pH, height - D, A, color);
    }
    const uint32_t top = __shfl(incl, kWave - 1, kWave);
    c.drawn = top > c.drawn ? top : c.drawn;
    c.height = __shfl(height, last, kWave);
    c.colLo = __shfl(lo, last, kWave);
    c.colHi = __shfl(hi, last, kWave);
}

// ---- the twister ---------------------------------------------------------

struct TwFrame {
    int fromX, fromY; // per row, 24:8
};

// vtwister_ray (code/torus-twister.cpp:39-85) for both rays of every row:
// wave 2*row+side; side 0 walks right from the centre, side 1 left.
// blockIdx.y is the frame: its rows and its render target (W x H, frames
// back to back from `targets`)
__global__ __launch_bounds__(256) void k_twister(uint32_t *__restrict__ targets, unsigned W,
                                                 unsigned H, const TwFrame *__restrict__ rows,
                                                 const uint32_t *__restrict__ tables,
                                                 const uint8_t *__restrict__ hmap,
                                                 const uint32_t *__restrict__ cmap)
{
    const unsigned lane = threadIdx.x & (kWave - 1);
    const unsigned ray = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (ray >= 2 * H) return; // whole waves
    uint32_t *target = targets + (size_t)blockIdx.y * W * H;
    rows += (size_t)blockIdx.y * H;
    const unsigned row = ray >> 1, side = ray & 1;
    const uint32_t *hp = tables, *hn = tables + kTwRay;
    const uint32_t dX = side ? (uint32_t)-512 : 512u;
    const int dir = side ? -1 : 1;
    const int fromX = side ? (int)((uint32_t)rows[row].fromX - 512u) : rows[row].fromX;
    const int curY = rows[row].fromY;
    uint32_t *base = target + (size_t)row * W + (W >> 1) - side;
    uint32_t c0[4];
    samp16(cmap, prep(fromX, curY), c0);
    RayCarry c;
    c.colLo = c0[0] | (c0[1] << 16);
    c.colHi = c0[2] | (c0[3] << 16);
    for (unsigned i0 = 0; i0 < kTwRay; i0 += kWave) {
        const unsigned i = i0 + lane;
        const Uv t = prep((int)((uint32_t)fromX - (i + 1) * dX), curY);
        const uint32_t h = samp8(hmap, t);
        uint32_t color[4];
        samp16(cmap, t, color);
        const uint32_t diffuse = (h * hn[i] >> 8) & 0xffffu;
#pragma unroll
        for (int k = 0; k < 4; ++k) color[k] = adds16(color[k], diffuse);
        chunk_spans(base, dir, lane, kWave - 1, true, h * hp[i] >> 8, color, c);
    }
}

// ---- the ball ------------------------------------------------------------

// the per-step tables of one radius and ray length (vball_precalc)
struct BallRays {
    uint32_t heightProj[kBallMaxRay];
    int32_t norm[kBallMaxRay][3];
};

// one frame's constants; `rays` indexes its BallRays in the staged batch
struct BallFrame {
    unsigned rayLength, beamAtten, lowLight;
    float beamAlphaMin;
    int fromX, fromY;
    unsigned rays;
};

// one frame of a batched Draw: its render target, map mixes and the index
// of its tables (BallFrame, fan deltas) in the staged batch
struct BallJob {
    uint32_t *target;
    const uint8_t *hmapMix;
    const uint32_t *cmap;
    const uint32_t *aux; // the beam map mix, or the environment map
    uint32_t frame;
};

// vball_ray_beams / vball_ray_no_beams (code/ball.cpp:77-261), wave per row;
// blockIdx.y picks the job.  Without beams the pixels past the ray's last
// span are cleared (the reference clears the whole target first).
template <bool kBeams>
__global__ __launch_bounds__(256) void k_ball(const BallJob *__restrict__ jobs, unsigned W,
                                              unsigned H, const BallFrame *__restrict__ frames,
                                              const BallRays *__restrict__ rays,
                                              const int2 *__restrict__ dXY)
{
    const unsigned lane = threadIdx.x & (kWave - 1);
    const unsigned row = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (row >= H) return; // whole waves
    const BallJob &J = jobs[blockIdx.y];
    const BallFrame *__restrict__ f = frames + J.frame;
    const BallRays *__restrict__ R = rays + f->rays;
    const uint8_t *__restrict__ hmapMix = J.hmapMix;
    const uint32_t *__restrict__ cmap = J.cmap;
    const uint32_t *__restrict__ aux = J.aux;
    uint32_t *dst = J.target + (size_t)row * W;
    const uint32_t dX = (uint32_t)dXY[row].x, dY = (uint32_t)dXY[row].y;
    const uint32_t fromX = (uint32_t)f->fromX, fromY = (uint32_t)f->fromY;
    const uint32_t env0 = (1024u >> 1) << 8;
    const unsigned L = f->rayLength;
    uint32_t c0[4];
    samp16(cmap, prep((int)fromX, (int)fromY), c0);
    RayCarry c;
    c.colLo = c0[0] | (c0[1] << 16);
    c.colHi = c0[2] | (c0[3] << 16);
    uint32_t beamSum[4] = {0, 0, 0, 0}; // unsaturated; at most 1024*255
    for (unsigned i0 = 0; i0 < L; i0 += kWave) {
        const unsigned i = i0 + lane;
        const bool active = i < L;
        uint32_t height = 0, color[4] = {0, 0, 0, 0}, extra[4], bl[4] = {0, 0, 0, 0}, lit = 0;
        if (active) {
            const uint32_t n = i + 1;
            const Uv t = prep((int)(fromX - n * dX), (int)(fromY - n * dY));
            const uint32_t h = samp8(hmapMix, t);
            samp16(cmap, t, color);
            if constexpr (kBeams) {
                samp16(aux, t, extra); // beam map
                const uint32_t hi = h * (uint32_t)R->norm[i][0] >> 8;
                const uint32_t lo = h * (uint32_t)R->norm[i][1] >> 8;
                lit = (hi + (((uint32_t)(int32_t)(lo - hi) * f->lowLight) >> 8)) & 0xffffu;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t b = ((extra[k] * f->beamAtten) & 0xffffu) >> 8;
                    bl[k] = ((b * lit) & 0xffffu) >> 8;
                }
            } else {
                const uint32_t eu = env0 - n * (dX << 1), ev = env0 - n * (dY << 1);
                samp16(aux, prep((int)(eu + h), (int)(ev + h)), extra); // environment map
                const uint32_t diffuse = h * (uint32_t)R->norm[i][2] >> 8;
                const uint32_t full = (32u + diffuse) < 255u ? 32u + diffuse : 255u;
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    color[k] = adds16(adds16(color[k], ((extra[k] * full) & 0xffffu) >> 8),
                                      diffuse & 0xffffu);
            }
            height = h * R->heightProj[i] >> 8;
        }
        if constexpr (kBeams) {
            // a chunk's sums stay below 64*255, so two channels share a register
            const uint32_t s01 = wave_incl_sum(bl[0] | (bl[1] << 16), lane);
            const uint32_t s23 = wave_incl_sum(bl[2] | (bl[3] << 16), lane);
            const uint32_t part[4] = {s01 & 0xffffu, s01 >> 16, s23 & 0xffffu, s23 >> 16};
            const uint32_t t01 = __shfl(s01, kWave - 1, kWave), t23 = __shfl(s23, kWave - 1, kWave);
            const uint32_t tot[4] = {t01 & 0xffffu, t01 >> 16, t23 & 0xffffu, t23 >> 16};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t s = beamSum[k] + part[k];
                const uint32_t acc = s < 65535u ? s : 65535u;
                color[k] = adds16(adds16(color[k], acc), lit);
                beamSum[k] += tot[k];
            }
        }
        const unsigned last = (L - i0 - 1) < (kWave - 1) ? (L - i0 - 1) : (kWave - 1);
        chunk_spans(dst, 1, lane, last, active, height, color, c);
    }
    if constexpr (!kBeams) {
        for (uint32_t x = c.drawn + lane; x < W; x += kWave) dst[x] = 0u;
    }
    if constexpr (kBeams) {
        // beam extrusion (code/ball.cpp:176-204): packus of the 16-bit sums
        uint32_t beamCol = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = beamSum[k] < 65535u ? beamSum[k] : 65535u;
            beamCol |= (v > 32767u ? 0u : (v > 255u ? 255u : v)) << (8 * k);
        }
        const uint32_t lastDrawn = c.drawn;
        const uint32_t remainder = (W - 1) - lastDrawn;
        const uint32_t bR = beamCol >> 16, bG = (beamCol >> 8) & 0xff, bB = beamCol & 0xff;
        // unsigned(0.0722f*65536.f), unsigned(0.7152f*65536.f), unsigned(0.2126f*65536.f)
        const uint32_t lum = ((bR * 4731u) >> 16) + ((bG * 46871u) >> 16) + ((bB * 13932u) >> 16);
        const float fLum = (float)lum;
        const float alphaStep = 1.f / (float)(remainder - 1);
        const float amin = f->beamAlphaMin;
        // the alpha ramp accumulates curStep += alphaStep pixel by pixel; every
        // lane runs the same chain and keeps the value of its own pixel
        float curStep = 0.f;
        for (uint32_t k0 = 0; k0 < remainder; k0 += kWave) {
            float mine = 0.f;
            for (unsigned j = 0; j < kWave; ++j) {
                mine = (j == lane) ? curStep : mine;
                curStep += alphaStep;
            }
            const uint32_t k = k0 + lane;
            if (k < remainder) {
                const float q = mine * mine * (3.f - 2.f * mine);
                const float a = amin + (fLum - amin) * q;
                dst[lastDrawn + k] = beamCol | ((uint32_t)a << 24);
            }
        }
    }
}

// ---- Polar_BlitA ------------------------------------------------------------

// Fetch16 through the map (code/polar.cpp:95-106), then the mix by the
// fetched alpha: d + a*(s-d)/256 per channel (code/polar.cpp:157-174)
__global__ __launch_bounds__(256) void k_polar_blit_a(uint32_t *__restrict__ dst,
                                                      const uint32_t *__restrict__ src,
                                                      const int2 *__restrict__ map, unsigned W,
                                                      unsigned H)
{
    const unsigned x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W || y >= H) return;
    const size_t i = (size_t)y * W + x;
    const int2 uv = map[i];
    const unsigned U0 = (unsigned)(uv.x >> 8), V0 = (unsigned)(uv.y >> 8) * W;
    const uint32_t s = ckd_bilerp32(src[U0 + V0], src[U0 + 1 + V0], src[U0 + V0 + W],
                                    src[U0 + 1 + V0 + W], (uint32_t)uv.x & 0xff,
                                    (uint32_t)uv.y & 0xff);
    const uint32_t a = s >> 24, d = dst[i];
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t dk = (d >> (8 * k)) & 0xff, sk = (s >> (8 * k)) & 0xff;
        r |= ((dk * (256 - a) + sk * a) >> 8) << (8 * k);
    }
    dst[i] = r;
}

// ---- batched Draw stages ----------------------------------------------------

// the per-frame map mixes of Ball_Draw (code/ball.cpp:448-472): the height
// map mix (Mix32 of the spike map over the base shape, four bytes per
// pixel) and, with beams, the beam map mix (memset32, then BlitAdd32A of
// each beam map with a non-zero alpha)
struct BallMix {
    const uint32_t *base; // height map 1..4 as 32-bit quads
    uint32_t *hmapMix;
    uint32_t spikes;      // Mix32 alpha; 0: copy
    uint32_t *beamMix;    // null without beams
    uint32_t lanes[3];    // 0x01010101 * unsigned(alpha*255) per beam map; 0: skipped
};

__global__ __launch_bounds__(256) void k_ball_mix(const BallMix *__restrict__ mixes,
                                                  const uint32_t *__restrict__ spikeMap,
                                                  const uint32_t *__restrict__ beamMaps,
                                                  unsigned mapPx)
{
    const BallMix &M = mixes[blockIdx.y];
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mapPx) return;
    if (i < mapPx / 4) {
        const uint32_t b = M.base[i];
        M.hmapMix[i] = M.spikes ? blend_px<OP_MIX>(b, spikeMap[i], M.spikes, 0) : b;
    }
    if (M.beamMix) {
        uint32_t acc = 0;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            if (!M.lanes[m]) continue;
            const uint32_t s = beamMaps[(size_t)m * mapPx + i];
            uint32_t t = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                t |= ((((s >> (8 * k)) & 0xff) * ((M.lanes[m] >> (8 * k)) & 0xff)) >> 8) << (8 * k);
            acc = ckd_adds_u8x4(acc, t);
        }
        M.beamMix[i] = acc;
    }
}

// the end of Ball_Draw / Twister_Draw: the background, then Polar_BlitA of
// the render target (the forward map), then SoftLight32A of the halo when
// halo != 0 (code/ball.cpp:489-502, code/torus-twister.cpp:177-183)
struct ComposeJob {
    const uint32_t *target;
    uint32_t *out;
    const uint32_t *background;
    uint32_t halo;
};

__global__ __launch_bounds__(256) void k_compose(const ComposeJob *__restrict__ jobs,
                                                 const int2 *__restrict__ map,
                                                 const uint32_t *__restrict__ halo, unsigned W,
                                                 unsigned H)
{
    const ComposeJob &J = jobs[blockIdx.z];
    const unsigned x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W || y >= H) return;
    const size_t i = (size_t)y * W + x;
    const uint32_t *src = J.target;
    const int2 uv = map[i];
    const unsigned U0 = (unsigned)(uv.x >> 8), V0 = (unsigned)(uv.y >> 8) * W;
    const uint32_t s = ckd_bilerp32(src[U0 + V0], src[U0 + 1 + V0], src[U0 + V0 + W],
                                    src[U0 + 1 + V0 + W], (uint32_t)uv.x & 0xff,
                                    (uint32_t)uv.y & 0xff);
    const uint32_t a = s >> 24, d = J.background[i];
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t dk = (d >> (8 * k)) & 0xff, sk = (s >> (8 * k)) & 0xff;
        r |= ((dk * (256 - a) + sk * a) >> 8) << (8 * k);
    }
    if (J.halo) r = blend_px<OP_SOFTLIGHT_A>(r, halo[i], 0, 0);
    J.out[i] = r;
}

// dst[y*dstStride] = src[y*srcStride] for y < H: the render target's last
// column, which the beam rays leave as it was
__global__ __launch_bounds__(256) void k_copy_column(uint32_t *__restrict__ dst, size_t dstStride,
                                                     const uint32_t *__restrict__ src,
                                                     size_t srcStride, unsigned H)
{
    const unsigned y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y < H) dst[(size_t)y * dstStride] = src[(size_t)y * srcStride];
}

// ---- host: the reference's FP32 set-up -------------------------------------

// float -> int as x86 cvttss2si (NaN and out of range give INT_MIN), and
// float -> unsigned through a 64-bit cvttss2si (the low 32 bits)
int cvt_i32(float v) { return (v == v && v < 2147483648.f && v >= -2147483648.f) ? (int)v : INT32_MIN; }
uint32_t cvt_u32(float v)
{
    return (v == v && v < 9.2233720368547758e18f && v >= -9.2233720368547758e18f)
               ? (uint32_t)(int64_t)v
               : 0u;
}
int fp24(float v) { return cvt_i32(v * 256.f); }
int geti(float v) { return cvt_i32(roundf(v)); }
int clampi_(int lo, int hi, int v) { return v < lo ? lo : (v > hi ? hi : v); }
float smax(float a, float b) { return (a < b) ? b : a; }
float smin(float a, float b) { return (b < a) ? b : a; }
float clampf_(float lo, float hi, float v) { return smax(lo, smin(hi, v)); }
float box_blur_scale(float s) { return s != 0.f ? smax(1.f, smin(100.f, s)) * 0.01f : s; }

// vtwister_precalc (code/torus-twister.cpp:118-134)
void twister_tables(uint32_t *hp, uint32_t *hn)
{
    for (unsigned i = 0; i < kTwRay; ++i) {
        const float angle = kPI / (float)(kTwRay - 1) * (float)i;
        hp[i] = cvt_u32(600.f * sinf(angle));
        const float cosine = cosf(angle * 0.99f);
        hn[i] = cosine > 0.f ? cvt_u32(255.f * powf(cosine, 2.f)) : 0u;
    }
}

// vball_precalc (code/ball.cpp:263-286) for a clamped radius and ray length
void ball_rays(float radius, unsigned rayLength, BallRays &t)
{
    const float angStepSin = kPI / (float)(rayLength - 1);
    const float angStepCos = angStepSin * 0.99f;
    for (unsigned i = 0; i < rayLength; ++i) {
        t.heightProj[i] = cvt_u32(radius * sinf(angStepSin * (float)i));
        const float cosine = cosf(angStepCos * (float)i);
        if (cosine >= 0.f) {
            t.norm[i][0] = cvt_i32(255.f * powf(cosine, kGoldenRatio));
            t.norm[i][1] = cvt_i32(255.f * powf(cosine, kGoldenAngle));
            t.norm[i][2] = cvt_i32(255.f * powf(cosine, kPI));
        } else {
            t.norm[i][0] = t.norm[i][1] = t.norm[i][2] = 0;
        }
    }
}

// the per-frame set-up of vball (code/ball.cpp:290-333) and Ball_Draw's
// clamped tracks; the radius comes back for the frame's BallRays
void ball_frame(const ckd_ball_params &p, BallFrame &f, float &radius)
{
    radius = clampf_(1.f, 1920.f, p.radius);
    f.rayLength = (unsigned)clampi_(1, (int)kBallMaxRay, geti(p.rayLength));
    f.beamAtten = (unsigned)clampi_(0, 255, geti(p.beamAttenuation));
    f.beamAlphaMin = clampf_(0.f, 255.f, p.beamAlphaMin);
    f.lowLight = (unsigned)clampi_(0, 255, geti(p.lowBeams));
    const float time = p.time * p.speed;
    const float timeScale = (float)f.rayLength * (0.25f / 1024.f);
    f.fromX = fp24(1024.f * sinf(time * timeScale) + 512.f + p.rotateOffsX);
    f.fromY = fp24(1024.f * cosf(time * timeScale) + 512.f + p.rotateOffsY);
}

// the fan of ray directions over H rows: voxel::calc_fandeltas
// (code/voxel-shared.h:18-31) at angle row * 2pi/(H-1), as 24:8
void ball_fans(unsigned H, std::vector<int2> &dXY)
{
    const float delta = (2.f * kPI) / (float)(H - 1);
    dXY.resize(H);
    for (unsigned r = 0; r < H; ++r) {
        float X = cosf((float)r * delta), Y = sinf((float)r * delta);
        if (fabsf(X + Y) > kEpsilon) {
            const float length = 1.f / sqrtf(X * X + Y * Y);
            X *= length;
            Y *= length;
        }
        dXY[r] = make_int2(fp24(X), fp24(Y));
    }
}

bool finite_all(const float *p, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (!(p[i] == p[i]) || p[i] > 3.0e38f || p[i] < -3.0e38f) return false;
    return true;
}

ckd_status ensure_polar(ckd_ctx *ctx, unsigned W, unsigned H)
{
    if (ckd::polar_maps(ctx, W, H)) return CKD_OK;
    return ckd_polar_create(ctx, W, H);
}

// frames a batched Draw renders together (each with its own render target)
constexpr unsigned kGroup = 32;

// the context's scratch for batched Draws, grown on demand
ckd_status v2_scratch(ckd_ctx *ctx, size_t bytes, uint32_t *&out)
{
    if (bytes > ctx->v2Cap) {
        // queued work may still use the old buffer
        CKD_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->v2Scratch) (void)hipFree(ctx->v2Scratch);
        ctx->v2Scratch = nullptr;
        ctx->v2Cap = 0;
        if (ckd::dev_alloc(&ctx->v2Scratch, bytes) != hipSuccess) {
            ctx->v2Scratch = nullptr;
            return ckd::fail(CKD_ENOMEM, "batched Draw: cannot allocate %zu bytes of scratch", bytes);
        }
        ctx->v2Cap = bytes;
    }
    out = reinterpret_cast<uint32_t *>(ctx->v2Scratch);
    return CKD_OK;
}

// the host image of one group's tables, staged with one copy: sections at
// 256-byte offsets
struct Blob {
    std::vector<uint8_t> bytes;
    void clear() { bytes.clear(); }
    size_t add(size_t n)
    {
        const size_t o = (bytes.size() + 255) & ~(size_t)255;
        bytes.resize(o + n);
        return o;
    }
    size_t put(const void *src, size_t n)
    {
        const size_t o = add(n);
        if (n) memcpy(bytes.data() + o, src, n);
        return o;
    }
    template <typename T> T *at(size_t o) { return reinterpret_cast<T *>(bytes.data() + o); }
    ckd_status stage(ckd_ctx *ctx, uint8_t *&dev)
    {
        void *d = nullptr;
        ckd_status s = ckd::stage(ctx, bytes.data(), bytes.size(), &d);
        dev = static_cast<uint8_t *>(d);
        return s;
    }
};

} // namespace

namespace ckd {
ckd_status launch_memset32(ckd_ctx *ctx, uint32_t *dst, uint32_t value, size_t n);
} // namespace ckd

extern "C" {

ckd_status ckd_polar_blit_a(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned W,
                            unsigned H, int inverse)
{
    CKD_REQUIRE(dst && src, "ckd_polar_blit_a: null buffer");
    CKD_REQUIRE(dst != src, "ckd_polar_blit_a: dst must differ from src");
    CKD_REQUIRE_CTX(ctx);
    const ckd_ctx::PolarMaps *pm = ckd::polar_maps(ctx, W, H);
    if (!pm)
        return ckd::fail(CKD_ESTATE, "ckd_polar_blit_a: no polar maps for %ux%u (ckd_polar_create)",
                         W, H);
    dim3 grid(ckd_div_up(W, 256), H);
    k_polar_blit_a<<<grid, 256, 0, ctx->stream>>>(
        dst, src, reinterpret_cast<const int2 *>(inverse ? pm->inv : pm->map), W, H);
    return ckd::check_launch("polar_blit_a");
}

ckd_status ckd_twister_create(ckd_ctx *ctx, const uint8_t *heightMap, const uint32_t *colorMap,
                              const uint32_t *background, unsigned W, unsigned H)
{
    CKD_REQUIRE(heightMap && colorMap && background, "ckd_twister_create: null map");
    // the spans reach 597 pixels from the centre (600 * 255/256)
    CKD_REQUIRE(W >= 1200 && W <= 8192 && H >= 2 && H <= 8192 && (W % 2) == 0,
                "ckd_twister_create: %ux%u: need an even width of 1200..8192 and 2..8192 rows", W, H);
    CKD_REQUIRE_CTX(ctx);
    ckd_status s = ensure_polar(ctx, W, H);
    if (s != CKD_OK) return s;
    // each map on its own, so a create that failed part way allocates the rest
    if ((!ctx->twHeight && ckd::dev_alloc(&ctx->twHeight, kMapPx) != hipSuccess) ||
        (!ctx->twColor && ckd::dev_alloc(&ctx->twColor, kMapPx * 4) != hipSuccess) ||
        (!ctx->twTables && ckd::dev_alloc(&ctx->twTables, 2 * kTwRay * sizeof(uint32_t)) != hipSuccess))
        return ckd::fail(CKD_ENOMEM, "ckd_twister_create: cannot allocate the maps");
    if (ctx->twBackground && (ctx->twW != W || ctx->twH != H)) {
        CKD_HIP(hipFree(ctx->twBackground));
        ctx->twBackground = nullptr;
    }
    if (!ctx->twBackground && ckd::dev_alloc(&ctx->twBackground, (size_t)W * H * 4) != hipSuccess)
        return ckd::fail(CKD_ENOMEM, "ckd_twister_create: cannot allocate the background");
    uint32_t tables[2 * kTwRay];
    twister_tables(tables, tables + kTwRay);
    CKD_HIP(hipMemcpy(ctx->twHeight, heightMap, kMapPx, hipMemcpyHostToDevice));
    CKD_HIP(hipMemcpy(ctx->twColor, colorMap, kMapPx * 4, hipMemcpyHostToDevice));
    CKD_HIP(hipMemcpy(ctx->twBackground, background, (size_t)W * H * 4, hipMemcpyHostToDevice));
    CKD_HIP(hipMemcpy(ctx->twTables, tables, sizeof(tables), hipMemcpyHostToDevice));
    ctx->twW = W;
    ctx->twH = H;
    return CKD_OK;
}

ckd_status ckd_twister_draw(ckd_ctx *ctx, uint32_t *dst, uint32_t *target, unsigned W, unsigned H,
                            const ckd_twister_params *frames, unsigned nFrames)
{
    CKD_REQUIRE(dst && target && frames, "ckd_twister_draw: null buffer");
    CKD_REQUIRE(dst != target, "ckd_twister_draw: dst must differ from target");
    for (unsigned k = 0; k < nFrames; ++k)
        CKD_REQUIRE(finite_all(&frames[k].time, sizeof(ckd_twister_params) / sizeof(float)),
                    "ckd_twister_draw: frame %u: parameters must be finite", k);
    CKD_REQUIRE_CTX(ctx);
    CKD_REQUIRE(ctx->twHeight && ctx->twW == W && ctx->twH == H,
                "ckd_twister_draw: no twister maps for %ux%u (ckd_twister_create)", W, H);
    if (nFrames == 0) return CKD_OK;
    ckd_status s = ensure_polar(ctx, W, H);
    if (s != CKD_OK) return s;
    const int2 *map = reinterpret_cast<const int2 *>(ckd::polar_maps(ctx, W, H)->map);
    const size_t px = (size_t)W * H;
    const unsigned G = nFrames < kGroup ? nFrames : kGroup;
    uint32_t *tg = nullptr;
    if ((s = v2_scratch(ctx, G * px * 4, tg)) != CKD_OK) return s;
    Blob blob;
    unsigned g = 0;
    for (unsigned f0 = 0; f0 < nFrames; f0 += g) {
        g = nFrames - f0 < G ? nFrames - f0 : G;
        // Twister_Draw (code/torus-twister.cpp:164-187) of g frames, each in
        // its own render target: the staged tables, then one launch per stage
        blob.clear();
        const size_t oRows = blob.add(sizeof(TwFrame) * H * g);
        std::vector<ckd::B07Job> blurs;
        std::vector<ComposeJob> comp(g);
        for (unsigned j = 0; j < g; ++j) {
            // vtwister (code/torus-twister.cpp:89-116)
            const ckd_twister_params &p = frames[f0 + j];
            TwFrame *rows = blob.at<TwFrame>(oRows) + (size_t)j * H;
            const float fMapSizeHH = (1024.f * 0.5f) - 0.5f, fMapSizeHHH = (1024.f * 0.25f) - 0.5f;
            const float mapStepY = 1024.f / (float)(H - 1), shearStep = (2.f * kPI) / (float)(H - 1);
            for (unsigned r = 0; r < H; ++r) {
                const float shearAngle = (float)r * shearStep;
                rows[r].fromX = fp24(fMapSizeHH + fMapSizeHHH * sinf(p.time * p.shearSpeed + shearAngle));
                rows[r].fromY = fp24((float)r * mapStepY + p.time * p.speed);
            }
            if (p.blur != 0.f) {
                ckd::B07Job job;
                if ((s = ckd::b07_h_job(tg + j * px, W, H, box_blur_scale(p.blur), job)) != CKD_OK) return s;
                blurs.push_back(job);
            }
            comp[j] = ComposeJob{tg + j * px, dst + (f0 + j) * px, ctx->twBackground, 0u};
        }
        const size_t oBlur = blob.put(blurs.data(), sizeof(ckd::B07Job) * blurs.size());
        const size_t oComp = blob.put(comp.data(), sizeof(ComposeJob) * g);
        uint8_t *dev = nullptr;
        if ((s = blob.stage(ctx, dev)) != CKD_OK) return s;
        if ((s = ckd::launch_memset32(ctx, tg, 0, g * px)) != CKD_OK) return s;
        k_twister<<<dim3(ckd_div_up((size_t)2 * H, 256 / kWave), g), 256, 0, ctx->stream>>>(
            tg, W, H, reinterpret_cast<const TwFrame *>(dev + oRows), ctx->twTables, ctx->twHeight,
            ctx->twColor);
        if ((s = ckd::check_launch("twister")) != CKD_OK) return s;
        if (!blurs.empty() &&
            (s = ckd::blur07_h_inplace_jobs(ctx, blurs.data(), reinterpret_cast<const ckd::B07Job *>(dev + oBlur),
                                            (unsigned)blurs.size(), W, H, px % 4 == 0)) != CKD_OK)
            return s;
        k_compose<<<dim3(ckd_div_up(W, 256), H, g), 256, 0, ctx->stream>>>(
            reinterpret_cast<const ComposeJob *>(dev + oComp), map, nullptr, W, H);
        if ((s = ckd::check_launch("twister compose")) != CKD_OK) return s;
    }
    // the render target ends as the last frame left it
    CKD_HIP(hipMemcpyAsync(target, tg + (size_t)(g - 1) * px, px * 4, hipMemcpyDeviceToDevice,
                           ctx->stream));
    return CKD_OK;
}

ckd_status ckd_ball_create(ckd_ctx *ctx, const uint8_t *heightMaps5, const uint32_t *colorMaps2,
                           const uint32_t *beamMaps3, const uint32_t *envMap,
                           const uint32_t *backgrounds2, const uint32_t *halo, unsigned W,
                           unsigned H)
{
    CKD_REQUIRE(heightMaps5 && colorMaps2 && beamMaps3 && envMap && backgrounds2 && halo,
                "ckd_ball_create: null map");
    CKD_REQUIRE(W >= 64 && W <= 8192 && H >= 2 && H <= 8192, "ckd_ball_create: bad size %ux%u", W, H);
    CKD_REQUIRE_CTX(ctx);
    ckd_status s = ensure_polar(ctx, W, H);
    if (s != CKD_OK) return s;
    // the largest map height bounds every projected height (radius <= 1920)
    uint32_t top = 0;
    for (size_t i = 0; i < 5 * kMapPx; ++i) top = heightMaps5[i] > top ? heightMaps5[i] : top;
    if (!ctx->ballMaps) {
        // heights 5 + mix, colour 2, beam 3 + mix, env
        const size_t bytes = 6 * kMapPx + (2 + 4 + 1) * kMapPx * 4;
        if (ckd::dev_alloc(&ctx->ballMaps, bytes) != hipSuccess)
            return ckd::fail(CKD_ENOMEM, "ckd_ball_create: cannot allocate the maps");
    }
    if (ctx->ballImages && (ctx->ballW != W || ctx->ballH != H)) {
        CKD_HIP(hipFree(ctx->ballImages));
        ctx->ballImages = nullptr;
    }
    // two backgrounds and the halo, then the fan deltas of the H rays
    const size_t fanOffset = ((size_t)3 * W * H * 4 + 255) & ~(size_t)255;
    if (!ctx->ballImages &&
        ckd::dev_alloc(&ctx->ballImages, fanOffset + (size_t)H * sizeof(int2)) != hipSuccess)
        return ckd::fail(CKD_ENOMEM, "ckd_ball_create: cannot allocate the backgrounds");
    std::vector<int2> fans;
    ball_fans(H, fans);
    CKD_HIP(hipMemcpy(reinterpret_cast<uint8_t *>(ctx->ballImages) + fanOffset, fans.data(),
                      (size_t)H * sizeof(int2), hipMemcpyHostToDevice));
    uint8_t *h = ctx->ballMaps;
    uint32_t *c = reinterpret_cast<uint32_t *>(h + 6 * kMapPx);
    CKD_HIP(hipMemcpy(h, heightMaps5, 5 * kMapPx, hipMemcpyHostToDevice));
    CKD_HIP(hipMemcpy(c, colorMaps2, 2 * kMapPx * 4, hipMemcpyHostToDevice));
    CKD_HIP(hipMemcpy(c + 2 * kMapPx, beamMaps3, 3 * kMapPx * 4, hipMemcpyHostToDevice));
    CKD_HIP(hipMemcpy(c + 6 * kMapPx, envMap, kMapPx * 4, hipMemcpyHostToDevice));
    CKD_HIP(hipMemcpy(ctx->ballImages, backgrounds2, (size_t)2 * W * H * 4, hipMemcpyHostToDevice));
    CKD_HIP(hipMemcpy(ctx->ballImages + (size_t)2 * W * H, halo, (size_t)W * H * 4,
                      hipMemcpyHostToDevice));
    ctx->ballW = W;
    ctx->ballH = H;
    ctx->ballTopHeight = top;
    return CKD_OK;
}

ckd_status ckd_ball_draw(ckd_ctx *ctx, uint32_t *dst, uint32_t *target, unsigned W, unsigned H,
                         const ckd_ball_params *frames, unsigned nFrames)
{
    CKD_REQUIRE(dst && target && frames, "ckd_ball_draw: null buffer");
    CKD_REQUIRE(dst != target, "ckd_ball_draw: dst must differ from target");
    for (unsigned k = 0; k < nFrames; ++k)
        CKD_REQUIRE(finite_all(&frames[k].time, sizeof(ckd_ball_params) / sizeof(float)),
                    "ckd_ball_draw: frame %u: parameters must be finite", k);
    CKD_REQUIRE_CTX(ctx);
    CKD_REQUIRE(ctx->ballMaps && ctx->ballW == W && ctx->ballH == H,
                "ckd_ball_draw: no ball maps for %ux%u (ckd_ball_create)", W, H);
    for (unsigned k = 0; k < nFrames; ++k) {
        // every span stays in its row: height <= top*radius/256 <= W-1
        const float radius = clampf_(1.f, 1920.f, frames[k].radius);
        CKD_REQUIRE((uint64_t)ctx->ballTopHeight * (uint64_t)radius <= (uint64_t)(W - 1) * 256u,
                    "ckd_ball_draw: frame %u: radius %g with map heights up to %u reaches past a "
                    "%u-pixel row", k, (double)radius, ctx->ballTopHeight, W);
    }
    if (nFrames == 0) return CKD_OK;
    ckd_status s = ensure_polar(ctx, W, H);
    if (s != CKD_OK) return s;
    const int2 *map = reinterpret_cast<const int2 *>(ckd::polar_maps(ctx, W, H)->map);
    const size_t px = (size_t)W * H;
    const uint8_t *heights = ctx->ballMaps;
    const uint32_t *cmaps = reinterpret_cast<const uint32_t *>(heights + 6 * kMapPx);
    const uint32_t *beamMaps = cmaps + 2 * kMapPx, *env = cmaps + 6 * kMapPx;
    const uint32_t *backgrounds = ctx->ballImages, *halo = ctx->ballImages + 2 * px;
    const int2 *fans = reinterpret_cast<const int2 *>(reinterpret_cast<const uint8_t *>(ctx->ballImages) +
                                                      ((3 * px * 4 + 255) & ~(size_t)255));
    // scratch per group of G frames: render targets, height map mixes, beam
    // map mixes, the split blurs' lane states, and the carried last column
    const unsigned G = nFrames < kGroup ? nFrames : kGroup;
    const size_t tgBytes = G * px * 4, hmBytes = G * kMapPx, bmBytes = G * kMapPx * 4;
    const size_t stWords = (size_t)H * 12; // per frame: 3 words per line and channel
    uint32_t *tg = nullptr;
    if ((s = v2_scratch(ctx, tgBytes + hmBytes + bmBytes + G * stWords * 4 + (size_t)H * 4, tg)) !=
        CKD_OK)
        return s;
    uint8_t *hm = reinterpret_cast<uint8_t *>(tg) + tgBytes;
    uint32_t *bm = reinterpret_cast<uint32_t *>(hm + hmBytes);
    uint32_t *states = bm + G * kMapPx;
    uint32_t *carry = states + G * stWords;
    // the column the first frame with beams keeps: the caller's target's
    const unsigned colBlocks = ckd_div_up(H, 256);
    k_copy_column<<<colBlocks, 256, 0, ctx->stream>>>(carry, 1, target + (W - 1), W, H);
    if ((s = ckd::check_launch("ball column")) != CKD_OK) return s;
    Blob blob;
    unsigned g = 0;
    for (unsigned f0 = 0; f0 < nFrames; f0 += g) {
        g = nFrames - f0 < G ? nFrames - f0 : G;
        // Ball_Draw (code/ball.cpp:442-511) of g frames, each in its own
        // render target
        blob.clear();
        std::vector<BallFrame> fr(g);
        std::vector<std::pair<float, unsigned>> rayKeys; // (radius, ray length) per BallRays
        std::vector<BallMix> mixes(g);
        std::vector<BallJob> jobsBeams, jobsEnv;
        std::vector<ckd::B07Job> blurEnv;
        std::vector<ckd::B07Job> blurBeams; // per frame with beams, in frame order
        std::vector<int> beamsBlur(g, -1);  // index into blurBeams
        std::vector<ComposeJob> comp(g);
        std::vector<bool> beams(g);
        for (unsigned j = 0; j < g; ++j) {
            const ckd_ball_params &p = frames[f0 + j];
            beams[j] = geti(p.hasBeams) != 0;
            const unsigned base = (unsigned)clampi_(1, 4, geti(p.baseShapeIndex));
            BallMix &M = mixes[j];
            M.base = reinterpret_cast<const uint32_t *>(heights + base * kMapPx);
            M.hmapMix = reinterpret_cast<uint32_t *>(hm + j * kMapPx);
            M.spikes = (uint8_t)geti(p.spikes);
            M.beamMix = beams[j] ? bm + j * kMapPx : nullptr;
            const float a[3] = {p.beams1, p.beams2, p.beams3};
            for (int m = 0; m < 3; ++m) {
                const float sa = smax(0.f, smin(1.f, a[m])); // saturatef
                M.lanes[m] = sa > 0.f ? 0x01010101u * (uint32_t)(sa * 255.f) : 0u;
            }
            // vball (code/ball.cpp:290-333); one BallRays per distinct radius
            // and ray length of the group
            float radius = 0.f;
            ball_frame(p, fr[j], radius);
            unsigned r = 0;
            while (r < rayKeys.size() &&
                   !(rayKeys[r].first == radius && rayKeys[r].second == fr[j].rayLength))
                ++r;
            if (r == rayKeys.size()) rayKeys.emplace_back(radius, fr[j].rayLength);
            fr[j].rays = r;
            uint32_t *t = tg + j * px;
            const BallJob job{t, hm + j * kMapPx, beams[j] ? cmaps : cmaps + kMapPx,
                              beams[j] ? bm + j * kMapPx : env, j};
            (beams[j] ? jobsBeams : jobsEnv).push_back(job);
            if (p.blur != 0.f) {
                ckd::B07Job bj;
                if ((s = ckd::b07_h_job(t, W, H, box_blur_scale(p.blur), bj)) != CKD_OK) return s;
                if (beams[j]) {
                    beamsBlur[j] = (int)blurBeams.size();
                    blurBeams.push_back(bj);
                } else {
                    blurEnv.push_back(bj);
                }
            }
            comp[j] = ComposeJob{t, dst + (f0 + j) * px, backgrounds + (beams[j] ? 0 : px),
                                 beams[j] ? 1u : 0u};
        }
        const size_t oFrames = blob.put(fr.data(), sizeof(BallFrame) * g);
        const size_t oRays = blob.add(sizeof(BallRays) * rayKeys.size());
        for (size_t r = 0; r < rayKeys.size(); ++r)
            ball_rays(rayKeys[r].first, rayKeys[r].second, blob.at<BallRays>(oRays)[r]);
        const size_t oMix = blob.put(mixes.data(), sizeof(BallMix) * g);
        const size_t oJB = blob.put(jobsBeams.data(), sizeof(BallJob) * jobsBeams.size());
        const size_t oJE = blob.put(jobsEnv.data(), sizeof(BallJob) * jobsEnv.size());
        const size_t oBE = blob.put(blurEnv.data(), sizeof(ckd::B07Job) * blurEnv.size());
        const size_t oBB = blob.put(blurBeams.data(), sizeof(ckd::B07Job) * blurBeams.size());
        const size_t oComp = blob.put(comp.data(), sizeof(ComposeJob) * g);
        // the carried column runs through the frames in order.  When every
        // blur of a frame with beams spans less than the row, those blurs
        // run together up to the step that adds the last column and one
        // chain launch finishes them in frame order; else each waits for
        // its predecessor.
        bool split = true;
        for (const ckd::B07Job &b : blurBeams) split = split && b.p.span < W;
        size_t oChain = 0;
        if (split) {
            std::vector<ckd::B07Chain> chain(g);
            for (unsigned j = 0; j < g; ++j) {
                ckd::B07Chain &c = chain[j];
                c.img = tg + j * px;
                c.state = nullptr;
                c.mode = beams[j] ? 1u : 0u;
                if (beamsBlur[j] >= 0) {
                    c.mode = 2;
                    c.state = states + beamsBlur[j] * stWords;
                    c.p = blurBeams[beamsBlur[j]].p;
                }
            }
            oChain = blob.put(chain.data(), sizeof(ckd::B07Chain) * g);
        }
        uint8_t *dev = nullptr;
        if ((s = blob.stage(ctx, dev)) != CKD_OK) return s;
        const BallFrame *dFrames = reinterpret_cast<const BallFrame *>(dev + oFrames);
        const BallRays *dRays = reinterpret_cast<const BallRays *>(dev + oRays);
        k_ball_mix<<<dim3(ckd_div_up(kMapPx, 256), g), 256, 0, ctx->stream>>>(
            reinterpret_cast<const BallMix *>(dev + oMix),
            reinterpret_cast<const uint32_t *>(heights), beamMaps, (unsigned)kMapPx);
        if ((s = ckd::check_launch("ball mix")) != CKD_OK) return s;
        const unsigned rowBlocks = ckd_div_up(H, 256 / kWave);
        if (!jobsBeams.empty()) {
            k_ball<true><<<dim3(rowBlocks, (unsigned)jobsBeams.size()), 256, 0, ctx->stream>>>(
                reinterpret_cast<const BallJob *>(dev + oJB), W, H, dFrames, dRays, fans);
            if ((s = ckd::check_launch("ball")) != CKD_OK) return s;
        }
        if (!jobsEnv.empty()) {
            k_ball<false><<<dim3(rowBlocks, (unsigned)jobsEnv.size()), 256, 0, ctx->stream>>>(
                reinterpret_cast<const BallJob *>(dev + oJE), W, H, dFrames, dRays, fans);
            if ((s = ckd::check_launch("ball")) != CKD_OK) return s;
        }
        const bool vec4 = px % 4 == 0;
        // frames without beams cleared their targets: their blurs go together
        if (!blurEnv.empty() &&
            (s = ckd::blur07_h_inplace_jobs(ctx, blurEnv.data(), reinterpret_cast<const ckd::B07Job *>(dev + oBE),
                                            (unsigned)blurEnv.size(), W, H, vec4)) != CKD_OK)
            return s;
        if (split) {
            if ((s = ckd::blur07_h_inplace_jobs(ctx, blurBeams.data(), reinterpret_cast<const ckd::B07Job *>(dev + oBB),
                                                (unsigned)blurBeams.size(), W, H, vec4, states)) !=
                    CKD_OK ||
                (s = ckd::blur07_h_chain(ctx, reinterpret_cast<const ckd::B07Chain *>(dev + oChain),
                                         g, carry, W, H)) != CKD_OK)
                return s;
        }
        // a frame with beams keeps its predecessor's last column (after that
        // frame's blur), so these go in frame order
        for (unsigned j = 0; j < g && !split; ++j) {
            if (!beams[j]) continue;
            const uint32_t *src = j == 0 ? carry : tg + (j - 1) * px + (W - 1);
            k_copy_column<<<colBlocks, 256, 0, ctx->stream>>>(tg + j * px + (W - 1), W, src,
                                                             j == 0 ? 1 : W, H);
            if ((s = ckd::check_launch("ball column")) != CKD_OK) return s;
            if (beamsBlur[j] >= 0 &&
                (s = ckd::blur07_h_inplace_jobs(
                     ctx, &blurBeams[beamsBlur[j]],
                     reinterpret_cast<const ckd::B07Job *>(dev + oBB) + beamsBlur[j], 1, W, H, vec4)) != CKD_OK)
                return s;
        }
        k_compose<<<dim3(ckd_div_up(W, 256), H, g), 256, 0, ctx->stream>>>(
            reinterpret_cast<const ComposeJob *>(dev + oComp), map, halo, W, H);
        if ((s = ckd::check_launch("ball compose")) != CKD_OK) return s;
        // the next group's first frame keeps this group's last column (the
        // chain launch left it in carry)
        if (!split) {
            k_copy_column<<<colBlocks, 256, 0, ctx->stream>>>(carry, 1, tg + (g - 1) * px + (W - 1),
                                                             W, H);
            if ((s = ckd::check_launch("ball column")) != CKD_OK) return s;
        }
    }
    // the render target ends as the last frame left it
    CKD_HIP(hipMemcpyAsync(target, tg + (size_t)(g - 1) * px, px * 4, hipMemcpyDeviceToDevice,
                           ctx->stream));
    return CKD_OK;
}

} // extern "C"
