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
#include "ckd_jobs.h"

#include <algorithm>
#include <cmath>


namespace {

// output pixels in the global address space
typedef __attribute__((address_space(1))) uint32_t global_px;

typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

// v_sqrt_f32 (1 ulp) for the distance field; the default sqrtf is the
// correctly rounded multi-instruction sequence
CKD_DEV float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

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
        const float l = fsqrt(dx * dx + dy * dy + dz * dz) - sc.sphere[i][3];
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
        const float l = fsqrt(dx * dx + dy * dy + dz * dz) - sc.sphere[i][3];
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
    const float il = __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz);
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
CKD_DEV uint32_t shade_lit(const ckd_sdf_scene &sc, const MarchResult &m, float nx, float ny,
                           float nz, int id)
{
    const float t3 = m.total * m.total * m.total;
    if (!kHalf) {
        const float nl = fsqrt(nx * nx + ny * ny + nz * nz);
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

template <bool kHalf>
CKD_DEV uint32_t shade(const ckd_sdf_scene &sc, const MarchResult &m)
{
    const float h = 1e-3f;
    float nx = scene_dist(sc, m.hx + h, m.hy, m.hz) - scene_dist(sc, m.hx - h, m.hy, m.hz);
    float ny = scene_dist(sc, m.hx, m.hy + h, m.hz) - scene_dist(sc, m.hx, m.hy - h, m.hz);
    float nz = scene_dist(sc, m.hx, m.hy, m.hz + h) - scene_dist(sc, m.hx, m.hy, m.hz - h);
    const int id = scene_id(sc, m.hx, m.hy, m.hz);
    return shade_lit<kHalf>(sc, m, nx, ny, nz, id);
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

// ---- persistent batched tracer ------------------------------------------
//
// One launch renders every SDF request of a batch.  Each wave pulls 32x32
// pixel chunks from a global queue (one atomic per chunk).  Inside a chunk
// the wave keeps its 64 lanes busy: lanes whose ray has finished are
// refilled with the chunk's next pixels (ballot + mbcnt ranks), so a
// long-marching ray no longer holds 63 idle lanes.  Finished rays are
// pushed to a per-wave LDS queue and shaded 64 at a time, so shading also
// runs with full waves.
//
// The march evaluates |o + t*d - c|^2 about the ray's closest approach to
// each centre: with m = d.(o-c) and w = (o-c) - m*d (the centre's offset
// perpendicular to the ray, |w| about the radius near a silhouette),
// |o + t*d - c|^2 = (t + m)^2 + |w|^2 for a unit d.  Per sphere and step
// that is an add, an FMA, a square root, a subtract and a min.  Both
// per-ray constants carry absolute errors of order 1e-6, as the direct
// form |p - c|^2 does; the expanded t^2 + t*B + |o-c|^2 cancels terms of
// several hundred and is off by up to 1e-4 near a surface, which decides
// hit or miss for rays that pass within that much of the 1e-3 threshold
// and moves rays that run out of steps.

struct ShadeEntry {
    uint32_t job, xy; // xy = x | y << 16 (W <= 4096, H <= 2160)
    float tHit, total;
    // the ray's direction as its set-up computed it: shading reads it rather
    // than computing it again from the job's camera
    float dx, dy, dz, pad;
};

constexpr int kQueueCap = 128;
constexpr int kWavesPerBlock = 4;

CKD_DEV uint32_t lane_id() { return __lane_id(); }

// the lane's bit of the wave mask m ? a : b: one v_cndmask that reads m from
// SGPRs (a select on a bool the compiler rebuilds from the mask costs it
// three compares again)
CKD_DEV float sel_lane(uint64_t m, float a, float b)
{
    float r;
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(b), "v"(a), "s"(m));
    return r;
}
// a + the lane's bit of m: one v_addc with m as the carry in
CKD_DEV int add_lane(uint64_t m, int a)
{
    int r;
    uint64_t carry;
    asm("v_addc_co_u32_e64 %0, %1, %2, 0, %3" : "=v"(r), "=s"(carry) : "v"(a), "s"(m));
    return r;
}

// v_cmp into a wave mask (LLVM predicates: 2 ordered >, 4 ordered <, 40
// signed <); a __ballot of a bool made of several compares costs a
// v_cndmask and a v_cmp more
constexpr int kCmpGt = 2, kCmpLt = 4, kCmpSlt = 40;

// a zero the compiler cannot see through: indexing LDS tables with it keeps
// their loads inside the loop instead of hoisting every table entry into
// registers for the whole kernel
CKD_DEV uint32_t opaque_zero()
{
    uint32_t z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    return z;
}

// a wave-uniform float kept in a scalar register
CKD_DEV float uniform_f(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// a / b rounded to nearest, from rb = 1 / b rounded to nearest: Markstein's
// correction of the quotient estimate a * rb (within an ulp) by the residual
// a - b * (a * rb), which the FMA gives exactly.  For the integers |a| <=
// 16384 and 1 <= b <= 16384 it equals the IEEE quotient bit for bit (checked
// over every pair); here |a| <= 4096 and b <= 2160.
CKD_DEV float div_rn(float a, float b, float rb)
{
    const float q = a * rb;
    return fmaf(fmaf(-b, q, a), rb, q);
}

// direction of pixel (x, y) through the camera basis; invH = 1 / H rounded
// to nearest
CKD_DEV void ray_dir(const float *cam, uint32_t W, uint32_t H, float invH, uint32_t x, uint32_t y,
                     float &dx, float &dy, float &dz)
{
    const float fH = (float)H;
    const float u = div_rn(2.f * (float)x - (float)W, fH, invH);
    const float v = div_rn(fH - 2.f * (float)y, fH, invH);
    dx = cam[3] + u * cam[6] + v * cam[9];
    dy = cam[4] + u * cam[7] + v * cam[10];
    dz = cam[5] + u * cam[8] + v * cam[11];
    const float il = __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz);
    dx *= il;
    dy *= il;
    dz *= il;
}

// shading of one finished ray: hit point, central-difference normal with
// the six offsets taken from one |p-c|^2 per sphere
// (|p +- h e_k - c|^2 = q + h^2 +- 2 h (p-c)_k), argmin id, Lambert, fog,
// gamma, pack
// Where the spheres' bounding sphere (centre C, radius R) is farther from
// the hit point p than the ground by more than 2h (an offset of h moves a
// sphere's distance by at most h) plus the rounding margin, every sphere
// term of the six offsets and of the argmin exceeds the ground's: the
// minima are the ground's values and the id is the ground's, exactly as the
// loop over the spheres gives them.  The loop runs only when some lane of
// the wave needs it.
template <bool kHalf, int kMaxSph>
CKD_DEV uint32_t shade_hit(const ckd_sdf_scene &sc, const float *cam, float dx, float dy, float dz,
                           float tHit, float total, const f4 &bound, uint32_t &evaluated)
{
    const float hx = cam[0] + dx * tHit, hy = cam[1] + dy * tHit, hz = cam[2] + dz * tHit;
    const float h = 1e-3f, hh = h * h, h2 = 2.f * h;
    const float pl = hy - sc.groundY;
    float mxp = pl, mxm = pl, myp = pl + h, mym = pl - h, mzp = pl, mzm = pl;
    float best = pl;
    int id = sc.nSpheres;
    const float bx = hx - bound.x, by = hy - bound.y, bz = hz - bound.z;
    const float db = fsqrt(bx * bx + by * by + bz * bz) - bound.w;
    const bool near = db < fmaf(fabsf(pl), 1e-3f, pl + (2.f * h + 1e-3f));
    const bool loop = __ballot(near) != 0;
    evaluated += loop ? 1u : 0u;
    if (loop)
#pragma unroll 4
    for (int i = 0; i < kMaxSph; ++i) {
        if (i < sc.nSpheres) {
            const float ex = hx - sc.sphere[i][0], ey = hy - sc.sphere[i][1], ez = hz - sc.sphere[i][2];
            const float r = sc.sphere[i][3];
            const float q = ex * ex + ey * ey + ez * ez;
            const float qh = q + hh;
            const float c = fsqrt(q) - r;
            if (c < best) {
                best = c;
                id = i;
            }
            mxp = fminf(mxp, fsqrt(qh + h2 * ex) - r);
            mxm = fminf(mxm, fsqrt(qh - h2 * ex) - r);
            myp = fminf(myp, fsqrt(qh + h2 * ey) - r);
            mym = fminf(mym, fsqrt(qh - h2 * ey) - r);
            mzp = fminf(mzp, fsqrt(qh + h2 * ez) - r);
            mzm = fminf(mzm, fsqrt(qh - h2 * ez) - r);
        }
    }
    const float nx = mxp - mxm, ny = myp - mym, nz = mzp - mzm;
    MarchResult m;
    m.hx = hx;
    m.hy = hy;
    m.hz = hz;
    m.total = total;
    m.steps = 0;
    return shade_lit<kHalf>(sc, m, nx, ny, nz, id);
}

// lanes that must be idle before a wave refills them (checked every
// kStepsPerRefill march steps)
constexpr int kRefillMinIdle = 32;

template <bool kHalf, int kMaxSph, int kStepsPerRefill>
__global__ __launch_bounds__(256, kMaxSph > 8 ? 4 : 5) void k_sdf_persistent(uint32_t *__restrict__ queueCounter,
                                                        uint32_t totalChunks,
                                                        const SdfJob *__restrict__ jobs, int nJobs,
                                                        const ckd_sdf_scene sc,
                                                        unsigned long long *__restrict__ stepCount,
                                                        uint32_t kChunkSide, f4 bound,
                                                        unsigned long long *__restrict__ work)
{
    __shared__ ShadeEntry queue[kWavesPerBlock][kQueueCap];
    // the radii, read every step with broadcast LDS loads instead of
    // occupying kMaxSph scalar registers
    __shared__ f4 sphR[kWavesPerBlock][kMaxSph / 4]; // {r_4k .. r_4k+3}
    // o - c_i of the chunk's camera, read when a lane starts a ray
    __shared__ f4 sphOC[kWavesPerBlock][kMaxSph];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    ShadeEntry *q = queue[wave];
    int qn = 0; // entries waiting in this wave's queue (uniform)

    // chunk state (uniform)
    int job = -1;
    uint32_t next = 0, end = 0, cx0 = 0, cy0 = 0, cw = 1, cwM = 1u << 20;
    float invH = 1.f;
    float cam[12];
    uint32_t W = 0, H = 0;

    // ray state (per lane)
    bool active = false;
    uint32_t xy = 0; // the ray's pixel, x | y << 16
    // The ray's last march step evaluated the field at tHit and found d; its
    // distance total is t = tHit + d, which is how the step formed it, so t
    // is recomputed from the two instead of kept.  A new ray has tHit = -1
    // and d = 1, so that t = 0.
    float d = 1.f, tHit = -1.f, dx = 0.f, dy = 0.f, dz = 0.f;
    int steps = 0;
    float M[kMaxSph], E2[kMaxSph]; // d.(o-c_i) and |w_i|^2 of the lane's ray
    float Mb = 0.f, E2b = 0.f;     // the same for the centre of the spheres' bounding sphere
    float bx = 0.f, by = 0.f, bz = 0.f; // o - C of the chunk's camera (uniform)
    unsigned long long mySteps = 0;
    // lane march steps of the wave that evaluated the spheres (uniform), and
    // shaded pixels of this lane that did
    unsigned long long evalSteps = 0;
    uint32_t evalShades = 0;

    while (true) {
        // refill finished lanes from the current chunk, once half the wave
        // is idle: a refill runs the new rays' set-up for the whole wave, so
        // it pays to refill many lanes at once (a wave whose lanes are all
        // idle always refills)
        const uint64_t idle = __ballot(!active);
        if (idle && next < end && __popcll(idle) >= kRefillMinIdle) {
            const uint32_t rank = __popcll(idle & ((1ull << lane) - 1));
            if (!active) {
                const uint32_t k = next + rank;
                if (k < end) {
                    // k / cw as (k * ceil(2^20 / cw)) >> 20, exact for k
                    // < 4096 and cw <= 64
                    const uint32_t row = __umul24(k, cwM) >> 20;
                    const uint32_t x = cx0 + k - __umul24(row, cw), y = cy0 + row;
                    xy = x | (y << 16);
                    ray_dir(cam, W, H, invH, x, y, dx, dy, dz);
                    const uint32_t z = opaque_zero();
#pragma unroll
                    for (int i = 0; i < kMaxSph; ++i) {
                        const f4 oc = sphOC[wave][i + z];
                        const float m = dx * oc.x + dy * oc.y + dz * oc.z;
                        const float wx = fmaf(-m, dx, oc.x);
                        const float wy = fmaf(-m, dy, oc.y);
                        const float wz = fmaf(-m, dz, oc.z);
                        M[i] = m;
                        E2[i] = wx * wx + wy * wy + wz * wz;
                    }
                    {
                        const float m = dx * bx + dy * by + dz * bz;
                        const float wx = fmaf(-m, dx, bx);
                        const float wy = fmaf(-m, dy, by);
                        const float wz = fmaf(-m, dz, bz);
                        Mb = m;
                        E2b = wx * wx + wy * wy + wz * wz;
                    }
                    d = 1.f;
                    tHit = -1.f;
                    steps = 0;
                    active = true;
                }
            }
            next = min(end, next + (uint32_t)__popcll(idle));
        }
        if (!__any(active)) {
            // chunk drained: take the next one
            uint32_t c = 0;
            if (lane == 0) c = atomicAdd(queueCounter, 1u);
            c = __shfl(c, 0, 64);
            if (c >= totalChunks) break;
            job = ckd_find_job(jobs, nJobs, c, &SdfJob::chunkBase);
            const SdfJob &J = jobs[job];
            W = J.W;
            H = J.H;
            invH = uniform_f(J.invH);
#pragma unroll
            for (int k = 0; k < 12; ++k) cam[k] = uniform_f(J.cam[k]);
            const uint32_t local = c - J.chunkBase;
            cx0 = (local % J.chunksX) * kChunkSide;
            cy0 = (local / J.chunksX) * kChunkSide;
            cw = min((uint32_t)kChunkSide, W - cx0);
            cwM = ((1u << 20) + cw - 1) / cw;
            const uint32_t chh = min((uint32_t)kChunkSide, H - cy0);
            next = 0;
            end = cw * chh;
            bx = cam[0] - bound.x;
            by = cam[1] - bound.y;
            bz = cam[2] - bound.z;
            // lane i stages sphere i; spheres past nSpheres become a neutral
            // sphere whose distance is never the minimum (a centre 1e15 away
            // from the camera, radius 0)
            if (lane < (uint32_t)kMaxSph) {
                float r = 0.f, ox = 1e15f, oy = 0.f, oz = 0.f;
                if ((int)lane < sc.nSpheres) {
                    ox = cam[0] - sc.sphere[lane][0];
                    oy = cam[1] - sc.sphere[lane][1];
                    oz = cam[2] - sc.sphere[lane][2];
                    r = sc.sphere[lane][3];
                }
                reinterpret_cast<float *>(&sphR[wave][lane >> 2])[lane & 3] = r;
                sphOC[wave][lane] = f4{ox, oy, oz, 0.f};
            }
            __builtin_amdgcn_wave_barrier();
            continue;
        }
        // march a few steps; the lanes' conditions are wave masks in SGPRs
        const float oyg = cam[1] - sc.groundY;
        const float Rm = bound.w + 1e-3f;
        const uint64_t activeM = __ballot(active);
#pragma unroll
        for (int k = 0; k < kStepsPerRefill; ++k) {
            const float t = tHit + d;
            const uint64_t marchM = activeM & __builtin_amdgcn_fcmpf(d, sc.eps, kCmpGt) &
                                    __builtin_amdgcn_sicmp(steps, sc.maxSteps, kCmpSlt) &
                                    __builtin_amdgcn_fcmpf(t, sc.tMax, kCmpLt);
            float dist = fmaf(t, dy, oyg);
            // Every sphere lies inside the bounding sphere (centre C, radius
            // R), so each sphere's distance is at least |p - C| - R.  Where
            // that bound exceeds the ground's distance by more than the
            // rounding of either (1e-3 + 1e-3 |ground|, against errors of
            // order 1e-5), no sphere can be the minimum: the ground's
            // distance is the step, exactly as the full minimum gives it.
            // The skip test is |p - C|^2 >= y^2 with y = ground + R +
            // margin, without a square root: for y > 0 that is |p - C| >= y;
            // for y <= 0 the test with the root skips too, and this one can
            // only evaluate the spheres where it need not, which changes
            // nothing.  The spheres are evaluated only when some marching
            // lane needs them.
            const float ub = t + Mb;
            const float yb = fmaf(fabsf(dist), 1e-3f, dist + Rm);
            const uint64_t nearM = __builtin_amdgcn_fcmpf(fmaf(ub, ub, E2b), yb * yb, kCmpLt);
            if (marchM & nearM) {
                evalSteps += (uint32_t)__popcll(marchM);
                const uint32_t z = opaque_zero();
#pragma unroll
                for (int i = 0; i < kMaxSph / 4; ++i) {
                    const f4 r = sphR[wave][i + z];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const float u = t + M[4 * i + j];
                        dist = fminf(dist, fsqrt(fmaf(u, u, E2[4 * i + j])) - r[j]);
                    }
                }
            }
            tHit = sel_lane(marchM, t, tHit);
            d = sel_lane(marchM, dist, d);
            steps = add_lane(marchM, steps);
        }
        // retire finished rays into the shade queue
        const float t = tHit + d;
        const bool fin = active && !(d > sc.eps && steps < sc.maxSteps && t < sc.tMax);
        const uint64_t finMask = __ballot(fin);
        if (finMask) {
            if (fin) {
                const uint32_t slot = qn + __popcll(finMask & ((1ull << lane) - 1));
                // a ray that took no step shades at its origin
                q[slot] = ShadeEntry{(uint32_t)job, xy, steps ? tHit : 0.f, t, dx, dy, dz, 0.f};
                mySteps += (unsigned long long)steps;
                active = false;
            }
            qn += __popcll(finMask);
            __builtin_amdgcn_wave_barrier();
            if (qn >= 64) {
                // shade 64 queued hits with the whole wave
                const ShadeEntry e = q[qn - 64 + lane];
                const SdfJob &J = jobs[e.job];
                const uint32_t x = e.xy & 0xffffu, y = e.xy >> 16;
                const float ex = e.dx, ey = e.dy, ez = e.dz;
                // a global store (the job's pointer would make a flat one,
                // whose completion the next LDS reads would wait for)
                ((global_px *)J.dst)[__umul24(y, J.W) + x] = shade_hit<kHalf, kMaxSph>(sc, J.cam, ex, ey, ez, e.tHit, e.total, bound, evalShades);
                qn -= 64;
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    // leftover hits
    __builtin_amdgcn_wave_barrier();
    if ((int)lane < qn) {
        const ShadeEntry e = q[lane];
        const SdfJob &J = jobs[e.job];
        const uint32_t x = e.xy & 0xffffu, y = e.xy >> 16;
        const float ex = e.dx, ey = e.dy, ez = e.dz;
        ((global_px *)J.dst)[__umul24(y, J.W) + x] = shade_hit<kHalf, kMaxSph>(sc, J.cam, ex, ey, ez, e.tHit, e.total, bound, evalShades);
    }
    if (stepCount) {
        unsigned long long s = mySteps;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) atomicAdd(stepCount, s);
    }
    if (work) {
        unsigned long long b = evalShades;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) b += __shfl_xor(b, o, 64);
        if (lane == 0) {
            atomicAdd(work, evalSteps);
            atomicAdd(work + 1, b);
        }
    }
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
    KernelTimer kt(ctx, CKD_KT_SDF_PLAIN);
    if (shading == CKD_SHADE_FP16)
        k_sdf<true><<<grid, 256, 0, ctx->stream>>>(dst, W, H, sc, cam, stepCount);
    else
        k_sdf<false><<<grid, 256, 0, ctx->stream>>>(dst, W, H, sc, cam, stepCount);
    return check_launch("sdf");
}

// blocks of one persistent kernel variant that fit on a CU at once
template <bool kHalf, int kMaxSph>
int persistent_blocks_per_cu(int steps)
{
    int n = 0;
    hipError_t e;
    if (steps <= 1)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_sdf_persistent<kHalf, kMaxSph, 1>, 256, 0);
    else if (steps == 2)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_sdf_persistent<kHalf, kMaxSph, 2>, 256, 0);
    else if (steps == 4)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_sdf_persistent<kHalf, kMaxSph, 4>, 256, 0);
    else if (steps == 8)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_sdf_persistent<kHalf, kMaxSph, 8>, 256, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_sdf_persistent<kHalf, kMaxSph, 16>, 256, 0);
    return e == hipSuccess ? n : 0;
}

template <bool kHalf, int kMaxSph>
void launch_persistent_steps(int steps, unsigned blocks, hipStream_t st, uint32_t *queueCounter,
                             uint32_t totalChunks, const SdfJob *jobs, int nJobs,
                             const ckd_sdf_scene &sc, unsigned long long *stepCount, uint32_t chunk,
                             f4 bound, unsigned long long *work)
{
    if (steps <= 1)
        k_sdf_persistent<kHalf, kMaxSph, 1><<<blocks, 256, 0, st>>>(queueCounter, totalChunks, jobs, nJobs, sc, stepCount, chunk, bound, work);
    else if (steps == 2)
        k_sdf_persistent<kHalf, kMaxSph, 2><<<blocks, 256, 0, st>>>(queueCounter, totalChunks, jobs, nJobs, sc, stepCount, chunk, bound, work);
    else if (steps == 4)
        k_sdf_persistent<kHalf, kMaxSph, 4><<<blocks, 256, 0, st>>>(queueCounter, totalChunks, jobs, nJobs, sc, stepCount, chunk, bound, work);
    else if (steps == 8)
        k_sdf_persistent<kHalf, kMaxSph, 8><<<blocks, 256, 0, st>>>(queueCounter, totalChunks, jobs, nJobs, sc, stepCount, chunk, bound, work);
    else
        k_sdf_persistent<kHalf, kMaxSph, 16><<<blocks, 256, 0, st>>>(queueCounter, totalChunks, jobs, nJobs, sc, stepCount, chunk, bound, work);
}

// A sphere containing every sphere of the scene: centre the centroid of the
// centres, radius the largest |c_i - C| + r_i, in double and rounded up.  A
// scene without spheres gets one far away (every step skips it).
f4 bounding_sphere(const ckd_sdf_scene &sc)
{
    if (sc.nSpheres <= 0) return f4{1e15f, 0.f, 0.f, 0.f};
    double cx = 0, cy = 0, cz = 0;
    for (int i = 0; i < sc.nSpheres; ++i) {
        cx += sc.sphere[i][0];
        cy += sc.sphere[i][1];
        cz += sc.sphere[i][2];
    }
    cx /= sc.nSpheres;
    cy /= sc.nSpheres;
    cz /= sc.nSpheres;
    const float fx = (float)cx, fy = (float)cy, fz = (float)cz;
    double R = 0;
    for (int i = 0; i < sc.nSpheres; ++i) {
        const double ex = sc.sphere[i][0] - (double)fx, ey = sc.sphere[i][1] - (double)fy,
                     ez = sc.sphere[i][2] - (double)fz;
        R = std::max(R, std::sqrt(ex * ex + ey * ey + ez * ez) + (double)sc.sphere[i][3]);
    }
    return f4{fx, fy, fz, std::nextafter((float)R, 3.4e38f)};
}

ckd_status launch_sdf_persistent(ckd_ctx *ctx, uint32_t *queueCounter, uint32_t totalChunks,
                                 const SdfJob *jobs, int nJobs, const ckd_sdf_scene &sc,
                                 int shading, unsigned long long *stepCount, bool helper, bool alone)
{
    if (totalChunks == 0) return CKD_OK;
    const bool big = sc.nSpheres > 8;
    const int steps = ctx->tuning.sdfStepsPerRefill;
    const uint32_t chunk = (uint32_t)ctx->tuning.sdfChunk;
    // persistent: sdfBlocksPerCU 4-wave blocks per CU; surplus blocks find
    // the queue empty and exit.  A helper launch (the batch's second stream,
    // once its voxel and tunnel work is done) adds the blocks that fill each
    // CU up to the kernel's occupancy and drains the same queue.  A batch
    // with nothing beside the march (alone) gets at least that occupancy:
    // the tuned count leaves room for work that is not there.
    int perCU = ctx->tuning.sdfBlocksPerCU;
    const int fit = shading == CKD_SHADE_FP16
                        ? (big ? persistent_blocks_per_cu<true, 32>(steps) : persistent_blocks_per_cu<true, 8>(steps))
                        : (big ? persistent_blocks_per_cu<false, 32>(steps) : persistent_blocks_per_cu<false, 8>(steps));
    if (helper) {
        perCU = fit - perCU;
        if (perCU <= 0) return CKD_OK;
    } else if (alone && perCU < fit) {
        perCU = fit;
    }
    unsigned blocks = (unsigned)ctx->cuCount * (unsigned)perCU;
    const unsigned useful = (totalChunks + kWavesPerBlock - 1) / kWavesPerBlock;
    if (blocks > useful) blocks = useful;
    if (blocks == 0) blocks = 1;
    // only the main launch is timed: it spans the whole march of the batch
    const f4 bound = bounding_sphere(sc);
    KernelTimer kt(ctx, helper ? -1 : CKD_KT_SDF_PERSISTENT);
    if (shading == CKD_SHADE_FP16) {
        if (big) launch_persistent_steps<true, 32>(steps, blocks, ctx->stream, queueCounter, totalChunks, jobs, nJobs, sc, stepCount, chunk, bound, ctx->sdfWork);
        else launch_persistent_steps<true, 8>(steps, blocks, ctx->stream, queueCounter, totalChunks, jobs, nJobs, sc, stepCount, chunk, bound, ctx->sdfWork);
    } else {
        if (big) launch_persistent_steps<false, 32>(steps, blocks, ctx->stream, queueCounter, totalChunks, jobs, nJobs, sc, stepCount, chunk, bound, ctx->sdfWork);
        else launch_persistent_steps<false, 8>(steps, blocks, ctx->stream, queueCounter, totalChunks, jobs, nJobs, sc, stepCount, chunk, bound, ctx->sdfWork);
    }
    return check_launch("sdf persistent");
}

} // namespace ckd

extern "C" {

ckd_status ckd_sdf_work_counter(ckd_ctx *ctx, unsigned long long *counters)
{
    CKD_REQUIRE_CTX(ctx);
    ctx->sdfWork = counters;
    return CKD_OK;
}

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
