// k_voxel.hip -- voxel landscape columns (code/landscape.cpp) and the
// tunnelscape rows (code/tunnelscape.cpp).
//
// One lane per column (rows for tunnelscape) marches the reference's 512
// front-to-back steps and draws the visible spans with the exact
// cspanISSE16 arithmetic.  The memset32 of the fog colour is fused: after
// the march every lane fills the part of its column no span reached, so
// each pixel is written exactly once.
#include "ckd_internal.h"
#include "ckd_pixel.h"
#include "ckd_jobs.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace {

constexpr unsigned kMapAnd = 1023, kMapShift = 10, kRayLength = 512;
// march steps whose texels are fetched ahead of the step being drawn
constexpr int kLsAhead = 4;

struct LsFrame {
    float X1, Y1, viewCos, viewSin;
    int mapTilt; // already scaled to H rows
};

// frames per launch, their parameters passed by value: 160 keeps the
// kernel's arguments inside the 4 KiB limit, and one launch of that many
// 1280-column frames is 3200 waves, enough to fill every SIMD
constexpr int kLsFramesPerLaunch = 160;

struct LsLaunch {
    float rayY, colScale, halfW;
    int mapScale;
    LsFrame f[kLsFramesPerLaunch];
};
static_assert(sizeof(LsLaunch) + 64 <= 4096, "k_landscape's arguments fit 4 KiB");

// bsamp32_16 (code/bilinear.h:58-83) followed by _mm_subs_epu16 of the fog
// colour (code/landscape.cpp:83): the interpolation runs on two channels per
// register (B,R and G,A in 16-bit lanes, exact as in ckd_bilerp32) and the
// saturating subtract is one packed u16 subtract per pair
CKD_DEV uint32_t ls_bilerp_fog(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3, uint32_t fu,
                               uint32_t fv, uint32_t fogBR, uint32_t fogGA)
{
    const uint32_t M = 0x00ff00ffu;
    const ckd_u16x2 f = ckd_splat16(fu), g = ckd_splat16(fv);
    const ckd_u16x2 nf = (unsigned short)256 - f, ng = (unsigned short)256 - g;
    const ckd_u16x2 br = ckd_lerp2x(ckd_lerp2x(ckd_u16x2_of(p0 & M), ckd_u16x2_of(p1 & M), nf, f),
                                    ckd_lerp2x(ckd_u16x2_of(p2 & M), ckd_u16x2_of(p3 & M), nf, f),
                                    ng, g);
    const ckd_u16x2 ga =
        ckd_lerp2x(ckd_lerp2x(ckd_u16x2_of(ckd_odd_bytes(p0)), ckd_u16x2_of(ckd_odd_bytes(p1)), nf, f),
                   ckd_lerp2x(ckd_u16x2_of(ckd_odd_bytes(p2)), ckd_u16x2_of(ckd_odd_bytes(p3)), nf, f),
                   ng, g);
    const ckd_u16x2 sbr = __builtin_elementwise_sub_sat(br, ckd_u16x2_of(fogBR));
    const ckd_u16x2 sga = __builtin_elementwise_sub_sat(ga, ckd_u16x2_of(fogGA));
    return ckd_u32_of(sbr) | (ckd_u32_of(sga) << 8);
}

// (255 - bsamp8(quad)) << 16 (code/bilinear.h:34-53, code/landscape.cpp:
// 76-86) for a height quad h4 holding (u,v), (u+1,v), (u,v+1), (u+1,v+1) in
// bytes 0..3: both horizontal lerps in the 16-bit lanes of (h0,h2) and
// (h1,h3), the vertical one as the sum of the lanes of (s01*(256-fv),
// s23*fv).  Every intermediate lies in [0, 65280], as in the scalar form.
CKD_DEV uint32_t ls_height_num(uint32_t h4, uint32_t fu, uint32_t fv)
{
    const ckd_u16x2 f = ckd_splat16(fu), nf = (unsigned short)256 - f;
    const ckd_u16x2 s = ckd_lerp2x(ckd_u16x2_of(ckd_perm(0u, h4, 0x0c020c00u)),  // h0, h2
                                   ckd_u16x2_of(ckd_perm(0u, h4, 0x0c030c01u)),  // h1, h3
                                   nf, f);                                       // s01, s23
    const ckd_u16x2 p = s * ckd_u16x2{(unsigned short)(256u - fv), (unsigned short)fv};
    const uint32_t sum = (uint32_t)p[0] + (uint32_t)p[1];
    return 0xff0000u - ckd_perm(0u, sum, 0x0c010c0cu); // (255 - (sum >> 8)) << 16
}

// exact floor(num/den) for num < 2^24 and 0 < den < 2^24 with a quotient
// below 2^22: num * rcp(den) is then within 0.5 of num/den, so the estimate
// floor(num * rcp(den) - 0.5) is the quotient or one less, and one upward
// correction fixes it.  q*den stays below 2^25, so the 24-bit multiply (full
// rate; the 32-bit one is quarter rate) is exact.  Checked on the host over
// every (num, den) of the march with fpFishMul >= 4 (W <= 64 H gives >= 8)
// and every span divisor, with the reciprocal rounded to nearest and one ulp
// off either way.
CKD_DEV uint32_t udiv24(uint32_t num, uint32_t den)
{
    // qf > -0.5: converting through a signed int truncates it to 0 there
    const float qf = __builtin_fmaf((float)num, __builtin_amdgcn_rcpf((float)den), -0.5f);
    const uint32_t q = (uint32_t)(int32_t)qf;
    return q + (num - ckd_mul_u24(q, den) >= den ? 1u : 0u);
}

// Per-column ray set-up of vscape (code/landscape.cpp:164-190) and
// voxel::vrot2D / vnorm2D (code/voxel-shared.h:8-24) in FP32 with no
// contraction and correctly rounded division and square root, so the
// integer ray parameters equal the host's bit for bit.
#pragma clang fp contract(off)
CKD_DEV void ls_setup(float rayY, float colScale, float halfW, float X1, float Y1, float viewCos,
                      float viewSin, unsigned iRay, int &dX, int &dY, int &fpFishMul)
{
    const float rayX = colScale * ((float)iRay - halfW);
    const float rotRayX = viewCos * rayX - viewSin * rayY;
    const float rotRayY = viewSin * rayX + viewCos * rayY;
    const float X2 = X1 + rotRayX;
    const float Y2 = Y1 + rotRayY;
    float fdX = X2 - X1;
    float fdY = Y2 - Y1;
    if (fabsf(fdX + fdY) > 1.1920928955078125e-07f) {
        const float length = 1.f / sqrtf(fdX * fdX + fdY * fdY);
        fdX *= length;
        fdY *= length;
    }
    const float fishMul = rayY / sqrtf(rotRayX * rotRayX + rotRayY * rotRayY);
    dX = (int)(fdX * 256.f);
    dY = (int)(fdY * 256.f);
    fpFishMul = (int)(fabsf(fishMul) * 256.f);
}
#pragma clang fp contract(on)

// frame pixels in the global address space (a job table's pointer would
// make flat stores, whose completion the fog table's scalar loads would
// wait for)
typedef __attribute__((address_space(1))) uint32_t global_u32;

// the 2x2 texel quads of the landscape maps: quadH[i] holds the heights of
// texels (u,v), (u+1,v), (u,v+1), (u+1,v+1) (indices wrapped into the map) in
// its bytes 0..3, quadC[i] their colours, for i = v*1024 + u
__global__ __launch_bounds__(256) void k_ls_quads(const uint8_t *__restrict__ hmap,
                                                  const uint32_t *__restrict__ cmap,
                                                  uint32_t *__restrict__ quadH,
                                                  uint4 *__restrict__ quadC)
{
    const unsigned i = blockIdx.x * 256 + threadIdx.x;
    if (i >= (kMapAnd + 1) * (kMapAnd + 1)) return;
    const unsigned u = i & kMapAnd, v = i >> kMapShift;
    const unsigned u1 = (u + 1) & kMapAnd, v1 = ((v + 1) & kMapAnd) << kMapShift;
    const unsigned v0 = v << kMapShift;
    quadH[i] = (uint32_t)hmap[u + v0] | ((uint32_t)hmap[u1 + v0] << 8) |
               ((uint32_t)hmap[u + v1] << 16) | ((uint32_t)hmap[u1 + v1] << 24);
    quadC[i] = make_uint4(cmap[u + v0], cmap[u1 + v0], cmap[u + v1], cmap[u1 + v1]);
}

// vscape_ray (code/landscape.cpp:56-107) for column iRay of one frame, march
// steps [s0, s1).  drawnIn() gives the highest row the steps before s0 drew
// (H when s0 is 0); it is called once the first texels are in flight.  The
// fused memset32(fog[0]) for the rows no span reached runs when s1 ends the
// march.  Returns the highest row drawn after step s1 - 1.
template <typename DrawnIn>
CKD_DEV int ls_column(global_u32 *__restrict__ col, unsigned W, unsigned H, unsigned iRay,
                      float rayY, float colScale, float halfW, int mapScale, float X1, float Y1,
                      float viewCos, float viewSin, int mapTilt,
                      const uint32_t *__restrict__ quadH, const uint4 *__restrict__ quadC,
                      const uint32_t *__restrict__ fog, unsigned s0, unsigned s1, DrawnIn drawnIn)
{
    int dX, dY, fpFishMul;
    ls_setup(rayY, colScale, halfW, X1, Y1, viewCos, viewSin, iRay, dX, dY, fpFishMul);
    int curX = (int)(X1 * 256.f), curY = (int)(Y1 * 256.f);

    // The march is software-pipelined: the texels of the next kLsAhead steps
    // are in flight while a step draws its span (only curX/curY carry from
    // step to step, and every fetch is masked into the maps, so fetches past
    // the last step are harmless).
    struct Texels {
        uint32_t h4, fu, fv;
        uint4 c;
    };
    auto fetch = [&](Texels &t) {
        curX += dX;
        curY += dY;
        const unsigned i = ((((unsigned)(curY >> 8)) & kMapAnd) << kMapShift) |
                           (((unsigned)(curX >> 8)) & kMapAnd);
        t.fu = (uint32_t)curX & 0xff;
        t.fv = (uint32_t)curY & 0xff;
        t.h4 = quadH[i];
        t.c = quadC[i];
    };
    // a step's projected row and fogged colour (f = fog[iStep >> 1], den =
    // fpFishMul * (iStep + 1))
    auto sample = [&](const Texels &t, uint32_t f, uint32_t den, int &hgt, uint32_t &color) {
        color = ls_bilerp_fog(t.c.x, t.c.y, t.c.z, t.c.w, t.fu, t.fv, f & 0x00ff00ffu, (f >> 8) & 0x00ff00ffu);

        // ((255-h)<<16) / (fpFishMul*(iStep+1)) is an unsigned division
        // fpFishMul <= 256 and iStep < 512: 24-bit multiplies are exact;
        // q < 2^17 (fpFishMul >= 191 at every width) and mapScale < 2^11
        const uint32_t q = udiv24(ls_height_num(t.h4, t.fu, t.fv), den);
        hgt = ((int)ckd_mul_u24(q, (uint32_t)mapScale) >> 8) + mapTilt;
    };

    int lastHeight = (int)H, lastDrawn;
    uint32_t lastColor;
    if (s0 == 0) {
        lastColor = quadC[(((unsigned)(curX >> 8)) & kMapAnd) | ((((unsigned)(curY >> 8)) & kMapAnd) << kMapShift)].x;
    } else {
        // the march's state after step s0 - 1: that step's row and colour
        // (s0 - 1 additions of dX wrap as one multiply-add does)
        curX = (int)((uint32_t)curX + (uint32_t)dX * (s0 - 1));
        curY = (int)((uint32_t)curY + (uint32_t)dY * (s0 - 1));
        Texels t;
        fetch(t);
        sample(t, fog[(s0 - 1) >> 1], ckd_mul_u24((uint32_t)fpFishMul, s0), lastHeight, lastColor);
    }

    // den = fpFishMul * (iStep + 1), one add per step (below 2^17)
    uint32_t den = ckd_mul_u24((uint32_t)fpFishMul, s0);
    auto step = [&](const Texels &t, uint32_t f) {
        den += (uint32_t)fpFishMul;
        int hgt;
        uint32_t color;
        sample(t, f, den, hgt, color);
        if (hgt < lastDrawn) {
            const uint32_t drawLength = (uint32_t)(lastDrawn - hgt);
            const uint32_t length = (uint32_t)(lastHeight - hgt);
            CkdSpan sp;
            ckd_span_setup_drawn(sp, color, lastColor, length, drawLength, udiv24(65536u, length));
            // hgt < H <= 2160 and W <= 4096: the row offset fits 24 bits
            global_u32 *p = col + ckd_mul_u24((uint32_t)hgt, W);
            for (uint32_t k = 0; k < drawLength; ++k) {
                *p = ckd_span_pixel(sp.acc);
                p += W;
#pragma unroll
                for (int c = 0; c < 4; ++c) sp.acc[c] += sp.step[c];
            }
            lastDrawn = hgt;
        }
        lastHeight = hgt;
        lastColor = color;
    };
    // two register sets, A and B, take turns: a set is refilled only after
    // it has been consumed, so each fetched set lands in the registers its
    // consumer reads and the loop carries no copies of values in flight
    // s0 and s1 are multiples of 2 * kLsAhead = 8, so a group's eight fog
    // words fog[iStep/2 .. iStep/2 + 3] are one aligned 16-byte load
    static_assert(kRayLength % (2 * kLsAhead) == 0 && 2 * kLsAhead == 8, "the march runs in whole groups");
    Texels A[kLsAhead], B[kLsAhead];
#pragma unroll
    for (int k = 0; k < kLsAhead; ++k) fetch(A[k]);
    lastDrawn = drawnIn();
    for (unsigned iStep = s0; iStep < s1; iStep += 2 * kLsAhead) {
        const uint4 fg = *reinterpret_cast<const uint4 *>(fog + (iStep >> 1));
        const uint32_t f8[4] = {fg.x, fg.y, fg.z, fg.w};
#pragma unroll
        for (int k = 0; k < kLsAhead; ++k) {
            fetch(B[k]);
            step(A[k], f8[k >> 1]);
        }
#pragma unroll
        for (int k = 0; k < kLsAhead; ++k) {
            fetch(A[k]);
            step(B[k], f8[(kLsAhead + k) >> 1]);
        }
    }
    if (s1 == kRayLength) {
        const uint32_t sky = fog[0];
        for (int y = 0; y < lastDrawn; ++y) col[(size_t)y * W] = sky;
    }
    return lastDrawn;
}

__global__ __launch_bounds__(64) void k_landscape(uint32_t *__restrict__ dst, unsigned W,
                                                  unsigned H, const uint32_t *__restrict__ quadH,
                                                  const uint4 *__restrict__ quadC,
                                                  const uint32_t *__restrict__ fog, LsLaunch P)
{
    const unsigned iRay = blockIdx.x * 64 + threadIdx.x;
    if (iRay >= W) return;
    const LsFrame &F = P.f[blockIdx.y];
    ls_column((global_u32 *)(dst + (size_t)blockIdx.y * W * H + iRay), W, H, iRay, P.rayY, P.colScale, P.halfW,
              P.mapScale, F.X1, F.Y1, F.viewCos, F.viewSin, F.mapTilt, quadH, quadC, fog, 0, kRayLength,
              [H] { return (int)H; });
}

// ---- the march split into step ranges, dequeued ---------------------------
//
// A launch of many frames has nGroups waves of 64 columns; when that is not
// close to a multiple of the SIMD count, the SIMDs that hold one wave more
// than the others set the launch time (120 frames of 1280 columns: 2400
// waves on 1024 SIMDs take as long as 3072).  Split each column's march into 2^c
// ranges of 512 >> c steps: unit u = k * nGroups + g is range k of group g.
// Persistent waves take units from a counter in that order.  Range k > 0
// needs only one value from range k - 1: the highest row drawn so far.  Range
// k - 1 publishes it per column, tagged (epoch, k), in a 64-bit state word;
// range k reads it (after its first texels are in flight) and waits while
// the tag is not there yet.  Range k - 1 was taken earlier, by a wave that
// runs, and depends only on ranges taken before it, so every wait ends; the
// wait is also bounded, and a wait that runs out counts in ctl[3] and
// carries on from row H.
//
// ctl: [0] next unit, [1] epoch, [2] waves finished, [3] waits that ran out.
// The last wave to finish resets [0] and [2] and moves the epoch on, so the
// next launch starts from a clean queue and tags of this one never match it.
constexpr int kLsMaxChunkLog2 = 3;
constexpr unsigned kLsWaitLimit = 1u << 18; // polls, with a sleep between them
constexpr size_t kLsStateWords = (size_t)kLsFramesPerLaunch * 4096;

CKD_DEV int ls_wait_drawn(const unsigned long long *w, unsigned long long want, int H,
                          uint32_t *timeouts)
{
    unsigned long long v = 0;
    bool ready = false;
    for (unsigned spin = 0; spin < kLsWaitLimit; ++spin) {
        if (!ready) {
            v = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ready = (v >> 16) == want;
        }
        if (__ballot(!ready) == 0) return (int)(v & 0xffffu);
        __builtin_amdgcn_s_sleep(2);
    }
    if (!ready) atomicAdd(timeouts, 1u);
    return ready ? (int)(v & 0xffffu) : H;
}

__global__ __launch_bounds__(64) void k_landscape_queue(uint32_t *__restrict__ dst, unsigned W,
                                                        unsigned H, const uint32_t *__restrict__ quadH,
                                                        const uint4 *__restrict__ quadC,
                                                        const uint32_t *__restrict__ fog, LsLaunch P,
                                                        unsigned groupsPerFrame, unsigned nGroups,
                                                        unsigned chunkLog2, unsigned nWaves,
                                                        uint32_t *__restrict__ ctl,
                                                        unsigned long long *__restrict__ state)
{
    const unsigned lane = threadIdx.x;
    const uint32_t epoch =
        __builtin_amdgcn_readfirstlane(__hip_atomic_load(&ctl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    const unsigned steps = kRayLength >> chunkLog2, nChunks = 1u << chunkLog2;
    const unsigned nUnits = nGroups << chunkLog2;
    for (;;) {
        uint32_t u = 0;
        if (lane == 0) u = atomicAdd(&ctl[0], 1u);
        u = __builtin_amdgcn_readfirstlane(u);
        if (u >= nUnits) break;
        const unsigned k = u / nGroups, g = u - k * nGroups;
        const unsigned f = g / groupsPerFrame;
        const unsigned iRay = (g - f * groupsPerFrame) * 64 + lane;
        if (iRay < W) {
            const LsFrame &F = P.f[f];
            unsigned long long *w = state + (size_t)f * W + iRay;
            const unsigned long long want = ((unsigned long long)epoch << 16) | k;
            const int drawn = ls_column(
                (global_u32 *)(dst + (size_t)f * W * H + iRay), W, H, iRay, P.rayY, P.colScale, P.halfW,
                P.mapScale, F.X1, F.Y1, F.viewCos, F.viewSin, F.mapTilt, quadH, quadC, fog, k * steps,
                (k + 1) * steps, [&] { return k == 0 ? (int)H : ls_wait_drawn(w, want, (int)H, &ctl[3]); });
            if (k + 1 < nChunks)
                __hip_atomic_store(w, ((unsigned long long)epoch << 32) | ((unsigned long long)(k + 1) << 16) |
                                          (unsigned)drawn,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    // relaxed: a wave's last dequeue returned before this increment was
    // issued, so the last wave to count sees every dequeue done (an
    // acquire-release here would write back and invalidate the L2 in every
    // wave, while frames are being written)
    if (lane == 0 &&
        __hip_atomic_fetch_add(&ctl[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == nWaves - 1) {
        __hip_atomic_store(&ctl[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl[2], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl[1], epoch + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// every voxel request of a batch in one launch: block -> job by binary search
__global__ __launch_bounds__(64) void k_landscape_jobs(const VoxJob *__restrict__ jobs, int nJobs,
                                                       const uint32_t *__restrict__ quadH,
                                                       const uint4 *__restrict__ quadC,
                                                       const uint32_t *__restrict__ fog)
{
    const int j = ckd_find_job(jobs, nJobs, blockIdx.x, &VoxJob::blockBase);
    const VoxJob &J = jobs[j];
    const unsigned iRay = (blockIdx.x - J.blockBase) * 64 + threadIdx.x;
    if (iRay >= J.W) return;
    ls_column((global_u32 *)(J.dst + iRay), J.W, J.H, iRay, J.rayY, J.colScale, J.halfW, J.mapScale, J.X1, J.Y1,
              J.viewCos, J.viewSin, J.mapTilt, quadH, quadC, fog, 0, kRayLength, [&J] { return (int)J.H; });
}

// ---- tunnelscape ---------------------------------------------------------

constexpr unsigned kTsAnd = 2047, kTsShift = 11;

// tscape_ray (code/tunnelscape.cpp:44-100), lane per output row.  The
// caller guarantees map heights <= 159, so every height is positive.
__global__ __launch_bounds__(64) void k_tunnelscape(uint32_t *__restrict__ dst, unsigned W,
                                                    unsigned H, const uint8_t *__restrict__ hmap,
                                                    const uint32_t *__restrict__ cmap,
                                                    const uint32_t *__restrict__ fog,
                                                    float mapStepX, float fromXAdd, int fpFromY,
                                                    int dX, int dY, float kMapViewLenScale)
{
    const unsigned iRay = blockIdx.x * 64 + threadIdx.x;
    if (iRay >= H) return;
    uint32_t *row = dst + (size_t)iRay * W;
    int curX, curY = fpFromY;
    {
#pragma clang fp contract(off)
        const float mapX = (float)iRay * mapStepX;
        const float fromX = mapX + fromXAdd;
        curX = (int)(fromX * 256.f);
    }
    int lastHeight = (int)W, lastDrawn = (int)W;
    uint32_t lastColor =
        cmap[(((unsigned)(curX >> 8)) & kTsAnd) | ((((unsigned)(curY >> 8)) & kTsAnd) << kTsShift)];
    unsigned written = 0;
    // pipelined as in ls_column: the texels of the next kLsAhead steps are in
    // flight while a step draws its span (fetches are masked into the maps)
    struct Texels {
        uint32_t h[4], c[4], fu, fv;
    };
    auto fetch = [&](Texels &t) {
        curX += dX;
        curY += dY;
        const unsigned u0 = ((unsigned)(curX >> 8)) & kTsAnd;
        const unsigned u1 = ((unsigned)(curX >> 8) + 1) & kTsAnd;
        const unsigned v0 = (((unsigned)(curY >> 8)) & kTsAnd) << kTsShift;
        const unsigned v1 = (((unsigned)(curY >> 8) + 1) & kTsAnd) << kTsShift;
        t.fu = (uint32_t)curX & 0xff;
        t.fv = (uint32_t)curY & 0xff;
        t.h[0] = hmap[u0 + v0];
        t.h[1] = hmap[u1 + v0];
        t.h[2] = hmap[u0 + v1];
        t.h[3] = hmap[u1 + v1];
        t.c[0] = cmap[u0 + v0];
        t.c[1] = cmap[u1 + v0];
        t.c[2] = cmap[u0 + v1];
        t.c[3] = cmap[u1 + v1];
    };
    auto step = [&](unsigned iStep, const Texels &t) {
        const uint32_t h = ckd_bilerp8(t.h[0], t.h[1], t.h[2], t.h[3], t.fu, t.fv);
        uint32_t color = ckd_bilerp32(t.c[0], t.c[1], t.c[2], t.c[3], t.fu, t.fv);
        color = ckd_subs_u8x4(color, fog[iStep >> 1]);

        int height = 255 - (int)h - 96;
        height <<= 8;
        height = (int)((float)height / kMapViewLenScale);
        height = (int)((unsigned)height / (iStep + 1));
        height *= 160;
        height >>= 8;
        height += 120;

        if (height < lastDrawn) {
            const uint32_t drawLength = (uint32_t)(lastDrawn - height);
            CkdSpan s;
            ckd_span_setup(s, color, lastColor, (uint32_t)(lastHeight - height), drawLength);
            for (uint32_t k = 0; k < drawLength; ++k) {
                row[written + k] = ckd_span_pixel(s.acc);
#pragma unroll
                for (int c = 0; c < 4; ++c) s.acc[c] += s.step[c];
            }
            written += drawLength;
            lastDrawn = height;
        }
        lastHeight = height;
        lastColor = color;
    };
    Texels A[kLsAhead], B[kLsAhead];
#pragma unroll
    for (int k = 0; k < kLsAhead; ++k) fetch(A[k]);
    for (unsigned iStep = 0; iStep < kRayLength; iStep += 2 * kLsAhead) {
#pragma unroll
        for (int k = 0; k < kLsAhead; ++k) {
            fetch(B[k]);
            step(iStep + k, A[k]);
        }
#pragma unroll
        for (int k = 0; k < kLsAhead; ++k) {
            fetch(A[k]);
            step(iStep + kLsAhead + k, B[k]);
        }
    }
    const uint32_t sky = fog[0];
    for (unsigned x = written; x < W; ++x) row[x] = sky;
}

ckd_status upload(const void *host, size_t bytes, void **dev)
{
    if (*dev) {
        (void)hipFree(*dev);
        *dev = nullptr;
    }
    if (ckd::dev_alloc(dev, bytes) != hipSuccess) {
        *dev = nullptr;
        return ckd::fail(CKD_ENOMEM, "map upload: ckd::dev_alloc(%zu) failed", bytes);
    }
    CKD_HIP(hipMemcpy(*dev, host, bytes, hipMemcpyHostToDevice));
    return CKD_OK;
}

float ls_rayY(unsigned W, unsigned H)
{
#pragma clang fp contract(off)
    const float kPI = 3.1415926535897932384626433832795f;
    const float aspect = (float)H / (float)W;
    const float viewLenScale = aspect * (kPI * 0.1f);
    return 1024.f * viewLenScale;
}

int scale_rows(int v720, unsigned H) { return (int)(((long long)v720 * (long long)H) / 720); }

} // namespace

namespace ckd {

void voxel_job_geometry(VoxJob &J, unsigned W, unsigned H, const ckd_landscape_cam &c)
{
    J.W = W;
    J.H = H;
    J.rayY = ls_rayY(W, H);
    J.colScale = 0.25f * (1280.f / (float)W);
    J.halfW = (float)W * 0.5f;
    J.mapScale = scale_rows(512, H);
    J.X1 = c.X1;
    J.Y1 = c.Y1;
    J.viewCos = c.viewCos;
    J.viewSin = c.viewSin;
    J.mapTilt = scale_rows(c.mapTilt, H);
}

ckd_status launch_landscape_jobs(ckd_ctx *ctx, const VoxJob *jobs, int nJobs, uint32_t totalBlocks)
{
    if (totalBlocks == 0) return CKD_OK;
    if (!ctx->lsHeight)
        return fail(CKD_ESTATE, "voxel requests need the landscape maps (ckd_landscape_create)");
    KernelTimer kt(ctx, CKD_KT_LANDSCAPE);
    k_landscape_jobs<<<totalBlocks, 64, 0, ctx->stream>>>(jobs, nJobs, ctx->lsQuadH, ctx->lsQuadC,
                                                          ctx->lsFog);
    return check_launch("landscape jobs");
}

} // namespace ckd

extern "C" {

ckd_status ckd_landscape_create(ckd_ctx *ctx, const uint8_t *heightMap, const uint32_t *colorMap,
                                const uint32_t *fog256)
{
    CKD_REQUIRE(heightMap && colorMap && fog256, "ckd_landscape_create: null map");
    CKD_REQUIRE_CTX(ctx);
    CKD_HIP(hipStreamSynchronize(ctx->stream));
    ckd_status s = upload(heightMap, 1024 * 1024, (void **)&ctx->lsHeight);
    if (s == CKD_OK) s = upload(colorMap, 1024 * 1024 * 4, (void **)&ctx->lsColor);
    if (s == CKD_OK) s = upload(fog256, 256 * 4, (void **)&ctx->lsFog);
    if (s != CKD_OK) return s;
    ctx->lsFog0 = fog256[0];
    constexpr size_t kTexels = 1024 * 1024;
    if ((!ctx->lsQuadH && ckd::dev_alloc(&ctx->lsQuadH, kTexels * 4) != hipSuccess) ||
        (!ctx->lsQuadC && ckd::dev_alloc(&ctx->lsQuadC, kTexels * 16) != hipSuccess))
        return ckd::fail(CKD_ENOMEM, "ckd_landscape_create: hipMalloc of the texel quads failed");
    if (!ctx->lsQueue) {
        if (ckd::dev_alloc(&ctx->lsQueue, 4 * sizeof(uint32_t)) != hipSuccess)
            return ckd::fail(CKD_ENOMEM, "ckd_landscape_create: hipMalloc of the work queue failed");
        CKD_HIP(hipMemsetAsync(ctx->lsQueue, 0, 4 * sizeof(uint32_t), ctx->stream));
    }
    if (!ctx->lsState) {
        if (ckd::dev_alloc(&ctx->lsState, kLsStateWords * sizeof(unsigned long long)) != hipSuccess)
            return ckd::fail(CKD_ENOMEM, "ckd_landscape_create: hipMalloc of the queue's state failed");
        CKD_HIP(hipMemsetAsync(ctx->lsState, 0, kLsStateWords * sizeof(unsigned long long), ctx->stream));
    }
    k_ls_quads<<<kTexels / 256, 256, 0, ctx->stream>>>(ctx->lsHeight, ctx->lsColor, ctx->lsQuadH,
                                                       ctx->lsQuadC);
    s = ckd::check_launch("landscape quads");
    if (s == CKD_OK) CKD_HIP(hipStreamSynchronize(ctx->stream));
    return s;
}

ckd_status ckd_landscape_draw(ckd_ctx *ctx, uint32_t *dst, unsigned W, unsigned H,
                              const ckd_landscape_cam *cams, unsigned nFrames)
{
    CKD_REQUIRE(dst && cams, "ckd_landscape_draw: null buffer");
    CKD_REQUIRE(W >= 8 && H >= 8 && W <= 4096 && H <= 2160, "ckd_landscape_draw: bad size %ux%u",
                W, H);
    // the edge columns' fpFishMul is at least 8 then (udiv24 needs 4)
    CKD_REQUIRE(W <= 64 * H, "ckd_landscape_draw: %ux%u is wider than 64 times its height", W, H);
    for (unsigned f = 0; f < nFrames; ++f)
        CKD_REQUIRE(cams[f].mapTilt >= 0 && cams[f].mapTilt <= 180,
                    "ckd_landscape_draw: frame %u: mapTilt %d outside 0..180", f, cams[f].mapTilt);
    CKD_REQUIRE_CTX(ctx);
    if (!ctx->lsHeight) return ckd::fail(CKD_ESTATE, "ckd_landscape_draw: no maps (ckd_landscape_create)");
    LsLaunch P;
    memset(&P, 0, sizeof(P));
    P.rayY = ls_rayY(W, H);
    P.colScale = 0.25f * (1280.f / (float)W);
    P.halfW = (float)W * 0.5f;
    P.mapScale = scale_rows(512, H);
    for (unsigned f0 = 0; f0 < nFrames; f0 += kLsFramesPerLaunch) {
        const unsigned n = nFrames - f0 < (unsigned)kLsFramesPerLaunch ? nFrames - f0
                                                                       : (unsigned)kLsFramesPerLaunch;
        for (unsigned k = 0; k < n; ++k) {
            const ckd_landscape_cam &c = cams[f0 + k];
            P.f[k] = LsFrame{c.X1, c.Y1, c.viewCos, c.viewSin, scale_rows(c.mapTilt, H)};
        }
        const unsigned groupsPerFrame = ckd_div_up(W, 64), nGroups = groupsPerFrame * n;
        const unsigned simds = 4u * (unsigned)ctx->cuCount;
        const int chunkLog2 = __builtin_ctz((unsigned)ctx->tuning.voxelChunks);
        // one wave per group puts ceil(nGroups / simds) waves on the busiest
        // SIMDs; the work queue pays when that is well above the average
        const unsigned busiest = ckd_div_up(nGroups, simds);
        const bool uneven = nGroups > simds && (double)nGroups < 0.9 * (double)busiest * simds;
        ckd::KernelTimer kt(ctx, CKD_KT_LANDSCAPE);
        if (chunkLog2 > 0 && uneven) {
            // the march ranges through the work queue
            static_assert(kLsMaxChunkLog2 == 3, "voxelChunks is at most 8");
            const unsigned nUnits = nGroups << chunkLog2;
            const unsigned nWaves = std::min(nUnits, simds * (unsigned)ctx->tuning.voxelWavesPerSIMD);
            k_landscape_queue<<<nWaves, 64, 0, ctx->stream>>>(
                dst + (size_t)f0 * W * H, W, H, ctx->lsQuadH, ctx->lsQuadC, ctx->lsFog, P, groupsPerFrame,
                nGroups, (unsigned)chunkLog2, nWaves, ctx->lsQueue, ctx->lsState);
        } else {
            dim3 grid(groupsPerFrame, n);
            k_landscape<<<grid, 64, 0, ctx->stream>>>(dst + (size_t)f0 * W * H, W, H, ctx->lsQuadH,
                                                      ctx->lsQuadC, ctx->lsFog, P);
        }
        ckd_status s = ckd::check_launch("landscape");
        if (s != CKD_OK) return s;
    }
    return CKD_OK;
}

ckd_status ckd_tunnelscape_create(ckd_ctx *ctx, const uint8_t *heightMap, const uint32_t *colorMap,
                                  const uint32_t *fog256)
{
    CKD_REQUIRE(heightMap && colorMap && fog256, "ckd_tunnelscape_create: null map");
    for (size_t i = 0; i < 2048u * 2048u; ++i)
        CKD_REQUIRE(heightMap[i] <= 159,
                    "ckd_tunnelscape_create: height %u at %zu > 159 (negative span heights)",
                    heightMap[i], i);
    CKD_REQUIRE_CTX(ctx);
    CKD_HIP(hipStreamSynchronize(ctx->stream));
    ckd_status s = upload(heightMap, 2048 * 2048, (void **)&ctx->tsHeight);
    if (s == CKD_OK) s = upload(colorMap, 2048 * 2048 * 4, (void **)&ctx->tsColor);
    if (s == CKD_OK) s = upload(fog256, 256 * 4, (void **)&ctx->tsFog);
    if (s == CKD_OK) ctx->tsFog0 = fog256[0];
    return s;
}

ckd_status ckd_tunnelscape_rows(ckd_ctx *ctx, uint32_t *dst, unsigned W, unsigned H, float stepU,
                                float stepV, float speed, float time)
{
    CKD_REQUIRE(dst, "ckd_tunnelscape_rows: null buffer");
    CKD_REQUIRE(W >= 8 && H >= 8 && W <= 4096 && H <= 4096, "ckd_tunnelscape_rows: bad size %ux%u",
                W, H);
    CKD_REQUIRE_CTX(ctx);
    if (!ctx->tsHeight)
        return ckd::fail(CKD_ESTATE, "ckd_tunnelscape_rows: no maps (ckd_tunnelscape_create)");
    // tscape (code/tunnelscape.cpp:103-118) frame constants, host FP32
    float mapStepX, fromXAdd, kMapViewLenScale;
    int dX, dY, fpFromY;
    {
#pragma clang fp contract(off)
        const float kGoldenRatio = 1.61803398875f;
        const float aspect = (float)H / (float)W;
        const float kOneOverAspect = 1.f / aspect;
        // kMapViewLenScale is the reference's 720p constant (code/tunnelscape.cpp:31)
        kMapViewLenScale = (720.f / 1280.f) * 0.5f;
        mapStepX = 2048.f / (float)(H - 1);
        const float speedMul = sqrtf(stepU * stepU + stepV * stepV) * speed;
        const float fromY = 1024.f + speedMul * time;
        dX = (int)(stepV * 256.f);
        dY = (int)((kOneOverAspect * stepU) * 256.f);
        fpFromY = (int)(fromY * 256.f);
        fromXAdd = stepU * time * kGoldenRatio;
    }
    k_tunnelscape<<<ckd_div_up(H, 64), 64, 0, ctx->stream>>>(dst, W, H, ctx->tsHeight, ctx->tsColor,
                                                            ctx->tsFog, mapStepX, fromXAdd, fpFromY,
                                                            dX, dY, kMapViewLenScale);
    return ckd::check_launch("tunnelscape");
}

} // extern "C"
