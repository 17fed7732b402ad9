// k_shadertoy.hip -- the reference's raymarched FX-map effects
// (code/shadertoy.cpp): Plasma, Nautilus, Spikey (distant, specular-only
// and close-up), Sinuses, Laura and the free-directional Tunnel, with the
// helpers of code/shadertoy-util.h.
//
// One lane per FX-map pixel; a 256-lane block covers 16x16 pixels as four
// 8x8 wave tiles (neighbouring rays march alike), blockIdx.z is the frame.
// Per-frame constants (palettes, fSpike_global, camera origin, the sines
// and cosines of per-frame angles) are computed on the host in the
// reference's FP32 order and staged once per call; the 2049-entry cosine
// LUT is copied into LDS by every block.
//
// Parity: every float operation is the reference's, in its order, in
// FP32 without contraction; sqrtf and division are correctly rounded.
// vNorm4's _mm_rsqrt_ps estimate is 1/sqrtf here (its table is specific to
// the CPU model), and GammaAdj's exp_ps/log_ps are the Cephes polynomials
// step for step, so the lane-level quirks of the reference survive: a
// channel <= 0 takes log's NaN path and packs to 0, a NaN channel takes
// log(FLT_MIN), and 255*c >= 32767.5 wraps to 0 in the packus pair.  The
// remaining libm calls (powf, expf, atan2f) are the device library's.
#include "ckd_internal.h"
#include "ckd_b07.h"
#include "ckd_pixel.h"

#include <cmath>
#include <cstring>

#pragma clang fp contract(off)

namespace {

constexpr float kPI = 3.1415926535897932384626433832795f;
constexpr float kGoldenRatio = 1.61803398875f;
constexpr float kGoldenAngle = 2.39996f;
constexpr float kEpsilon = 1.1920928955078125e-07f;
constexpr unsigned kLutN = 2049;
constexpr int kTile = 16; // block tile side (pixels)

enum StFx { FX_PLASMA, FX_NAUTILUS, FX_SPIKEY, FX_SINUSES, FX_LAURA, FX_FDTUNNEL };

// ---- per-frame constants (host-computed) --------------------------------

struct PlasmaF {
    float colA[3], colB[3], time, dirCos, dirSin, gamma;
};
struct NautilusF {
    float g[3], diff[3], time, rollCos, rollSin, cosHitOffs, funkCos;
};
struct SpikeyF {
    int mode; // 0 distant, 1 distant specular-only, 2 close-up
    float g[3], diff[4], origin[3], rollCos, rollSin, specPow, gamma, xOffs, yOffs, zDir, warmup,
        nOffs;
};
struct SinusesF {
    float origin[3], diff[4], specPow, rollCos, rollSin, offsX, gamma;
};
struct LauraF {
    float origin[3], diff[4], yawCos, yawSin, pitchCos, pitchSin, rollCos, rollSin;
};
struct TunnelF {
    float boxT, flowerScale, flowerFreq, flowerPhase, pitchCos, pitchSin, rollCos, rollSin,
        radius, uMul, vMul, vShift, fog0, fog1;
};
union StFrame {
    PlasmaF p;
    NautilusF n;
    SpikeyF s;
    SinusesF si;
    LauraF l;
    TunnelF t;
};

struct StGeom {
    unsigned mapW, mapH;
    float invW, invH, aspect, oneOverAspect, lutScale, halfPi;
};

// ---- scalar helpers shared by host set-up and device code ---------------

struct F4 {
    float x, y, z, w;
};

// lutcosf (code/sincos-lut.h:13-19) over a LUT in any address space; the
// index of an angle at or above 2^31 is 0, as x86 cvttss2si makes it
__host__ __device__ __forceinline__ float lut_cos(const float *lut, float angle, float scale)
{
    angle = fabsf(angle);
    angle *= scale;
    const int index = angle < 2147483648.f ? ((int)angle & 2047) : 0;
    const float frac = angle - truncf(angle);
    const float l0 = lut[index], l1 = lut[index + 1];
    return l0 + (l1 - l0) * frac;
}

// lutsinf, ARRESTED_DEV_LEGACY form (code/sincos-lut.h:21-26)
__host__ __device__ __forceinline__ float lut_sin(const float *lut, float angle, float scale,
                                                  float halfPi)
{
    return lut_cos(lut, angle + halfPi, scale);
}

// Desaturate (code/shadertoy-util.h:244-249) of a Vector3 (padding 0):
// the padding lane becomes amount*luma
__host__ __device__ __forceinline__ F4 desaturate(F4 c, float amount)
{
    const float luma = (0.0722f * c.x + 0.7152f * c.y) + (0.2126f * c.z + 0.f * c.w);
    return F4{c.x + amount * (luma - c.x), c.y + amount * (luma - c.y), c.z + amount * (luma - c.z),
              c.w + amount * (luma - c.w)};
}

// ---- device helpers ------------------------------------------------------

struct Lut {
    const float *t;
    float scale, halfPi;
    __device__ __forceinline__ float cos(float a) const { return lut_cos(t, a, scale); }
    __device__ __forceinline__ float sin(float a) const { return lut_sin(t, a, scale, halfPi); }
};

// vFastLen3 of a Vector3 (padding 0): _mm_dp_ps(p, p, 0xff) is
// (x*x + y*y) + (z*z + 0), and z*z + 0 == z*z for every z
__device__ __forceinline__ float vlen3(float x, float y, float z)
{
    return sqrtf((x * x + y * y) + z * z);
}

// vNorm4 of a Vector3 (padding 0), the padding lane kept: 0 * (1/|v|)
__device__ __forceinline__ F4 vnorm3(float x, float y, float z)
{
    const float r = 1.f / sqrtf((x * x + y * y) + z * z);
    return F4{x * r, y * r, z * r, 0.f * r};
}

__device__ __forceinline__ float dot3(const F4 &a, const F4 &b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

__device__ __forceinline__ float dp4(const F4 &a, const F4 &b)
{
    return (a.x * b.x + a.y * b.y) + (a.z * b.z + a.w * b.w);
}

// std::max<float>(a, b) / std::min<float>(a, b)
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }

// ExpFog (code/shadertoy-util.h:237-241)
__device__ __forceinline__ float exp_fog(float distance, float scale)
{
    return 1.f - expf(-scale * distance * distance * distance);
}

// the Cephes constants as sse_mathfun.h has them: double literals rounded
// to float
#define F(x) ((float)(x))

// log_ps of 3rdparty/sse_mathfun.h (Cephes logf), one lane: x <= 0 gives
// NaN; a NaN input is replaced by FLT_MIN by the maxps that cuts off
// denormals
__device__ __forceinline__ float cephes_log(float x)
{
    const bool invalid = x <= 0.f;
    x = x > 1.17549435e-38f ? x : 1.17549435e-38f;
    const uint32_t bits = __float_as_uint(x);
    float e = (float)((int)(bits >> 23) - 0x7f);
    e = e + 1.f;
    x = __uint_as_float((bits & ~0x7f800000u) | 0x3f000000u);
    const bool small = x < F(0.707106781186547524);
    const float t0 = small ? x : 0.f;
    x = x - 1.f;
    e = e - (small ? 1.f : 0.f);
    x = x + t0;
    const float z = x * x;
    float y = F(7.0376836292E-2);
    y = y * x + F(-1.1514610310E-1);
    y = y * x + F(1.1676998740E-1);
    y = y * x + F(-1.2420140846E-1);
    y = y * x + F(1.4249322787E-1);
    y = y * x + F(-1.6668057665E-1);
    y = y * x + F(2.0000714765E-1);
    y = y * x + F(-2.4999993993E-1);
    y = y * x + F(3.3333331174E-1);
    y = y * x;
    y = y * z;
    y = y + e * F(-2.12194440e-4);
    y = y - z * 0.5f;
    x = x + y;
    x = x + e * F(0.693359375);
    return invalid ? __uint_as_float(0xffffffffu) : x;
}

// exp_ps of 3rdparty/sse_mathfun.h (Cephes expf), one lane: the argument
// is clamped to +-88.3762626647949 with minps/maxps (NaN -> the bound)
__device__ __forceinline__ float cephes_exp(float x)
{
    x = x < 88.3762626647949f ? x : 88.3762626647949f;
    x = x > -88.3762626647949f ? x : -88.3762626647949f;
    float fx = x * F(1.44269504088896341);
    fx = fx + 0.5f;
    const float tr = (float)(int)fx;
    fx = tr - (tr > fx ? 1.f : 0.f);
    x = x - fx * F(0.693359375);
    x = x - fx * F(-2.12194440e-4);
    const float z = x * x;
    float y = F(1.9875691500E-4);
    y = y * x + F(1.3981999507E-3);
    y = y * x + F(8.3334519073E-3);
    y = y * x + F(4.1665795894E-2);
    y = y * x + F(1.6666665459E-1);
    y = y * x + F(5.0000001201E-1);
    y = y * z;
    y = y + x;
    y = y + 1.f;
    const float pow2n = __uint_as_float((uint32_t)((int)fx + 0x7f) << 23);
    return y * pow2n;
}

// GammaAdj (code/shadertoy-util.h:186-190) of one lane
__device__ __forceinline__ float gamma_lane(float c, float gamma)
{
    return cephes_exp(gamma * cephes_log(c));
}

// cvtps_epi32 (nearest even; NaN / out of range -> INT_MIN), then the
// max(0) and the packus_epi32 -> packus_epi16 pair of ToPixel4
// (code/shadertoy-util.h:136-146) or ToPixel4_NoConv (:148-157)
__device__ __forceinline__ uint32_t pack_lane(float v)
{
    const bool inRange = v < 2147483648.f && v >= -2147483648.f; // false for NaN
    int i = inRange ? (int)rintf(v) : 0;
    i = i < 0 ? 0 : (i > 65535 ? 65535 : i);
    const int s16 = (int)(int16_t)(uint16_t)i;
    return (uint32_t)(s16 < 0 ? 0 : (s16 > 255 ? 255 : s16));
}

__device__ __forceinline__ uint32_t to_pixel(const F4 &c)
{
    return pack_lane(255.f * c.x) | (pack_lane(255.f * c.y) << 8) | (pack_lane(255.f * c.z) << 16) |
           (pack_lane(255.f * c.w) << 24);
}

__device__ __forceinline__ uint32_t to_pixel_noconv(const F4 &c)
{
    return pack_lane(c.x) | (pack_lane(c.y) << 8) | (pack_lane(c.z) << 16) | (pack_lane(c.w) << 24);
}

__device__ __forceinline__ F4 gamma_adj(const F4 &c, float gamma)
{
    return F4{gamma_lane(c.x, gamma), gamma_lane(c.y, gamma), gamma_lane(c.z, gamma),
              gamma_lane(c.w, gamma)};
}

// vLerp4(A, B, f) with B a broadcast scalar
__device__ __forceinline__ F4 lerp_to(const F4 &a, float b, float f)
{
    return F4{a.x + f * (b - a.x), a.y + f * (b - a.y), a.z + f * (b - a.z), a.w + f * (b - a.w)};
}

// ToUV_FxMap (code/shadertoy-util.h:123-130)
__device__ __forceinline__ void to_uv(const StGeom &g, unsigned iX, unsigned iY, float scale,
                                      float &u, float &v)
{
    const float fX = (float)iX * g.invW, fY = (float)iY * g.invH;
    u = (fX - 0.5f) * scale * g.oneOverAspect;
    v = (fY - 0.5f) * scale;
}

// rotZ / rotX / rotY (code/shadertoy-util.h:31-59) with the per-frame
// cosine and sine of the angle
__device__ __forceinline__ void rot(float cs, float sn, float &a, float &b)
{
    const float ra = cs * a + sn * b;
    const float rb = -sn * a + cs * b;
    a = ra;
    b = rb;
}
__device__ __forceinline__ void rot_x(float cs, float sn, float &y, float &z)
{
    const float ry = cs * y + -sn * z;
    const float rz = sn * y + cs * z;
    y = ry;
    z = rz;
}

// ---- the effects, per pixel -----------------------------------------------

// fPlasma (code/shadertoy.cpp:200-208)
__device__ __forceinline__ float f_plasma(const Lut &L, float x, float y, float z, float time)
{
    const float sine = 0.2f * L.sin(x - y);
    const float fX = sine + L.cos(x * 0.33f);
    const float fY = sine + L.cos(y * 0.43f);
    const float fZ = sine + L.cos((5.f * time + z) * 0.53f);
    return sqrtf(fX * fX + fY * fY + fZ * fZ) - 0.8f;
}

// RenderPlasmaMap per pixel (code/shadertoy.cpp:236-268)
__device__ uint32_t plasma_pixel(const Lut &L, const StGeom &g, const PlasmaF &f, unsigned iX,
                                 unsigned iY)
{
    float u, v;
    to_uv(g, iX, iY, 4.f, u, v);
    const float dx = f.dirCos * u * g.aspect - f.dirSin * 0.75f;
    const float dy = v;
    const float dz = f.dirSin * u + f.dirCos * 0.75f;
    float total = 0.f, march = 0.f, hx = 0.f, hy = 0.f, hz = 0.f;
    for (int step = 0; step < 24; ++step) {
        march = f_plasma(L, hx, hy, hz, f.time);
        total += march * (0.5f * kGoldenRatio);
        hx = dx * total;
        hy = dy * total;
        hz = dz * total;
    }
    const float f2 = f_plasma(L, hx * 0.5f, hy * 0.5f, hz * 0.5f, f.time);
    const float mul = 8.f - dx * 0.5f;
    const F4 c{(f.colA[0] * march + f.colB[0] * f2) * mul, (f.colA[1] * march + f.colB[1] * f2) * mul,
               (f.colA[2] * march + f.colB[2] * f2) * mul, 0.f};
    return to_pixel(gamma_adj(c, f.gamma));
}

// fNautilus (code/shadertoy.cpp:289-298)
__device__ __forceinline__ float f_nautilus(const Lut &L, const NautilusF &f, float x, float y,
                                            float z)
{
    const float cosX = L.cos(L.cos(x + f.g[0]) * x - L.cos(y + f.g[1]) * y);
    const float cosY = L.cos(z * 0.33f * x - f.g[2] * y);
    const float cosZ = L.cos(x + y + z * 0.8f + f.time);
    const float dotted = cosX * cosX + cosY * cosY + cosZ * cosZ;
    return dotted * 0.5f - .7f;
}

__device__ __forceinline__ F4 nautilus_grad(const Lut &L, const NautilusF &f, float x, float y,
                                            float z, float march)
{
    const float nOffs = 0.15f;
    const float nx = march - f_nautilus(L, f, x + nOffs, y, z);
    const float ny = march - f_nautilus(L, f, x, y + nOffs, z);
    const float nz = march - f_nautilus(L, f, x, y, z + nOffs);
    return vnorm3(nx, ny, nz);
}

// RenderNautilusMap_2x2 per pixel (code/shadertoy.cpp:338-378)
__device__ uint32_t nautilus_pixel(const Lut &L, const StGeom &g, const NautilusF &f, unsigned iX,
                                   unsigned iY, int &steps)
{
    float u, v;
    to_uv(g, iX, iY, 2.f, u, v);
    float dx = u * g.aspect, dy = v;
    rot(f.rollCos, f.rollSin, dx, dy);
    const F4 d = vnorm3(dx, dy, 1.f);
    float total = 0.01f, march = 1.f, hx = 0.f, hy = 0.f, hz = 0.f;
    int step = 0;
    for (; march > 0.01f && step < 48; ++step) {
        hx = d.x * total;
        hy = d.y * total;
        hz = d.z * total;
        march = f_nautilus(L, f, hx, hy, hz);
        total += march * 0.628f;
    }
    steps = step;
    const F4 n = nautilus_grad(L, f, hx, hy, hz, march);
    float diffuse = n.z * 0.1f;
    const float spec = powf(smax(0.f, dot3(n, d)), 16.f);
    const F4 funk = nautilus_grad(L, f, hx + f.cosHitOffs, hy + f.cosHitOffs, hz + f.cosHitOffs, march);
    const float m = hy * 0.3f + funk.x * 0.628f + funk.y * f.funkCos;
    const float yMod = m - truncf(m);
    diffuse *= yMod * yMod * yMod;
    const float s = 1.56f * total + spec;
    const float b = spec * kGoldenRatio * 0.2f;
    const F4 c{(diffuse + f.diff[0] * s) + b, (diffuse + f.diff[1] * s) + b,
               (diffuse + f.diff[2] * s) + b, 0.f};
    return to_pixel(gamma_adj(c, 1.44f));
}

// fSpikey1 / fSpikey2 (code/shadertoy.cpp:418-430)
template <int kScaleSel>
__device__ __forceinline__ float f_spikey(const Lut &L, const SpikeyF &f, float x, float y, float z)
{
    const float scale = kScaleSel == 1 ? kGoldenAngle * 0.1f : kGoldenRatio * 0.1f;
    const float radius = 1.35f + scale * L.cos(f.g[1] * y - f.g[0]) + scale * L.cos(f.g[2] * x + f.g[0]);
    return vlen3(x, y, z) - radius;
}

template <int kScaleSel>
__device__ __forceinline__ F4 spikey_grad(const Lut &L, const SpikeyF &f, float x, float y, float z,
                                          float march, float nOffs)
{
    const float nx = march - f_spikey<kScaleSel>(L, f, x + nOffs, y, z);
    const float ny = march - f_spikey<kScaleSel>(L, f, x, y + nOffs, z);
    const float nz = march - f_spikey<kScaleSel>(L, f, x, y, z + nOffs);
    return vnorm3(nx, ny, nz);
}

// RenderSpikeyMap_2x2_Distant (code/shadertoy.cpp:558-593),
// _Distant_SpecularOnly (:620-653) and _Close (:468-517) per pixel
__device__ uint32_t spikey_pixel(const Lut &L, const StGeom &g, const SpikeyF &f, unsigned iX,
                                 unsigned iY, int &steps)
{
    const float ox = f.origin[0], oy = f.origin[1], oz = f.origin[2];
    float u, v, march = 1.f, total = 0.f, hx = 0.f, hy = 0.f, hz = 0.f;
    F4 c;
    if (f.mode == 0) {
        to_uv(g, iX, iY, 2.f, u, v);
        float dx = u + f.xOffs, dy = v + f.yOffs;
        rot(f.rollCos, f.rollSin, dx, dy);
        const F4 d = vnorm3(dx, dy, 1.f);
        int step = 0;
        for (; march > 0.001f && step < 48; ++step) {
            hx = ox + d.x * total;
            hy = oy + d.y * total;
            hz = oz + d.z * total;
            march = f_spikey<2>(L, f, hx, hy, hz);
            march *= 0.314f;
            total += march;
        }
        steps = step;
        const F4 n = spikey_grad<2>(L, f, hx, hy, hz, march, kPI * 0.02f);
        const float diffuse = smax(0.f, n.z * 0.8f + n.y * 0.2f);
        const float spec = powf(dot3(n, d), f.specPow);
        const F4 a{(f.diff[0] + spec) * diffuse, (f.diff[1] + spec) * diffuse,
                   (f.diff[2] + spec) * diffuse, (f.diff[3] + spec) * diffuse};
        c = gamma_adj(lerp_to(a, 1.f, exp_fog(hz - oz, 0.133f)), f.gamma);
    } else if (f.mode == 1) {
        to_uv(g, iX, iY, kGoldenRatio, u, v);
        float dx = u * g.aspect, dy = v;
        rot(f.rollCos, f.rollSin, dx, dy);
        const F4 d = vnorm3(dx, dy, 1.f);
        for (int step = 0; step < 36; ++step) {
            hx = ox + d.x * total;
            hy = oy + d.y * total;
            hz = oz + d.z * total;
            march = f_spikey<2>(L, f, hx, hy, hz);
            total += march * 0.075f * kGoldenRatio;
        }
        steps = 36;
        const F4 n = spikey_grad<2>(L, f, hx, hy, hz, march, 0.01f);
        const float spec = f.warmup * powf(smax(0.f, dot3(n, d)), f.specPow);
        const float fog = exp_fog(hz - oz, 0.0133f);
        const float s = spec + fog * (0.f - spec);
        c = F4{s, s, s, s};
    } else {
        to_uv(g, iX, iY, 2.f, u, v);
        float dx = (u + f.xOffs) * g.aspect, dy = v + f.yOffs;
        rot(f.rollCos, f.rollSin, dx, dy);
        const F4 d = vnorm3(dx, dy, f.zDir);
        int step = 0;
        for (; march > 0.0001f && step < 32; ++step) {
            hx = ox + d.x * total;
            hy = oy + d.y * total;
            hz = oz + d.z * total;
            march = f_spikey<1>(L, f, hx, hy, hz);
            total += march * (0.05f * kPI);
        }
        steps = step;
        const F4 n = spikey_grad<1>(L, f, hx, hy, hz, march, f.nOffs);
        // the rim factor max(1, min(0, x)) is 1 for every x
        const float diffuse = n.z;
        const float spec = powf(smax(0.f, dot3(n, d)), f.specPow);
        const F4 a{(f.diff[0] + spec) * diffuse, (f.diff[1] + spec) * diffuse,
                   (f.diff[2] + spec) * diffuse, (f.diff[3] + spec) * diffuse};
        c = gamma_adj(lerp_to(a, 1.f, exp_fog(hz - oz, kGoldenRatio * 0.1f)), f.gamma);
    }
    return to_pixel(c);
}

// fSinMap (code/shadertoy.cpp:878-898)
__device__ __forceinline__ float f_sinmap(const Lut &L, float x, float y, float z)
{
    const float zMod = z * 0.314f;
    const float pathCos = L.cos(zMod);
    const float pathCos2 = L.cos(zMod + ((2.f * kPI) / 4.f)) * kGoldenRatio;
    const float pX = x - (pathCos2 * 2.f - pathCos * 1.5f);
    const float pY = y - (pathCos * 3.14f + pathCos2);
    const float aX = pX * 0.315f * 1.25f + L.sin(z * (0.814f * 1.25f));
    const float aY = pY * 0.315f * 1.25f + L.sin(pX * (0.814f * 1.25f));
    const float aZ = z * 0.315f * 1.25f + L.sin(pY * (0.814f * 1.25f));
    const float cX = L.cos(aX), cY = L.cos(aY), cZ = L.cos(aZ);
    return (sqrtf(cX * cX + cY * cY + cZ * cZ) - 1.025f) * 1.33f;
}

// RenderSinMap_2x2 per pixel (code/shadertoy.cpp:935-970)
__device__ uint32_t sinuses_pixel(const Lut &L, const StGeom &g, const SinusesF &f, unsigned iX,
                                  unsigned iY, int &steps)
{
    const float ox = f.origin[0], oy = f.origin[1], oz = f.origin[2];
    float u, v;
    to_uv(g, iX, iY, 2.f, u, v);
    float dx = (u + f.offsX) * g.aspect, dy = v;
    rot(f.rollCos, f.rollSin, dx, dy);
    const F4 d = vnorm3(dx, dy, 0.314f);
    float march = 1.f, total = 0.f, hx = 0.f, hy = 0.f, hz = 0.f;
    int step = 0;
    for (; march > 0.01f && step < 32; ++step) {
        hx = ox + d.x * total;
        hy = oy + d.y * total;
        hz = oz + d.z * total;
        march = f_sinmap(L, hx, hy, hz);
        total += march * 0.814f;
    }
    steps = step;
    const float nx = march - f_sinmap(L, hx + 0.2f, hy, hz);
    const float ny = march - f_sinmap(L, hx, hy + 0.2f, hz);
    const float nz = march - f_sinmap(L, hx, hy, hz + 0.2f);
    const F4 n = vnorm3(nx, ny, nz);
    float diffuse = n.z * 0.7f + 0.3f * n.y;
    diffuse = 0.2f + 0.8f * diffuse;
    const float spec = powf(dot3(n, d), f.specPow);
    const F4 a{(f.diff[0] + spec) * diffuse, (f.diff[1] + spec) * diffuse,
               (f.diff[2] + spec) * diffuse, (f.diff[3] + spec) * diffuse};
    return to_pixel(gamma_adj(lerp_to(a, 1.f, exp_fog(hz - oz, 0.03f)), f.gamma));
}

// fLaura (code/shadertoy.cpp:993-996)
__device__ __forceinline__ float f_laura(const Lut &L, float x, float y, float z)
{
    return L.cos(x) + L.cos(y) + L.cos(z) + 1.f;
}

// RenderLaura_2x2 per pixel (code/shadertoy.cpp:1040-1095)
__device__ uint32_t laura_pixel(const Lut &L, const StGeom &g, const LauraF &f, unsigned iX,
                                unsigned iY)
{
    const float ox = f.origin[0], oy = f.origin[1], oz = f.origin[2];
    float u, v;
    to_uv(g, iX, iY, 2.f, u, v);
    float dx = u * g.aspect, dy = v, dz = kPI;
    rot(f.yawCos, f.yawSin, dx, dz);
    rot_x(f.pitchCos, f.pitchSin, dy, dz);
    rot(f.rollCos, f.rollSin, dx, dy);
    const F4 d = vnorm3(dx, dy, dz);
    float march = 0.f, total = 0.f, hx = 0.f, hy = 0.f, hz = 0.f;
    for (int step = 0; step < 32; ++step) {
        hx = ox + d.x * total;
        hy = oy + d.y * total;
        hz = oz + d.z * total;
        march = f_laura(L, hx, hy, hz);
        total += march * 0.5f;
    }
    // LauraNormal (code/shadertoy.cpp:998-1010)
    const float nOffs = 0.1628f;
    const float nx = f_laura(L, hx + nOffs, hy, hz) - march;
    const float ny = f_laura(L, hx, hy + nOffs, hz) - march;
    const float nz = f_laura(L, hx, hy, hz + nOffs) - march;
    const F4 n = vnorm3(nx, ny, nz);
    const float lx = ox - d.x, ly = oy - d.y, lz = oz - d.z;
    const F4 ld = vnorm3(lx - hx, ly - hy, lz - hz);
    const float diffuse = smax(0.3f, dot3(n, ld));
    // Specular (code/shadertoy-util.h:251-279): four-lane V and H
    F4 V{ox - hx, oy - hy, oz - hz, 0.f - 0.f};
    const float rV = 1.f / sqrtf(dp4(V, V));
    V = F4{V.x * rV, V.y * rV, V.z * rV, V.w * rV};
    F4 H{ld.x + V.x, ld.y + V.y, ld.z + V.z, ld.w + V.w};
    const float rH = 1.f / sqrtf(dp4(H, H));
    H = F4{H.x * rH, H.y * rH, H.z * rH, H.w * rH};
    const float cosAng = dp4(n, H);
    const float spec = (__float_as_uint(cosAng) >> 31) == 0 ? powf(cosAng, 4.f) : 0.f;
    // the rim factor max(1, min(0, x)) is 1 for every x
    const float ds = diffuse + spec;
    // Q3_rsqrtf<2> (code/q3-rsqrt.h:23-41)
    const float sd = spec + diffuse;
    const float half = 0.5f * sd;
    float q = __int_as_float(0x5f3759df - (__float_as_int(sd) >> 1));
    q = q * (1.5f - half * q * q);
    q = q * (1.5f - half * q * q);
    const F4 a{f.diff[0] * ds, f.diff[1] * ds, f.diff[2] * ds, f.diff[3] * ds};
    return to_pixel(gamma_adj(lerp_to(a, q, exp_fog(hz - oz, 0.001f)), 1.44f));
}

// ftofp24 (code/util.h:195-197) with x86 cvttss2si's INT_MIN for NaN and
// out-of-range values
__device__ __forceinline__ int fp24(float v)
{
    const float s = v * 256.f;
    return (s < 2147483648.f && s >= -2147483648.f) ? (int)s : INT_MIN;
}

// bsamp_prepUVs(.., 1023, 10, ..) + bsamp32_32f (code/bilinear.h:10-31,
// 86-134) on a 1024x1024 texture, as four channel floats
__device__ __forceinline__ F4 sample1024(const uint32_t *__restrict__ tex, int U, int V)
{
    const unsigned u0 = (unsigned)(U >> 8) & 1023u, u1 = ((unsigned)(U >> 8) + 1u) & 1023u;
    const unsigned v0 = ((unsigned)(V >> 8) & 1023u) << 10;
    const unsigned v1 = (((unsigned)(V >> 8) + 1u) & 1023u) << 10;
    const uint32_t p = ckd_bilerp32(tex[u0 + v0], tex[u1 + v0], tex[u0 + v1], tex[u1 + v1],
                                    (uint32_t)U & 0xff, (uint32_t)V & 0xff);
    return F4{(float)(p & 0xff), (float)((p >> 8) & 0xff), (float)((p >> 16) & 0xff),
              (float)(p >> 24)};
}

// RenderTunnelMap_2x2 per pixel (code/shadertoy.cpp:780-825)
__device__ void fdtunnel_pixel(const Lut &L, const StGeom &g, const TunnelF &f,
                               const uint32_t *__restrict__ tex, const uint32_t *__restrict__ texHi,
                               unsigned iX, unsigned iY, uint32_t &color, uint32_t &glow)
{
    float u, v;
    to_uv(g, iX, iY, 2.f, u, v);
    float dx = u, dy = v, dz = 1.f;
    rot_x(f.pitchCos, f.pitchSin, dy, dz);
    rot(f.rollCos, f.rollSin, dx, dy);
    const F4 d = vnorm3(dx, dy, dz);
    float A = d.x * d.x + d.y * d.y;
    A += f.flowerScale * L.cos(atan2f(d.y, d.x) * f.flowerFreq + f.flowerPhase);
    const float absX = fabsf(d.x), absY = fabsf(d.y);
    const float box = absX > absY ? absX : absY;
    A = A + (box - A) * f.boxT; // smoothstepf (code/Std3DMath-stripped/Math.h:59-63)
    A += kEpsilon;
    A = 1.f / A;
    const float T = f.radius * A;
    const float T2 = T * 0.912f;
    const float U = atan2f(d.y * T, d.x * T) / kPI;
    const float V = d.z * T + f.vShift;
    const float U2 = atan2f(d.y * T2, d.x * T2) / kPI;
    const float V2 = d.z * T2 + f.vShift;
    const float shade = smax(0.f, smin(1.f, 1.f - expf(-0.006f * T * T)));
    const F4 c = sample1024(tex, fp24(U * f.uMul), fp24(V * f.vMul));
    const F4 h = sample1024(texHi, fp24(U2 * f.uMul), fp24(V2 * f.vMul));
    color = to_pixel_noconv(lerp_to(c, f.fog0, shade));
    glow = to_pixel_noconv(lerp_to(h, f.fog1, shade));
}

// steps: while non-null, every pixel's march steps (loop iterations; the
// fixed-length marches count their length, the tunnel none) are added to
// *steps, one atomic per wave
template <int kFx>
__global__ __launch_bounds__(256) void k_st_map(const StFrame *__restrict__ frames,
                                                uint32_t *__restrict__ maps,
                                                uint32_t *__restrict__ glows, StGeom g,
                                                const float *__restrict__ lutG,
                                                const uint32_t *__restrict__ tex,
                                                const uint32_t *__restrict__ texHi,
                                                unsigned long long *__restrict__ steps)
{
    __shared__ float lut[kLutN + 3];
    for (unsigned i = threadIdx.x; i < kLutN; i += 256) lut[i] = lutG[i];
    __syncthreads();
    const Lut L{lut, g.lutScale, g.halfPi};
    // four 8x8 wave tiles per 16x16 block
    const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned iX = blockIdx.x * kTile + (wave & 1) * 8 + (lane & 7);
    const unsigned iY = blockIdx.y * kTile + (wave >> 1) * 8 + (lane >> 3);
    int n = 0;
    if (iX < g.mapW && iY < g.mapH) {
        const StFrame &F = frames[blockIdx.z];
        const size_t at = (size_t)blockIdx.z * g.mapW * g.mapH + (size_t)iY * g.mapW + iX;
        if constexpr (kFx == FX_PLASMA) {
            maps[at] = plasma_pixel(L, g, F.p, iX, iY);
            n = 24;
        }
        if constexpr (kFx == FX_NAUTILUS) maps[at] = nautilus_pixel(L, g, F.n, iX, iY, n);
        if constexpr (kFx == FX_SPIKEY) maps[at] = spikey_pixel(L, g, F.s, iX, iY, n);
        if constexpr (kFx == FX_SINUSES) maps[at] = sinuses_pixel(L, g, F.si, iX, iY, n);
        if constexpr (kFx == FX_LAURA) {
            maps[at] = laura_pixel(L, g, F.l, iX, iY);
            n = 32;
        }
        if constexpr (kFx == FX_FDTUNNEL) {
            uint32_t c, h;
            fdtunnel_pixel(L, g, F.t, tex, texHi, iX, iY, c, h);
            maps[at] = c;
            glows[at] = h;
        }
    }
    if (steps) {
        unsigned long long v = (unsigned long long)n;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) atomicAdd(steps, v);
    }
}

// ---- host: per-frame set-up in the reference's FP32 order ---------------

struct HostLut {
    const float *t;
    float scale, halfPi;
    float cos(float a) const { return lut_cos(t, a, scale); }
    float sin(float a) const { return lut_sin(t, a, scale, halfPi); }
};

// MichielPal (code/shadertoy-util.h:193-198)
F4 michiel_pal(const HostLut &L, float phase)
{
    return F4{.1f - L.cos(phase / 3.f) / (19.f * 0.5f), .1f, .1f + L.cos(phase / 14.f) / 4.f, 0.f};
}

// easeInOutElasticf (code/synth-math-easings.h:190-202)
float ease_in_out_elastic(float x)
{
    const float c5 = (2.f * kPI) / 4.5f;
    return (0.f == x)   ? 0.f
           : (1.f == x) ? 1.f
           : (x < 0.5f) ? -(powf(2.f, 20.f * x - 10.f) * sinf((20.f * x - 11.125f) * c5)) * 0.5f
                        : (powf(2.f, -20.f * x + 10.f) * sinf((20.f * x - 11.125f) * c5)) * 0.5f + 1.f;
}

void set3(float *d, float x, float y, float z)
{
    d[0] = x;
    d[1] = y;
    d[2] = z;
}

// code/shadertoy.cpp:213-228
PlasmaF plasma_frame(const HostLut &L, const ckd_plasma_params &p)
{
    PlasmaF f{};
    const F4 a = desaturate(michiel_pal(L, p.hue), p.desaturation);
    const F4 b = desaturate(a, 0.8f);
    set3(f.colA, a.x, a.y, a.z);
    set3(f.colB, b.x, b.y, b.z);
    f.time = p.time * p.speed;
    const float angle = f.time * 0.314f * 0.5f;
    f.dirCos = L.cos(angle);
    f.dirSin = L.sin(angle);
    f.gamma = p.gamma;
    return f;
}

// code/shadertoy.cpp:303-327
NautilusF nautilus_frame(const HostLut &L, const ckd_nautilus_params &p)
{
    NautilusF f{};
    f.time = p.time * p.speed;
    set3(f.g, f.time * 0.125f, f.time / 9.f, L.cos(f.time * 0.1428f));
    const F4 col{.1f - L.cos(p.hue / 3.f) / 19.f, .1f, .1f + L.cos(p.hue / 14.f) / 8.f, 0.f};
    const F4 d = desaturate(col, p.desaturation);
    set3(f.diff, d.x, d.y, d.z);
    const float rollT = p.roll * f.time;
    f.rollCos = L.cos(rollT);
    f.rollSin = L.sin(rollT);
    f.cosHitOffs = L.cos(f.time * 0.314f * 0.5f);
    f.funkCos = L.cos(f.time * kGoldenRatio * 0.1f);
    return f;
}

// code/shadertoy.cpp:436-459 (close), :527-547 (distant), :600-610 and
// :712-725 (specular-only, warmup = 1 + track)
SpikeyF spikey_frame(const HostLut &L, const ckd_spikey_params &p, bool close, float aspect)
{
    SpikeyF f{};
    const F4 d = desaturate(michiel_pal(L, p.hue), p.desaturation);
    f.diff[0] = d.x;
    f.diff[1] = d.y;
    f.diff[2] = d.z;
    f.diff[3] = d.w;
    f.rollCos = L.cos(p.roll);
    f.rollSin = L.sin(p.roll);
    f.specPow = p.specular;
    f.gamma = p.gamma;
    const float t = p.speed * p.time;
    if (close) {
        f.mode = 2;
        f.xOffs = p.closeXOffs;
        f.yOffs = p.closeYOffs;
        f.zDir = 1.f + ease_in_out_elastic(p.closeZOffs) * p.closeZOffsScale;
        f.nOffs = p.closeNormalGrain;
        const bool aspectMul = 0 != (int)roundf(p.closeAspectMul);
        set3(f.g, t, 16.f * p.closeScale,
             aspectMul ? aspect * 22.f * p.closeScale : 22.f * p.closeScale);
        set3(f.origin, 0.2f, 0.f, -2.23f);
    } else if (0.f == p.distWarmup) {
        f.mode = 0;
        f.xOffs = p.distXOffs;
        f.yOffs = p.distYOffs;
        set3(f.g, t, 16.f, 16.f);
        set3(f.origin, 0.f, 0.f, -2.614f + p.distZOffs);
    } else {
        f.mode = 1;
        f.warmup = 1.f + p.distWarmup;
        set3(f.g, t, 8.f, 16.f);
        set3(f.origin, 0.f, 0.f, -3.314f);
    }
    return f;
}

// code/shadertoy.cpp:904-924 with fSinPath (:869-875)
SinusesF sinuses_frame(const HostLut &L, const ckd_sinuses_params &p)
{
    SinusesF f{};
    f.specPow = 1.f + p.specular;
    f.rollCos = L.cos(p.roll);
    f.rollSin = L.sin(p.roll);
    f.offsX = p.offsX;
    f.gamma = p.gamma;
    const F4 d = desaturate(michiel_pal(L, p.hue), p.desaturation);
    f.diff[0] = d.x;
    f.diff[1] = d.y;
    f.diff[2] = d.z;
    f.diff[3] = d.w;
    const float time = p.time * p.speed;
    const float timeMod = time * 0.314f;
    const float sine = L.sin(timeMod), cosine = L.cos(timeMod);
    set3(f.origin, sine * 2.f * kGoldenRatio - cosine * 1.5f, cosine * 3.14f + sine * kGoldenRatio,
         time);
    return f;
}

// code/shadertoy.cpp:1015-1029
LauraF laura_frame(const HostLut &L, const ckd_laura_params &p)
{
    LauraF f{};
    const F4 d = desaturate(michiel_pal(L, p.hue), p.saturate);
    f.diff[0] = d.x;
    f.diff[1] = d.y;
    f.diff[2] = d.z;
    f.diff[3] = d.w;
    set3(f.origin, 0.f, 0.f, p.speed * p.time);
    f.yawCos = L.cos(p.yaw);
    f.yawSin = L.sin(p.yaw);
    f.pitchCos = L.cos(p.pitch);
    f.pitchSin = L.sin(p.pitch);
    const float rollT = p.roll * p.time;
    f.rollCos = L.cos(rollT);
    f.rollSin = L.sin(rollT);
    return f;
}

// code/shadertoy.cpp:752-768
TunnelF fdtunnel_frame(const HostLut &L, const ckd_fdtunnel_params &p)
{
    TunnelF f{};
    float t = p.boxy;
    t = t * t * (3.f - 2.f * t);
    f.boxT = t;
    f.flowerScale = p.flowerScale;
    f.flowerFreq = p.flowerFreq;
    f.flowerPhase = p.flowerPhase * p.time;
    const float roll = p.roll * p.time, pitch = p.pitch * p.time;
    f.pitchCos = L.cos(pitch);
    f.pitchSin = L.sin(pitch);
    f.rollCos = L.cos(roll);
    f.rollSin = L.sin(roll);
    f.radius = p.radius;
    f.uMul = p.mulU;
    f.vMul = p.mulV;
    const float time = p.time * p.speed;
    f.vShift = time * p.speed;
    f.fog0 = p.fog1;
    f.fog1 = p.fog2;
    return f;
}

// BoxBlurScale (code/deprecated/boxblur.h:13-19)
float box_blur_scale(float strength)
{
    if (strength != 0.f) {
        const float lo = (strength < 100.f) ? strength : 100.f; // std::min<float>(100, s)
        strength = ((1.f < lo) ? lo : 1.f) * 0.01f;             // std::max<float>(1, .)
    }
    return strength;
}

bool finite_params(const float *p, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (!(p[i] == p[i]) || p[i] > 3.0e38f || p[i] < -3.0e38f) return false;
    return true;
}

} // namespace

namespace ckd {

// compositing blends on device buffers (k_composite.hip)
ckd_status comp_mix32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned n, uint32_t alpha);
ckd_status comp_softlight32aa(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned n,
                              uint32_t iA);
ckd_status comp_overlay32a(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned n);
uint32_t softlight_alpha(float alpha);

// Renders nFrames FX maps of one effect (frames: host array of per-frame
// constants, already converted).
static ckd_status st_launch(ckd_ctx *ctx, int fx, const StFrame *frames, unsigned nFrames,
                            uint32_t *maps, uint32_t *glows, unsigned resX, unsigned resY)
{
    if (nFrames == 0) return CKD_OK;
    void *dFrames = nullptr;
    ckd_status s = stage(ctx, frames, sizeof(StFrame) * nFrames, &dFrames);
    if (s != CKD_OK) return s;
    StGeom g{};
    g.mapW = resX / 2 + 4;
    g.mapH = resY / 2 + 4;
    g.invW = 1.f / (float)g.mapW;
    g.invH = 1.f / (float)g.mapH;
    g.aspect = (float)resY / (float)resX;
    g.oneOverAspect = 1.f / g.aspect;
    g.lutScale = (1.f / (2.f * kPI)) * 2048.f;
    g.halfPi = kPI * 0.5f;
    const dim3 grid(ckd_div_up(g.mapW, kTile), ckd_div_up(g.mapH, kTile), nFrames);
    const StFrame *df = static_cast<const StFrame *>(dFrames);
    KernelTimer timer(ctx, CKD_KT_SHADERTOY);
    switch (fx) {
    case FX_PLASMA:
        k_st_map<FX_PLASMA><<<grid, 256, 0, ctx->stream>>>(df, maps, glows, g, ctx->cosLut, nullptr, nullptr, ctx->stSteps);
        break;
    case FX_NAUTILUS:
        k_st_map<FX_NAUTILUS><<<grid, 256, 0, ctx->stream>>>(df, maps, glows, g, ctx->cosLut, nullptr, nullptr, ctx->stSteps);
        break;
    case FX_SPIKEY:
        k_st_map<FX_SPIKEY><<<grid, 256, 0, ctx->stream>>>(df, maps, glows, g, ctx->cosLut, nullptr, nullptr, ctx->stSteps);
        break;
    case FX_SINUSES:
        k_st_map<FX_SINUSES><<<grid, 256, 0, ctx->stream>>>(df, maps, glows, g, ctx->cosLut, nullptr, nullptr, ctx->stSteps);
        break;
    case FX_LAURA:
        k_st_map<FX_LAURA><<<grid, 256, 0, ctx->stream>>>(df, maps, glows, g, ctx->cosLut, nullptr, nullptr, ctx->stSteps);
        break;
    case FX_FDTUNNEL:
        k_st_map<FX_FDTUNNEL><<<grid, 256, 0, ctx->stream>>>(df, maps, glows, g, ctx->cosLut, ctx->fdTex,
                                                              ctx->fdTexHi, ctx->stSteps);
        break;
    default:
        return fail(CKD_EINVAL, "shadertoy: unknown effect %d", fx);
    }
    return check_launch("st_map");
}

// FX-map scratch of the Draw calls (maps, then glow maps), grown on demand
static ckd_status st_scratch(ckd_ctx *ctx, size_t pixels, uint32_t **out)
{
    if (ctx->stScratchCap < pixels) {
        if (ctx->stScratch) {
            CKD_HIP(hipStreamSynchronize(ctx->stream));
            CKD_HIP(hipFree(ctx->stScratch));
            ctx->stScratch = nullptr;
            ctx->stScratchCap = 0;
        }
        if (ckd::dev_alloc(&ctx->stScratch, pixels * sizeof(uint32_t)) != hipSuccess)
            return fail(CKD_ENOMEM, "shadertoy: cannot allocate %zu map pixels", pixels);
        ctx->stScratchCap = pixels;
    }
    *out = ctx->stScratch;
    return CKD_OK;
}

} // namespace ckd

namespace {

ckd_status check_geometry(const char *fn, unsigned resX, unsigned resY, unsigned nFrames,
                          const void *out, const void *frames)
{
    CKD_REQUIRE(out && frames, "%s: null buffer", fn);
    CKD_REQUIRE(resX >= 64 && resY >= 64 && resX <= 8192 && resY <= 8192 && resX % 8 == 0 &&
                    resY % 2 == 0,
                "%s: output %ux%u: need 64..8192, width a multiple of 8, even height", fn, resX,
                resY);
    CKD_REQUIRE(nFrames <= 65535, "%s: at most 65535 frames per call", fn);
    return CKD_OK;
}

// every parameter of every frame finite
template <typename P>
ckd_status check_params(const char *fn, const P *frames, unsigned nFrames)
{
    for (unsigned k = 0; k < nFrames; ++k)
        CKD_REQUIRE(finite_params(reinterpret_cast<const float *>(&frames[k]), sizeof(P) / sizeof(float)),
                    "%s: frame %u: parameters must be finite", fn, k);
    return CKD_OK;
}

HostLut host_lut(const ckd_ctx *ctx)
{
    return HostLut{ctx->hostLut, (1.f / (2.f * kPI)) * 2048.f, kPI * 0.5f};
}

// Converts the parameter structs to per-frame constants and renders the
// maps; `close` selects the spike's close-up renderer.
template <typename P>
ckd_status st_maps(ckd_ctx *ctx, int fx, const P *params, unsigned nFrames, uint32_t *maps,
                   uint32_t *glows, unsigned resX, unsigned resY, bool close)
{
    std::vector<StFrame> frames(nFrames);
    const HostLut L = host_lut(ctx);
    const float aspect = (float)resY / (float)resX;
    for (unsigned k = 0; k < nFrames; ++k) {
        const P &p = params[k];
        std::memset(&frames[k], 0, sizeof(StFrame));
        if constexpr (std::is_same<P, ckd_plasma_params>::value) frames[k].p = plasma_frame(L, p);
        if constexpr (std::is_same<P, ckd_nautilus_params>::value) frames[k].n = nautilus_frame(L, p);
        if constexpr (std::is_same<P, ckd_spikey_params>::value) frames[k].s = spikey_frame(L, p, close, aspect);
        if constexpr (std::is_same<P, ckd_sinuses_params>::value) frames[k].si = sinuses_frame(L, p);
        if constexpr (std::is_same<P, ckd_laura_params>::value) frames[k].l = laura_frame(L, p);
        if constexpr (std::is_same<P, ckd_fdtunnel_params>::value) frames[k].t = fdtunnel_frame(L, p);
    }
    return ckd::st_launch(ctx, fx, frames.data(), nFrames, maps, glows, resX, resY);
}

// clampf(lo, hi, v) = std::max(lo, std::min(hi, v)) (code/Std3DMath-stripped/Math.h:37-40)
float clampf_(float lo, float hi, float v)
{
    const float m = (v < hi) ? v : hi;
    return (lo < m) ? m : lo;
}

// the close-up's mix-blur layer on one FX map (code/shadertoy.cpp:665-701):
// blend of the two blur maps (Mix32), box blurs in place, SoftLight32AA of
// a blurred copy of the map by the blur map, Overlay32A back onto the map;
// work = two FX-map scratch images
ckd_status spikey_mix_blur(ckd_ctx *ctx, const ckd_spikey_params &p, uint32_t *map,
                           unsigned mapW, unsigned mapH, uint32_t *blurMap, uint32_t *fx1)
{
    const float mbOpacity = clampf_(0.f, 1.f, p.closeMixBlurOpacity);
    if (!(mbOpacity > 0.f)) return CKD_OK;
    const unsigned n = mapW * mapH;
    const size_t bytes = (size_t)n * sizeof(uint32_t);
    const float mbMap = clampf_(0.f, 1.f, p.closeMixBlurMap);
    const float mbBlur = clampf_(0.f, 100.f, p.closeMixBlur);
    const float mbMapBlur = clampf_(0.f, 100.f, p.closeMixMapBlur);
    const uint32_t *maps = ctx->spikeBlurMaps;
    ckd_status s;
    CKD_HIP(hipMemcpyAsync(blurMap, mbMap == 1.f ? maps + n : maps, bytes, hipMemcpyDeviceToDevice,
                           ctx->stream));
    if (mbMap != 0.f && mbMap != 1.f &&
        (s = ckd::comp_mix32(ctx, blurMap, maps + n, n, (uint8_t)(unsigned)(mbMap * 255.f))) != CKD_OK)
        return s;
    CKD_HIP(hipMemcpyAsync(fx1, map, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    if (mbMapBlur >= 1.f &&
        (s = ckd_blur07(ctx, blurMap, blurMap, mapW, mapH, box_blur_scale(mbMapBlur))) != CKD_OK)
        return s;
    if (mbBlur >= 1.f && (s = ckd_blur07(ctx, fx1, fx1, mapW, mapH, box_blur_scale(mbBlur))) != CKD_OK)
        return s;
    if ((s = ckd::comp_softlight32aa(ctx, fx1, blurMap, n,
                                     ckd::softlight_alpha(tanhf(mbBlur + mbOpacity)))) != CKD_OK)
        return s;
    return ckd::comp_overlay32a(ctx, map, fx1, n);
}

// Map(s) into context scratch, then each frame's post-processing and
// Fx_Blit_2x2 into dst, as the effect's Draw function.
template <typename P>
ckd_status st_draw(ckd_ctx *ctx, int fx, const P *params, unsigned nFrames, uint32_t *dst,
                   unsigned resX, unsigned resY, bool close)
{
    const unsigned mapW = resX / 2 + 4, mapH = resY / 2 + 4;
    const size_t mapPx = (size_t)mapW * mapH, outPx = (size_t)resX * resY;
    const bool glow = fx == FX_FDTUNNEL;
    // maps, then glow maps (tunnel) or the mix-blur layer's two work images
    uint32_t *scratch = nullptr;
    ckd_status s = ckd::st_scratch(ctx, mapPx * nFrames * (glow ? 2 : 1) + 2 * mapPx, &scratch);
    if (s != CKD_OK) return s;
    uint32_t *maps = scratch, *glows = glow ? scratch + mapPx * nFrames : nullptr;
    uint32_t *work = scratch + mapPx * nFrames * (glow ? 2 : 1);
    if ((s = st_maps(ctx, fx, params, nFrames, maps, glows, resX, resY, close)) != CKD_OK) return s;
    CKD_REQUIRE(((uintptr_t)dst & 7) == 0, "st_draw: dst must be 8-byte aligned");
    // The frames are independent, so each stage runs for all of them at once:
    // the map's own post-processing, one Fx_Blit_2x2 launch, the output's
    // blur.  In-place blurs of many frames are one launch per pass.
    std::vector<ckd::B07Job> blurs;
    auto run_blurs = [&](unsigned W, unsigned H, bool columns) -> ckd_status {
        if (blurs.empty()) return CKD_OK;
        bool vec4 = true;
        for (const ckd::B07Job &j : blurs) vec4 = vec4 && (reinterpret_cast<uintptr_t>(j.img) & 15) == 0;
        void *dev = nullptr;
        ckd_status r = ckd::stage(ctx, blurs.data(), blurs.size() * sizeof(ckd::B07Job), &dev);
        const ckd::B07Job *jobs = static_cast<const ckd::B07Job *>(dev);
        const unsigned n = (unsigned)blurs.size();
        if (r == CKD_OK) r = ckd::blur07_inplace_jobs(ctx, blurs.data(), jobs, n, W, H, false, vec4);
        if (r == CKD_OK && columns) r = ckd::blur07_inplace_jobs(ctx, blurs.data(), jobs, n, W, H, true, false);
        blurs.clear();
        return r;
    };
    if constexpr (std::is_same<P, ckd_spikey_params>::value) {
        // Spikey_Draw, code/shadertoy.cpp:663-725: the close-up's mix-blur
        // layer (its two work images are shared, so frame after frame), or
        // the distant spike's HorizontalBoxBlur32 of the map in place
        for (unsigned k = 0; k < nFrames; ++k) {
            uint32_t *map = maps + mapPx * k;
            if (close) {
                if ((s = spikey_mix_blur(ctx, params[k], map, mapW, mapH, work, work + mapPx)) != CKD_OK)
                    return s;
            } else if (params[k].distWarmup != 0.f) {
                blurs.emplace_back();
                if ((s = ckd::b07_h_job(map, mapW, mapH, box_blur_scale(1.f + params[k].distWarmup),
                                        blurs.back())) != CKD_OK)
                    return s;
            }
        }
        if ((s = run_blurs(mapW, mapH, false)) != CKD_OK) return s;
    }
    if constexpr (std::is_same<P, ckd_fdtunnel_params>::value) {
        // Tunnel_Draw, code/shadertoy.cpp:842-853: the glow map's BoxBlur32,
        // then Add32 of the glow onto the map
        bool allLit = true;
        for (unsigned k = 0; k < nFrames; ++k) {
            const float lb = params[k].litBlur;
            const float litBlur = (0.f < ((lb < 100.f) ? lb : 100.f)) ? ((lb < 100.f) ? lb : 100.f) : 0.f;
            const bool lit = 0 != (int)roundf(params[k].litTiles);
            allLit = allLit && lit;
            if (lit && litBlur >= 1.f) {
                blurs.emplace_back();
                if ((s = ckd::b07_job(glows + mapPx * k, mapW, mapH, box_blur_scale(litBlur),
                                      blurs.back())) != CKD_OK)
                    return s;
            }
        }
        if ((s = run_blurs(mapW, mapH, true)) != CKD_OK) return s;
        const size_t all = mapPx * nFrames;
        if (allLit && all <= 0xffffffffu) {
            // maps and glow maps are both back to back: one Add32
            if ((s = ckd::launch_blit_add(ctx, maps, glows, (unsigned)all, (unsigned)all, (unsigned)all,
                                          1)) != CKD_OK)
                return s;
        } else {
            for (unsigned k = 0; k < nFrames; ++k)
                if (0 != (int)roundf(params[k].litTiles) &&
                    (s = ckd_add32(ctx, maps + mapPx * k, glows + mapPx * k, (unsigned)mapPx)) != CKD_OK)
                    return s;
        }
    }
    if ((s = ckd::fx_blit_frames(ctx, dst, maps, resX, resY, nFrames)) != CKD_OK) return s;
    if constexpr (std::is_same<P, ckd_nautilus_params>::value) {
        // Nautilus_Draw, code/shadertoy.cpp:385-396: BoxBlur32 of the output
        // in place
        for (unsigned k = 0; k < nFrames; ++k) {
            const float blur = box_blur_scale(params[k].blur);
            if (0.f != blur) {
                blurs.emplace_back();
                if ((s = ckd::b07_job(dst + outPx * k, resX, resY, blur, blurs.back())) != CKD_OK) return s;
            }
        }
        return run_blurs(resX, resY, true);
    }
    return CKD_OK;
}

} // namespace

extern "C" {

ckd_status ckd_shadertoy_step_counter(ckd_ctx *ctx, unsigned long long *counter)
{
    CKD_REQUIRE_CTX(ctx);
    ctx->stSteps = counter;
    return CKD_OK;
}

#define CKD_ST_ENTRY(name, P, FX)                                                               \
    ckd_status ckd_##name##_map(ckd_ctx *ctx, uint32_t *maps, unsigned resX, unsigned resY,     \
                                const P *frames, unsigned nFrames)                              \
    {                                                                                           \
        ckd_status s = check_geometry("ckd_" #name "_map", resX, resY, nFrames, maps, frames);  \
        if (s != CKD_OK || (s = check_params("ckd_" #name "_map", frames, nFrames)) != CKD_OK)  \
            return s;                                                                           \
        CKD_REQUIRE_CTX(ctx);                                                                   \
        return st_maps(ctx, FX, frames, nFrames, maps, nullptr, resX, resY, false);             \
    }                                                                                           \
    ckd_status ckd_##name##_draw(ckd_ctx *ctx, uint32_t *dst, unsigned resX, unsigned resY,     \
                                 const P *frames, unsigned nFrames)                             \
    {                                                                                           \
        ckd_status s = check_geometry("ckd_" #name "_draw", resX, resY, nFrames, dst, frames);  \
        if (s != CKD_OK || (s = check_params("ckd_" #name "_draw", frames, nFrames)) != CKD_OK) \
            return s;                                                                           \
        CKD_REQUIRE_CTX(ctx);                                                                   \
        return st_draw(ctx, FX, frames, nFrames, dst, resX, resY, false);                       \
    }

CKD_ST_ENTRY(plasma, ckd_plasma_params, FX_PLASMA)
CKD_ST_ENTRY(nautilus, ckd_nautilus_params, FX_NAUTILUS)
CKD_ST_ENTRY(sinuses, ckd_sinuses_params, FX_SINUSES)
CKD_ST_ENTRY(laura, ckd_laura_params, FX_LAURA)

#undef CKD_ST_ENTRY

ckd_status ckd_spikey_map(ckd_ctx *ctx, uint32_t *maps, unsigned resX, unsigned resY,
                          const ckd_spikey_params *frames, unsigned nFrames, int close)
{
    ckd_status s = check_geometry("ckd_spikey_map", resX, resY, nFrames, maps, frames);
    if (s != CKD_OK || (s = check_params("ckd_spikey_map", frames, nFrames)) != CKD_OK) return s;
    CKD_REQUIRE_CTX(ctx);
    return st_maps(ctx, FX_SPIKEY, frames, nFrames, maps, nullptr, resX, resY, close != 0);
}

ckd_status ckd_spikey_draw(ckd_ctx *ctx, uint32_t *dst, unsigned resX, unsigned resY,
                           const ckd_spikey_params *frames, unsigned nFrames, int close)
{
    ckd_status s = check_geometry("ckd_spikey_draw", resX, resY, nFrames, dst, frames);
    if (s != CKD_OK || (s = check_params("ckd_spikey_draw", frames, nFrames)) != CKD_OK) return s;
    CKD_REQUIRE_CTX(ctx);
    if (close)
        for (unsigned k = 0; k < nFrames; ++k) {
            // saturatef(opacity) > 0 turns the layer on (code/shadertoy.cpp:668-671)
            CKD_REQUIRE(!(frames[k].closeMixBlurOpacity > 0.f) ||
                            (ctx->spikeBlurMaps && ctx->spikeBlurW == resX && ctx->spikeBlurH == resY),
                        "ckd_spikey_draw: frame %u: the close-up's mix-blur layer needs blur maps "
                        "for %ux%u (ckd_spikey_blurmaps)", k, resX, resY);
        }
    return st_draw(ctx, FX_SPIKEY, frames, nFrames, dst, resX, resY, close != 0);
}

ckd_status ckd_spikey_blurmaps(ckd_ctx *ctx, const uint32_t *map1, const uint32_t *map2,
                               unsigned resX, unsigned resY)
{
    CKD_REQUIRE(map1 && map2, "ckd_spikey_blurmaps: null map");
    CKD_REQUIRE(resX >= 64 && resY >= 64 && resX <= 8192 && resY <= 8192 && resX % 8 == 0 &&
                    resY % 2 == 0,
                "ckd_spikey_blurmaps: output %ux%u: need 64..8192, width a multiple of 8, even height",
                resX, resY);
    CKD_REQUIRE_CTX(ctx);
    const size_t px = (size_t)(resX / 2 + 4) * (resY / 2 + 4), bytes = px * sizeof(uint32_t);
    if (ctx->spikeBlurMaps) {
        CKD_HIP(hipStreamSynchronize(ctx->stream));
        CKD_HIP(hipFree(ctx->spikeBlurMaps));
        ctx->spikeBlurMaps = nullptr;
    }
    if (ckd::dev_alloc(&ctx->spikeBlurMaps, 2 * bytes) != hipSuccess)
        return ckd::fail(CKD_ENOMEM, "ckd_spikey_blurmaps: cannot allocate the maps");
    CKD_HIP(hipMemcpy(ctx->spikeBlurMaps, map1, bytes, hipMemcpyHostToDevice));
    CKD_HIP(hipMemcpy(ctx->spikeBlurMaps + px, map2, bytes, hipMemcpyHostToDevice));
    ctx->spikeBlurW = resX;
    ctx->spikeBlurH = resY;
    return CKD_OK;
}

ckd_status ckd_fdtunnel_create(ckd_ctx *ctx, const uint32_t *tex, const uint32_t *texHighlights)
{
    CKD_REQUIRE(tex && texHighlights, "ckd_fdtunnel_create: null texture");
    CKD_REQUIRE_CTX(ctx);
    const size_t bytes = 1024 * 1024 * sizeof(uint32_t);
    if ((!ctx->fdTex && ckd::dev_alloc(&ctx->fdTex, bytes) != hipSuccess) ||
        (!ctx->fdTexHi && ckd::dev_alloc(&ctx->fdTexHi, bytes) != hipSuccess))
        return ckd::fail(CKD_ENOMEM, "ckd_fdtunnel_create: cannot allocate the textures");
    CKD_HIP(hipMemcpy(ctx->fdTex, tex, bytes, hipMemcpyHostToDevice));
    CKD_HIP(hipMemcpy(ctx->fdTexHi, texHighlights, bytes, hipMemcpyHostToDevice));
    return CKD_OK;
}

ckd_status ckd_fdtunnel_map(ckd_ctx *ctx, uint32_t *maps, uint32_t *glowMaps, unsigned resX,
                            unsigned resY, const ckd_fdtunnel_params *frames, unsigned nFrames)
{
    ckd_status s = check_geometry("ckd_fdtunnel_map", resX, resY, nFrames, maps, frames);
    if (s != CKD_OK || (s = check_params("ckd_fdtunnel_map", frames, nFrames)) != CKD_OK) return s;
    CKD_REQUIRE(glowMaps, "ckd_fdtunnel_map: null glow maps");
    CKD_REQUIRE_CTX(ctx);
    CKD_REQUIRE(ctx->fdTex, "ckd_fdtunnel_map: no textures (ckd_fdtunnel_create)");
    return st_maps(ctx, FX_FDTUNNEL, frames, nFrames, maps, glowMaps, resX, resY, false);
}

ckd_status ckd_fdtunnel_draw(ckd_ctx *ctx, uint32_t *dst, unsigned resX, unsigned resY,
                             const ckd_fdtunnel_params *frames, unsigned nFrames)
{
    ckd_status s = check_geometry("ckd_fdtunnel_draw", resX, resY, nFrames, dst, frames);
    if (s != CKD_OK || (s = check_params("ckd_fdtunnel_draw", frames, nFrames)) != CKD_OK) return s;
    CKD_REQUIRE_CTX(ctx);
    CKD_REQUIRE(ctx->fdTex, "ckd_fdtunnel_draw: no textures (ckd_fdtunnel_create)");
    return st_draw(ctx, FX_FDTUNNEL, frames, nFrames, dst, resX, resY, false);
}

} // extern "C"
