// k_blur26.hip -- the 2026 multi-pass box blur (code/boxblur.cpp).
//
// HorzBlur32 (code/boxblur.cpp:87-212) keeps a running i32 sum per channel
// that adds the lerp term T(j) = p[j] + (((p[j+1]-p[j])*alpha) >> 16) at
// the head and subtracts it at the tail.  Integer addition is associative,
// so the sum at output o has the closed form
//   S(o) = S0 + sum_{j=iSpan+1}^{min(iSpan+o, L)} T(j) - sum_{j<max(o-iSpan,0)} T(j)
//   S0   = sum_{j<iSpan} p[j] + ((p[iSpan]*alpha) >> 16)
// and each pass becomes prefix sums over T in LDS, one block per line.  The
// reference reads p[L] and p[L+1] past every line: the next line's first
// pixels, or (after the last line) two pixels defined here as zero.
// Output: iDiv (mul_epu32 of the low dword, >> 22, hadd of the two halves)
// and v2cISSE32 (code/boxblur.cpp:69-79, code/util.h:152-154).
#include "ckd_internal.h"
#include "ckd_pixel.h"

#include <cstring>

namespace {

constexpr int kMaxRes = 2048, kMaxRadius = 500, kBlock = 256;

struct B26Params {
    int32_t iSpan, iAlpha;
    uint32_t iScale; // low dword of the 10:22 scale (what _mm_mul_epu32 uses)
    uint32_t spanScales[kMaxRadius];
};

// code/boxblur.cpp:98-124, host FP32 without contraction
#pragma clang fp contract(off)
void b26_setup(unsigned xRes, float strength, float gain, B26Params &P)
{
    strength *= 0.01f;
    const float r0 = strength * (float)((xRes - 2) / 2);
    const float radius = r0 < (float)kMaxRadius ? r0 : (float)kMaxRadius;
    const unsigned iSpan = (unsigned)radius;
    const float scale = 1.f / ((2.f - gain) * radius + 1.f);
    const float alpha = radius - (float)iSpan;
    P.iSpan = (int32_t)iSpan;
    P.iScale = (uint32_t)(uint64_t)(int64_t)(scale * (float)(1 << 22));
    P.iAlpha = (int32_t)(65536.f * alpha);
    const float halfScale = scale * 0.5f;
    const float dScale = halfScale / (float)iSpan;
    memset(P.spanScales, 0, sizeof(P.spanScales));
    for (unsigned i = 0; i < iSpan; ++i)
        P.spanScales[i] = (uint32_t)(uint64_t)(int64_t)((halfScale + (float)i * dScale) * (float)(1 << 22));
}
#pragma clang fp contract(on)

CKD_DEV int32_t lerp_term(int32_t A, int32_t B, int32_t alpha)
{
    return A + ((int32_t)((uint32_t)(B - A) * (uint32_t)alpha) >> 16);
}

// two channels (c, c+2 of byte order B,G,R,A -> fields lo/hi) per u64; the
// prefix sums of T stay below 2^32 per field, so fields never carry
CKD_DEV uint64_t pack2(int32_t lo, int32_t hi) { return (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32); }

CKD_DEV uint32_t div_lane(int32_t sum, uint32_t k)
{
    const uint64_t r = ((uint64_t)(uint32_t)sum * (uint64_t)k) >> 22;
    const int32_t lane = (int32_t)((uint32_t)r + (uint32_t)(r >> 32));
    const int32_t u16 = lane < 0 ? 0 : (lane > 65535 ? 65535 : lane);
    const int32_t s16 = (int32_t)(int16_t)(uint16_t)u16;
    return (uint32_t)(s16 < 0 ? 0 : (s16 > 255 ? 255 : s16));
}

// inclusive block scan of a pair of u64 values (256 threads, 4 waves)
CKD_DEV void block_scan2(uint64_t &a, uint64_t &b, uint64_t *sh /* 8 entries */)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t ya = __shfl_up(a, o, 64);
        const uint64_t yb = __shfl_up(b, o, 64);
        if (lane >= o) {
            a += ya;
            b += yb;
        }
    }
    if (lane == 63) {
        sh[wave] = a;
        sh[4 + wave] = b;
    }
    __syncthreads();
    for (int w = 0; w < wave; ++w) {
        a += sh[w];
        b += sh[4 + w];
    }
}

__global__ __launch_bounds__(kBlock) void k_b26_pass(uint32_t *__restrict__ out,
                                                     const uint32_t *__restrict__ in, int L,
                                                     int rows, int colStride, int rowStride,
                                                     B26Params P)
{
    extern __shared__ uint64_t lds[];
    uint64_t *PT0 = lds;                 // L+2 prefix sums, channels B,R
    uint64_t *PT1 = lds + (L + 2);       // L+2 prefix sums, channels G,A
    uint32_t *px = reinterpret_cast<uint32_t *>(lds + 2 * (L + 2)); // L+2 pixels
    __shared__ uint64_t scanSh[8];
    __shared__ int32_t s0Sh[4][4];

    const int row = blockIdx.x;
    const size_t n = (size_t)L * rows;
    const size_t base = (size_t)row * L;
    for (int i = threadIdx.x; i < L + 2; i += kBlock)
        px[i] = (base + i < n) ? in[base + i] : 0u;
    __syncthreads();

    const int iSpan = P.iSpan, alpha = P.iAlpha;
    // T(j), j in [0, L], prefix over chunks
    const int nT = L + 1;
    const int chunk = (nT + kBlock - 1) / kBlock;
    const int c0 = threadIdx.x * chunk, c1 = min(c0 + chunk, nT);
    uint64_t l0 = 0, l1 = 0;
    int32_t s0[4] = {0, 0, 0, 0};
    for (int j = c0; j < c1; ++j) {
        const uint32_t a = px[j], b = px[j + 1];
        int32_t t[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            t[k] = lerp_term((int32_t)((a >> (8 * k)) & 0xff), (int32_t)((b >> (8 * k)) & 0xff), alpha);
        l0 += pack2(t[0], t[2]);
        l1 += pack2(t[1], t[3]);
    }
    // S0 partial sums over p[j < iSpan] (any thread's chunk of pixels)
    for (int j = threadIdx.x; j < iSpan; j += kBlock) {
        const uint32_t a = px[j];
#pragma unroll
        for (int k = 0; k < 4; ++k) s0[k] += (int32_t)((a >> (8 * k)) & 0xff);
    }
    uint64_t i0 = l0, i1 = l1;
    block_scan2(i0, i1, scanSh);
    uint64_t run0 = i0 - l0, run1 = i1 - l1;
    if (threadIdx.x == 0) PT0[0] = PT1[0] = 0;
    for (int j = c0; j < c1; ++j) {
        const uint32_t a = px[j], b = px[j + 1];
        int32_t t[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            t[k] = lerp_term((int32_t)((a >> (8 * k)) & 0xff), (int32_t)((b >> (8 * k)) & 0xff), alpha);
        run0 += pack2(t[0], t[2]);
        run1 += pack2(t[1], t[3]);
        PT0[j + 1] = run0;
        PT1[j + 1] = run1;
    }
    // reduce S0 over the block
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int32_t v = s0[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if ((threadIdx.x & 63) == 0) s0Sh[threadIdx.x >> 6][k] = v;
    }
    __syncthreads();
    int32_t S0[4];
    {
        const uint32_t a = px[iSpan];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            S0[k] = s0Sh[0][k] + s0Sh[1][k] + s0Sh[2][k] + s0Sh[3][k];
            S0[k] += (int32_t)((uint32_t)((a >> (8 * k)) & 0xff) * (uint32_t)alpha) >> 16;
        }
    }
    const uint64_t Hbase0 = PT0[iSpan + 1], Hbase1 = PT1[iSpan + 1];
    for (int o = threadIdx.x; o < L; o += kBlock) {
        const int hEnd = min(iSpan + o, L) + 1;
        const int tEnd = max(o - iSpan, 0);
        // heads minus tails can be negative, so subtract field by field (a
        // u64 difference would borrow across the 32-bit fields)
        const uint64_t a0 = PT0[hEnd], a1 = PT1[hEnd], t0 = PT0[tEnd], t1 = PT1[tEnd];
        const int32_t sum[4] = {
            S0[0] + (int32_t)((uint32_t)a0 - (uint32_t)Hbase0 - (uint32_t)t0),
            S0[1] + (int32_t)((uint32_t)a1 - (uint32_t)Hbase1 - (uint32_t)t1),
            S0[2] + (int32_t)((uint32_t)(a0 >> 32) - (uint32_t)(Hbase0 >> 32) - (uint32_t)(t0 >> 32)),
            S0[3] + (int32_t)((uint32_t)(a1 >> 32) - (uint32_t)(Hbase1 >> 32) - (uint32_t)(t1 >> 32))};
        const uint32_t k = o < iSpan ? P.spanScales[o]
                                     : (o >= L - iSpan ? P.spanScales[L - 1 - o] : P.iScale);
        uint32_t pix = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) pix |= div_lane(sum[c], k) << (8 * c);
        out[(size_t)row * colStride + (size_t)o * rowStride] = pix;
    }
}

__global__ __launch_bounds__(256) void k_transpose(uint32_t *__restrict__ dst,
                                                   const uint32_t *__restrict__ src, int xRes,
                                                   int yRes)
{
    __shared__ uint32_t tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int x = bx + tx, y = by + r;
        if (x < xRes && y < yRes) tile[r][tx] = src[(size_t)y * xRes + x];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int y = by + tx, x = bx + r; // dst row = x, dst col = y
        if (x < xRes && y < yRes) dst[(size_t)x * yRes + y] = tile[tx][r];
    }
}

// HorzBlur32's pass loop (code/boxblur.cpp:126-211) over device buffers
ckd_status horz(ckd_ctx *ctx, uint32_t *pDest, uint32_t *pScratch, const uint32_t *pSrc,
                unsigned xRes, unsigned yRes, unsigned wCol, unsigned wRow, float strength,
                float gain, unsigned numPasses)
{
    B26Params P;
    b26_setup(xRes, strength, gain, P);
    const uint32_t *pRead = pSrc;
    if ((numPasses & 1) == 0) {
        uint32_t *t = pDest;
        pDest = pScratch;
        pScratch = t;
    }
    const size_t lds = (size_t)(xRes + 2) * 2 * sizeof(uint64_t) + (size_t)(xRes + 2) * sizeof(uint32_t);
    for (unsigned pass = 0; pass < numPasses; ++pass) {
        int cs = (int)xRes, rs = 1;
        if (pass == numPasses - 1) {
            cs = (int)wCol;
            rs = (int)wRow;
        }
        k_b26_pass<<<yRes, kBlock, lds, ctx->stream>>>(pDest, pRead, (int)xRes, (int)yRes, cs, rs, P);
        ckd_status s = ckd::check_launch("blur26 pass");
        if (s != CKD_OK) return s;
        pRead = pDest;
        uint32_t *t = pDest;
        pDest = pScratch;
        pScratch = t;
    }
    return CKD_OK;
}

ckd_status ensure_scratch(ckd_ctx *ctx, size_t pixels)
{
    if (ctx->b26Cap >= pixels) return CKD_OK;
    CKD_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < 2; ++i) {
        if (ctx->b26Scratch[i]) (void)hipFree(ctx->b26Scratch[i]);
        ctx->b26Scratch[i] = nullptr;
    }
    ctx->b26Cap = 0;
    for (int i = 0; i < 2; ++i)
        if (ckd::dev_alloc(&ctx->b26Scratch[i], pixels * 4) != hipSuccess) {
            ctx->b26Scratch[i] = nullptr;
            return ckd::fail(CKD_ENOMEM, "blur26: scratch allocation failed");
        }
    ctx->b26Cap = pixels;
    return CKD_OK;
}

ckd_status validate(const char *fn, const uint32_t *dst, const uint32_t *src, unsigned xRes,
                    unsigned yRes, float strength, float gain, unsigned passes, bool mult4)
{
    CKD_REQUIRE(dst && src, "%s: null buffer", fn);
    CKD_REQUIRE(dst != src, "%s: dst must differ from src", fn);
    CKD_REQUIRE(xRes >= 4 && yRes >= 4 && xRes <= (unsigned)kMaxRes && yRes <= (unsigned)kMaxRes,
                "%s: size %ux%u outside 4..%d", fn, xRes, yRes, kMaxRes);
    CKD_REQUIRE(!mult4 || ((xRes % 4) == 0 && (yRes % 4) == 0), "%s: sizes must be multiples of 4",
                fn);
    CKD_REQUIRE(strength >= 0.f && strength <= 100.f, "%s: strength %g outside [0,100]", fn,
                (double)strength);
    CKD_REQUIRE(gain >= 0.f && gain <= 1.f, "%s: gain %g outside [0,1]", fn, (double)gain);
    CKD_REQUIRE(passes >= 1 && passes <= 64, "%s: numPasses %u outside 1..64", fn, passes);
    return CKD_OK;
}

} // namespace

extern "C" {

ckd_status ckd_blur26_h(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned xRes,
                        unsigned yRes, float strength, float gain, unsigned numPasses)
{
    ckd_status s = validate("ckd_blur26_h", dst, src, xRes, yRes, strength, gain, numPasses, false);
    if (s != CKD_OK) return s;
    CKD_REQUIRE_CTX(ctx);
    s = ensure_scratch(ctx, (size_t)xRes * yRes);
    if (s != CKD_OK) return s;
    return horz(ctx, dst, ctx->b26Scratch[0], src, xRes, yRes, xRes, 1, strength, gain, numPasses);
}

ckd_status ckd_blur26_v(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned xRes,
                        unsigned yRes, float strength, float gain, unsigned numPasses)
{
    ckd_status s = validate("ckd_blur26_v", dst, src, xRes, yRes, strength, gain, numPasses, true);
    if (s != CKD_OK) return s;
    CKD_REQUIRE_CTX(ctx);
    s = ensure_scratch(ctx, (size_t)xRes * yRes);
    if (s != CKD_OK) return s;
    dim3 g(ckd_div_up(xRes, 32), ckd_div_up(yRes, 32));
    k_transpose<<<g, 256, 0, ctx->stream>>>(ctx->b26Scratch[1], src, (int)xRes, (int)yRes);
    s = ckd::check_launch("blur26 transpose");
    if (s != CKD_OK) return s;
    return horz(ctx, dst, ctx->b26Scratch[0], ctx->b26Scratch[1], yRes, xRes, 1, xRes, strength,
                gain, numPasses);
}

ckd_status ckd_blur26(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned xRes,
                      unsigned yRes, float strength, float gain, unsigned numPasses)
{
    ckd_status s = validate("ckd_blur26", dst, src, xRes, yRes, strength, gain, numPasses, true);
    if (s != CKD_OK) return s;
    CKD_REQUIRE_CTX(ctx);
    s = ensure_scratch(ctx, (size_t)xRes * yRes);
    if (s != CKD_OK) return s;
    s = horz(ctx, ctx->b26Scratch[1], ctx->b26Scratch[0], src, xRes, yRes, 1, yRes, strength, gain,
             numPasses);
    if (s != CKD_OK) return s;
    return horz(ctx, dst, ctx->b26Scratch[0], ctx->b26Scratch[1], yRes, xRes, 1, xRes, strength,
                gain, numPasses);
}

} // extern "C"
