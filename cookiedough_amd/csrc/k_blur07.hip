// k_blur07.hip -- the 2007 box blur (code/deprecated/boxblur.cpp).
//
// Out of place the reference's sliding window is an exact window sum (its
// u16 accumulator never saturates for spans <= 255), so each output pixel
// is computed independently:
//   odd span 2e+1:  acc = sum p[x-e .. x+e] (clamped to the line)
//   even span 2e:   the same window with half weights at both ends,
//                   ceil(p[x+e]/2) at the head and floor(p[x-e]/2) at the
//                   tail (code/deprecated/boxblur.cpp:17-36 split each pixel
//                   into p-(p>>1) now and p>>1 one step later)
//   out = packus16(mulhi_u16(acc, div)) with div = (2^24/w16)>>4 truncated
//   to 16 bits, w16 = 16*weight (+8 for even spans' half pixel).
// Four 8-bit channels are summed at once in the 16-bit fields of a u64;
// window sums stay below 2^16, so differences of u64 prefix sums are exact.
//
// In place (dst == src) the reference's subtract reads pixels it has
// already overwritten, which makes every line a sequential recurrence with
// u16 saturation; those kernels run one lane per line and channel and
// replay it, on lines staged in LDS (k_blur07_lines) or with an LDS ring of
// the recent outputs per lane, reading the image directly
// (k_blur07_ring_*).
#include "ckd_internal.h"
#include "ckd_pixel.h"
#include "ckd_b07.h"

#include <vector>

namespace {

using ckd::B07Job;
using ckd::B07Params;

// code/deprecated/boxblur.cpp:11-14 (stored through _mm_set1_epi16)
uint16_t weight_to_div(unsigned weight) { return (uint16_t)(((65536u * 256u) / weight) >> 4); }

// code/deprecated/boxblur.cpp:52-76
unsigned b07_span(float strength)
{
    const float f = strength * 255.f;
    unsigned k = f >= 4294967040.f ? 0xffffffffu : (unsigned)f;
    int ki = (int)k;
    return (unsigned)(ki > 255 ? 255 : (ki < 1 ? 1 : ki));
}

void b07_setup(float strength, B07Params &p)
{
    p.span = b07_span(strength);
    p.subEdges = (p.span & 1) == 0;
    p.edgeSpan = p.span >> 1;
    p.shift = 1 + (p.subEdges ? 0 : 7);
    p.median = p.edgeSpan + (p.subEdges ? 0 : 1);
    unsigned w = (p.median << 4) + (p.subEdges << 3);
    for (unsigned i = 0; i < 256; ++i) p.edgeDivs[i] = 0;
    for (unsigned i = 0; i < p.median; ++i, w += 16) p.edgeDivs[i] = weight_to_div(w);
    p.fullDiv = weight_to_div(p.span << 4);
}

CKD_DEV uint64_t spread(uint32_t px)
{
    return (uint64_t)(px & 0xffu) | ((uint64_t)(px & 0xff00u) << 8) |
           ((uint64_t)(px & 0xff0000u) << 16) | ((uint64_t)(px & 0xff000000u) << 24);
}

// per-channel floor(p/2) and ceil(p/2), spread
CKD_DEV uint64_t floor_half(uint32_t px) { return spread((px >> 1) & 0x7f7f7f7fu); }
CKD_DEV uint64_t ceil_half(uint32_t px) { return spread(px) - floor_half(px); }

CKD_DEV uint32_t div_pack(uint64_t acc, uint32_t div)
{
    uint32_t px = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t a = (uint32_t)(acc >> (16 * k)) & 0xffffu;
        const uint32_t q = (a * div) >> 16;
        px |= (q > 32767u ? 0u : (q > 255u ? 255u : q)) << (8 * k);
    }
    return px;
}

CKD_DEV uint32_t divisor_at(const B07Params &p, int x, int len)
{
    if (x < (int)p.median) return p.edgeDivs[x];
    if (x >= len - (int)p.edgeSpan) return p.edgeDivs[len - 1 - x];
    return p.fullDiv;
}

// horizontal, out of place: one 256-thread block per row, LDS prefix sums
constexpr int kHBlock = 256;
constexpr int kHMaxW = 4096; // (W+1)*8 + W*4 bytes of LDS stay below 64 KiB

__global__ __launch_bounds__(kHBlock) void k_blur07_h(uint32_t *__restrict__ dst,
                                                      const uint32_t *__restrict__ src,
                                                      int W, B07Params p)
{
    extern __shared__ uint64_t lds[];
    uint64_t *P = lds;                                        // W+1 prefix sums
    uint32_t *px = reinterpret_cast<uint32_t *>(lds + W + 1); // W pixels
    __shared__ uint64_t waveTot[kHBlock / 64];

    const int row = blockIdx.x;
    const uint32_t *s = src + (size_t)row * W;
    for (int i = threadIdx.x; i < W; i += kHBlock) px[i] = s[i];
    __syncthreads();

    const int chunk = (W + kHBlock - 1) / kHBlock;
    const int c0 = threadIdx.x * chunk;
    const int c1 = min(c0 + chunk, W);
    uint64_t local = 0;
    for (int i = c0; i < c1; ++i) local += spread(px[i]);

    // block exclusive scan of the chunk totals
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t incl = local;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) waveTot[wave] = incl;
    __syncthreads();
    uint64_t base = 0;
    for (int w = 0; w < wave; ++w) base += waveTot[w];
    uint64_t run = base + incl - local;
    if (threadIdx.x == 0) P[0] = 0;
    for (int i = c0; i < c1; ++i) {
        run += spread(px[i]);
        P[i + 1] = run;
    }
    __syncthreads();

    const int e = (int)p.edgeSpan;
    uint32_t *d = dst + (size_t)row * W;
    for (int x = threadIdx.x; x < W; x += kHBlock) {
        const int lo = x - e, hi = x + e;
        const int loC = lo < 0 ? 0 : lo, hiC = hi > W - 1 ? W - 1 : hi;
        uint64_t acc = P[hiC + 1] - P[loC];
        if (p.subEdges) {
            if (hi <= W - 1) acc -= floor_half(px[hi]);
            if (lo >= 0) acc -= ceil_half(px[lo]);
        }
        d[x] = div_pack(acc, divisor_at(p, x, W));
    }
}

// vertical, out of place: lane per column, each lane slides over a segment
// of kVSeg rows (reads are coalesced across the wave)
constexpr int kVSeg = 32;

__global__ __launch_bounds__(256) void k_blur07_v(uint32_t *__restrict__ dst,
                                                  const uint32_t *__restrict__ src, int W,
                                                  int H, B07Params p)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y0 = blockIdx.y * kVSeg;
    if (x >= W) return;
    const int e = (int)p.edgeSpan;
    const int y1 = min(y0 + kVSeg, H);
    const uint32_t *col = src + x;
    uint64_t F = 0;
    for (int r = max(0, y0 - e); r <= min(H - 1, y0 + e); ++r) F += spread(col[(size_t)r * W]);
    for (int y = y0; y < y1; ++y) {
        const int lo = y - e, hi = y + e;
        uint64_t acc = F;
        if (p.subEdges) {
            if (hi <= H - 1) acc -= floor_half(col[(size_t)hi * W]);
            if (lo >= 0) acc -= ceil_half(col[(size_t)lo * W]);
        }
        dst[(size_t)y * W + x] = div_pack(acc, divisor_at(p, y, H));
        if (hi + 1 <= H - 1) F += spread(col[(size_t)(hi + 1) * W]);
        if (lo >= 0) F -= spread(col[(size_t)lo * W]);
    }
}


// in place with the lines staged in LDS: a one-wave block copies NL lines
// into LDS (one dword per pixel, each line padded by a dword so the NL
// lines fall in different banks), runs the recurrence there -- lane
// (line, channel) reads and overwrites byte `channel` of its line's pixels,
// exactly as the reference's loop reads and overwrites the image -- and
// writes the lines back.  No global access is on the recurrence's path.
// Every phase takes K steps at a time with their LDS reads issued first:
// inputs are always read before their position is overwritten, and the
// outgoing pixel of step j was written at step j - median, so K <= median
// keeps a chunk's reads clear of its own writes.
// One channel of one line.  Odd spans have shift 8, so the split of each
// pixel into p - p>>shift and p>>shift is (p, 0): a step is one saturating
// u16 add and one u16 subtract clamped at 0.  Even spans carry the
// half-pixel remainders between steps.
template <bool kEven> struct B07Lane {
    uint8_t *px; // byte `channel` of pixel 0; pixels are 4 bytes apart
    int addPos, subPos, dest;
    uint32_t acc, addRem, subRem;
    CKD_DEV void add(uint32_t v)
    {
        uint32_t a = v;
        if constexpr (kEven) {
            const uint32_t rem = v >> 1;
            a = addRem + (v - rem);
            addRem = rem;
        }
        acc = __builtin_elementwise_add_sat((uint16_t)acc, (uint16_t)a);
    }
    CKD_DEV void sub(uint32_t v)
    {
        uint32_t d = v;
        if constexpr (kEven) {
            const uint32_t rem = v >> 1;
            d = subRem + (v - rem);
            subRem = rem;
        }
        acc = __builtin_elementwise_sub_sat((uint16_t)acc, (uint16_t)d);
    }
    // Div + packus_epi16 (code/deprecated/boxblur.cpp:39-42)
    CKD_DEV uint8_t out(uint32_t div) const
    {
        const int16_t q = (int16_t)(__umul24(acc, div) >> 16);
        return (uint8_t)__builtin_elementwise_min(__builtin_elementwise_max(q, (int16_t)0),
                                                  (int16_t)255);
    }
};

template <int K, bool kAdd, bool kSub, bool kEmit, bool kEven, typename Div>
CKD_DEV void b07_run(B07Lane<kEven> &s, int n, Div div)
{
    int i = 0;
    for (; i + K <= n; i += K) {
        uint32_t vin[K], vsub[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if constexpr (kAdd) vin[j] = s.px[(s.addPos + j) * 4];
            if constexpr (kSub) vsub[j] = s.px[(s.subPos + j) * 4];
        }
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if constexpr (kAdd) s.add(vin[j]);
            if constexpr (kSub) s.sub(vsub[j]);
            if constexpr (kEmit) s.px[(s.dest + j) * 4] = s.out(div(i + j));
        }
        if constexpr (kAdd) s.addPos += K;
        if constexpr (kSub) s.subPos += K;
        if constexpr (kEmit) s.dest += K;
    }
    for (; i < n; ++i) {
        if constexpr (kAdd) s.add(s.px[(s.addPos++) * 4]);
        if constexpr (kSub) s.sub(s.px[(s.subPos++) * 4]);
        if constexpr (kEmit) s.px[(s.dest++) * 4] = s.out(div(i));
    }
}

// the phases that subtract, with the largest chunk the median allows
template <bool kAdd, bool kEven, typename Div>
CKD_DEV void b07_run_sub(B07Lane<kEven> &s, int n, int median, Div div)
{
    if (median >= 8) b07_run<8, kAdd, true, true>(s, n, div);
    else if (median >= 4) b07_run<4, kAdd, true, true>(s, n, div);
    else if (median >= 2) b07_run<2, kAdd, true, true>(s, n, div);
    else b07_run<1, kAdd, true, true>(s, n, div);
}

// the whole line (code/deprecated/boxblur.cpp:81-125).  kSplit: stop just
// before the step that adds the line's last pixel (the middle run's last
// step; the caller guarantees full >= 1) and save the lane's state to
// state[0..2]; k_blur07_chain (b07_chain_tail) finishes the line later.
template <bool kEven, bool kSplit = false>
CKD_DEV void b07_lds_line(uint8_t *px, int len, const B07Params &p, const uint16_t *edgeDivs,
                          uint32_t *state = nullptr)
{
    B07Lane<kEven> s;
    s.px = px;
    s.addPos = s.subPos = s.dest = 0;
    s.acc = s.addRem = s.subRem = 0;
    const int e = (int)p.edgeSpan, m = (int)p.median;
    const int full = len - (m + e);
    const uint32_t fullDiv = p.fullDiv;
    b07_run<8, true, false, false>(s, e, [](int) { return 0u; });
    b07_run<8, true, false, true>(s, m, [&](int i) { return (uint32_t)edgeDivs[i]; });
    if constexpr (kSplit) {
        b07_run_sub<true>(s, full - 1, m, [&](int) { return fullDiv; });
        state[0] = s.acc;
        state[1] = s.addRem;
        state[2] = s.subRem;
        return;
    }
    b07_run_sub<true>(s, full, m, [&](int) { return fullDiv; });
    if constexpr (kEven) s.acc = __builtin_elementwise_add_sat((uint16_t)s.acc, (uint16_t)s.addRem);
    b07_run_sub<false>(s, e, m, [&](int i) { return (uint32_t)edgeDivs[e - 1 - i]; });
}

// lines of `len` pixels: line l, pixel x at img[l*lineStride + x*pixStride]
constexpr int kLdsBatch = 16;

// images are global memory: pointers in that address space make the
// staging global_* accesses rather than flat ones (which LDS waits would
// also wait for); the 16-byte form is a clang vector, which can be copied
// through such a pointer
typedef unsigned int b07_u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) b07_u32x4 global_u32x4;
typedef __attribute__((address_space(1))) uint32_t global_px32;

template <bool kStore>
CKD_DEV void b07_stage(uint32_t *image, uint32_t *lds, int line0, int nl, int len, int stride,
                       size_t lineStride, size_t pixStride, bool vec4)
{
    const int tid = threadIdx.x;
    global_px32 *img = (global_px32 *)image;
    if (pixStride == 1 && vec4) {
        // rows, 16-byte accesses along each row
        const int perLine = len >> 2, total = nl * perLine;
        for (int base = 0; base < total; base += 64 * kLdsBatch) {
            b07_u32x4 r[kLdsBatch];
#pragma unroll
            for (int b = 0; b < kLdsBatch; ++b) {
                const int idx = base + b * 64 + tid;
                if (idx < total) {
                    const int l = idx / perLine, q = idx - l * perLine;
                    global_u32x4 *g = (global_u32x4 *)(img + (size_t)(line0 + l) * lineStride) + q;
                    uint32_t *d = lds + l * stride + 4 * q;
                    if constexpr (kStore) *g = b07_u32x4{d[0], d[1], d[2], d[3]};
                    else r[b] = *g;
                }
            }
            if constexpr (!kStore) {
#pragma unroll
                for (int b = 0; b < kLdsBatch; ++b) {
                    const int idx = base + b * 64 + tid;
                    if (idx < total) {
                        const int l = idx / perLine, q = idx - l * perLine;
                        uint32_t *d = lds + l * stride + 4 * q;
                        d[0] = r[b].x;
                        d[1] = r[b].y;
                        d[2] = r[b].z;
                        d[3] = r[b].w;
                    }
                }
            }
        }
        return;
    }
    // one pixel per access; lines fastest across lanes (coalesced for columns)
    const int total = nl * len;
    for (int base = 0; base < total; base += 64 * kLdsBatch) {
        uint32_t r[kLdsBatch];
#pragma unroll
        for (int b = 0; b < kLdsBatch; ++b) {
            const int idx = base + b * 64 + tid;
            if (idx < total) {
                const int x = idx / nl, l = idx - x * nl;
                global_px32 *g = img + (size_t)(line0 + l) * lineStride + (size_t)x * pixStride;
                if constexpr (kStore) *g = lds[l * stride + x];
                else r[b] = *g;
            }
        }
        if constexpr (!kStore) {
#pragma unroll
            for (int b = 0; b < kLdsBatch; ++b) {
                const int idx = base + b * 64 + tid;
                if (idx < total) {
                    const int x = idx / nl, l = idx - x * nl;
                    lds[l * stride + x] = r[b];
                }
            }
        }
    }
}

// one block: NL lines of img, staged, blurred in place, written back; with
// kSplit each lane stops before the last pixel's add and saves its state at
// state + ((line * 4) + channel) * 3
template <int NL, bool kSplit = false>
CKD_DEV void b07_lines_block(uint32_t *img, int nLines, int len, size_t lineStride,
                             size_t pixStride, bool vec4, const B07Params &p, uint32_t *lineBuf,
                             uint16_t *edgeDivs, uint32_t *state = nullptr)
{
    const int stride = len + 1;
    const int line0 = blockIdx.x * NL;
    const int nl = min(NL, nLines - line0);
    for (int i = threadIdx.x; i < 256; i += 64) edgeDivs[i] = p.edgeDivs[i];
    b07_stage<false>(img, lineBuf, line0, nl, len, stride, lineStride, pixStride, vec4);
    __syncthreads();
    const int l = threadIdx.x >> 2, k = threadIdx.x & 3;
    if (l < nl) {
        uint8_t *px = reinterpret_cast<uint8_t *>(lineBuf + l * stride) + k;
        if constexpr (kSplit) {
            uint32_t *st = state + ((size_t)(line0 + l) * 4 + k) * 3;
            if (p.subEdges) b07_lds_line<true, true>(px, len, p, edgeDivs, st);
            else b07_lds_line<false, true>(px, len, p, edgeDivs, st);
        } else {
            if (p.subEdges) b07_lds_line<true>(px, len, p, edgeDivs);
            else b07_lds_line<false>(px, len, p, edgeDivs);
        }
    }
    __syncthreads();
    b07_stage<true>(img, lineBuf, line0, nl, len, stride, lineStride, pixStride, vec4);
}

template <int NL>
__global__ __launch_bounds__(64) void k_blur07_lines(uint32_t *img, int nLines, int len,
                                                     size_t lineStride, size_t pixStride,
                                                     int vec4, B07Params p)
{
    extern __shared__ uint32_t lineBuf[];
    __shared__ uint16_t edgeDivs[256]; // read per step in the edge phases
    b07_lines_block<NL>(img, nLines, len, lineStride, pixStride, vec4 != 0, p, lineBuf, edgeDivs);
}

// the in-place blur of several images of one size, each with its own
// strength: blockIdx.y picks the image.  Rows (lineStride len, pixStride 1)
// or columns (lineStride 1, pixStride nLines).
template <int NL, bool kSplit>
__global__ __launch_bounds__(64) void k_blur07_lines_jobs(const B07Job *__restrict__ jobs,
                                                          int nLines, int len, int columns,
                                                          int vec4, uint32_t *__restrict__ states)
{
    extern __shared__ uint32_t lineBuf[];
    __shared__ uint16_t edgeDivs[256];
    const B07Job &J = jobs[blockIdx.y];
    const size_t lineStride = columns ? 1 : (size_t)len, pixStride = columns ? (size_t)nLines : 1;
    b07_lines_block<NL, kSplit>(J.img, nLines, len, lineStride, pixStride, vec4 != 0, J.p, lineBuf,
                                edgeDivs, kSplit ? states + (size_t)blockIdx.y * nLines * 12 : nullptr);
}

// The same in-place blur over a job array with an LDS ring per lane instead
// of whole lines in LDS, so a batch's lines are all resident at once: one
// lane per line and channel (32 lines per 128-lane block, lines of
// blockIdx.y's image).  The recurrence adds position p at step p and emits
// dest = p - edge.  Outputs go to the ring only (ringLen >= max(median + 8,
// 64) entries per lane, a power of two; each lane's entries are contiguous,
// lanes ringLen + 4 bytes apart, an odd number of dwords, so the lanes of a
// wave hit distinct banks): the subtract reads the output written `median`
// steps earlier from it, and every completed aligned group of four outputs
// is flushed to the image, the quad of a line exchanging bytes so that each
// of its lanes stores one whole pixel.  Inputs come straight from the image: every input pixel is read
// before its own output is stored (dest <= p), so each lane loads one pixel
// of every group of four, 24..31 steps ahead, and the quad exchange hands
// each lane its channel of the four.  Per eight steps a lane thus issues two
// loads and two stores instead of sixteen byte accesses.
// kSplit: stop before the line's last pixel is added, store every output
// made and save the lane's state, as k_blur07_lines_jobs does.

// the image in the global address space, addressed by 32-bit pixel offsets
// (images stay below 4 GiB)
typedef __attribute__((address_space(1))) uint32_t global_u32;

// 4 x 4 byte transpose across the four lanes of a quad: lane k holds bytes
// x[k][0..3] and receives x[0..3][k].  Two DPP exchanges (xor 2, then xor
// 1), each followed by a rotate of the partner's value and a bit-field
// insert under a per-lane mask, so no lane branches: after the first lanes
// 0,1 hold [x0 x1 | partner's x0 x1] and lanes 2,3 [partner's x2 x3 | x2 x3];
// the second does the same with single bytes.
struct QuadT {
    uint32_t m1, m2, r2;
    CKD_DEV explicit QuadT(int k)
        : m1((k & 2) ? 0xffff0000u : 0x0000ffffu), m2((k & 1) ? 0xff00ff00u : 0x00ff00ffu),
          r2((k & 1) ? 8u : 24u)
    {
    }
    CKD_DEV static uint32_t rotr(uint32_t x, uint32_t r) { return __builtin_amdgcn_alignbit(x, x, r); }
    CKD_DEV uint32_t operator()(uint32_t v) const
    {
        uint32_t o = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
        v = (v & m1) | (rotr(o, 16) & ~m1);
        o = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
        return (v & m2) | (rotr(o, r2) & ~m2);
    }
};

template <int K, bool kEven, bool kSplit>
CKD_DEV void b07_ring_line(global_u32 *img, uint32_t lineOff, uint32_t pstep, int len, int k,
                           const B07Params &p, const uint16_t *divs, uint8_t *lane, int rmask,
                           uint32_t *state)
{
    auto slot = [&](int entry) -> uint8_t & { return lane[entry & rmask]; };
    const int e = (int)p.edgeSpan, m = (int)p.median, em = e + m;
    const int nAdd = kSplit ? len - 1 : len;
    const uint32_t fullDiv = p.fullDiv;
    B07Lane<kEven> s;
    s.acc = s.addRem = s.subRem = 0;
    const QuadT quad_transpose(k);
    // outputs 0..flushed-1 are in the image
    int flushed = 0;
    auto flush_group = [&](int made) {
        const uint32_t x = *reinterpret_cast<const uint32_t *>(lane + (flushed & rmask));
        const uint32_t px = quad_transpose(x);
        if (flushed + k < made) img[lineOff + (uint32_t)(flushed + k) * pstep] = px;
        flushed += 4;
    };
    // after at most eight new outputs: fewer than 4 were pending before, so
    // at most two groups are complete (no loop, so the waits on the
    // prefetched loads stay exact)
    auto flush_to = [&](int made) {
        if (flushed + 4 <= made) flush_group(made);
        if (flushed + 4 <= made) flush_group(made);
    };
    // every output made, the last group partly
    auto flush_all = [&](int made) {
        while (flushed < made) flush_group(made);
    };
    // this lane's pixel of the groups at b and b + 4
    auto load2 = [&](uint32_t (&q)[2], int b) {
        q[0] = img[lineOff + (uint32_t)min(b + k, len - 1) * pstep];
        q[1] = img[lineOff + (uint32_t)min(b + 4 + k, len - 1) * pstep];
    };
    // positions b..b+7
    auto chunk = [&](const uint32_t (&q)[2], int b) {
        const uint32_t t0 = quad_transpose(q[0]), t1 = quad_transpose(q[1]);
        auto in = [&](int i) -> uint32_t { return ((i < 4 ? t0 : t1) >> (8 * (i & 3))) & 0xffu; };
        const int rd0 = (b - em) & rmask, wr0 = (b - e) & rmask;
        if (b >= em && b + 8 <= nAdd && rd0 <= rmask - 7 && wr0 <= rmask - 7) {
            // the middle run, neither ring run wrapping: fixed offsets
            const uint8_t *rd = lane + rd0;
            uint8_t *wr = lane + wr0;
#pragma unroll
            for (int i = 0; i < 8; i += K) {
                uint32_t v[K];
#pragma unroll
                for (int j = 0; j < K; ++j) v[j] = rd[i + j];
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    s.add(in(i + j));
                    s.sub(v[j]);
                    wr[i + j] = s.out(fullDiv);
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int pos = b + i;
                if (pos < nAdd) {
                    s.add(in(i));
                    if (pos >= e) {
                        if (pos >= em) s.sub(slot(pos - em));
                        const int d = pos - e;
                        slot(d) = s.out(d < m ? (uint32_t)divs[d] : fullDiv);
                    }
                }
            }
        }
        flush_to(max(0, min(b + 8, nAdd) - e));
    };
    // the first four chunks' loads in chunk order (the barriers keep the
    // scheduler from reordering them, which would make the first chunk of
    // every pass wait for the youngest loads)
    uint32_t q0[2], q1[2], q2[2], q3[2];
    load2(q0, 0);
    __builtin_amdgcn_sched_barrier(0);
    load2(q1, 8);
    __builtin_amdgcn_sched_barrier(0);
    load2(q2, 16);
    __builtin_amdgcn_sched_barrier(0);
    load2(q3, 24);
    __builtin_amdgcn_sched_barrier(0);
    // four chunks per pass, always (chunks past the end do nothing): one
    // exit, so the waits on the loads stay exact across the back edge
    for (int b = 0; b < nAdd; b += 32) {
        chunk(q0, b);
        load2(q0, b + 32);
        chunk(q1, b + 8);
        load2(q1, b + 40);
        chunk(q2, b + 16);
        load2(q2, b + 48);
        chunk(q3, b + 24);
        load2(q3, b + 56);
    }
    if constexpr (kSplit) {
        flush_all(max(0, nAdd - e));
        state[0] = s.acc;
        state[1] = s.addRem;
        state[2] = s.subRem;
        return;
    }
    if constexpr (kEven) s.acc = __builtin_elementwise_add_sat((uint16_t)s.acc, (uint16_t)s.addRem);
    // the end phase: e steps of subtract + emit, descending weights
    int d = len - e, i = 0;
    for (; i + K <= e; i += K, d += K) {
        uint32_t v[K];
#pragma unroll
        for (int j = 0; j < K; ++j) v[j] = slot(d + j - m);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            s.sub(v[j]);
            slot(d + j) = s.out(divs[e - 1 - (i + j)]);
        }
        flush_to(d + K);
    }
    for (; i < e; ++i, ++d) {
        s.sub(slot(d - m));
        slot(d) = s.out(divs[e - 1 - i]);
    }
    flush_all(len);
}

constexpr int kRingBlock = 128; // lanes per block of the ring kernel (32 lines)

template <int K, bool kSplit>
CKD_DEV void b07_ring_block(const B07Job &J, int nLines, int len, int columns, int ringLen,
                            uint32_t *states, uint8_t *ringBuf, uint16_t *divs)
{
    for (int i = threadIdx.x; i < 256; i += kRingBlock) divs[i] = J.p.edgeDivs[i];
    __syncthreads();
    const int t = blockIdx.x * kRingBlock + threadIdx.x;
    const int line = t >> 2, k = t & 3;
    if (line >= nLines) return;
    const uint32_t lineOff = (uint32_t)line * (columns ? 1u : (uint32_t)len);
    const uint32_t pstep = columns ? (uint32_t)nLines : 1u;
    global_u32 *img = (global_u32 *)J.img;
    uint8_t *lane = ringBuf + threadIdx.x * (ringLen + 4);
    uint32_t *st = kSplit ? states + (((size_t)blockIdx.y * nLines + line) * 4 + k) * 3 : nullptr;
    if (J.p.subEdges)
        b07_ring_line<K, true, kSplit>(img, lineOff, pstep, len, k, J.p, divs, lane, ringLen - 1, st);
    else
        b07_ring_line<K, false, kSplit>(img, lineOff, pstep, len, k, J.p, divs, lane, ringLen - 1, st);
}

// blockIdx.y picks the job
template <int K, bool kSplit>
__global__ __launch_bounds__(kRingBlock) void k_blur07_ring_jobs(const B07Job *__restrict__ jobs,
                                                          int nLines, int len, int columns,
                                                          int ringLen, uint32_t *__restrict__ states)
{
    extern __shared__ uint8_t ringBuf[];
    __shared__ uint16_t divs[256];
    b07_ring_block<K, kSplit>(jobs[blockIdx.y], nLines, len, columns, ringLen, states, ringBuf, divs);
}

// one image, its job passed by value
template <int K>
__global__ __launch_bounds__(kRingBlock) void k_blur07_ring_one(B07Job job, int nLines, int len,
                                                         int columns, int ringLen)
{
    extern __shared__ uint8_t ringBuf[];
    __shared__ uint16_t divs[256];
    b07_ring_block<K, false>(job, nLines, len, columns, ringLen, nullptr, ringBuf, divs);
}

// the chain of a batch of render targets whose last column carries from one
// image to the next (the ball's frames with beams): one lane per line and
// channel walks the images in order.  mode 0: the image's last column is
// final and becomes the carry; mode 1: the image takes the carry as its last
// column; mode 2: the line's split blur ends with the carry as its last
// pixel (the middle run's last step, then the end phase) and its blurred
// last pixel becomes the carry.  carry (one pixel per line) holds the
// incoming column and receives the outgoing one.
//
// The tail's subtracted values are the split blur's outputs at
// full-1 .. full-1+e (all written before the tail, except, for even spans,
// the last: the tail's own first output), so they are read into the lane's
// LDS slots first and only the recurrence itself waits on the carry.
constexpr int kChainLanes = 64;

typedef __attribute__((address_space(1))) uint8_t global_b8;
typedef __attribute__((address_space(1))) const uint32_t global_cu32;

template <bool kEven>
CKD_DEV uint32_t b07_chain_tail(global_b8 *px, int len, const B07Params &p, global_cu32 *state,
                                uint32_t c, uint8_t *slots)
{
    const int e = (int)p.edgeSpan, m = (int)p.median;
    const int full = len - (m + e);
    const int nPre = kEven ? e : e + 1;
    for (int i = 0; i < nPre; ++i) slots[i * kChainLanes] = px[(full - 1 + i) * 4];
    B07Lane<kEven> s;
    s.acc = state[0];
    s.addRem = state[1];
    s.subRem = state[2];
    s.add(c);
    s.sub(slots[0]);
    const uint8_t o0 = s.out(p.fullDiv);
    px[(m + full - 1) * 4] = o0;
    uint32_t last = o0;
    if constexpr (kEven) s.acc = __builtin_elementwise_add_sat((uint16_t)s.acc, (uint16_t)s.addRem);
    global_b8 *out = px + (m + full) * 4;
    int i = 0;
    for (; i + 8 <= e; i += 8) {
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = (kEven && i + j == e - 1) ? o0 : slots[(i + j + 1) * kChainLanes];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            s.sub(v[j]);
            last = s.out(p.edgeDivs[e - 1 - (i + j)]);
            out[(i + j) * 4] = (uint8_t)last;
        }
    }
    for (; i < e; ++i) {
        s.sub((kEven && i == e - 1) ? (uint32_t)o0 : (uint32_t)slots[(i + 1) * kChainLanes]);
        last = s.out(p.edgeDivs[e - 1 - i]);
        out[i * 4] = (uint8_t)last;
    }
    return last;
}

__global__ __launch_bounds__(kChainLanes) void k_blur07_chain(const ckd::B07Chain *__restrict__ chain,
                                                              int n, uint32_t *__restrict__ carry,
                                                              int nLines, int len)
{
    __shared__ uint8_t slots[256 * kChainLanes];
    const int t = blockIdx.x * kChainLanes + threadIdx.x;
    const int line = t >> 2, k = t & 3;
    if (line >= nLines) return;
    uint8_t *cp = reinterpret_cast<uint8_t *>(carry + line) + k;
    uint32_t c = *cp;
    for (int f = 0; f < n; ++f) {
        const ckd::B07Chain &J = chain[f];
        // image bytes and saved states through global-address-space
        // pointers: global accesses, which the LDS slots' waits do not
        // wait for
        global_b8 *px = (global_b8 *)(reinterpret_cast<uint8_t *>(J.img + (size_t)line * len) + k);
        if (J.mode == 0) {
            c = px[(size_t)(len - 1) * 4];
        } else if (J.mode == 1) {
            px[(size_t)(len - 1) * 4] = (uint8_t)c;
        } else {
            global_cu32 *st = (global_cu32 *)(J.state + ((size_t)line * 4 + k) * 3);
            c = J.p.subEdges ? b07_chain_tail<true>(px, len, J.p, st, c, slots + threadIdx.x)
                             : b07_chain_tail<false>(px, len, J.p, st, c, slots + threadIdx.x);
        }
    }
    *cp = (uint8_t)c;
}

ckd_status validate(const char *fn, const uint32_t *dst, const uint32_t *src, unsigned xRes,
                    unsigned yRes, float strength, unsigned passLen, B07Params &p)
{
    CKD_REQUIRE(dst && src, "%s: null buffer", fn);
    CKD_REQUIRE(xRes > 0 && yRes > 0 && xRes <= 16384 && yRes <= 16384,
                "%s: bad size %ux%u", fn, xRes, yRes);
    CKD_REQUIRE(strength >= 0.f && strength <= 1e30f, "%s: strength %g must be finite and >= 0",
                fn, (double)strength);
    b07_setup(strength, p);
    CKD_REQUIRE(p.span <= passLen, "%s: kernel span %u exceeds the pass length %u", fn, p.span,
                passLen);
    return CKD_OK;
}


} // namespace

namespace {

// lines per block of the LDS-staged kernel: the most of 16/8/4/2/1 whose
// padded lines fit in 64 KiB; 0 when a single line does not fit
int lines_per_block(int len)
{
    for (int nl = 16; nl >= 1; nl >>= 1)
        if ((size_t)nl * (size_t)(len + 1) * 4 <= 65536) return nl;
    return 0;
}

// the ring kernel's ring length (a power of two >= max(median + 8, 64)), its LDS and its
// read-ahead (eight ring entries per chunk while the median allows)
int ring_len(unsigned mHi)
{
    int r = 64;
    while (r < (int)mHi + 8) r <<= 1;
    return r;
}
size_t ring_lds(int ringLen) { return (size_t)kRingBlock * (size_t)(ringLen + 4); }
int ring_k(unsigned mLo) { return mLo >= 8 ? 8 : mLo >= 4 ? 4 : mLo >= 2 ? 2 : 1; }

ckd_status launch_inplace(ckd_ctx *ctx, uint32_t *img, int nLines, int len, size_t lineStride,
                          size_t pixStride, const B07Params &p, const char *what)
{
    const int nl = lines_per_block(len);
    if (pixStride != 1 || nl == 0) {
        // columns (the ring kernel reads them coalesced across the wave,
        // where staging strided columns in LDS is slow), or lines longer
        // than LDS holds
        B07Job job;
        job.img = img;
        job.p = p;
        const int ringLen = ring_len(p.median);
        const size_t lds = ring_lds(ringLen);
        const int columns = pixStride != 1 ? 1 : 0;
        const unsigned grid = ckd_div_up((size_t)nLines * 4, kRingBlock);
        switch (ring_k(p.median)) {
        case 8: k_blur07_ring_one<8><<<grid, kRingBlock, lds, ctx->stream>>>(job, nLines, len, columns, ringLen); break;
        case 4: k_blur07_ring_one<4><<<grid, kRingBlock, lds, ctx->stream>>>(job, nLines, len, columns, ringLen); break;
        case 2: k_blur07_ring_one<2><<<grid, kRingBlock, lds, ctx->stream>>>(job, nLines, len, columns, ringLen); break;
        default: k_blur07_ring_one<1><<<grid, kRingBlock, lds, ctx->stream>>>(job, nLines, len, columns, ringLen); break;
        }
        return ckd::check_launch(what);
    }
    const size_t lds = (size_t)nl * (size_t)(len + 1) * 4;
    const int vec4 = pixStride == 1 && (len % 4) == 0 && (lineStride % 4) == 0 &&
                     ((uintptr_t)img % 16) == 0;
    const unsigned grid = ckd_div_up((size_t)nLines, (size_t)nl);
    switch (nl) {
    case 16: k_blur07_lines<16><<<grid, 64, lds, ctx->stream>>>(img, nLines, len, lineStride, pixStride, vec4, p); break;
    case 8: k_blur07_lines<8><<<grid, 64, lds, ctx->stream>>>(img, nLines, len, lineStride, pixStride, vec4, p); break;
    case 4: k_blur07_lines<4><<<grid, 64, lds, ctx->stream>>>(img, nLines, len, lineStride, pixStride, vec4, p); break;
    case 2: k_blur07_lines<2><<<grid, 64, lds, ctx->stream>>>(img, nLines, len, lineStride, pixStride, vec4, p); break;
    default: k_blur07_lines<1><<<grid, 64, lds, ctx->stream>>>(img, nLines, len, lineStride, pixStride, vec4, p); break;
    }
    return ckd::check_launch(what);
}

ckd_status launch_h(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned xRes,
                    unsigned yRes, const B07Params &p)
{
    if (dst == src) return launch_inplace(ctx, dst, (int)yRes, (int)xRes, xRes, 1, p, "blur07_h (in place)");
    if (xRes > (unsigned)kHMaxW) return ckd::fail(CKD_EINVAL, "blur07_h: xRes %u > %d", xRes, kHMaxW);
    const size_t lds = (size_t)(xRes + 1) * sizeof(uint64_t) + (size_t)xRes * sizeof(uint32_t);
    k_blur07_h<<<yRes, kHBlock, lds, ctx->stream>>>(dst, src, (int)xRes, p);
    return ckd::check_launch("blur07_h");
}

ckd_status launch_v(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned xRes,
                    unsigned yRes, const B07Params &p)
{
    if (dst == src) return launch_inplace(ctx, dst, (int)xRes, (int)yRes, 1, xRes, p, "blur07_v (in place)");
    dim3 grid(ckd_div_up(xRes, 64), ckd_div_up(yRes, kVSeg));
    k_blur07_v<<<grid, 64, 0, ctx->stream>>>(dst, src, (int)xRes, (int)yRes, p);
    return ckd::check_launch("blur07_v");
}

} // namespace

namespace ckd {

ckd_status b07_h_job(uint32_t *img, unsigned W, unsigned H, float strength, B07Job &job)
{
    job.img = img;
    return validate("blur07_h (in place)", img, img, W, H, strength, W, job.p);
}

ckd_status b07_job(uint32_t *img, unsigned W, unsigned H, float strength, B07Job &job)
{
    job.img = img;
    return validate("blur07 (in place)", img, img, W, H, strength, W < H ? W : H, job.p);
}

template <bool kSplit>
void launch_lines_jobs(int nl, dim3 grid, size_t lds, hipStream_t st, const B07Job *devJobs,
                       int nLines, int len, int columns, int v4, uint32_t *states)
{
    switch (nl) {
    case 16: k_blur07_lines_jobs<16, kSplit><<<grid, 64, lds, st>>>(devJobs, nLines, len, columns, v4, states); break;
    case 8: k_blur07_lines_jobs<8, kSplit><<<grid, 64, lds, st>>>(devJobs, nLines, len, columns, v4, states); break;
    case 4: k_blur07_lines_jobs<4, kSplit><<<grid, 64, lds, st>>>(devJobs, nLines, len, columns, v4, states); break;
    case 2: k_blur07_lines_jobs<2, kSplit><<<grid, 64, lds, st>>>(devJobs, nLines, len, columns, v4, states); break;
    default: k_blur07_lines_jobs<1, kSplit><<<grid, 64, lds, st>>>(devJobs, nLines, len, columns, v4, states); break;
    }
}

template <bool kSplit>
void launch_ring_jobs(int K, dim3 grid, size_t lds, hipStream_t st, const B07Job *devJobs,
                      int nLines, int len, int columns, int ringLen, uint32_t *states)
{
    switch (K) {
    case 8: k_blur07_ring_jobs<8, kSplit><<<grid, kRingBlock, lds, st>>>(devJobs, nLines, len, columns, ringLen, states); break;
    case 4: k_blur07_ring_jobs<4, kSplit><<<grid, kRingBlock, lds, st>>>(devJobs, nLines, len, columns, ringLen, states); break;
    case 2: k_blur07_ring_jobs<2, kSplit><<<grid, kRingBlock, lds, st>>>(devJobs, nLines, len, columns, ringLen, states); break;
    default: k_blur07_ring_jobs<1, kSplit><<<grid, kRingBlock, lds, st>>>(devJobs, nLines, len, columns, ringLen, states); break;
    }
}

ckd_status blur07_inplace_jobs(ckd_ctx *ctx, const B07Job *hostJobs, const B07Job *devJobs,
                               unsigned n, unsigned W, unsigned H, bool columns, bool vec4,
                               uint32_t *splitStates)
{
    if (n == 0) return CKD_OK;
    const unsigned len = columns ? H : W, nLines = columns ? W : H;
    CKD_REQUIRE(!(splitStates && columns), "blur07 (in place, batch): split blurs run along rows");
    unsigned mLo = 256, mHi = 0;
    for (unsigned j = 0; j < n; ++j) {
        mLo = hostJobs[j].p.median < mLo ? hostJobs[j].p.median : mLo;
        mHi = hostJobs[j].p.median > mHi ? hostJobs[j].p.median : mHi;
    }
    const int nl = lines_per_block((int)len);
    // the staged-lines kernel holds whole lines in LDS; the ring kernel only
    // the last outputs per lane, so its lines all run at once, but one line
    // takes it longer (about 85 against 50 ns a step at 1280 px, measured).
    // Automatic choice: the ring kernel for columns (staging strided
    // columns is slow) and for rows once the staged kernel would need more
    // than one round of blocks on the device.
    bool ring = ctx->tuning.blurKernel == 2 || nl == 0;
    if (ctx->tuning.blurKernel == 0 && nl > 0) {
        const size_t ldsLines = (size_t)nl * (size_t)(len + 1) * 4;
        const size_t perCU = (160u * 1024u) / ldsLines;
        const size_t blocks = ckd_div_up((size_t)nLines, (size_t)nl) * n;
        ring = columns || blocks > perCU * (size_t)ctx->cuCount;
    }
    const char *what = columns ? "blur07_v (in place, batch)" : "blur07_h (in place, batch)";
    if (ring) {
        const int ringLen = ring_len(mHi);
        const int K = ring_k(mLo);
        const size_t lds = ring_lds(ringLen);
        const dim3 grid(ckd_div_up((size_t)nLines * 4, kRingBlock), n);
        if (splitStates)
            launch_ring_jobs<true>(K, grid, lds, ctx->stream, devJobs, (int)nLines, (int)len, 0, ringLen,
                                   splitStates);
        else
            launch_ring_jobs<false>(K, grid, lds, ctx->stream, devJobs, (int)nLines, (int)len, columns ? 1 : 0,
                                    ringLen, nullptr);
        return check_launch(what);
    }
    const size_t lds = (size_t)nl * (size_t)(len + 1) * 4;
    const int v4 = vec4 && !columns && (W % 4) == 0;
    const dim3 grid(ckd_div_up((size_t)nLines, (size_t)nl), n);
    if (splitStates)
        launch_lines_jobs<true>(nl, grid, lds, ctx->stream, devJobs, (int)nLines, (int)len, 0, v4, splitStates);
    else
        launch_lines_jobs<false>(nl, grid, lds, ctx->stream, devJobs, (int)nLines, (int)len, columns ? 1 : 0, v4,
                                 nullptr);
    return check_launch(what);
}

ckd_status blur07_h_inplace_jobs(ckd_ctx *ctx, const B07Job *hostJobs, const B07Job *devJobs,
                                 unsigned n, unsigned W, unsigned H, bool vec4, uint32_t *splitStates)
{
    return blur07_inplace_jobs(ctx, hostJobs, devJobs, n, W, H, false, vec4, splitStates);
}

ckd_status blur07_h_chain(ckd_ctx *ctx, const B07Chain *devChain, unsigned n, uint32_t *carry,
                          unsigned W, unsigned H)
{
    if (n == 0) return CKD_OK;
    k_blur07_chain<<<ckd_div_up((size_t)H * 4, kChainLanes), kChainLanes, 0, ctx->stream>>>(
        devChain, (int)n, carry, (int)H, (int)W);
    return check_launch("blur07_h chain");
}

// used by the tunnel pipeline (span 2r+1, out of place)
ckd_status blur07_hv_oop(ckd_ctx *ctx, uint32_t *tmp, uint32_t *dst, const uint32_t *src,
                         unsigned W, unsigned H, float strength)
{
    B07Params p;
    b07_setup(strength, p);
    if (p.span > W || p.span > H)
        return fail(CKD_EINVAL, "tunnel blur: span %u exceeds %ux%u", p.span, W, H);
    ckd_status s = launch_h(ctx, tmp, src, W, H, p);
    if (s != CKD_OK) return s;
    return launch_v(ctx, dst, tmp, W, H, p);
}

} // namespace ckd

extern "C" {

ckd_status ckd_blur07_h(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned xRes,
                        unsigned yRes, float strength)
{
    B07Params p;
    ckd_status s = validate("ckd_blur07_h", dst, src, xRes, yRes, strength, xRes, p);
    if (s != CKD_OK) return s;
    CKD_REQUIRE_CTX(ctx);
    return launch_h(ctx, dst, src, xRes, yRes, p);
}

ckd_status ckd_blur07_v(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned xRes,
                        unsigned yRes, float strength)
{
    B07Params p;
    ckd_status s = validate("ckd_blur07_v", dst, src, xRes, yRes, strength, yRes, p);
    if (s != CKD_OK) return s;
    CKD_REQUIRE_CTX(ctx);
    return launch_v(ctx, dst, src, xRes, yRes, p);
}

ckd_status ckd_blur07(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned xRes,
                      unsigned yRes, float strength)
{
    B07Params p;
    ckd_status s = validate("ckd_blur07", dst, src, xRes, yRes, strength,
                            xRes < yRes ? xRes : yRes, p);
    if (s != CKD_OK) return s;
    CKD_REQUIRE_CTX(ctx);
    s = launch_h(ctx, dst, src, xRes, yRes, p);
    if (s != CKD_OK) return s;
    return launch_v(ctx, dst, dst, xRes, yRes, p);
}

ckd_status ckd_blur07_inplace_batch(ckd_ctx *ctx, uint32_t *const *imgs, const float *strengths,
                                    unsigned n, unsigned xRes, unsigned yRes, int passes)
{
    CKD_REQUIRE(n == 0 || (imgs && strengths), "ckd_blur07_inplace_batch: null array");
    CKD_REQUIRE(passes >= 1 && passes <= 3, "ckd_blur07_inplace_batch: passes %d not in 1..3", passes);
    const unsigned passLen = passes == 1 ? xRes : passes == 2 ? yRes : (xRes < yRes ? xRes : yRes);
    std::vector<B07Job> jobs(n);
    bool vec4 = true;
    for (unsigned j = 0; j < n; ++j) {
        jobs[j].img = imgs[j];
        ckd_status s = validate("ckd_blur07_inplace_batch", imgs[j], imgs[j], xRes, yRes,
                                strengths[j], passLen, jobs[j].p);
        if (s != CKD_OK) return s;
        vec4 = vec4 && (reinterpret_cast<uintptr_t>(imgs[j]) & 15) == 0;
    }
    CKD_REQUIRE_CTX(ctx);
    if (n == 0) return CKD_OK;
    void *dev = nullptr;
    ckd_status s = ckd::stage(ctx, jobs.data(), n * sizeof(B07Job), &dev);
    const B07Job *d = static_cast<const B07Job *>(dev);
    if (s == CKD_OK && (passes & 1))
        s = ckd::blur07_inplace_jobs(ctx, jobs.data(), d, n, xRes, yRes, false, vec4);
    if (s == CKD_OK && (passes & 2))
        s = ckd::blur07_inplace_jobs(ctx, jobs.data(), d, n, xRes, yRes, true, false);
    return s;
}

} // extern "C"
