// k_tunnel.hip -- the tunnel frame pipeline: per-8x8-block grid remap, the
// 2007 box blur with odd span 2e+1 (horizontal then vertical, out of place)
// and BlitAdd32 of the blur onto the remap (code/deprecated/boxblur.cpp,
// code/util.cpp:754-773), in two fused kernels:
//
//   T1  one 128-thread block per row: interpolate the row's grid corners
//       once per 8-pixel block, remap the row into dst and (spread into
//       16-bit fields) into LDS, prefix-sum it there in place, write the
//       horizontal blur to tmp.  Its stores are non-temporal: the rows it
//       writes are read once, later, and would otherwise push the texture
//       out of L2
//   T2  one wave per 64 columns x one segment of rows: vertical blur of tmp
//       as a running window sum and the saturating add into dst; the
//       window's rows stay in an LDS ring (radius <= 32), so each row of
//       tmp is read once per segment
//
// For an odd span the window of x is [x-e, x+e] clipped to the line and the
// reference's divisor is WeightToDiv(16*count) of that window's pixel count
// (edge ramps and full window alike), so out = div(sum, count).
#include "ckd_internal.h"
#include "ckd_pixel.h"
#include "ckd_grid.h"
#include "ckd_jobs.h"

#include <type_traits>

namespace {

// threads per row of T1: 128 gives a 1280-pixel row 10 pixels per thread,
// which spreads the per-thread fixed work (grid set-up, scan) thinner
constexpr int kT1Block = 128;
// texel quads (16 bytes per texel) for textures up to 2^11 x 2^11 (64 MB)
constexpr int kTexQuadMaxLog2 = 11;

// ckd_b07_count_div of every window count 0..255 (radius <= 127), evaluated
// by the compiler: blocks copy it instead of dividing
struct alignas(16) B07DivTable {
    uint16_t d[256];
    constexpr B07DivTable() : d{}
    {
        for (uint32_t c = 0; c < 256; ++c) d[c] = (uint16_t)ckd_b07_count_div(c);
    }
};
__constant__ const B07DivTable kB07Div = B07DivTable();

// the table into a block's LDS copy D (16-byte aligned) with one load per
// thread: 8 bytes each from 64 threads, 4 from 128
template <int kThreads> CKD_DEV void copy_div_table(uint16_t *D)
{
    static_assert(kThreads == 64 || kThreads == 128, "512 bytes over 64 or 128 threads");
    if constexpr (kThreads == 64)
        reinterpret_cast<uint2 *>(D)[threadIdx.x] = reinterpret_cast<const uint2 *>(kB07Div.d)[threadIdx.x];
    else
        reinterpret_cast<uint32_t *>(D)[threadIdx.x] = reinterpret_cast<const uint32_t *>(kB07Div.d)[threadIdx.x];
}

// the job's images and grid are global memory: pointers in that address
// space make their accesses global_* instructions rather than flat ones,
// which also wait on LDS traffic
typedef __attribute__((address_space(1))) const uint32_t global_u32c;
typedef __attribute__((address_space(1))) uint32_t global_u32;

// A block's job is the same for all its lanes.  readfirstlane says so to the
// compiler, which then keeps the job's fields, and the row arithmetic built
// on them, in scalar registers: a row's address is a scalar base plus the
// lane's 32-bit offset.
CKD_DEV uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
template <typename T> CKD_DEV T *uniform(T *p)
{
    const uint64_t a = (uint64_t)(uintptr_t)p;
    return (T *)(uintptr_t)(((uint64_t)uniform((uint32_t)(a >> 32)) << 32) | uniform((uint32_t)a));
}

CKD_DEV TunJob pick(const TunJob *jobs, int nJobs, const TunJob &single, uint32_t unit,
                    uint32_t TunJob::*base)
{
    TunJob J = jobs ? jobs[ckd_find_job(jobs, nJobs, unit, base)] : single;
    J.dst = uniform(J.dst);
    J.tmp = uniform(J.tmp);
    J.grid = uniform(J.grid);
    J.W = uniform(J.W);
    J.H = uniform(J.H);
    J.rowBase = uniform(J.rowBase);
    J.vBase = uniform(J.vBase);
    J.edge = uniform(J.edge);
    J.vBlocksX = uniform(J.vBlocksX);
    return J;
}

// the texel quads of a 2^L x 2^L texture (grid_sample_quad)
__global__ __launch_bounds__(256) void k_tex_quads(const uint32_t *__restrict__ tex, unsigned L,
                                                   uint4 *__restrict__ quads)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= (1u << (2 * L))) return;
    const uint32_t mask = (1u << L) - 1, u = i & mask, v = i >> L;
    const uint32_t u1 = (u + 1) & mask, v0 = v << L, v1 = ((v + 1) & mask) << L;
    quads[i] = make_uint4(tex[v0 + u], tex[v0 + u1], tex[v1 + u], tex[v1 + u1]);
}

// Inclusive prefix sum of a 64-bit value over the wave's 64 lanes (mod
// 2^64), with DPP moves rather than LDS permutes: shifts by 1, 2, 4 and 8
// within each row of 16 lanes, then lane 15's sum into rows 1 and 3 and
// lane 31's into rows 2 and 3.  A lane with no source lane, or in a row the
// step leaves out, adds 0.
template <int kCtrl, int kRowMask> CKD_DEV void dpp_add(uint32_t &lo, uint32_t &hi)
{
    const uint32_t dl = __builtin_amdgcn_update_dpp(0u, lo, kCtrl, kRowMask, 0xf, true);
    const uint32_t dh = __builtin_amdgcn_update_dpp(0u, hi, kCtrl, kRowMask, 0xf, true);
    const uint32_t s = lo + dl;
    hi += dh + (s < lo ? 1u : 0u);
    lo = s;
}
CKD_DEV uint64_t wave_incl_scan(uint64_t x)
{
    uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    dpp_add<0x111, 0xf>(lo, hi); // row_shr:1
    dpp_add<0x112, 0xf>(lo, hi); // row_shr:2
    dpp_add<0x114, 0xf>(lo, hi); // row_shr:4
    dpp_add<0x118, 0xf>(lo, hi); // row_shr:8
    dpp_add<0x142, 0xa>(lo, hi); // row_bcast:15
    dpp_add<0x143, 0xc>(lo, hi); // row_bcast:31
    return ((uint64_t)hi << 32) | lo;
}

// kQuads: the texture's texels come from its quads (one load per pixel)
template <bool kQuads>
__global__ __launch_bounds__(kT1Block) void k_tunnel_h(const TunJob *__restrict__ jobs, int nJobs,
                                                       const TunJob single,
                                                       const uint32_t *__restrict__ tex,
                                                       const uint4 *__restrict__ quads,
                                                       unsigned texLog2)
{
    const uint32_t row = blockIdx.x;
    extern __shared__ uint4 lds[];
    __shared__ __align__(16) uint16_t D[256];
    __shared__ uint64_t waveTot[kT1Block / 64];
    const TunJob J = pick(jobs, nJobs, single, row, &TunJob::rowBase);
    const int W = (int)J.W;
    const unsigned y = row - J.rowBase;
    const int NB = (W + 7) / 8;
    // per 8-pixel block b of the row: E[b] = (8*U_b, 8*V_b, U_b+1 - U_b,
    // V_b+1 - V_b), where (U_e, V_e) is grid corner column e interpolated
    // vertically to this row.  Pixel 8b+i then has U = (8*U_b + i*(U_b+1 -
    // U_b)) >> 6, equal mod 2^32 to (U_b*(8-i) + U_b+1*i) >> 6.
    uint4 *E = lds;
    // P[x + 1] first holds pixel x spread into 16-bit fields, then (scanned
    // in place) the prefix sum of pixels 0..x; P[0] = 0.  P starts 8 bytes
    // into its 16-byte slot, so that P[1] is 16-byte aligned.
    uint64_t *P = reinterpret_cast<uint64_t *>(lds + NB) + 1;

    copy_div_table<kT1Block>(D);
    const unsigned GW = (unsigned)NB + 1;
    const uint32_t j = y & 7;
    const global_u32c *g0 = (const global_u32c *)(J.grid) + 2 * (size_t)(y >> 3) * GW;
    const global_u32c *g1 = g0 + 2 * (size_t)GW;
    // whether every step fits a signed 24-bit multiply (exact then)
    bool wide = false;
    for (int b = threadIdx.x; b < NB; b += kT1Block) {
        const uint32_t u0 = g0[2 * b] * (8 - j) + g1[2 * b] * j;
        const uint32_t v0 = g0[2 * b + 1] * (8 - j) + g1[2 * b + 1] * j;
        const uint32_t u1 = g0[2 * b + 2] * (8 - j) + g1[2 * b + 2] * j;
        const uint32_t v1 = g0[2 * b + 3] * (8 - j) + g1[2 * b + 3] * j;
        const uint32_t du = u1 - u0, dv = v1 - v0;
        wide |= (du + (1u << 23)) >= (1u << 24) || (dv + (1u << 23)) >= (1u << 24);
        E[b] = make_uint4(8 * u0, 8 * v0, du, dv);
    }
    __shared__ uint32_t waveWide[kT1Block / 64];
    const bool anyWide = __ballot(wide) != 0;
    if ((threadIdx.x & 63) == 0) waveWide[threadIdx.x >> 6] = anyWide;
    __syncthreads();
    bool narrow = true;
#pragma unroll
    for (int w = 0; w < kT1Block / 64; ++w) narrow &= waveWide[w] == 0;
    global_u32 *drow = (global_u32 *)(J.dst + (size_t)y * W);
    // the remap of the row into LDS and dst: four pixels per thread at a time
    // while all four are in the row, all their loads before any store; then
    // the rest one at a time.  A thread's pixels share i = x & 7 (the stride
    // is a multiple of 8).
    auto remap_row = [&](auto narrowTag) {
        constexpr bool kNarrow = decltype(narrowTag)::value;
        uint32_t i = threadIdx.x & 7;
        // opaque to the compiler, so that __mul24 stays v_mad_i32_i24 rather
        // than becoming a 64-bit multiply-add on i's known small range
        asm("" : "+v"(i));
        // pixel x of the row, b = x >> 3 its block
        auto remap = [&](int b) {
            const uint4 e = E[b];
            const uint32_t iu = kNarrow ? (uint32_t)__mul24((int)i, (int)e.z) : i * e.z;
            const uint32_t iv = kNarrow ? (uint32_t)__mul24((int)i, (int)e.w) : i * e.w;
            if constexpr (kQuads)
                return ckd::grid_sample_quad(quads, texLog2, e.x + iu, e.y + iv);
            else
                return ckd::grid_sample_q14(tex, texLog2, e.x + iu, e.y + iv);
        };
        int x0 = threadIdx.x;
        for (; x0 + 3 * kT1Block < W; x0 += 4 * kT1Block) {
            // the four pixels' blocks are 16 apart: one base, constant offsets
            const int b0 = x0 >> 3;
            uint32_t v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k] = remap(b0 + k * (kT1Block / 8));
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                P[x0 + k * kT1Block + 1] = ckd_spread(v[k]);
                __builtin_nontemporal_store(v[k], &drow[x0 + k * kT1Block]);
            }
        }
        for (; x0 < W; x0 += kT1Block) {
            const uint32_t v = remap(x0 >> 3);
            P[x0 + 1] = ckd_spread(v);
            __builtin_nontemporal_store(v, &drow[x0]);
        }
    };
    if (narrow)
        remap_row(std::true_type{});
    else
        remap_row(std::false_type{});
    __syncthreads();

    // thread t scans pixels [c0, c1), an even run: P[c0 + 1] is 16-byte
    // aligned, so the run's entries move in pairs
    const int chunk = ((W + kT1Block - 1) / kT1Block + 1) & ~1;
    const int c0 = min((int)threadIdx.x * chunk, W), c1 = min(c0 + chunk, W);
    uint4 *Q = reinterpret_cast<uint4 *>(P + 1); // Q[j] = (P[2j + 1], P[2j + 2])
    auto pair_lo = [](const uint4 &q) { return ((uint64_t)q.y << 32) | q.x; };
    auto pair_hi = [](const uint4 &q) { return ((uint64_t)q.w << 32) | q.z; };
    // the run's pairs from one base at constant offsets
    const int np = (c1 - c0) >> 1;
    uint4 *Qc = Q + (c0 >> 1);
    uint64_t local = 0;
#pragma unroll 4
    for (int k = 0; k < np; ++k) {
        const uint4 q = Qc[k];
        local += pair_lo(q) + pair_hi(q);
    }
    if ((c1 - c0) & 1) local += P[c1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t incl = wave_incl_scan(local);
    if (lane == 63) waveTot[wave] = incl;
    __syncthreads();
    uint64_t run = incl - local;
    for (int w = 0; w < wave; ++w) run += waveTot[w];
    if (threadIdx.x == 0) P[0] = 0;
#pragma unroll 4
    for (int k = 0; k < np; ++k) {
        const uint4 q = Qc[k];
        const uint64_t a = run + pair_lo(q), b = a + pair_hi(q);
        Qc[k] = make_uint4((uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32));
        run = b;
    }
    if ((c1 - c0) & 1) P[c1] += run;
    __syncthreads();

    const int e = (int)J.edge;
    const uint32_t divI = D[2 * e + 1];
    global_u32 *trow = (global_u32 *)(J.tmp + (size_t)y * W);
    auto blur_px = [&](int x) {
        // the wave's 64 pixels [xw, xw + 63] (lane 0 is active whenever
        // any lane is): a full window for all of them, or clipped ones
        const int xw = __builtin_amdgcn_readfirstlane(x);
        if (xw >= e && xw + 63 + e <= W - 1) {
            __builtin_nontemporal_store(ckd_b07_div_pack(P[x + e + 1] - P[x - e], divI), &trow[x]);
        } else {
            const int lo = max(x - e, 0), hi = min(x + e, W - 1);
            __builtin_nontemporal_store(ckd_b07_div_pack(P[hi + 1] - P[lo], D[hi - lo + 1]), &trow[x]);
        }
    };
    // four pixels at a time while all four are in the row: their LDS reads
    // and stores at constant offsets from one base
    int x = threadIdx.x;
    for (; x + 3 * kT1Block < W; x += 4 * kT1Block) {
#pragma unroll
        for (int k = 0; k < 4; ++k) blur_px(x + k * kT1Block);
    }
    for (; x < W; x += kT1Block) blur_px(x);
}

__global__ __launch_bounds__(64) void k_tunnel_v(const TunJob *__restrict__ jobs, int nJobs,
                                                 const TunJob single, int kT2Seg)
{
    const uint32_t unit = blockIdx.x;
    __shared__ __align__(16) uint16_t D[256];
    const TunJob J = pick(jobs, nJobs, single, unit, &TunJob::vBase);
    copy_div_table<64>(D);
    __syncthreads();
    const uint32_t local = unit - J.vBase;
    const int x = (int)((local % J.vBlocksX) * 64 + threadIdx.x);
    const int y0 = (int)((local / J.vBlocksX) * kT2Seg);
    const int W = (int)J.W, H = (int)J.H, e = (int)J.edge;
    // lanes past the right edge take the last column: they compute and store
    // exactly what its own lane does, so every lane runs the same
    // (wave-uniform, scalar) row loop without a branch around the stores
    const int xc = min(x, W - 1);
    const int y1 = min(y0 + kT2Seg, H);
    // rows are a wave-uniform base plus the lane's column: the row
    // arithmetic stays scalar, each access one 32-bit lane offset
    const global_u32c *__restrict__ tmp = (const global_u32c *)J.tmp;
    global_u32 *__restrict__ dst = (global_u32 *)J.dst;
    auto T = [&](int r) { return (tmp + (size_t)r * W)[xc]; };
    auto Drow = [&](int r) { return dst + (size_t)r * W; };

    uint64_t F = 0;
    for (int r = max(0, y0 - e); r <= min(H - 1, y0 + e); ++r) F += ckd_spread(T(r));
    // a row whose window is clipped by the image: its own count, and the
    // rows entering and leaving only where they exist
    auto edge_row = [&](int y) {
        const int lo = max(y - e, 0), hi = min(y + e, H - 1);
        global_u32 *d = Drow(y);
        d[xc] = ckd_adds_u8x4(d[xc], ckd_b07_div_pack(F, D[hi - lo + 1]));
        if (y + e + 1 <= H - 1) F += ckd_spread(T(y + e + 1));
        if (y - e >= 0) F -= ckd_spread(T(y - e));
    };
    // rows y in [e, H - 2 - e]: full window of 2e+1 rows, row y+e+1 enters
    // and row y-e leaves
    const int iBeg = max(y0, e), iEnd = min(y1, H - 1 - e);
    int y = y0;
    for (; y < min(iBeg, y1); ++y) edge_row(y);
    const uint32_t div = D[2 * e + 1];
    for (; y + 4 <= iEnd; y += 4) {
        uint32_t dv[4], av[4], sv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            dv[j] = Drow(y + j)[xc];
            av[j] = T(y + j + e + 1);
            sv[j] = T(y + j - e);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            Drow(y + j)[xc] = ckd_adds_u8x4(dv[j], ckd_b07_div_pack(F, div));
            F += ckd_spread(av[j]) - ckd_spread(sv[j]);
        }
    }
    for (; y < iEnd; ++y) {
        global_u32 *d = Drow(y);
        d[xc] = ckd_adds_u8x4(d[xc], ckd_b07_div_pack(F, div));
        F += ckd_spread(T(y + e + 1)) - ckd_spread(T(y - e));
    }
    for (; y < y1; ++y) edge_row(y);
}

// T2 with the window in LDS: each wave keeps the 2e+1 rows of its 64
// columns that the window spans, row r in slot r mod (2e+1), so the row
// leaving the window comes from LDS instead of a second read of tmp (at
// 2e+1 rows' distance it has usually left L2).  The row entering the
// window takes the slot of the row leaving it.  U rows per group: their
// loads of dst and tmp are in flight together.
constexpr int kRingMaxEdge = 32;
template <int U>
__global__ __launch_bounds__(64) void k_tunnel_v_ring(const TunJob *__restrict__ jobs, int nJobs,
                                                      const TunJob single, int kT2Seg)
{
    const uint32_t unit = blockIdx.x;
    extern __shared__ uint32_t ring[]; // (2 * maxEdge + 1) slots of 64 lanes
    __shared__ __align__(16) uint16_t D[256];
    const TunJob J = pick(jobs, nJobs, single, unit, &TunJob::vBase);
    copy_div_table<64>(D);
    __syncthreads();
    const uint32_t local = unit - J.vBase;
    const int x = (int)((local % J.vBlocksX) * 64 + threadIdx.x);
    const int y0 = (int)((local / J.vBlocksX) * kT2Seg);
    const int W = (int)J.W, H = (int)J.H, e = (int)J.edge;
    const int xc = min(x, W - 1);
    const int y1 = min(y0 + kT2Seg, H);
    const int R = 2 * e + 1;
    const global_u32c *__restrict__ tmp = (const global_u32c *)J.tmp;
    global_u32 *__restrict__ dst = (global_u32 *)J.dst;
    auto T = [&](int r) { return (tmp + (size_t)r * W)[xc]; };
    auto Drow = [&](int r) { return dst + (size_t)r * W; };
    uint32_t *mine = ring + threadIdx.x;

    // the window of row y0, rows clipped to the image, into F and the ring;
    // U rows' loads at a time are in flight together (past the window's last
    // row they read that row again, and are not used)
    uint64_t F = 0;
    {
        const int r0 = max(0, y0 - e), r1 = min(H - 1, y0 + e);
        int s = r0 % R;
        for (int r = r0; r <= r1; r += U) {
            const int n = min(U, r1 + 1 - r);
            uint32_t v[U];
#pragma unroll
            for (int j = 0; j < U; ++j) v[j] = T(min(r + j, r1));
            // keeps the compiler from moving each load into the branch that
            // uses it, behind the previous one's wait
#pragma unroll
            for (int j = 0; j < U; ++j) asm volatile("" : "+v"(v[j]));
#pragma unroll
            for (int j = 0; j < U; ++j) {
                if (j < n) {
                    F += ckd_spread(v[j]);
                    mine[s * 64] = v[j];
                    s = s + 1 == R ? 0 : s + 1;
                }
            }
        }
    }
    // slot of row y - e (taken mod R also where y - e < 0): the leaving row
    // at row y, and the slot the entering row y + e + 1 takes
    int sl = ((y0 - e) % R + R) % R;
    auto advance = [&]() { sl = sl + 1 == R ? 0 : sl + 1; };
    // rows [ya, yb) whose window the image clips: their own counts, and the
    // rows entering and leaving only where they exist; U rows' loads at a
    // time in flight together (loads past the group's last row or the
    // image's read those rows again, and are not used)
    auto edge_rows = [&](int ya, int yb) {
        for (int g = ya; g < yb; g += U) {
            const int n = min(U, yb - g);
            uint32_t dg[U], ag[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                dg[j] = Drow(min(g + j, yb - 1))[xc];
                ag[j] = T(min(g + j + e + 1, H - 1));
            }
#pragma unroll
            for (int j = 0; j < U; ++j) {
                if (j < n) {
                    const int yy = g + j;
                    const int lo = max(yy - e, 0), hi = min(yy + e, H - 1);
                    Drow(yy)[xc] = ckd_adds_u8x4(dg[j], ckd_b07_div_pack(F, D[hi - lo + 1]));
                    if (yy - e >= 0) F -= ckd_spread(mine[sl * 64]);
                    if (yy + e + 1 <= H - 1) {
                        F += ckd_spread(ag[j]);
                        mine[sl * 64] = ag[j];
                    }
                    advance();
                }
            }
        }
    };
    const int iBeg = max(y0, e), iEnd = min(y1, H - 1 - e);
    int y = y0;
    edge_rows(y, min(iBeg, y1));
    y = max(y, min(iBeg, y1));
    const uint32_t div = D[2 * e + 1];
    // groups of U rows, the next group's loads in flight while this one is
    // blended; with R >= U the group's U leaving rows are distinct slots and
    // are read before any entering row is written
    if (y + U <= iEnd) {
        uint32_t dv[U], av[U];
        auto load = [&](int g, uint32_t *d, uint32_t *a) {
#pragma unroll
            for (int j = 0; j < U; ++j) {
                d[j] = Drow(g + j)[xc];
                a[j] = T(g + j + e + 1);
            }
        };
        auto blend = [&](int g, const uint32_t *d, const uint32_t *a) {
            uint32_t leaving[U];
            if (R >= U) {
                int t = sl;
#pragma unroll
                for (int j = 0; j < U; ++j) {
                    leaving[j] = mine[t * 64];
                    t = t + 1 == R ? 0 : t + 1;
                }
#pragma unroll
                for (int j = 0; j < U; ++j) {
                    mine[sl * 64] = a[j];
                    advance();
                }
            } else {
#pragma unroll
                for (int j = 0; j < U; ++j) {
                    leaving[j] = mine[sl * 64];
                    mine[sl * 64] = a[j];
                    advance();
                }
            }
#pragma unroll
            for (int j = 0; j < U; ++j) {
                Drow(g + j)[xc] = ckd_adds_u8x4(d[j], ckd_b07_div_pack(F, div));
                F += ckd_spread(a[j]) - ckd_spread(leaving[j]);
            }
        };
        load(y, dv, av);
        for (; y + 2 * U <= iEnd; y += U) {
            uint32_t dn[U], an[U];
            load(y + U, dn, an);
            blend(y, dv, av);
#pragma unroll
            for (int j = 0; j < U; ++j) {
                dv[j] = dn[j];
                av[j] = an[j];
            }
        }
        blend(y, dv, av);
        y += U;
    }
    for (; y < iEnd; ++y) {
        global_u32 *d = Drow(y);
        d[xc] = ckd_adds_u8x4(d[xc], ckd_b07_div_pack(F, div));
        const uint32_t a = T(y + e + 1), leaving = mine[sl * 64];
        mine[sl * 64] = a;
        F += ckd_spread(a) - ckd_spread(leaving);
        advance();
    }
    edge_rows(y, y1);
}

} // namespace

namespace ckd {

void tunnel_job_geometry(TunJob &J, uint32_t *dst, uint32_t *tmp, const int32_t *grid, unsigned W,
                         unsigned H, int radius)
{
    J.dst = dst;
    J.tmp = tmp;
    J.grid = grid;
    J.W = W;
    J.H = H;
    J.edge = (uint32_t)radius;
    J.vBlocksX = (W + 63) / 64;
}

uint32_t tunnel_vblocks(unsigned W, unsigned H, int seg) { return ((W + 63) / 64) * ((H + seg - 1) / seg); }

unsigned tunnel_ring_rows(unsigned maxEdge)
{
    return maxEdge <= (unsigned)kRingMaxEdge ? 2 * maxEdge + 1 : 0;
}

size_t tex_quad_bytes(unsigned texLog2)
{
    return texLog2 <= (unsigned)kTexQuadMaxLog2 ? sizeof(uint4) << (2 * texLog2) : 0;
}

ckd_status launch_tunnel_jobs(ckd_ctx *ctx, const TunJob *jobs, int nJobs, const TunJob &single,
                              uint32_t totalRows, uint32_t totalVBlocks, unsigned maxW,
                              unsigned maxEdge, const uint32_t *tex, uint4 *quads, unsigned texLog2,
                              bool timed)
{
    if (totalRows == 0) return CKD_OK;
    // the row's per-block interpolation terms, its spread pixels / prefix sums
    const size_t lds = (size_t)((maxW + 7) / 8) * sizeof(uint4) + (size_t)(maxW + 2) * sizeof(uint64_t);
    {
        KernelTimer kt(ctx, timed ? CKD_KT_TUNNEL_H : -1);
        if (quads) {
            // rebuilt on every call (and graph replay): the texture's
            // contents may have changed since the last one
            k_tex_quads<<<ckd_div_up((size_t)1 << (2 * texLog2), 256), 256, 0, ctx->stream>>>(tex, texLog2,
                                                                                             quads);
            k_tunnel_h<true><<<totalRows, kT1Block, lds, ctx->stream>>>(jobs, nJobs, single, tex, quads,
                                                                        texLog2);
        } else {
            k_tunnel_h<false><<<totalRows, kT1Block, lds, ctx->stream>>>(jobs, nJobs, single, tex, nullptr,
                                                                         texLog2);
        }
    }
    ckd_status s = check_launch("tunnel remap+H");
    if (s != CKD_OK) return s;
    KernelTimer kt(ctx, timed ? CKD_KT_TUNNEL_V : -1);
    // the window ring holds up to 2 * kRingMaxEdge + 1 rows (16.6 KB per
    // wave); wider windows read their leaving rows again from memory
    const unsigned ringRows = tunnel_ring_rows(maxEdge);
    if (ringRows)
        k_tunnel_v_ring<8><<<totalVBlocks, 64, (size_t)ringRows * 64 * sizeof(uint32_t), ctx->stream>>>(
            jobs, nJobs, single, ctx->tuning.tunnelVSeg);
    else
        k_tunnel_v<<<totalVBlocks, 64, 0, ctx->stream>>>(jobs, nJobs, single, ctx->tuning.tunnelVSeg);
    return check_launch("tunnel V+add");
}

} // namespace ckd
