// k_batch.hip -- the batched renderer: many frame requests of mixed effects
// and resolutions.
//
// The host sorts the requests into three job tables (SDF, voxel, tunnel)
// and stages them in one device buffer together with the SDF chunk-queue
// counter.  One persistent SDF kernel runs on the context stream; the voxel
// kernel (every voxel column of every request) and the two fused tunnel
// kernels run beside it on the auxiliary stream.  The same launch sequence
// is replayed eagerly (ckd_render_batch), timed over candidate launch
// shapes (ckd_autotune_batch) or captured once as a hipGraph
// (ckd_batch_graph_create / ckd_graph_launch).
#include "ckd_internal.h"
#include "ckd_jobs.h"

#include <cstring>
#include <vector>

namespace ckd {
ckd_status launch_sdf_persistent(ckd_ctx *ctx, uint32_t *queueCounter, uint32_t totalChunks,
                                 const SdfJob *jobs, int nJobs, const ckd_sdf_scene &sc,
                                 int shading, unsigned long long *stepCount, bool helper,
                                 bool alone);
void voxel_job_geometry(VoxJob &J, unsigned W, unsigned H, const ckd_landscape_cam &c);
ckd_status launch_landscape_jobs(ckd_ctx *ctx, const VoxJob *jobs, int nJobs, uint32_t totalBlocks);
void tunnel_job_geometry(TunJob &J, uint32_t *dst, uint32_t *tmp, const int32_t *grid, unsigned W,
                         unsigned H, int radius);
uint32_t tunnel_vblocks(unsigned W, unsigned H, int seg);
unsigned tunnel_ring_rows(unsigned maxEdge);
ckd_status launch_tunnel_jobs(ckd_ctx *ctx, const TunJob *jobs, int nJobs, const TunJob &single,
                              uint32_t totalRows, uint32_t totalVBlocks, unsigned maxW,
                              unsigned maxEdge, const uint32_t *tex, uint4 *quads, unsigned texLog2,
                              bool timed);
size_t tex_quad_bytes(unsigned texLog2);
} // namespace ckd

// per request, what fixes the captured launch geometry: the effect and the
// size
struct ReqShape {
    int effect;
    unsigned W, H;
};

struct ckd_graph {
    ckd_ctx *ctx = nullptr;
    void *tables = nullptr;
    size_t tableBytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool hasSdf = false;
    // what ckd_graph_update must keep: the launch shapes the tables were
    // laid out with, the scratch image, the requests' shapes
    ckd_tuning tuning{};
    uint32_t *tmp = nullptr;
    size_t tmpPixels = 0;
    std::vector<ReqShape> shapes;
    // window rows of the captured vertical-blur ring (0: no ring): the
    // largest radius a replay can take
    unsigned tunRingRows = 0;
    uint4 *quads = nullptr;
};

namespace {

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// host image of the staged tables: [queue counter][SDF][voxel][tunnel]
struct Plan {
    std::vector<unsigned char> host;
    size_t oSdf = 256, oVox = 0, oTun = 0;
    int nSdf = 0, nVox = 0, nTun = 0;
    uint32_t chunks = 0, voxBlocks = 0, tunRows = 0, tunV = 0;
    unsigned tunMaxW = 0;
    unsigned tunMaxE = 0;
};

__global__ void k_reset_counter(uint32_t *c) { *c = 0u; }

ckd_status validate(const char *fn, const ckd_request *requests, unsigned n,
                    const ckd_sdf_scene *scene, int shading, const uint32_t *tex, unsigned texLog2,
                    const uint32_t *tmp, size_t tmpPixels)
{
    CKD_REQUIRE(n == 0 || requests, "%s: null request array", fn);
    size_t need = 0;
    for (unsigned i = 0; i < n; ++i) {
        const ckd_request &r = requests[i];
        CKD_REQUIRE(r.dst != nullptr, "%s: request %u: null dst", fn, i);
        CKD_REQUIRE(r.W >= 8 && r.H >= 8 && r.W <= 4096 && r.H <= 2160,
                    "%s: request %u: bad size %ux%u", fn, i, r.W, r.H);
        switch (r.effect) {
        case CKD_FX_SDF:
            CKD_REQUIRE(scene != nullptr, "%s: request %u: SDF needs a scene", fn, i);
            break;
        case CKD_FX_VOXEL:
            CKD_REQUIRE(r.vcam.mapTilt >= 0 && r.vcam.mapTilt <= 180,
                        "%s: request %u: mapTilt %d outside 0..180", fn, i, r.vcam.mapTilt);
            CKD_REQUIRE(r.W <= 64 * r.H, "%s: request %u: voxel frame %ux%u wider than 64 times its height",
                        fn, i, r.W, r.H);
            break;
        case CKD_FX_TUNNEL:
            CKD_REQUIRE(r.grid && tex, "%s: request %u: TUNNEL needs grid and texture", fn, i);
            CKD_REQUIRE(texLog2 >= 1 && texLog2 <= 14, "%s: texLog2 %u out of 1..14", fn, texLog2);
            CKD_REQUIRE(r.radius >= 1 && r.radius <= 127 &&
                            (unsigned)(2 * r.radius + 1) <= (r.W < r.H ? r.W : r.H),
                        "%s: request %u: radius %d out of range", fn, i, r.radius);
            need += (size_t)r.W * r.H;
            break;
        default:
            return ckd::fail(CKD_EINVAL, "%s: request %u: unknown effect %d", fn, i, r.effect);
        }
    }
    CKD_REQUIRE(need <= tmpPixels && (need == 0 || tmp),
                "%s: TUNNEL requests need %zu scratch pixels, got %zu", fn, need, tmpPixels);
    CKD_REQUIRE(shading == CKD_SHADE_FP32 || shading == CKD_SHADE_FP16,
                "%s: unknown shading mode %d", fn, shading);
    if (scene)
        CKD_REQUIRE(scene->nSpheres >= 0 && scene->nSpheres <= 32 && scene->maxSteps >= 1 &&
                        scene->maxSteps <= 1024,
                    "%s: bad scene (nSpheres %d, maxSteps %d)", fn, scene->nSpheres, scene->maxSteps);
    return CKD_OK;
}

void make_plan(const ckd_tuning &T, const ckd_request *requests, unsigned n, uint32_t *tmp,
               Plan &P)
{
    std::vector<SdfJob> sdf;
    std::vector<VoxJob> vox;
    std::vector<TunJob> tun;
    const uint32_t chunk = (uint32_t)T.sdfChunk;
    size_t off = 0;
    for (unsigned i = 0; i < n; ++i) {
        const ckd_request &r = requests[i];
        if (r.effect == CKD_FX_SDF) {
            SdfJob J;
            J.dst = r.dst;
            J.W = r.W;
            J.H = r.H;
            J.chunksX = (r.W + chunk - 1) / chunk;
            J.chunkBase = P.chunks;
            memcpy(J.cam, r.cam, sizeof(J.cam));
            J.invH = 1.0f / (float)r.H;
            P.chunks += J.chunksX * ((r.H + chunk - 1) / chunk);
            sdf.push_back(J);
        } else if (r.effect == CKD_FX_VOXEL) {
            VoxJob J;
            J.dst = r.dst;
            ckd::voxel_job_geometry(J, r.W, r.H, r.vcam);
            J.blockBase = P.voxBlocks;
            P.voxBlocks += (r.W + 63) / 64;
            vox.push_back(J);
        } else {
            TunJob J;
            ckd::tunnel_job_geometry(J, r.dst, tmp + off, r.grid, r.W, r.H, r.radius);
            off += (size_t)r.W * r.H;
            J.rowBase = P.tunRows;
            J.vBase = P.tunV;
            P.tunRows += r.H;
            P.tunV += ckd::tunnel_vblocks(r.W, r.H, T.tunnelVSeg);
            if (r.W > P.tunMaxW) P.tunMaxW = r.W;
            if ((unsigned)r.radius > P.tunMaxE) P.tunMaxE = (unsigned)r.radius;
            tun.push_back(J);
        }
    }
    P.nSdf = (int)sdf.size();
    P.nVox = (int)vox.size();
    P.nTun = (int)tun.size();
    P.oVox = P.oSdf + align256(sdf.size() * sizeof(SdfJob));
    P.oTun = P.oVox + align256(vox.size() * sizeof(VoxJob));
    P.host.assign(P.oTun + align256(tun.size() * sizeof(TunJob)), 0);
    if (!sdf.empty()) memcpy(P.host.data() + P.oSdf, sdf.data(), sdf.size() * sizeof(SdfJob));
    if (!vox.empty()) memcpy(P.host.data() + P.oVox, vox.data(), vox.size() * sizeof(VoxJob));
    if (!tun.empty()) memcpy(P.host.data() + P.oTun, tun.data(), tun.size() * sizeof(TunJob));
}

// the launch sequence over tables staged at `dev`
ckd_status launch_plan(ckd_ctx *ctx, const Plan &P, unsigned char *dev, const ckd_sdf_scene *scene,
                       int shading, const uint32_t *tex, uint4 *quads, unsigned texLog2,
                       unsigned long long *stepCount)
{
    ckd_status s = CKD_OK;
    const bool fork = P.nSdf > 0 && (P.nVox > 0 || P.nTun > 0);
    if (fork) {
        CKD_HIP(hipEventRecord(ctx->forkEv, ctx->stream));
        CKD_HIP(hipStreamWaitEvent(ctx->aux, ctx->forkEv, 0));
    }
    if (P.nSdf > 0) {
        s = ckd::launch_sdf_persistent(ctx, reinterpret_cast<uint32_t *>(dev), P.chunks,
                                       reinterpret_cast<const SdfJob *>(dev + P.oSdf), P.nSdf,
                                       *scene, shading, stepCount, false, !fork);
        if (s != CKD_OK) return s;
    }
    // with both voxel and tunnel work beside the march, the tunnel work runs
    // on a third stream, beside the voxel work
    const bool split = fork && P.nVox > 0 && P.nTun > 0;
    if (split) {
        CKD_HIP(hipStreamWaitEvent(ctx->aux2, ctx->forkEv, 0));
        ckd::StreamSwap swap(ctx, ctx->aux2);
        TunJob none;
        memset(&none, 0, sizeof(none));
        s = ckd::launch_tunnel_jobs(ctx, reinterpret_cast<const TunJob *>(dev + P.oTun), P.nTun, none,
                                    P.tunRows, P.tunV, P.tunMaxW, P.tunMaxE, tex, quads, texLog2, true);
        if (s != CKD_OK) return s;
        CKD_HIP(hipEventRecord(ctx->tunEv, ctx->aux2));
    }
    {
        ckd::StreamSwap swap(ctx, fork ? ctx->aux : ctx->stream);
        if (P.nVox > 0) {
            s = ckd::launch_landscape_jobs(ctx, reinterpret_cast<const VoxJob *>(dev + P.oVox),
                                           P.nVox, P.voxBlocks);
            if (s != CKD_OK) return s;
        }
        if (P.nTun > 0 && !split) {
            TunJob none;
            memset(&none, 0, sizeof(none));
            s = ckd::launch_tunnel_jobs(ctx, reinterpret_cast<const TunJob *>(dev + P.oTun), P.nTun,
                                        none, P.tunRows, P.tunV, P.tunMaxW, P.tunMaxE, tex, quads, texLog2, true);
            if (s != CKD_OK) return s;
        }
        if (split) CKD_HIP(hipStreamWaitEvent(ctx->aux, ctx->tunEv, 0));
        if (fork) {
            // the voxel / tunnel work is done: its CUs join the SDF march
            s = ckd::launch_sdf_persistent(ctx, reinterpret_cast<uint32_t *>(dev), P.chunks,
                                           reinterpret_cast<const SdfJob *>(dev + P.oSdf), P.nSdf,
                                           *scene, shading, stepCount, true, false);
            if (s != CKD_OK) return s;
        }
    }
    if (fork) {
        CKD_HIP(hipEventRecord(ctx->joinEv, ctx->aux));
        CKD_HIP(hipStreamWaitEvent(ctx->stream, ctx->joinEv, 0));
    }
    return CKD_OK;
}

ckd_status run_eager(ckd_ctx *ctx, const ckd_request *requests, unsigned n,
                     const ckd_sdf_scene *scene, int shading, const uint32_t *tex,
                     unsigned texLog2, uint32_t *tmp, unsigned long long *stepCount)
{
    Plan P;
    make_plan(ctx->tuning, requests, n, tmp, P);
    void *dev = nullptr;
    ckd_status s = ckd::stage(ctx, P.host.data(), P.host.size(), &dev);
    if (s != CKD_OK) return s;
    // the texture's texel quads, in a buffer of the context's that grows
    // to the largest texture seen (freed only after the work queued on it)
    uint4 *quads = nullptr;
    const size_t qb = P.nTun ? ckd::tex_quad_bytes(texLog2) : 0;
    if (qb) {
        if (ctx->texQuadBytes < qb) {
            if (ctx->texQuads) {
                CKD_HIP(hipStreamSynchronize(ctx->stream));
                (void)hipFree(ctx->texQuads);
                ctx->texQuads = nullptr;
                ctx->texQuadBytes = 0;
            }
            if (ckd::dev_alloc(&ctx->texQuads, qb) != hipSuccess)
                return ckd::fail(CKD_ENOMEM, "ckd_render_batch: texel quad allocation failed");
            ctx->texQuadBytes = qb;
        }
        quads = ctx->texQuads;
    }
    return launch_plan(ctx, P, static_cast<unsigned char *>(dev), scene, shading, tex, quads, texLog2,
                       stepCount);
}

ckd_status valid_tuning(const ckd_tuning *t)
{
    CKD_REQUIRE(t != nullptr, "tuning: null pointer");
    CKD_REQUIRE(t->sdfStepsPerRefill == 1 || t->sdfStepsPerRefill == 2 || t->sdfStepsPerRefill == 4 ||
                    t->sdfStepsPerRefill == 8 || t->sdfStepsPerRefill == 16,
                "tuning: sdfStepsPerRefill %d not in {1,2,4,8,16}", t->sdfStepsPerRefill);
    CKD_REQUIRE(t->sdfChunk == 16 || t->sdfChunk == 32, "tuning: sdfChunk %d not in {16,32}",
                t->sdfChunk);
    CKD_REQUIRE(t->sdfBlocksPerCU >= 1 && t->sdfBlocksPerCU <= 32,
                "tuning: sdfBlocksPerCU %d outside 1..32", t->sdfBlocksPerCU);
    CKD_REQUIRE(t->tunnelVSeg >= 8 && t->tunnelVSeg <= 512, "tuning: tunnelVSeg %d outside 8..512",
                t->tunnelVSeg);
    CKD_REQUIRE(t->blurKernel >= 0 && t->blurKernel <= 2, "tuning: blurKernel %d not in {0,1,2}",
                t->blurKernel);
    CKD_REQUIRE(t->voxelChunks == 0 || t->voxelChunks == 1 || t->voxelChunks == 2 || t->voxelChunks == 4 ||
                    t->voxelChunks == 8,
                "tuning: voxelChunks %d not in {0,1,2,4,8}", t->voxelChunks);
    CKD_REQUIRE(t->voxelWavesPerSIMD >= 0 && t->voxelWavesPerSIMD <= 8,
                "tuning: voxelWavesPerSIMD %d outside 0..8", t->voxelWavesPerSIMD);
    return CKD_OK;
}

} // namespace

extern "C" {

ckd_status ckd_get_tuning(ckd_ctx *ctx, ckd_tuning *out)
{
    CKD_REQUIRE(out != nullptr, "ckd_get_tuning: null output");
    CKD_REQUIRE_CTX(ctx);
    *out = ctx->tuning;
    return CKD_OK;
}

ckd_status ckd_set_tuning(ckd_ctx *ctx, const ckd_tuning *t)
{
    ckd_status s = valid_tuning(t);
    if (s != CKD_OK) return s;
    CKD_REQUIRE_CTX(ctx);
    ctx->tuning = *t;
    if (t->voxelChunks == 0) ctx->tuning.voxelChunks = kCkdDefaultTuning.voxelChunks;
    if (t->voxelWavesPerSIMD == 0) ctx->tuning.voxelWavesPerSIMD = kCkdDefaultTuning.voxelWavesPerSIMD;
    return CKD_OK;
}

ckd_status ckd_render_batch(ckd_ctx *ctx, const ckd_request *requests, unsigned n,
                            const ckd_sdf_scene *scene, int shading, const uint32_t *tex,
                            unsigned texLog2, uint32_t *tmp, size_t tmpPixels,
                            unsigned long long *stepCount)
{
    ckd_status s = validate("ckd_render_batch", requests, n, scene, shading, tex, texLog2, tmp,
                            tmpPixels);
    if (s != CKD_OK) return s;
    CKD_REQUIRE_CTX(ctx);
    if (n == 0) return CKD_OK;
    return run_eager(ctx, requests, n, scene, shading, tex, texLog2, tmp, stepCount);
}

ckd_status ckd_autotune_batch(ckd_ctx *ctx, const ckd_request *requests, unsigned n,
                              const ckd_sdf_scene *scene, int shading, const uint32_t *tex,
                              unsigned texLog2, uint32_t *tmp, size_t tmpPixels, ckd_tuning *best)
{
    ckd_status s = validate("ckd_autotune_batch", requests, n, scene, shading, tex, texLog2, tmp,
                            tmpPixels);
    if (s != CKD_OK) return s;
    CKD_REQUIRE_CTX(ctx);
    if (n == 0) return CKD_OK;
    // fewer persistent SDF blocks per CU leave room for the voxel / tunnel
    // branch on the second stream, which otherwise waits for them to drain
    const int steps[] = {4, 8, 16}, chunks[] = {16, 32}, bpc[] = {2, 3, 4, 5, 6, 8},
              segs[] = {16, 32, 64, 128, 256};
    const ckd_tuning saved = ctx->tuning;
    const bool timing = ctx->timing;
    ctx->timing = false;
    ckd_tuning bestT = saved;
    float bestMs = 3.4e38f;
    hipEvent_t a = ctx->events[CKD_NUM_EVENTS - 2], b = ctx->events[CKD_NUM_EVENTS - 1];
    for (int st : steps)
        for (int ch : chunks)
            for (int bp : bpc)
                for (int sg : segs) {
                    ctx->tuning = ckd_tuning{st, ch, bp, sg, saved.blurKernel, saved.voxelChunks, saved.voxelWavesPerSIMD};
                    // warm once, then time the second run
                    s = run_eager(ctx, requests, n, scene, shading, tex, texLog2, tmp, nullptr);
                    if (s == CKD_OK) {
                        if (hipEventRecord(a, ctx->stream) != hipSuccess) s = CKD_EDEVICE;
                    }
                    if (s == CKD_OK)
                        s = run_eager(ctx, requests, n, scene, shading, tex, texLog2, tmp, nullptr);
                    float ms = 0.f;
                    if (s == CKD_OK && (hipEventRecord(b, ctx->stream) != hipSuccess ||
                                        hipEventSynchronize(b) != hipSuccess ||
                                        hipEventElapsedTime(&ms, a, b) != hipSuccess))
                        s = ckd::fail(CKD_EDEVICE, "ckd_autotune_batch: event timing failed");
                    if (s != CKD_OK) {
                        ctx->tuning = saved;
                        ctx->timing = timing;
                        return s;
                    }
                    if (ms < bestMs) {
                        bestMs = ms;
                        bestT = ctx->tuning;
                    }
                }
    ctx->tuning = bestT;
    ctx->timing = timing;
    if (best) *best = bestT;
    return CKD_OK;
}

ckd_status ckd_batch_graph_create(ckd_ctx *ctx, const ckd_request *requests, unsigned n,
                                  const ckd_sdf_scene *scene, int shading, const uint32_t *tex,
                                  unsigned texLog2, uint32_t *tmp, size_t tmpPixels,
                                  unsigned long long *stepCount, ckd_graph **out)
{
    CKD_REQUIRE(out != nullptr, "ckd_batch_graph_create: null output");
    *out = nullptr;
    ckd_status s = validate("ckd_batch_graph_create", requests, n, scene, shading, tex, texLog2,
                            tmp, tmpPixels);
    if (s != CKD_OK) return s;
    CKD_REQUIRE(n > 0, "ckd_batch_graph_create: empty batch");
    CKD_REQUIRE_CTX(ctx);
    Plan P;
    make_plan(ctx->tuning, requests, n, tmp, P);
    ckd_graph *g = new ckd_graph();
    g->ctx = ctx;
    g->tableBytes = P.host.size();
    g->tuning = ctx->tuning;
    g->tmp = tmp;
    g->tmpPixels = tmpPixels;
    g->tunRingRows = P.nTun ? ckd::tunnel_ring_rows(P.tunMaxE) : 0;
    for (unsigned i = 0; i < n; ++i)
        g->shapes.push_back({requests[i].effect, requests[i].W, requests[i].H});
    if (ckd::dev_alloc(&g->tables, P.host.size()) != hipSuccess) {
        delete g;
        return ckd::fail(CKD_ENOMEM, "ckd_batch_graph_create: table allocation failed");
    }
    if (hipMemcpy(g->tables, P.host.data(), P.host.size(), hipMemcpyHostToDevice) != hipSuccess) {
        ckd_graph_destroy(g);
        return ckd::fail(CKD_EDEVICE, "ckd_batch_graph_create: table upload failed");
    }
    // the graph's own texel quads: rebuilt from the texture by every replay
    const size_t qb = P.nTun ? ckd::tex_quad_bytes(texLog2) : 0;
    if (qb && ckd::dev_alloc(&g->quads, qb) != hipSuccess) {
        ckd_graph_destroy(g);
        return ckd::fail(CKD_ENOMEM, "ckd_batch_graph_create: texel quad allocation failed");
    }
    // capture on a private stream so nothing else on the context stream is
    // recorded; the kernel-timing events are not part of the graph
    hipStream_t cap = nullptr;
    if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) {
        ckd_graph_destroy(g);
        return ckd::fail(CKD_EDEVICE, "ckd_batch_graph_create: stream creation failed");
    }
    g->hasSdf = P.nSdf > 0;
    const bool timing = ctx->timing;
    ctx->timing = false;
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        ckd::StreamSwap swap(ctx, cap);
        unsigned char *dev = static_cast<unsigned char *>(g->tables);
        // the queue counter is reset by a kernel node on every replay
        k_reset_counter<<<1, 1, 0, cap>>>(reinterpret_cast<uint32_t *>(dev));
        s = launch_plan(ctx, P, dev, scene, shading, tex, g->quads, texLog2, stepCount);
        const hipError_t e2 = hipStreamEndCapture(cap, &g->graph);
        if (s == CKD_OK && e2 != hipSuccess)
            s = ckd::fail(CKD_EDEVICE, "ckd_batch_graph_create: capture failed: %s",
                          hipGetErrorString(e2));
    } else {
        s = ckd::fail(CKD_EDEVICE, "ckd_batch_graph_create: hipStreamBeginCapture: %s",
                      hipGetErrorString(e));
    }
    ctx->timing = timing;
    (void)hipStreamDestroy(cap);
    if (s == CKD_OK && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
        s = ckd::fail(CKD_EDEVICE, "ckd_batch_graph_create: hipGraphInstantiate failed");
    if (s != CKD_OK) {
        ckd_graph_destroy(g);
        return s;
    }
    *out = g;
    return CKD_OK;
}

ckd_status ckd_graph_update(ckd_graph *g, const ckd_request *requests, unsigned n)
{
    CKD_REQUIRE(g && g->exec, "ckd_graph_update: null graph");
    CKD_REQUIRE(n == g->shapes.size(), "ckd_graph_update: %u requests, the graph holds %zu", n,
                g->shapes.size());
    CKD_REQUIRE(requests, "ckd_graph_update: null request array");
    for (unsigned i = 0; i < n; ++i) {
        const ckd_request &r = requests[i];
        const ReqShape &c = g->shapes[i];
        CKD_REQUIRE(r.effect == c.effect && r.W == c.W && r.H == c.H,
                    "ckd_graph_update: request %u is effect %d at %ux%u, the graph's is %d at %ux%u", i,
                    r.effect, r.W, r.H, c.effect, c.W, c.H);
        CKD_REQUIRE(r.dst != nullptr, "ckd_graph_update: request %u: null dst", i);
        if (r.effect == CKD_FX_VOXEL)
            CKD_REQUIRE(r.vcam.mapTilt >= 0 && r.vcam.mapTilt <= 180,
                        "ckd_graph_update: request %u: mapTilt %d outside 0..180", i, r.vcam.mapTilt);
        if (r.effect == CKD_FX_TUNNEL) {
            CKD_REQUIRE(r.grid != nullptr, "ckd_graph_update: request %u: null grid", i);
            CKD_REQUIRE(r.radius >= 1 && r.radius <= 127 &&
                            (unsigned)(2 * r.radius + 1) <= (r.W < r.H ? r.W : r.H),
                        "ckd_graph_update: request %u: radius %d out of range", i, r.radius);
        }
    }
    Plan P;
    make_plan(g->tuning, requests, n, g->tmp, P);
    if (P.host.size() != g->tableBytes)
        return ckd::fail(CKD_EINVAL, "ckd_graph_update: the requests' tables do not fit the graph");
    if (g->tunRingRows && 2 * P.tunMaxE + 1 > g->tunRingRows)
        return ckd::fail(CKD_EINVAL,
                         "ckd_graph_update: radius %u exceeds the captured batch's largest (%u)",
                         P.tunMaxE, (g->tunRingRows - 1) / 2);
    // a replay still in flight reads the tables: the copy waits for it, and
    // the next replay is queued after the copy
    ckd_ctx *ctx = g->ctx;
    CKD_HIP(hipStreamSynchronize(ctx->stream));
    CKD_HIP(hipMemcpy(g->tables, P.host.data(), P.host.size(), hipMemcpyHostToDevice));
    return CKD_OK;
}

ckd_status ckd_graph_launch(ckd_graph *g)
{
    CKD_REQUIRE(g && g->exec, "ckd_graph_launch: null graph");
    ckd_ctx *ctx = g->ctx;
    // with kernel timing on, a replay is timed as one persistent-SDF launch:
    // events on the context stream around the whole graph, whose critical
    // path is the SDF kernel (the counter reset and the fork/join add
    // microseconds; the voxel and tunnel branch runs beside it)
    ckd::KernelTimer kt(ctx, g->hasSdf ? CKD_KT_SDF_PERSISTENT : -1);
    CKD_HIP(hipGraphLaunch(g->exec, ctx->stream));
    return CKD_OK;
}

void ckd_graph_destroy(ckd_graph *g)
{
    if (!g) return;
    if (g->ctx) (void)hipStreamSynchronize(g->ctx->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->tables) (void)hipFree(g->tables);
    if (g->quads) (void)hipFree(g->quads);
    delete g;
}

} // extern "C"
