// k_batch.hip -- ckd_render_batch: many frame requests of mixed effects and
// resolutions.  The host sorts the requests into three job tables (SDF,
// voxel, tunnel), stages them in one device buffer together with the SDF
// chunk-queue counter, and launches one persistent SDF kernel, one voxel
// kernel over every voxel column and the two fused tunnel kernels.
#include "ckd_internal.h"
#include "ckd_jobs.h"

#include <cstring>
#include <vector>

namespace ckd {
ckd_status launch_sdf_persistent(ckd_ctx *ctx, uint32_t *queueCounter, uint32_t totalChunks,
                                 const SdfJob *jobs, int nJobs, const ckd_sdf_scene &sc,
                                 int shading, unsigned long long *stepCount);
void voxel_job_geometry(VoxJob &J, unsigned W, unsigned H, const ckd_landscape_cam &c);
ckd_status launch_landscape_jobs(ckd_ctx *ctx, const VoxJob *jobs, int nJobs, uint32_t totalBlocks);
void tunnel_job_geometry(TunJob &J, uint32_t *dst, uint32_t *tmp, const int32_t *grid, unsigned W,
                         unsigned H, int radius);
uint32_t tunnel_vblocks(unsigned W, unsigned H);
ckd_status launch_tunnel_jobs(ckd_ctx *ctx, const TunJob *jobs, int nJobs, const TunJob &single,
                              uint32_t totalRows, uint32_t totalVBlocks, unsigned maxW,
                              const uint32_t *tex, unsigned texLog2);
} // namespace ckd

namespace {

constexpr int kSdfChunk = 32;

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

} // namespace

extern "C" {

ckd_status ckd_render_batch(ckd_ctx *ctx, const ckd_request *requests, unsigned n,
                            const ckd_sdf_scene *scene, int shading, const uint32_t *tex,
                            unsigned texLog2, uint32_t *tmp, size_t tmpPixels,
                            unsigned long long *stepCount)
{
    CKD_REQUIRE(n == 0 || requests, "ckd_render_batch: null request array");
    size_t need = 0;
    for (unsigned i = 0; i < n; ++i) {
        const ckd_request &r = requests[i];
        CKD_REQUIRE(r.dst != nullptr, "ckd_render_batch: request %u: null dst", i);
        CKD_REQUIRE(r.W >= 8 && r.H >= 8 && r.W <= 4096 && r.H <= 2160,
                    "ckd_render_batch: request %u: bad size %ux%u", i, r.W, r.H);
        switch (r.effect) {
        case CKD_FX_SDF:
            CKD_REQUIRE(scene != nullptr, "ckd_render_batch: request %u: SDF needs a scene", i);
            break;
        case CKD_FX_VOXEL:
            CKD_REQUIRE(r.vcam.mapTilt >= 0 && r.vcam.mapTilt <= 180,
                        "ckd_render_batch: request %u: mapTilt %d outside 0..180", i, r.vcam.mapTilt);
            break;
        case CKD_FX_TUNNEL:
            CKD_REQUIRE(r.grid && tex, "ckd_render_batch: request %u: TUNNEL needs grid and texture", i);
            CKD_REQUIRE(texLog2 >= 1 && texLog2 <= 14, "ckd_render_batch: texLog2 %u out of 1..14",
                        texLog2);
            CKD_REQUIRE(r.radius >= 1 && r.radius <= 127 &&
                            (unsigned)(2 * r.radius + 1) <= (r.W < r.H ? r.W : r.H),
                        "ckd_render_batch: request %u: radius %d out of range", i, r.radius);
            need += (size_t)r.W * r.H;
            break;
        default:
            return ckd::fail(CKD_EINVAL, "ckd_render_batch: request %u: unknown effect %d", i, r.effect);
        }
    }
    CKD_REQUIRE(need <= tmpPixels && (need == 0 || tmp),
                "ckd_render_batch: TUNNEL requests need %zu scratch pixels, got %zu", need, tmpPixels);
    CKD_REQUIRE(shading == CKD_SHADE_FP32 || shading == CKD_SHADE_FP16,
                "ckd_render_batch: unknown shading mode %d", shading);
    if (scene) {
        CKD_REQUIRE(scene->nSpheres >= 0 && scene->nSpheres <= 32 && scene->maxSteps >= 1 &&
                        scene->maxSteps <= 1024,
                    "ckd_render_batch: bad scene (nSpheres %d, maxSteps %d)", scene->nSpheres,
                    scene->maxSteps);
    }
    CKD_REQUIRE_CTX(ctx);
    if (n == 0) return CKD_OK;

    std::vector<SdfJob> sdf;
    std::vector<VoxJob> vox;
    std::vector<TunJob> tun;
    uint32_t chunks = 0, voxBlocks = 0, tunRows = 0, tunV = 0;
    unsigned tunMaxW = 0;
    size_t off = 0;
    for (unsigned i = 0; i < n; ++i) {
        const ckd_request &r = requests[i];
        if (r.effect == CKD_FX_SDF) {
            SdfJob J;
            J.dst = r.dst;
            J.W = r.W;
            J.H = r.H;
            J.chunksX = (r.W + kSdfChunk - 1) / kSdfChunk;
            J.chunkBase = chunks;
            memcpy(J.cam, r.cam, sizeof(J.cam));
            chunks += J.chunksX * ((r.H + kSdfChunk - 1) / kSdfChunk);
            sdf.push_back(J);
        } else if (r.effect == CKD_FX_VOXEL) {
            VoxJob J;
            J.dst = r.dst;
            ckd::voxel_job_geometry(J, r.W, r.H, r.vcam);
            J.blockBase = voxBlocks;
            voxBlocks += (r.W + 63) / 64;
            vox.push_back(J);
        } else {
            TunJob J;
            ckd::tunnel_job_geometry(J, r.dst, tmp + off, r.grid, r.W, r.H, r.radius);
            off += (size_t)r.W * r.H;
            J.rowBase = tunRows;
            J.vBase = tunV;
            tunRows += r.H;
            tunV += ckd::tunnel_vblocks(r.W, r.H);
            if (r.W > tunMaxW) tunMaxW = r.W;
            tun.push_back(J);
        }
    }
    // one staging buffer: [queue counter][SDF jobs][voxel jobs][tunnel jobs]
    const size_t oSdf = 256;
    const size_t oVox = oSdf + align256(sdf.size() * sizeof(SdfJob));
    const size_t oTun = oVox + align256(vox.size() * sizeof(VoxJob));
    const size_t bytes = oTun + align256(tun.size() * sizeof(TunJob));
    std::vector<unsigned char> host(bytes, 0);
    if (!sdf.empty()) memcpy(host.data() + oSdf, sdf.data(), sdf.size() * sizeof(SdfJob));
    if (!vox.empty()) memcpy(host.data() + oVox, vox.data(), vox.size() * sizeof(VoxJob));
    if (!tun.empty()) memcpy(host.data() + oTun, tun.data(), tun.size() * sizeof(TunJob));
    void *dev = nullptr;
    ckd_status s = ckd::stage(ctx, host.data(), bytes, &dev);
    if (s != CKD_OK) return s;
    unsigned char *d = static_cast<unsigned char *>(dev);

    if (!vox.empty()) {
        s = ckd::launch_landscape_jobs(ctx, reinterpret_cast<const VoxJob *>(d + oVox), (int)vox.size(),
                                       voxBlocks);
        if (s != CKD_OK) return s;
    }
    if (!tun.empty()) {
        TunJob none;
        memset(&none, 0, sizeof(none));
        s = ckd::launch_tunnel_jobs(ctx, reinterpret_cast<const TunJob *>(d + oTun), (int)tun.size(),
                                    none, tunRows, tunV, tunMaxW, tex, texLog2);
        if (s != CKD_OK) return s;
    }
    if (!sdf.empty()) {
        s = ckd::launch_sdf_persistent(ctx, reinterpret_cast<uint32_t *>(d), chunks,
                                       reinterpret_cast<const SdfJob *>(d + oSdf), (int)sdf.size(),
                                       *scene, shading, stepCount);
        if (s != CKD_OK) return s;
    }
    return CKD_OK;
}

} // extern "C"
