// k_batch.hip -- ckd_render_batch: many frame requests of mixed effects and
// resolutions.
#include "ckd_internal.h"
#include "ckd_pixel.h"

namespace ckd {
ckd_status launch_sdf(ckd_ctx *ctx, uint32_t *dst, unsigned W, unsigned H,
                      const ckd_sdf_scene &sc, const float *cam12, int shading,
                      unsigned long long *stepCount);
}

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
            CKD_REQUIRE(r.radius >= 1 && r.radius <= 127 &&
                            (unsigned)(2 * r.radius + 1) <= (r.W < r.H ? r.W : r.H),
                        "ckd_render_batch: request %u: radius %d out of range", i, r.radius);
            need += 2 * (size_t)r.W * r.H;
            break;
        default:
            return ckd::fail(CKD_EINVAL, "ckd_render_batch: request %u: unknown effect %d", i, r.effect);
        }
    }
    CKD_REQUIRE(need <= tmpPixels && (need == 0 || tmp),
                "ckd_render_batch: TUNNEL requests need %zu scratch pixels, got %zu", need, tmpPixels);
    CKD_REQUIRE(shading == CKD_SHADE_FP32 || shading == CKD_SHADE_FP16,
                "ckd_render_batch: unknown shading mode %d", shading);
    CKD_REQUIRE_CTX(ctx);
    size_t off = 0;
    for (unsigned i = 0; i < n; ++i) {
        const ckd_request &r = requests[i];
        ckd_status s = CKD_OK;
        if (r.effect == CKD_FX_SDF) {
            s = ckd::launch_sdf(ctx, r.dst, r.W, r.H, *scene, r.cam, shading, stepCount);
        } else if (r.effect == CKD_FX_VOXEL) {
            s = ckd_landscape_draw(ctx, r.dst, r.W, r.H, &r.vcam, 1);
        } else {
            uint32_t *t0 = tmp + off, *t1 = tmp + off + (size_t)r.W * r.H;
            off += 2 * (size_t)r.W * r.H;
            s = ckd_tunnel_frame(ctx, r.dst, t0, t1, r.W, r.H, r.grid, tex, texLog2, r.radius);
        }
        if (s != CKD_OK) return s;
    }
    return CKD_OK;
}

} // extern "C"
