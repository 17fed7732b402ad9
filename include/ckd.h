/*
 * ckd.h -- C ABI of the MI355X (gfx950) renderer for cookiedough's per-pixel
 * effect loops.
 *
 * All entry points are extern "C", take plain pointers and sizes, and return
 * a ckd_status; nothing throws across this boundary.  Pixel buffers are
 * device pointers to uint32_t 0xAARRGGBB pixels (memory byte order B,G,R,A),
 * rows contiguous.  Work is enqueued on the context's stream (hipStream_t
 * passed as void*, default: the null stream) and is asynchronous unless the
 * comment says otherwise.  A context is used by one host thread at a time.
 *
 * Each entry point names the reference interface it replaces as file:line
 * of the reference repository (visualizersdotnl/cookiedough).
 */
#ifndef CKD_H
#define CKD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CKD_API __attribute__((visibility("default")))

typedef enum ckd_status {
    CKD_OK = 0,
    CKD_EINVAL = 1,  /* bad argument (size, pointer, parameter range)       */
    CKD_ENOMEM = 2,  /* device allocation failed                            */
    CKD_EDEVICE = 3, /* no usable device / HIP error                        */
    CKD_ESTATE = 4   /* missing set-up (maps not created, ...)              */
} ckd_status;

typedef struct ckd_ctx ckd_ctx;

/* ---- context ------------------------------------------------------------
 * ckd_create replaces the set-up chain of code/main.cpp:263-279
 * (CalculateCosLUT, Polar_Create, FxBlitter_Create, BoxBlur_Create) and
 * ckd_destroy the teardown at code/main.cpp:372-374.  Errors are reported
 * through ckd_last_error, the counterpart of SetLastError (code/main.h:49). */
CKD_API ckd_status ckd_create(int device, ckd_ctx **out);
CKD_API void ckd_destroy(ckd_ctx *ctx);
/* message of the last failing call on this host thread ("" if none) */
CKD_API const char *ckd_last_error(void);
CKD_API ckd_status ckd_set_stream(ckd_ctx *ctx, void *hipStream);
CKD_API ckd_status ckd_synchronize(ckd_ctx *ctx);

/* ---- device memory and timing plumbing (no reference counterpart: the
 * reference's buffers are host memory, code/main.cpp:307) ---------------- */
CKD_API ckd_status ckd_malloc(ckd_ctx *ctx, void **dptr, size_t bytes);
CKD_API ckd_status ckd_free(ckd_ctx *ctx, void *dptr);
CKD_API ckd_status ckd_memcpy_h2d(ckd_ctx *ctx, void *dst, const void *src,
                                  size_t bytes); /* stream-ordered        */
CKD_API ckd_status ckd_memcpy_d2h(ckd_ctx *ctx, void *dst, const void *src,
                                  size_t bytes); /* synchronous            */
CKD_API ckd_status ckd_memcpy_d2d(ckd_ctx *ctx, void *dst, const void *src,
                                  size_t bytes);
/* HIP events on the context stream; slots 0..15 */
CKD_API ckd_status ckd_event_record(ckd_ctx *ctx, int slot);
CKD_API ckd_status ckd_event_elapsed_ms(ckd_ctx *ctx, int slotA, int slotB,
                                        float *ms);

/* Per-kernel timing: with timing on, every launch of the kernels below is
 * bracketed by HIP events on the context stream; ckd_kernel_timing_read
 * waits for them and returns the total milliseconds and launch count since
 * timing was switched on (enable = 1 resets the totals, 0 stops). */
enum {
    CKD_KT_SDF_PLAIN = 0,      /* ckd_sdf_render                           */
    CKD_KT_SDF_PERSISTENT = 1, /* SDF requests of ckd_render_batch         */
    CKD_KT_LANDSCAPE = 2,      /* ckd_landscape_draw and voxel requests    */
    CKD_KT_TUNNEL_H = 3,       /* tunnel remap + horizontal blur           */
    CKD_KT_TUNNEL_V = 4,       /* tunnel vertical blur + additive blit     */
    CKD_KT_SHADERTOY = 5,      /* FX maps of the raymarched effects        */
    CKD_KT_COUNT = 6
};
CKD_API ckd_status ckd_kernel_timing(ckd_ctx *ctx, int enable);
CKD_API ckd_status ckd_kernel_timing_read(ckd_ctx *ctx, int kernel,
                                          double *totalMs,
                                          unsigned long long *launches);

/* Test hook, no reference counterpart (the reference's failures are a
 * Create returning false, code/main.h:49): the n-th next device allocation
 * the library makes (n = 0: the next one) fails as if the device were out
 * of memory, so tests can drive every CKD_ENOMEM path; n = -1 disarms.
 * Process-wide. */
CKD_API ckd_status ckd_debug_fail_alloc(int n);

/* ---- pixel primitives ---------------------------------------------------- */
/* memset32, code/util.h:57-67 */
CKD_API ckd_status ckd_memset32(ckd_ctx *ctx, uint32_t *dst, uint32_t value,
                                size_t numInts);
/* Fx_Blit_2x2, code/fx-blitter.cpp:27-75.  fxMap is (dstW/2+4) x (dstH/2+4)
 * pixels (code/fx-blitter.h:16-17 at any even size); dstW/2 % 4 == 0. */
CKD_API ckd_status ckd_fx_blit_2x2(ckd_ctx *ctx, uint32_t *dst,
                                   const uint32_t *fxMap, unsigned dstW,
                                   unsigned dstH);
/* Add32, code/util.cpp:132-145 */
CKD_API ckd_status ckd_add32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                             unsigned numPixels);
/* BlitAdd32, code/util.cpp:754-773 */
CKD_API ckd_status ckd_blit_add32(ckd_ctx *ctx, uint32_t *dst,
                                  const uint32_t *src, unsigned destResX,
                                  unsigned srcResX, unsigned yRes);
/* BlitAdd32A, code/util.cpp:775-796 */
CKD_API ckd_status ckd_blit_add32a(ckd_ctx *ctx, uint32_t *dst,
                                   const uint32_t *src, unsigned destResX,
                                   unsigned srcResX, unsigned yRes, float alpha);

/* TapeWarp32, code/util.cpp:552-603: per-pixel FP32 offsets from the
 * interpolated cosine LUT (code/sincos-lut.h:13-26, built by ckd_create as
 * CalculateCosLUT does, code/sincos-lut.cpp:9-16), clamped, converted to
 * 24:8 and sampled bilinearly.  The row stride of the fetch is xRes (the
 * reference uses kResX = 1280, the only width it calls this with). */
CKD_API ckd_status ckd_tape_warp32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                   unsigned xRes, unsigned yRes, float strength,
                                   float speed);

/* ---- 2007 box blur, code/deprecated/boxblur.cpp -------------------------
 * dst == src is allowed and reproduces the reference's in-place recurrence
 * (the sliding-window subtract then reads already-blurred pixels).  The
 * kernel span clampi(1,255,unsigned(strength*255)) must not exceed the pass
 * length (xRes for H, yRes for V); strength must be finite and >= 0. */
/* HorizontalBoxBlur32, code/deprecated/boxblur.cpp:44-126 */
CKD_API ckd_status ckd_blur07_h(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                unsigned xRes, unsigned yRes, float strength);
/* VerticalBoxBlur32, code/deprecated/boxblur.cpp:128-220 */
CKD_API ckd_status ckd_blur07_v(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                unsigned xRes, unsigned yRes, float strength);
/* BoxBlur32 (H, then V in place on dst), code/deprecated/boxblur.cpp:222-231 */
CKD_API ckd_status ckd_blur07(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                              unsigned xRes, unsigned yRes, float strength);
/* The in-place forms (pDest == pSrc) of HorizontalBoxBlur32 (passes 1),
 * VerticalBoxBlur32 (2) or BoxBlur32 (3) on n device images of one size,
 * each with its own strength: imgs and strengths are host arrays of n.  One
 * launch per pass for all images; the kernel is the tuning's blurKernel.
 * Each span must fit every pass it runs. */
CKD_API ckd_status ckd_blur07_inplace_batch(ckd_ctx *ctx, uint32_t *const *imgs,
                                            const float *strengths, unsigned n, unsigned xRes,
                                            unsigned yRes, int passes);

/* ---- 2026 box blur, code/boxblur.cpp -------------------------------------
 * xRes, yRes <= 2048; strength in [0,100]; gain in [0,1]; passes >= 1;
 * dst != src.  The two pixels the reference reads past the end of the last
 * row of a pass (code/boxblur.cpp:185,195) are taken as zero. */
/* BoxBlur_Horz32, code/boxblur.cpp:270-279 */
CKD_API ckd_status ckd_blur26_h(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                unsigned xRes, unsigned yRes, float strength,
                                float gain, unsigned numPasses);
/* BoxBlur_Vert32, code/boxblur.cpp:281-297 (xRes, yRes multiples of 4) */
CKD_API ckd_status ckd_blur26_v(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                unsigned xRes, unsigned yRes, float strength,
                                float gain, unsigned numPasses);
/* BoxBlur_32, code/boxblur.cpp:299-318 (xRes, yRes multiples of 4) */
CKD_API ckd_status ckd_blur26(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                              unsigned xRes, unsigned yRes, float strength,
                              float gain, unsigned numPasses);

/* ---- polar UV-map remap, code/polar.cpp ----------------------------------
 * ckd_polar_create computes the 24:8 maps of CalculateMaps
 * (code/polar.cpp:18-59) on the host in FP32 for a W x H target and uploads
 * them (Polar_Create, code/polar.cpp:61-72, which makes them for the
 * 1280 x 720 output and the 644 x 364 FX map).  The context keeps the maps
 * of up to four sizes; a fifth replaces the oldest.  Polar_Blit
 * (code/polar.cpp:135-154) then remaps src (W x H) into dst (W x H);
 * Polar_Blit_2x2 (code/polar.cpp:200-218) is the same call at the FX
 * map's size. */
CKD_API ckd_status ckd_polar_create(ckd_ctx *ctx, unsigned W, unsigned H);
CKD_API ckd_status ckd_polar_blit(ckd_ctx *ctx, uint32_t *dst,
                                  const uint32_t *src, unsigned W, unsigned H,
                                  int inverse);

/* ---- per-8x8-block UV grid remap (no reference counterpart: the grid law
 * of the tunnel benchmark; the per-pixel fetch is bsamp_prepUVs +
 * bsamp32_32, code/bilinear.h:10-125).  grid holds (ceil(W/8)+1) x
 * (ceil(H/8)+1) corner (U,V) pairs in 24:8; tex is 2^texLog2 square. */
CKD_API ckd_status ckd_grid_remap(ckd_ctx *ctx, uint32_t *dst, unsigned W,
                                  unsigned H, const int32_t *grid,
                                  const uint32_t *tex, unsigned texLog2);

/* ---- voxel landscape, code/landscape.cpp ---------------------------------
 * Landscape_Create (code/landscape.cpp:196-222) with the maps passed in:
 * 1024x1024 u8 heights, 1024x1024 ARGB colours and the 256-entry fog LUT,
 * all host pointers, copied to device memory the context owns. */
CKD_API ckd_status ckd_landscape_create(ckd_ctx *ctx, const uint8_t *heightMap,
                                        const uint32_t *colorMap,
                                        const uint32_t *fog256);
typedef struct ckd_landscape_cam {
    float X1, Y1;           /* origin in map texels (moveX+strafeX, ...)    */
    float viewCos, viewSin; /* cosf/sinf of the view angle                  */
    int mapTilt;            /* s_mapTilt at 720 rows (90 + int(tilt))       */
} ckd_landscape_cam;
/* Landscape_Draw without TapeWarp32 (code/landscape.cpp:228-243): fills dst
 * (W x H, W <= 4096, H <= 2160) for each of nFrames cameras; frame f is
 * written at dst + f*W*H.  The 720-row constants kMapScale and s_mapTilt
 * are scaled by H/720 and the column spacing by 1280/W.  Up to 160 frames
 * run in one launch; a launch whose 64-column groups do not spread evenly
 * over the GPU's SIMDs splits the march into ranges taken from a work queue
 * the context holds (ckd_tuning.voxelChunks), so two draws of one context
 * must not overlap (calls on its one stream never do). */
CKD_API ckd_status ckd_landscape_draw(ckd_ctx *ctx, uint32_t *dst, unsigned W,
                                      unsigned H, const ckd_landscape_cam *cams,
                                      unsigned nFrames);

/* ---- tunnelscape rows, code/tunnelscape.cpp ------------------------------
 * Tunnelscape_Create (code/tunnelscape.cpp:138-162) with 2048x2048 maps (all
 * heights <= 159, see code/tunnelscape.cpp:75-84) and the fog LUT. */
CKD_API ckd_status ckd_tunnelscape_create(ckd_ctx *ctx, const uint8_t *heightMap,
                                          const uint32_t *colorMap,
                                          const uint32_t *fog256);
/* memset32 + tscape (code/tunnelscape.cpp:103-134, :170-171) into dst */
CKD_API ckd_status ckd_tunnelscape_rows(ckd_ctx *ctx, uint32_t *dst, unsigned W,
                                        unsigned H, float stepU, float stepV,
                                        float speed, float time);

/* ---- synthetic SDF scene (no reference counterpart: the sphere-tracing
 * benchmark scene, structured like RenderSpikeyMap_2x2_Distant,
 * code/shadertoy.cpp:525-598) -------------------------------------------- */
typedef struct ckd_sdf_scene {
    int nSpheres;           /* 1..32                                        */
    float sphere[32][4];    /* cx, cy, cz, radius                           */
    float albedo[33][3];    /* B,G,R of sphere i; entry nSpheres = ground   */
    float groundY;
    float light[3];         /* unit vector towards the light                */
    float fogColor[3];      /* B,G,R                                        */
    float fogDensity;       /* ExpFog scale: fog = 1-exp(-s*d^3)            */
    float gamma;
    float ambient;
    float eps;              /* march stops when the distance <= eps         */
    float tMax;             /* far clip                                     */
    int maxSteps;           /* <= 1024                                      */
    float camRadius, camHeight, camSpeed, focal;
} ckd_sdf_scene;

enum { CKD_SHADE_FP32 = 0, CKD_SHADE_FP16 = 1 };

/* one frame per call ("plain"): dst W x H; cam is the float camera basis of
 * the frame (origin, forward*focal, right, up: 12 floats, host pointer);
 * stepCount (device, optional) accumulates the total number of march steps */
CKD_API ckd_status ckd_sdf_render(ckd_ctx *ctx, uint32_t *dst, unsigned W,
                                  unsigned H, const ckd_sdf_scene *scene,
                                  const float *cam, int shading,
                                  unsigned long long *stepCount);

/* Measurement hook, no reference counterpart: while counters (device memory,
 * two u64) is non-null, the batched SDF kernel adds to counters[0] the march
 * steps that evaluated the spheres and to counters[1] the shaded pixels
 * whose normal evaluated them.  The others took the ground's distance after
 * a bounding-sphere test showed no sphere could be nearer (the result is the
 * same either way).  NULL switches it off. */
CKD_API ckd_status ckd_sdf_work_counter(ckd_ctx *ctx, unsigned long long *counters);

/* ---- tunnel frame (no reference counterpart as a whole: the tunnel
 * benchmark's pipeline of grid remap, 2007 blur H then V out of place with
 * span 2r+1 -- HorizontalBoxBlur32 / VerticalBoxBlur32,
 * code/deprecated/boxblur.cpp:44-220 -- and BlitAdd32 of the blur onto the
 * remap, code/util.cpp:754-773).  tmp is a W*H scratch image; radius in
 * 1..127 with 2r+1 <= min(W, H). */
CKD_API ckd_status ckd_tunnel_frame(ckd_ctx *ctx, uint32_t *dst, uint32_t *tmp,
                                    unsigned W, unsigned H, const int32_t *grid,
                                    const uint32_t *tex, unsigned texLog2,
                                    int radius);

/* No reference counterpart: ckd_tunnel_frame for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The W*H scratch image is a fixed array of device
 * memory the library keeps on every GPU it runs on (14.7 MB, until the
 * process ends), so W*H is limited to 2560*1440; larger frames get
 * CKD_EINVAL.  The call allocates nothing (device or host) and is untimed by
 * ckd_kernel_timing.  Calls on one GPU share that array: calls that overlap
 * on one GPU (other streams, other contexts, other threads) race. */
CKD_API ckd_status ckd_tunnel_frame_noalloc(ckd_ctx *ctx, uint32_t *dst, unsigned W,
                                            unsigned H, const int32_t *grid,
                                            const uint32_t *tex, unsigned texLog2,
                                            int radius);

/* ---- batched renderer (no reference counterpart: many frames, effects and
 * resolutions in one persistent tile-queue launch) ------------------------ */
enum { CKD_FX_SDF = 0, CKD_FX_VOXEL = 1, CKD_FX_TUNNEL = 2 };

typedef struct ckd_request {
    int effect;             /* CKD_FX_*                                     */
    unsigned W, H;
    uint32_t *dst;          /* W*H output (device)                          */
    float cam[12];          /* SDF: camera basis                            */
    ckd_landscape_cam vcam; /* VOXEL: camera                                */
    const int32_t *grid;    /* TUNNEL: corner grid (device)                 */
    int radius;             /* TUNNEL: blur radius 1..127                   */
} ckd_request;

/* Launch shapes of the batched renderer, chosen by ckd_autotune_batch (or
 * set directly), the kernel of the batched in-place 2007 blurs of the Draw
 * calls and the split of ckd_landscape_draw's launches (ckd_autotune_batch
 * leaves the last three as they are).  Defaults: {16, 32, 8, 32, 0, 2, 2}. */
typedef struct ckd_tuning {
    int sdfStepsPerRefill; /* march steps between lane refills: 1, 2, 4, 8
                              or 16                                         */
    int sdfChunk;          /* side of the square pixel chunks: 16 or 32     */
    int sdfBlocksPerCU;    /* persistent 4-wave blocks per CU: 1..32 while
                              voxel or tunnel work runs beside the march; a
                              batch of SDF requests only gets at least the
                              kernel's occupancy                            */
    int tunnelVSeg;        /* rows per lane in the vertical blur: 8..512     */
    int blurKernel;        /* in-place blur batches: 0 automatic, 1 lines
                              staged in LDS, 2 an LDS ring per lane          */
    int voxelChunks;       /* ckd_landscape_draw launches whose 64-column
                              groups (one wave each) would load the busiest
                              SIMDs over 1/0.9 times the average: each
                              column's 512-step march in this many ranges
                              taken from a work queue, so the SIMDs share
                              the work evenly; 1 (no queue), 2, 4 or 8; 0
                              sets the default                              */
    int voxelWavesPerSIMD; /* persistent waves of that queue per SIMD: 1..8;
                              0 sets the default                            */
} ckd_tuning;
CKD_API ckd_status ckd_get_tuning(ckd_ctx *ctx, ckd_tuning *out);
CKD_API ckd_status ckd_set_tuning(ckd_ctx *ctx, const ckd_tuning *t);

/* Renders requests[0..n) (host array).  SDF requests use `scene` and the
 * given shading precision, VOXEL requests the landscape maps, TUNNEL
 * requests `tex` (2^texLog2 square) and need scratch: tmp (device) of at
 * least sum over TUNNEL requests of W*H pixels.  stepCount as above.
 * SDF requests run in one persistent launch over a queue of 32x32-pixel
 * chunks, VOXEL requests in one launch, TUNNEL requests in two. */
CKD_API ckd_status ckd_render_batch(ckd_ctx *ctx, const ckd_request *requests,
                                    unsigned n, const ckd_sdf_scene *scene,
                                    int shading, const uint32_t *tex,
                                    unsigned texLog2, uint32_t *tmp,
                                    size_t tmpPixels,
                                    unsigned long long *stepCount);

/* Times every candidate launch shape on this batch (each run twice, HIP
 * events), keeps the fastest in the context and returns it in *best (may
 * be NULL).  The outputs of the batch are overwritten. */
CKD_API ckd_status ckd_autotune_batch(ckd_ctx *ctx, const ckd_request *requests,
                                      unsigned n, const ckd_sdf_scene *scene,
                                      int shading, const uint32_t *tex,
                                      unsigned texLog2, uint32_t *tmp,
                                      size_t tmpPixels, ckd_tuning *best);

/* A batch captured once as a hipGraph (job tables staged in device memory
 * the graph owns, the chunk-queue reset as a kernel node, the voxel and
 * tunnel branch on the auxiliary stream) and replayed with ckd_graph_launch
 * on the context stream.  With kernel timing on, each replay counts as one
 * CKD_KT_SDF_PERSISTENT launch timed around the whole graph.  The request array, scene, texture and scratch
 * must stay valid and unchanged while the graph lives; the launch shapes
 * are those of the context at capture time. */
typedef struct ckd_graph ckd_graph;
CKD_API ckd_status ckd_batch_graph_create(ckd_ctx *ctx, const ckd_request *requests,
                                          unsigned n, const ckd_sdf_scene *scene,
                                          int shading, const uint32_t *tex,
                                          unsigned texLog2, uint32_t *tmp,
                                          size_t tmpPixels,
                                          unsigned long long *stepCount,
                                          ckd_graph **out);
CKD_API ckd_status ckd_graph_launch(ckd_graph *graph);
/* New parameters for the next replays of a captured batch (no reference
 * counterpart; the next frames of the animation that code/main.cpp:314-351
 * renders frame by frame): the requests must have the captured effects and
 * sizes, in the captured order; cameras, times, grids, radii and dst
 * pointers may change.  A tunnel radius may not exceed the largest of the
 * captured batch where that largest is at most 32 (the vertical blur's
 * window rows are sized for it); CKD_EINVAL otherwise.  The scene, shading
 * mode, texture and scratch stay the captured ones.  The graph's job tables are rewritten in place (after any replay
 * still queued has finished with them), so the graph itself is not
 * re-instantiated. */
CKD_API ckd_status ckd_graph_update(ckd_graph *graph, const ckd_request *requests,
                                    unsigned n);
CKD_API void ckd_graph_destroy(ckd_graph *graph);

/* ---- compositing blends (code/util.cpp:58-811) ---------------------------
 * In place on dst, device buffers, per channel of 0xAARRGGBB pixels, bit for
 * bit the reference's results (including the unsigned wrap-around of the
 * scalar blends, whose wrapped bits reach the packed pixel). */
/* Mix32, code/util.cpp:83-96: dst + alpha*(src-dst)/256, all four channels */
CKD_API ckd_status ckd_mix32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                             unsigned numPixels, uint8_t alpha);
/* MixOver32, code/util.cpp:148-177 (alpha channel of the result 0) */
CKD_API ckd_status ckd_mixover32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                 unsigned numPixels);
/* Sub32, code/util.cpp:179-192: saturating subtract */
CKD_API ckd_status ckd_sub32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                             unsigned numPixels);
/* Excl32, code/util.cpp:194-224: exclusion, destination alpha kept */
CKD_API ckd_status ckd_excl32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                              unsigned numPixels);
/* SoftLight32 / SoftLight32A / SoftLight32AA, code/util.cpp:227-346 */
CKD_API ckd_status ckd_softlight32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                   unsigned numPixels);
CKD_API ckd_status ckd_softlight32a(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                    unsigned numPixels);
CKD_API ckd_status ckd_softlight32aa(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                     unsigned numPixels, float alpha);
/* Overlay32 / Overlay32A (the integer forms), code/util.cpp:434-515 */
CKD_API ckd_status ckd_overlay32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                 unsigned numPixels);
CKD_API ckd_status ckd_overlay32a(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                  unsigned numPixels);
/* Darken32_50, code/util.cpp:518-549 */
CKD_API ckd_status ckd_darken32_50(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                   unsigned numPixels);
/* MulSrc32 / MulSrc32A, code/util.cpp:605-634 */
CKD_API ckd_status ckd_mulsrc32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                unsigned numPixels);
CKD_API ckd_status ckd_mulsrc32a(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                 unsigned numPixels);
/* MixSrc32 / MixSrc32S, code/util.cpp:636-676: blend by the source alpha;
 * the S form reads source rows srcStride pixels apart */
CKD_API ckd_status ckd_mixsrc32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                unsigned numPixels);
CKD_API ckd_status ckd_mixsrc32s(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                 unsigned resX, unsigned resY, unsigned srcStride);
/* BlitSrc32 / BlitSrc32A, code/util.cpp:707-752 (alpha in [0, 255]) */
CKD_API ckd_status ckd_blitsrc32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                 unsigned destResX, unsigned srcResX, unsigned yRes);
CKD_API ckd_status ckd_blitsrc32a(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src,
                                  unsigned destResX, unsigned srcResX, unsigned yRes,
                                  float alpha);
/* Fade32, code/util.cpp:798-811: towards the colour RGB by alpha */
CKD_API ckd_status ckd_fade32(ckd_ctx *ctx, uint32_t *dst, unsigned numPixels, uint32_t RGB,
                              uint8_t alpha);
/* Zoom32, code/util.cpp:58-81: nearest-neighbour zoom in, 0 < scale <= 1 */
CKD_API ckd_status ckd_zoom32(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned xRes,
                              unsigned yRes, float scale);

/* ---- the reference's raymarched effects (code/shadertoy.cpp) ------------
 *
 * Each effect renders an FX map of (resX/2+4) x (resY/2+4) pixels for an
 * output of resX x resY (code/fx-blitter.h:13-17; aspect resY/resX as
 * kAspect, code/main.h:43-44), then the Draw form post-processes it and
 * blits it 2x2 into the output as the reference's *_Draw does.  resX is a
 * multiple of 8 and resY even, both 64..8192; 1280x720 gives the
 * reference's 644x364 map.
 *
 * Parameters: one struct per frame (host memory, copied per call), holding
 * the frame's `time` (the Draw functions' time argument) and the Rocket
 * track values the reference reads, under the track's name.  All must be
 * finite.  _map writes nFrames consecutive maps, _draw nFrames consecutive
 * frames; both run asynchronously on the context stream.
 *
 * Results equal the reference's with two exceptions: vNorm4's
 * _mm_rsqrt_ps estimate is a correctly rounded 1/sqrtf, and powf, expf
 * and atan2f are the device library's (within 2 ulp of the host libm). */

/* RenderPlasmaMap / Plasma_Draw (code/shadertoy.cpp:200-280) */
typedef struct ckd_plasma_params {
    float time;
    float speed;        /* plasma:Speed        */
    float hue;          /* plasma:Hue          */
    float gamma;        /* plasma:Gamma        */
    float desaturation; /* plasma:Desaturation */
} ckd_plasma_params;
CKD_API ckd_status ckd_plasma_map(ckd_ctx *ctx, uint32_t *maps, unsigned resX, unsigned resY,
                                  const ckd_plasma_params *frames, unsigned nFrames);
CKD_API ckd_status ckd_plasma_draw(ckd_ctx *ctx, uint32_t *dst, unsigned resX, unsigned resY,
                                   const ckd_plasma_params *frames, unsigned nFrames);

/* RenderNautilusMap_2x2 / Nautilus_Draw (code/shadertoy.cpp:288-396); the
 * Draw form box-blurs the output in place when BoxBlurScale(blur) != 0 */
typedef struct ckd_nautilus_params {
    float time;
    float roll;         /* nautilus:Roll  */
    float hue;          /* nautilus:Hue   */
    float speed;        /* nautilus:Speed */
    float desaturation; /* nautilus:Desat */
    float blur;         /* nautilus:Blur  */
} ckd_nautilus_params;
CKD_API ckd_status ckd_nautilus_map(ckd_ctx *ctx, uint32_t *maps, unsigned resX, unsigned resY,
                                    const ckd_nautilus_params *frames, unsigned nFrames);
CKD_API ckd_status ckd_nautilus_draw(ckd_ctx *ctx, uint32_t *dst, unsigned resX, unsigned resY,
                                     const ckd_nautilus_params *frames, unsigned nFrames);

/* Spikey_Draw (code/shadertoy.cpp:661-727) and its renderers:
 * close = 0: RenderSpikeyMap_2x2_Distant (:525-598), or when distWarmup is
 *   non-zero _Distant_SpecularOnly (:600-659) followed by the in-place
 *   horizontal box blur of the map;
 * close = 1: RenderSpikeyMap_2x2_Close (:432-523), and in the Draw form,
 *   when closeMixBlurOpacity > 0, the mix-blur layer (:665-701) over the two
 *   blur maps given to ckd_spikey_blurmaps for the same output size
 *   (CKD_EINVAL without them). */
typedef struct ckd_spikey_params {
    float time;
    float speed;               /* spike:Speed               */
    float roll;                /* spike:Roll                */
    float specular;            /* spike:Specular            */
    float desaturation;        /* spike:Desaturation        */
    float hue;                 /* spike:Hue                 */
    float gamma;               /* spike:Gamma               */
    float distXOffs;           /* distSpike:xOffs           */
    float distYOffs;           /* distSpike:yOffs           */
    float distZOffs;           /* distSpike:zOffs           */
    float distWarmup;          /* distSpike:Warmup          */
    float closeXOffs;          /* closeSpike:xOffs          */
    float closeYOffs;          /* closeSpike:yOffs          */
    float closeZOffs;          /* closeSpike:zOffs          */
    float closeZOffsScale;     /* closeSpike:zOffsScale     */
    float closeNormalGrain;    /* closeSpike:NormalGrain    */
    float closeScale;          /* closeSpike:Scale          */
    float closeRim;            /* closeSpike:Rim            */
    float closeAspectMul;      /* closeSpike:AspectMul      */
    float closeMixBlurMap;     /* closeSpike:MixBlurMap     */
    float closeMixBlur;        /* closeSpike:MixBlur        */
    float closeMixMapBlur;     /* closeSpike:MixMapBlur     */
    float closeMixBlurOpacity; /* closeSpike:MixBlurOpacity */
} ckd_spikey_params;
CKD_API ckd_status ckd_spikey_map(ckd_ctx *ctx, uint32_t *maps, unsigned resX, unsigned resY,
                                  const ckd_spikey_params *frames, unsigned nFrames, int close);
CKD_API ckd_status ckd_spikey_draw(ckd_ctx *ctx, uint32_t *dst, unsigned resX, unsigned resY,
                                   const ckd_spikey_params *frames, unsigned nFrames, int close);
/* the close-up's two blur maps (Shadertoy_Create, code/shadertoy.cpp:180-185),
 * host memory, each the FX map's size for an output of resX x resY */
CKD_API ckd_status ckd_spikey_blurmaps(ckd_ctx *ctx, const uint32_t *map1, const uint32_t *map2,
                                       unsigned resX, unsigned resY);

/* RenderSinMap_2x2 / Sinuses_Draw (code/shadertoy.cpp:862-988) */
typedef struct ckd_sinuses_params {
    float time;
    float specular;     /* sinusesTunnel:Specular     */
    float roll;         /* sinusesTunnel:Roll         */
    float speed;        /* sinusesTunnel:Speed        */
    float offsX;        /* sinusesTunnel:OffsX        */
    float gamma;        /* sinusesTunnel:Gamma        */
    float hue;          /* sinusesTunnel:Hue          */
    float desaturation; /* sinusesTunnel:Desaturation */
} ckd_sinuses_params;
CKD_API ckd_status ckd_sinuses_map(ckd_ctx *ctx, uint32_t *maps, unsigned resX, unsigned resY,
                                   const ckd_sinuses_params *frames, unsigned nFrames);
CKD_API ckd_status ckd_sinuses_draw(ckd_ctx *ctx, uint32_t *dst, unsigned resX, unsigned resY,
                                    const ckd_sinuses_params *frames, unsigned nFrames);

/* RenderLaura_2x2 / Laura_Draw (code/shadertoy.cpp:990-1109) */
typedef struct ckd_laura_params {
    float time;
    float speed;    /* laura:Speed    */
    float yaw;      /* laura:Yaw      */
    float pitch;    /* laura:Pitch    */
    float roll;     /* laura:Roll     */
    float hue;      /* laura:Hue      */
    float saturate; /* laura:Saturate */
} ckd_laura_params;
CKD_API ckd_status ckd_laura_map(ckd_ctx *ctx, uint32_t *maps, unsigned resX, unsigned resY,
                                 const ckd_laura_params *frames, unsigned nFrames);
CKD_API ckd_status ckd_laura_draw(ckd_ctx *ctx, uint32_t *dst, unsigned resX, unsigned resY,
                                  const ckd_laura_params *frames, unsigned nFrames);

/* RenderTunnelMap_2x2 / Tunnel_Draw (code/shadertoy.cpp:735-861), the
 * free-directional tunnel.  Its two 1024x1024 ARGB textures (host memory,
 * copied) are those Shadertoy_Create loads (code/shadertoy.cpp:174-178).
 * _map writes the colour maps and the glow maps; the Draw form box-blurs
 * the glow in place when litBlur >= 1 and adds it when litTiles != 0. */
typedef struct ckd_fdtunnel_params {
    float time;
    float boxy;        /* tunnel:Boxy        */
    float flowerScale; /* tunnel:FlowerScale */
    float flowerFreq;  /* tunnel:FlowerFreq  */
    float flowerPhase; /* tunnel:FlowerPhase */
    float speed;       /* tunnel:Speed       */
    float roll;        /* tunnel:Roll        */
    float pitch;       /* tunnel:Pitch       */
    float radius;      /* tunnel:Radius      */
    float mulU;        /* tunnel:MulU        */
    float mulV;        /* tunnel:MulV        */
    float litTiles;    /* tunnel:LitTiles    */
    float litBlur;     /* tunnel:LitBlur     */
    float fog1;        /* tunnel:Fog1        */
    float fog2;        /* tunnel:Fog2        */
} ckd_fdtunnel_params;
CKD_API ckd_status ckd_fdtunnel_create(ckd_ctx *ctx, const uint32_t *tex1024,
                                       const uint32_t *texHighlights1024);
CKD_API ckd_status ckd_fdtunnel_map(ckd_ctx *ctx, uint32_t *maps, uint32_t *glowMaps,
                                    unsigned resX, unsigned resY,
                                    const ckd_fdtunnel_params *frames, unsigned nFrames);
CKD_API ckd_status ckd_fdtunnel_draw(ckd_ctx *ctx, uint32_t *dst, unsigned resX, unsigned resY,
                                     const ckd_fdtunnel_params *frames, unsigned nFrames);
/* Measurement hook, no reference counterpart: while counter (device memory)
 * is non-null, the FX-map kernels of the effects above add every map
 * pixel's march steps to *counter (loop iterations of its march; Plasma and
 * Laura count their fixed 24 and 32, Spikey's specular-only form its 36,
 * the tunnel none).  NULL switches it off. */
CKD_API ckd_status ckd_shadertoy_step_counter(ckd_ctx *ctx, unsigned long long *counter);

/* ---- voxel ball, torus twister, Polar_BlitA ------------------------------ */

/* Polar_BlitA (code/polar.cpp:157-198): dst = dst + a*(s - dst)/256 per
 * channel, s the bilinear fetch of src through the polar map (inverse = 1:
 * the inverse map) and a its alpha.  W x H must match ckd_polar_create. */
CKD_API ckd_status ckd_polar_blit_a(ckd_ctx *ctx, uint32_t *dst, const uint32_t *src, unsigned W,
                                    unsigned H, int inverse);

/* Twister_Create / Twister_Draw (code/torus-twister.cpp:136-187).  _create
 * copies the 1024x1024 height map (bytes), the 1024x1024 colour map and the
 * W x H background from host memory (the reference's 1280 x 720; an even W
 * of at least 1200, so that both 597-pixel rays of a row fit) and builds the
 * polar maps for W x H.  _draw renders nFrames frames into dst
 * (nFrames * W * H pixels): the twister into `target` (W x H, device
 * scratch: the reference's g_renderTarget[0]), its optional box blur, then
 * the background with Polar_BlitA of the target. */
typedef struct ckd_twister_params {
    float time;
    float speed;      /* twister:Speed        */
    float shearSpeed; /* twister::ShearSpeed  */
    float blur;       /* twister:Blur         */
} ckd_twister_params;
CKD_API ckd_status ckd_twister_create(ckd_ctx *ctx, const uint8_t *heightMap,
                                      const uint32_t *colorMap, const uint32_t *background,
                                      unsigned W, unsigned H);
CKD_API ckd_status ckd_twister_draw(ckd_ctx *ctx, uint32_t *dst, uint32_t *target, unsigned W,
                                    unsigned H, const ckd_twister_params *frames,
                                    unsigned nFrames);

/* Ball_Create / Ball_Draw (code/ball.cpp:337-511).  _create copies, from
 * host memory: five 1024x1024 height maps (bytes; map 0 is the spike map
 * mixed over the base shape 1..4), two colour maps (beams, no beams), three
 * beam maps, the environment map, two W x H backgrounds (beams, no beams)
 * and the W x H halo, and builds the polar maps.  _draw renders nFrames
 * frames into dst through `target` (W x H device scratch, the reference's
 * g_renderTarget[0]; the beam rays leave its last column as it was).  A
 * frame whose radius would project a map height past the end of a row
 * (height * radius / 256 > W - 1, where the reference writes into the next
 * row) is rejected. */
typedef struct ckd_ball_params {
    float time;
    float blur;            /* ball:Blur             */
    float radius;          /* ball:Radius           */
    float rayLength;       /* ball:RayLength        */
    float spikes;          /* ball:Spikes           */
    float hasBeams;        /* ball:HasBeams         */
    float baseShapeIndex;  /* ball:BaseShapeIndex   */
    float speed;           /* ball:Speed            */
    float beamAttenuation; /* ball:BeamAttenuation  */
    float beamAlphaMin;    /* ball:BeamAlphaMin     */
    float rotateOffsX;     /* ball:RotateOffsX      */
    float rotateOffsY;     /* ball:RotateOffsY      */
    float beams1;          /* ball:Beams1           */
    float beams2;          /* ball:Beams2           */
    float beams3;          /* ball:Beams3           */
    float lowBeams;        /* ball:BallLowBeams     */
} ckd_ball_params;
CKD_API ckd_status ckd_ball_create(ckd_ctx *ctx, const uint8_t *heightMaps5,
                                   const uint32_t *colorMaps2, const uint32_t *beamMaps3,
                                   const uint32_t *envMap, const uint32_t *backgrounds2,
                                   const uint32_t *halo, unsigned W, unsigned H);
CKD_API ckd_status ckd_ball_draw(ckd_ctx *ctx, uint32_t *dst, uint32_t *target, unsigned W,
                                 unsigned H, const ckd_ball_params *frames, unsigned nFrames);

/* ---- GNU Rocket track playback (host only) ------------------------------- */

/* A demo's sync tracks, read as the Rocket player reads them
 * (3rdparty/rocket-stripped/lib/device.c:41-75,309-331): track `name` is
 * <dir>/sync/_<name with every char outside [A-Za-z0-9._/] as -XX>.track; a
 * missing file is a track without keys.  Replaces sync_create_device +
 * sync_get_track (code/rocket.cpp:26-46) for callers that drive the ABI's
 * parameter structs from the production tracks. */
typedef struct ckd_sync ckd_sync;
CKD_API ckd_status ckd_sync_open(const char *dir, ckd_sync **out);
CKD_API void ckd_sync_close(ckd_sync *sync);
/* loads (once) and returns the index of track `name`; CKD_EINVAL for a file
 * that is not a track file */
CKD_API ckd_status ckd_sync_track(ckd_sync *sync, const char *name, int *index);
/* sync_get_val (3rdparty/rocket-stripped/lib/track.c:29-60): the track's
 * value at a fractional row, in double */
CKD_API double ckd_sync_get(const ckd_sync *sync, int index, double row);
/* Audio_Rocket_Sync (code/audio.cpp:18,159-179): the row at a time in
 * seconds (170 BPM as 174, 16 rows a beat) */
CKD_API double ckd_sync_row(double seconds);
/* the Rocket track names of an effect's parameter struct, in field order
 * after `time`: "plasma", "nautilus", "spikey", "sinuses", "laura",
 * "fdtunnel", "ball", "twister" */
CKD_API ckd_status ckd_effect_tracks(const char *effect, const char *const **names,
                                     unsigned *count);
/* fills nFrames parameter structs of `effect` (ckd_<effect>_params,
 * contiguous): time = times[k], every other field (float) of its track at
 * rows[k] (Rocket::getf, code/rocket.h:23-25) */
CKD_API ckd_status ckd_sync_params(ckd_sync *sync, const char *effect, const double *rows,
                                   const float *times, unsigned nFrames, void *params);

#ifdef __cplusplus
}
#endif
#endif
