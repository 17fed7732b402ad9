/*
 * ckd_oracle.h -- CPU restatement of cookiedough's per-pixel effect path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * called only by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py.  The product (cookiedough_amd/libckd.so) never calls it.
 *
 * Every function restates one reference function and cites it as
 * file:line relative to the reference repository root.  Conventions:
 *   - pixels are uint32_t 0xAARRGGBB (memory byte order B,G,R,A);
 *   - "24:8" is a signed int with 8 fractional bits;
 *   - "u16 lanes" are the four channels widened to uint16_t, index 0 = B.
 *
 * Pinning status of each function (what checks it besides this restatement)
 * is stated next to its declaration.
 */
#ifndef CKD_ORACLE_H
#define CKD_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- a1..a3 bilinear samplers (code/bilinear.h) ---------------------- */
/* pinned: golden vectors from the compiled reference + exhaustive closed form */
void orc_bsamp_prep(int U, int V, unsigned mapAnd, unsigned mapShift,
                    unsigned *U0, unsigned *V0, unsigned *U1, unsigned *V1,
                    unsigned *fracU, unsigned *fracV);
unsigned orc_bsamp8(const uint8_t *tex, unsigned U0, unsigned V0, unsigned U1,
                    unsigned V1, unsigned fracU, unsigned fracV);
void orc_bsamp32_16(const uint32_t *tex, unsigned U0, unsigned V0, unsigned U1,
                    unsigned V1, unsigned fracU, unsigned fracV, uint16_t out[4]);
uint32_t orc_bsamp32_32_packed(const uint32_t *tex, unsigned U0, unsigned V0,
                               unsigned U1, unsigned V1, unsigned fracU,
                               unsigned fracV);

/* ---- a4 colored span (code/cspan.h:47-78) ----------------------------- */
/* pinned: golden vectors (lane quirks) from the compiled reference */
void orc_cspan16(uint32_t *dst, ptrdiff_t destIncr, unsigned length,
                 unsigned drawLength, const uint16_t A[4], const uint16_t B[4]);

/* ---- a5..a7 voxel landscape (code/landscape.cpp) ---------------------- */
typedef struct orc_landscape_cam {
    float X1, Y1;          /* camera origin in map texels (float)           */
    float viewCos, viewSin;/* cosf/sinf of the view angle (host libm)       */
    int   mapTilt;         /* s_mapTilt at 720 rows (kMapTilt + int(tilt))  */
} orc_landscape_cam;

/* one column, reference vscape_ray (code/landscape.cpp:56-107) with the
 * frame height, map scale and tilt as parameters (720 / 512 / s_mapTilt at
 * the reference resolution).  pinned: golden frames from the compiled ref. */
void orc_vscape_ray(uint32_t *pColumn, unsigned W, unsigned H,
                    const uint8_t *heightMap, const uint32_t *colorMap,
                    const uint32_t *fog256, int mapScale, int mapTilt,
                    int curX, int curY, int dX, int dY, float fishMul);

/* per-column FP32 ray set-up of vscape (code/landscape.cpp:164-190);
 * outputs fpX1,fpY1 once and dX,dY,fishMul per column.
 * pinned: straight restatement (the reference vscape() reads gamepad and
 * Rocket state and cannot be called); invariants in tests. */
void orc_vscape_setup(unsigned W, unsigned H, const orc_landscape_cam *cam,
                      int *fpX1, int *fpY1, int *dX, int *dY, float *fishMul);

/* whole frame: memset32(fog[0]) + vscape (code/landscape.cpp:228-243). */
void orc_landscape_frame(uint32_t *dst, unsigned W, unsigned H,
                         const uint8_t *heightMap, const uint32_t *colorMap,
                         const uint32_t *fog256, const orc_landscape_cam *cam,
                         int nthreads);

/* ---- a8 tunnelscape rows (code/tunnelscape.cpp:44-134) ---------------- */
void orc_tscape_ray(uint32_t *pRow, unsigned W, const uint8_t *heightMap,
                    const uint32_t *colorMap, const uint32_t *fog256,
                    int curX, int curY, int dX, int dY);
void orc_tscape_frame(uint32_t *dst, unsigned W, unsigned H,
                      const uint8_t *heightMap2048, const uint32_t *colorMap2048,
                      const uint32_t *fog256, float stepU, float stepV,
                      float speed, float time, int nthreads);

/* ---- a9/a10 polar maps and blit (code/polar.cpp) ----------------------- */
void orc_polar_maps(int *map, int *invMap, unsigned srcResX, unsigned srcResY,
                    unsigned destResX, unsigned destResY);
void orc_polar_blit(uint32_t *dst, const uint32_t *src, const int *map,
                    unsigned W, unsigned H, unsigned srcStride, int nthreads);

/* ---- a11 per-8x8-block UV grid remap (build-defined spec) -------------- */
void orc_grid_remap(uint32_t *dst, unsigned W, unsigned H, const int *gridUV,
                    const uint32_t *tex, unsigned texLog2, int nthreads);

/* ---- a12 2007 box blur (code/deprecated/boxblur.cpp) ------------------- */
/* dst may equal src (in-place recurrence, exactly as the reference). */
int  orc_blur07_span(float strength);
void orc_blur07_h(uint32_t *dst, const uint32_t *src, unsigned xRes,
                  unsigned yRes, float strength, int nthreads);
void orc_blur07_v(uint32_t *dst, const uint32_t *src, unsigned xRes,
                  unsigned yRes, float strength, int nthreads);
void orc_blur07(uint32_t *dst, const uint32_t *src, unsigned xRes,
                unsigned yRes, float strength, int nthreads);

/* ---- a13 2026 box blur (code/boxblur.cpp) ------------------------------ */
/* scratch0/scratch1 each hold at least xRes*yRes pixels; src must be
 * readable 2 pixels past the end of every row (the reference reads them). */
void orc_blur26_h(uint32_t *dst, const uint32_t *src, unsigned xRes,
                  unsigned yRes, float strength, float gain, unsigned passes,
                  uint32_t *scratch0, uint32_t *scratch1);
void orc_blur26_v(uint32_t *dst, const uint32_t *src, unsigned xRes,
                  unsigned yRes, float strength, float gain, unsigned passes,
                  uint32_t *scratch0, uint32_t *scratch1);
void orc_blur26(uint32_t *dst, const uint32_t *src, unsigned xRes,
                unsigned yRes, float strength, float gain, unsigned passes,
                uint32_t *scratch0, uint32_t *scratch1);

/* ---- a14..a16 blitters (code/fx-blitter.cpp, code/util.cpp, util.h) ---- */
void orc_fx_blit_2x2(uint32_t *dst, const uint32_t *src, unsigned dstW,
                     unsigned dstH);
void orc_add32(uint32_t *dst, const uint32_t *src, unsigned n);
void orc_blit_add32(uint32_t *dst, const uint32_t *src, unsigned destResX,
                    unsigned srcResX, unsigned yRes);
void orc_blit_add32a(uint32_t *dst, const uint32_t *src, unsigned destResX,
                     unsigned srcResX, unsigned yRes, float alpha);
void orc_memset32(uint32_t *dst, uint32_t value, size_t n);
int  orc_ftofp24(float v);

/* ---- c2 frame: grid remap -> blur07 H,V (out of place) -> BlitAdd32 ---- */
void orc_tunnel_frame(uint32_t *dst, uint32_t *tmp0, uint32_t *tmp1,
                      unsigned W, unsigned H, const int *gridUV,
                      const uint32_t *tex, unsigned texLog2, int radius,
                      int nthreads);

/* ---- a20 synthetic SDF scene (build-defined spec) ----------------------- */
typedef struct orc_sdf_scene {
    int   nSpheres;
    float sphere[32][4];     /* cx, cy, cz, radius                          */
    float albedo[33][3];     /* B, G, R per sphere, last entry = ground     */
    float groundY;
    float light[3];          /* normalized direction towards the light      */
    float fogColor[3];       /* B, G, R                                     */
    float fogDensity;        /* ExpFog scale (cubic in distance)            */
    float gamma;
    float ambient;
    float eps;               /* hit threshold                               */
    float tMax;              /* far clip                                    */
    int   maxSteps;
    float camRadius, camHeight, camSpeed, focal;
} orc_sdf_scene;

/* render one frame in double (prec=64) or float (prec=32) arithmetic;
 * steps (optional) receives the per-pixel march step count. */
void orc_sdf_frame(uint32_t *dst, unsigned W, unsigned H,
                   const orc_sdf_scene *scene, float t, int prec,
                   uint16_t *steps, int nthreads);
/* rows rowStart, rowStart+rowStep, ... only, compacted into dst */
void orc_sdf_rows(uint32_t *dst, unsigned W, unsigned H,
                  const orc_sdf_scene *scene, float t, int prec,
                  unsigned rowStart, unsigned rowStep, uint16_t *steps,
                  int nthreads);
/* float camera basis of frame time t: origin[3], forward*focal[3],
 * right[3], up[3] (computed in double, rounded to float) */
void orc_sdf_rows_ex(uint32_t *dst, unsigned W, unsigned H,
                     const orc_sdf_scene *scene, float t, int prec,
                     unsigned rowStart, unsigned rowStep, uint16_t *steps,
                     uint8_t *cond, int nthreads);
void orc_sdf_camera(const orc_sdf_scene *scene, float t, float cam[12]);

/* ---- a19 LUT trig (code/sincos-lut.h:13-33, code/sincos-lut.cpp:9-16) --- */
void  orc_cos_lut(float lut[2049]);
float orc_lutcosf(const float *lut, float angle);
float orc_lutsinf(const float *lut, float angle);
/* code/util.cpp:552-603; pinned: reference harness (TapeWarp32 itself) */
void  orc_tape_warp32(uint32_t *dst, const uint32_t *src, unsigned xRes, unsigned yRes,
                      float strength, float speed, const float *lut);

/* ---- a17-a19 the reference's raymarched effects (code/shadertoy.cpp) ---- */
enum { ORC_ST_PLASMA = 0, ORC_ST_NAUTILUS = 1, ORC_ST_SPIKEY_DISTANT = 2,
       ORC_ST_SPIKEY_CLOSE = 3, ORC_ST_SINUSES = 4, ORC_ST_LAURA = 5,
       ORC_ST_FDTUNNEL = 6 };
/* One FX map ((resX/2+4) x (resY/2+4)) of an effect.  params: the Rocket
 * track values in the order of the C ABI's parameter structs (time first);
 * glow, tex and texHi are used by ORC_ST_FDTUNNEL only (1024x1024 ARGB).
 * rsqrtMode 0 = 1/sqrtf, 1 = x86 rsqrtss (as the reference on this CPU).
 * pinned: the compiled reference driven by its own Rocket player. */
void orc_st_map(int effect, uint32_t *dst, uint32_t *glow, unsigned resX, unsigned resY,
                const float *params, const float *lut, const uint32_t *tex,
                const uint32_t *texHi, int rsqrtMode, int nthreads);
/* the effect's Draw function into a resX x resY frame; blurMaps = the
 * close-up's two blur maps (FX-map size each, back to back; NULL allowed
 * when its mix-blur layer is off).  0, or -1 for the layer without maps. */
int orc_st_draw(int effect, uint32_t *dst, unsigned resX, unsigned resY, const float *params,
                const float *lut, const uint32_t *tex, const uint32_t *texHi,
                const uint32_t *blurMaps, int rsqrtMode, int nthreads);
/* one lane of exp_ps / log_ps (3rdparty/sse_mathfun.h) and ToPixel4 of one
 * pixel (code/shadertoy-util.h:136-146) */
float orc_log_ps(float x);
float orc_exp_ps(float x);
uint32_t orc_topixel(const float c[4]);
/* x86 rsqrtss of this CPU */
float orc_rsqrtss(float x);

/* ---- compositing blends (code/util.cpp:58-811) ------------------------- */
/* pinned: the compiled reference, bit for bit, on random images */
void orc_mix32(uint32_t *dst, const uint32_t *src, unsigned n, uint8_t alpha);
void orc_mixover32(uint32_t *dst, const uint32_t *src, unsigned n);
void orc_sub32(uint32_t *dst, const uint32_t *src, unsigned n);
void orc_excl32(uint32_t *dst, const uint32_t *src, unsigned n);
void orc_softlight32(uint32_t *dst, const uint32_t *src, unsigned n);
void orc_softlight32a(uint32_t *dst, const uint32_t *src, unsigned n);
void orc_softlight32aa(uint32_t *dst, const uint32_t *src, unsigned n, float alpha);
void orc_overlay32(uint32_t *dst, const uint32_t *src, unsigned n);
void orc_overlay32a(uint32_t *dst, const uint32_t *src, unsigned n);
void orc_darken32_50(uint32_t *dst, const uint32_t *src, unsigned n);
void orc_mulsrc32(uint32_t *dst, const uint32_t *src, unsigned n);
void orc_mulsrc32a(uint32_t *dst, const uint32_t *src, unsigned n);
void orc_mixsrc32(uint32_t *dst, const uint32_t *src, unsigned n);
void orc_mixsrc32s(uint32_t *dst, const uint32_t *src, unsigned resX, unsigned resY,
                   unsigned srcStride);
void orc_blitsrc32(uint32_t *dst, const uint32_t *src, unsigned destResX, unsigned srcResX,
                   unsigned yRes);
void orc_blitsrc32a(uint32_t *dst, const uint32_t *src, unsigned destResX, unsigned srcResX,
                    unsigned yRes, float alpha);
void orc_fade32(uint32_t *dst, unsigned n, uint32_t RGB, uint8_t alpha);
void orc_zoom32(uint32_t *dst, const uint32_t *src, unsigned xRes, unsigned yRes, float scale);

/* ---- voxel ball, torus twister, Polar_BlitA (code/ball.cpp,
 * code/torus-twister.cpp, code/polar.cpp:157-198) ----------------------- */
/* pinned: the compiled reference's Ball_Draw / Twister_Draw / Polar_BlitA */
void orc_polar_blit_a(uint32_t *dst, const uint32_t *src, const int *map, unsigned W,
                      unsigned H, unsigned srcStride, int nthreads);
void orc_twister_tables(uint32_t heightProj[512], uint32_t heightProjNorm[512]);
void orc_twister_frame(uint32_t *target, unsigned W, unsigned H, const uint8_t *hmap,
                       const uint32_t *cmap, float time, float speed, float shearSpeed,
                       int nthreads);
int orc_twister_draw(uint32_t *dst, uint32_t *target, unsigned W, unsigned H,
                     const uint8_t *hmap, const uint32_t *cmap, const uint32_t *background,
                     const int *polarMap, const float *p, int nthreads);
void orc_ball_rays(unsigned H, float time, float rayLengthTrack, float rotateOffsX,
                   float rotateOffsY, int *fromXY, int *dXY);
int orc_ball_draw(uint32_t *dst, uint32_t *target, unsigned W, unsigned H,
                  const uint8_t *heightMaps, const uint32_t *colorMaps, const uint32_t *beamMaps,
                  const uint32_t *envMap, const uint32_t *backgrounds, const uint32_t *halo,
                  const int *polarMap, const float *p, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
