// ckd_shim.h -- the reference's effect, filter and blit entry points,
// implemented over the libckd C ABI (include/ckd.h).  Same names and
// signatures as the reference headers, so code/demo.cpp links against this
// shim in place of landscape.cpp, tunnelscape.cpp, polar.cpp,
// fx-blitter.cpp, boxblur.cpp, deprecated/boxblur.cpp, shadertoy.cpp,
// ball.cpp, torus-twister.cpp and the blits of util.cpp.  Buffers are host memory, as in the reference (1280x720 frames,
// 644x364 FX maps); every call uploads its inputs, renders on the GPU and
// downloads the result before returning.
#pragma once

#include <cstddef>
#include <cstdint>

// code/landscape.h:7-9
bool Landscape_Create();
void Landscape_Destroy();
void Landscape_Draw(uint32_t *pDest, float time, float delta);

// code/tunnelscape.h:7-9
bool Tunnelscape_Create();
void Tunnelscape_Destroy();
void Tunnelscape_Draw(uint32_t *pDest, float time, float delta);

// code/polar.h:7-10
bool Polar_Create();
void Polar_Destroy();
void Polar_Blit(uint32_t *pDest, const uint32_t *pSrc, bool inverse = false);
// code/polar.h:14-17
void Polar_BlitA(uint32_t *pDest, const uint32_t *pSrc, bool inverse = false);
void Polar_Blit_2x2(uint32_t *pDest, const uint32_t *pSrc, bool inverse = false);

// code/ball.h:7-13
bool Ball_Create();
void Ball_Destroy();
void Ball_Draw(uint32_t *pDest, float time, float delta);
uint32_t *Ball_GetBackground();
bool Ball_HasBeams();

// code/torus-twister.h:7-9
bool Twister_Create();
void Twister_Destroy();
void Twister_Draw(uint32_t *pDest, float time, float delta);

// code/fx-blitter.h:26-33
constexpr size_t kShimNumFxMaps = 4;
extern uint32_t *g_pFxMap[kShimNumFxMaps];
bool FxBlitter_Create();
void FxBlitter_Destroy();
void Fx_Blit_2x2(uint32_t *pDest, const uint32_t *pSrc);
void FxBlitter_DrawTestPattern(uint32_t *pDest);

// code/boxblur.h:7-20
bool BoxBlur_Create();
void BoxBlur_Destroy();
void BoxBlur_Horz32(uint32_t *pDest, const uint32_t *pSrc, unsigned xRes, unsigned yRes,
                    float strength, float gain, unsigned numPasses);
void BoxBlur_Vert32(uint32_t *pDest, const uint32_t *pSrc, unsigned xRes, unsigned yRes,
                    float strength, float gain, unsigned numPasses);
void BoxBlur_32(uint32_t *pDest, const uint32_t *pSrc, unsigned xRes, unsigned yRes,
                float strength, float gain, unsigned numPasses);

// code/deprecated/boxblur.h:21-41
void HorizontalBoxBlur32(uint32_t *pDest, const uint32_t *pSrc, unsigned int xRes,
                         unsigned int yRes, float strength);
void VerticalBoxBlur32(uint32_t *pDest, const uint32_t *pSrc, unsigned int xRes,
                       unsigned int yRes, float strength);
void BoxBlur32(uint32_t *pDest, const uint32_t *pSrc, unsigned int xRes, unsigned int yRes,
               float strength);

// code/shadertoy.h:5-15
bool Shadertoy_Create();
void Shadertoy_Destroy();
void Nautilus_Draw(uint32_t *pDest, float time, float delta);
void Spikey_Draw(uint32_t *pDest, float time, float delta, bool close = true);
void Sinuses_Draw(uint32_t *pDest, float time, float delta);
void Laura_Draw(uint32_t *pDest, float time, float delta);
void Plasma_Draw(uint32_t *pDest, float time, float delta);
void Tunnel_Draw(uint32_t *pDest, float time, float delta);

// code/util.h:79,118-119,93
void Add32(uint32_t *pDest, const uint32_t *pSrc, unsigned numPixels);
void BlitAdd32(uint32_t *pDest, const uint32_t *pSrc, unsigned destResX, unsigned srcResX,
               unsigned yRes);
void BlitAdd32A(uint32_t *pDest, const uint32_t *pSrc, unsigned destResX, unsigned srcResX,
                unsigned yRes, float alpha);
void TapeWarp32(uint32_t *pDest, const uint32_t *pSrc, unsigned xRes, unsigned yRes,
                float strength, float speed);

// code/util.h:70-122, the compositing blends
void Zoom32(uint32_t *pDest, const uint32_t *pSrc, unsigned xRes, unsigned yRes, float scale);
void Mix32(uint32_t *pDest, const uint32_t *pSrc, unsigned numPixels, uint8_t alpha);
void MixOver32(uint32_t *pDest, const uint32_t *pSrc, unsigned numPixels);
void Sub32(uint32_t *pDest, const uint32_t *pSrc, unsigned numPixels);
void Excl32(uint32_t *pDest, const uint32_t *pSrc, unsigned numPixels);
void SoftLight32(uint32_t *pDest, const uint32_t *pSrc, unsigned numPixels);
void SoftLight32A(uint32_t *pDest, const uint32_t *pSrc, unsigned numPixels);
void SoftLight32AA(uint32_t *pDest, const uint32_t *pSrc, unsigned numPixels, float alpha);
void Overlay32(uint32_t *pDest, const uint32_t *pSrc, unsigned numPixels);
void Overlay32A(uint32_t *pDest, const uint32_t *pSrc, unsigned numPixels);
void Darken32_50(uint32_t *pDest, const uint32_t *pSrc, unsigned numPixels);
void MulSrc32(uint32_t *pDest, const uint32_t *pSrc, unsigned int numPixels);
void MulSrc32A(uint32_t *pDest, const uint32_t *pSrc, unsigned int numPixels);
void MixSrc32(uint32_t *pDest, const uint32_t *pSrc, unsigned int numPixels);
void MixSrc32S(uint32_t *pDest, const uint32_t *pSrc, unsigned destResX, unsigned destResY,
               unsigned srcStride);
void BlitSrc32(uint32_t *pDest, const uint32_t *pSrc, unsigned destResX, unsigned srcResX,
               unsigned yRes);
void BlitSrc32A(uint32_t *pDest, const uint32_t *pSrc, unsigned destResX, unsigned srcResX,
                unsigned yRes, float alpha);
void Fade32(uint32_t *pDest, unsigned int numPixels, uint32_t RGB, uint8_t alpha);

// No reference counterpart: the number of references to the shim's one GPU
// context that Create calls (and the first call of a free function) hold;
// 0 once every part is destroyed.  For tests of the parts' lifetimes.
int ckd_shim_context_refs();
