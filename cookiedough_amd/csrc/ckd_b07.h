// ckd_b07.h -- the 2007 box blur's parameters (code/deprecated/boxblur.cpp:
// 52-76) and its batched in-place horizontal form, for the effects whose
// Draw functions blur their render target in place.
#pragma once

#include "ckd_internal.h"

namespace ckd {

struct B07Params {
    uint32_t span, edgeSpan, median, shift, subEdges;
    uint32_t fullDiv;
    uint16_t edgeDivs[256];
};

// one image of a batched in-place horizontal blur
struct B07Job {
    uint32_t *img;
    B07Params p;
};

// the job of HorizontalBoxBlur32(img, img, W, H, strength); the checks of
// ckd_blur07_h
ckd_status b07_h_job(uint32_t *img, unsigned W, unsigned H, float strength, B07Job &job);

// a job for the in-place 2D blur (BoxBlur32 with pDest == pSrc) of a W x H
// image: the span must fit both passes
ckd_status b07_job(uint32_t *img, unsigned W, unsigned H, float strength, B07Job &job);

// the in-place blurs of n jobs, all on W x H images, one launch: along the
// rows, or along the columns with `columns`.  devJobs is the device copy of
// hostJobs (the host copy picks the kernel and its ring length); vec4: every
// image 16-byte aligned (rows only).  The kernel is the context tuning's
// blurKernel: staged lines (lines up to 16383 pixels), a ring per lane, or
// (0) the ring once the staged lines would need more than one round of blocks.
ckd_status blur07_inplace_jobs(ckd_ctx *ctx, const B07Job *hostJobs, const B07Job *devJobs,
                               unsigned n, unsigned W, unsigned H, bool columns, bool vec4,
                               uint32_t *splitStates = nullptr);

// the in-place horizontal blurs of n jobs already in device memory, all on
// W x H images (W <= 16383), one launch; vec4: every image 16-byte aligned.
// With splitStates (12 * 4 * H bytes per job) each line stops just before
// its last pixel is added (every job needs W > span) and leaves its state
// there for blur07_h_chain.
ckd_status blur07_h_inplace_jobs(ckd_ctx *ctx, const B07Job *hostJobs, const B07Job *devJobs,
                                 unsigned n, unsigned W, unsigned H, bool vec4,
                                 uint32_t *splitStates = nullptr);

// one image of a chain whose last column carries from image to image
struct B07Chain {
    uint32_t *img;
    const uint32_t *state; // mode 2: its split blur's states
    uint32_t mode;         // 0: its last column is final; 1: takes the carried
                           // column; 2: takes it, then its split blur ends
    B07Params p;           // mode 2
};

// walks n chained images (device array) in order; carry (H pixels) holds
// the column coming in and receives the one going out
ckd_status blur07_h_chain(ckd_ctx *ctx, const B07Chain *devChain, unsigned n, uint32_t *carry,
                          unsigned W, unsigned H);

} // namespace ckd
