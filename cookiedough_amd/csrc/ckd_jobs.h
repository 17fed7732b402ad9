// ckd_jobs.h -- device-side job tables of the batched renderer, shared by
// the host code that builds them (k_batch.hip) and the kernels that read
// them.  Each table is sorted by its `base` prefix so a block or wave finds
// its job by binary search.
#pragma once

#include <cstdint>

// SDF request: 32x32-pixel chunks are the units of the persistent queue
struct SdfJob {
    uint32_t *dst;
    uint32_t W, H;
    uint32_t chunksX;   // chunk columns
    uint32_t chunkBase; // first global chunk index of this request
    float cam[12];      // origin, forward*focal, right, up
    float invH;         // 1 / H, rounded to nearest
};

// voxel request: 64-column blocks, one lane per column
struct VoxJob {
    uint32_t *dst;
    uint32_t W, H;
    uint32_t blockBase; // first global block index
    int mapScale;       // kMapScale scaled to H rows
    float rayY, colScale, halfW;
    float X1, Y1, viewCos, viewSin;
    int mapTilt; // scaled to H rows
};

// tunnel request (grid remap + 2007 blur with span 2e+1 + additive blit)
struct TunJob {
    uint32_t *dst;
    uint32_t *tmp;      // W*H scratch (horizontal blur output)
    const int32_t *grid;
    uint32_t W, H;
    uint32_t rowBase;   // first global row (remap + horizontal pass)
    uint32_t vBase;     // first global block of the vertical pass
    uint32_t edge;      // e = radius
    uint32_t vBlocksX;  // 64-column blocks per row band
};

template <typename Job>
__host__ __device__ inline int ckd_find_job(const Job *jobs, int n, uint32_t unit,
                                            uint32_t Job::*base)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].*base <= unit) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
