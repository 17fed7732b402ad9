"""ckd_landscape_draw's work queue: a launch whose 64-column groups do not
spread evenly over the 1024 SIMDs (more than 1024 groups, and at most 90% of
ceil(groups / 1024) * 1024) splits each column's 512-step march into 2, 4 or
8 ranges that persistent waves take from a queue, a range waiting for the
highest row its predecessor drew.  Frames must be bit-identical to the one-wave-per-group
kernel (voxelChunks 1) and to the oracle (code/landscape.cpp:56-107), for
every split and queue width, across launches and repeated calls."""
import ctypes as C

import numpy as np
import pytest

from cookiedough_amd import workloads as wl
from oracles import P, orc_cam

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def lmaps(ctx):
    m = wl.landscape_maps(0)
    ctx.landscape_create(m.height, m.color, m.fog)
    return m


def _draw(ctx, d, W, H, cams, chunks, waves):
    saved = ctx.get_tuning()
    t = ctx.get_tuning()
    t.voxelChunks, t.voxelWavesPerSIMD = chunks, waves
    ctx.set_tuning(t)
    try:
        _say("landscape queue", W, H, len(cams), "frames, chunks", chunks, "waves/SIMD", waves)
        ctx.landscape_draw(d, W, H, cams)
        ctx.sync()
    finally:
        ctx.set_tuning(saved)
    return d.numpy().reshape(len(cams), W * H)


def _oracle(orc, lmaps, W, H, cam):
    ref = np.zeros(W * H, np.uint32)
    orc.orc_landscape_frame(P(ref), W, H, P(lmaps.height), P(lmaps.color), P(lmaps.fog),
                            C.byref(orc_cam(cam)), 8)
    return ref


# 1280x720 x 60: 1200 groups; 1000x560 x 80: 1280 groups of which the last of
# each frame is partly outside the row; 640x360 x 200: a 160-frame launch of
# 1600 groups (queue) and a 40-frame launch of 400 (one wave per group)
@pytest.mark.parametrize("W,H,n,check", [(1280, 720, 60, (0, 31, 59)), (1000, 560, 80, (0, 40, 79)),
                                         (640, 360, 200, (0, 159, 160, 199))])
def test_queue_frames_equal_oracle(ctx, orc, lmaps, W, H, n, check):
    path = wl.landscape_path(0)
    cams = [wl.landscape_cam(path, 0.11 * k) for k in range(n)]
    d = ctx.empty(n * W * H)
    base = _draw(ctx, d, W, H, cams, 1, 4)
    for k in check:
        assert np.array_equal(base[k], _oracle(orc, lmaps, W, H, cams[k])), (W, H, k)
    for chunks in (2, 4, 8):
        ctx.memset32(d, 0x5a5a5a5a, n * W * H)
        out = _draw(ctx, d, W, H, cams, chunks, 4)
        bad = np.nonzero((out != base).any(axis=1))[0]
        assert bad.size == 0, (W, H, chunks, bad[:8])


@pytest.mark.parametrize("waves", [1, 2, 8])
def test_queue_widths(ctx, orc, lmaps, waves):
    """1, 2 and 8 persistent waves per SIMD, twice each in a row (the queue
    and its epoch are reset by the previous launch's last wave)"""
    W, H, n = 1280, 720, 56
    path = wl.landscape_path(1)
    cams = [wl.landscape_cam(path, 0.37 * k) for k in range(n)]
    d = ctx.empty(n * W * H)
    base = _draw(ctx, d, W, H, cams, 1, 4).copy()
    assert np.array_equal(base[n - 1], _oracle(orc, lmaps, W, H, cams[n - 1]))
    for rep in range(2):
        ctx.memset32(d, 0, n * W * H)
        out = _draw(ctx, d, W, H, cams, 8, waves)
        assert np.array_equal(out, base), (waves, rep)
