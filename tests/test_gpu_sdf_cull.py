"""The batched SDF tracer's bounding-sphere test (k_sdf_persistent): on a
step or a normal where the bounding sphere of all spheres is farther than
the ground by more than the rounding margin, the spheres are not evaluated.
That must give the same bytes as evaluating them.

A scene with one more sphere, far away, has a bounding sphere that contains
everything, so the test never skips there; the far sphere itself is never
the nearest.  Both scenes must render identical frames.  The work counter
(ckd_sdf_work_counter) must not change the frames and must count no more
than every step and every pixel."""
import ctypes as C

import numpy as np
import pytest

from cookiedough_amd import workloads as wl
from cookiedough_amd._lib import CKD_FX_SDF, CKD_SHADE_FP16, CKD_SHADE_FP32, SdfScene

pytestmark = pytest.mark.gpu


def _render(ctx, sc, W, H, times, shading, steps=None):
    path = wl.landscape_path(0)
    out = ctx.empty(len(times) * W * H)
    reqs = ctx.request_array([wl.make_request(wl.BatchItem(CKD_FX_SDF, W, H, t), out.ptr + 4 * k * W * H,
                                              sc, path, None) for k, t in enumerate(times)])
    ctx.memset32(out, 0x12345678, len(times) * W * H)
    ctx.render_batch(reqs, sc, shading, None, 9, None, 0, steps)
    ctx.sync()
    return out.numpy()


@pytest.mark.parametrize("n_spheres,max_steps", [(8, 64), (20, 128)])
def test_skipping_spheres_changes_nothing(ctx, n_spheres, max_steps):
    sc = wl.sdf_scene(0, n_spheres, max_steps)
    far = SdfScene.from_buffer_copy(sc)
    far.nSpheres = n_spheres + 1
    far.sphere[n_spheres][0], far.sphere[n_spheres][1] = 4000.0, 2500.0
    far.sphere[n_spheres][2], far.sphere[n_spheres][3] = -3000.0, 1.0
    for c in range(3):
        far.albedo[n_spheres + 1][c] = sc.albedo[n_spheres][c]  # the ground's albedo follows the spheres
    times = [0.0, 7.3, 19.1]
    for shading in (CKD_SHADE_FP32, CKD_SHADE_FP16):
        print("cull", n_spheres, shading, flush=True)
        a = _render(ctx, sc, 640, 360, times, shading)
        b = _render(ctx, far, 640, 360, times, shading)
        assert np.array_equal(a, b), (n_spheres, shading, int((a != b).sum()))


def test_work_counter(ctx):
    sc = wl.sdf_scene(0, 8, 64)
    times = [0.0, 3.0, 11.0, 30.0]
    W, H = 640, 360
    plain = _render(ctx, sc, W, H, times, CKD_SHADE_FP16)
    work = ctx.upload(np.zeros(2, np.uint64))
    steps = ctx.upload(np.zeros(1, np.uint64))
    ctx.sdf_work_counter(work)
    try:
        counted = _render(ctx, sc, W, H, times, CKD_SHADE_FP16, steps)
    finally:
        ctx.sdf_work_counter(None)
    assert np.array_equal(plain, counted)
    ev_steps, ev_shades = (int(v) for v in work.numpy())
    n_steps, px = int(steps.numpy()[0]), len(times) * W * H
    print("work", ev_steps, n_steps, ev_shades, px, flush=True)
    assert 0 < ev_steps <= n_steps and 0 < ev_shades <= px
    # no spheres: nothing to evaluate
    empty = SdfScene.from_buffer_copy(sc)
    empty.nSpheres = 0
    empty.albedo[0][0], empty.albedo[0][1], empty.albedo[0][2] = sc.albedo[8][0], sc.albedo[8][1], sc.albedo[8][2]
    ctx.memset32(work, 0, 4)
    ctx.sdf_work_counter(work)
    try:
        _render(ctx, empty, W, H, times, CKD_SHADE_FP16)
    finally:
        ctx.sdf_work_counter(None)
    assert [int(v) for v in work.numpy()] == [0, 0]


@pytest.mark.parametrize("n_spheres", [8, 20])
def test_steps_per_refill_change_nothing(ctx, n_spheres):
    """the lanes are refilled every 1, 2, 4, 8 or 16 march steps: the order of
    work changes, the frames and the total march step count do not"""
    sc = wl.sdf_scene(1, n_spheres, 96)
    times = [0.7, 13.0, 41.5]
    saved = ctx.get_tuning()
    outs, counts = [], []
    try:
        for k in (1, 2, 4, 8, 16):
            t = ctx.get_tuning()
            t.sdfStepsPerRefill = k
            ctx.set_tuning(t)
            cnt = ctx.upload(np.zeros(1, np.uint64))
            print("steps per refill", k, flush=True)
            outs.append(_render(ctx, sc, 640, 360, times, CKD_SHADE_FP16, cnt))
            counts.append(int(cnt.numpy()[0]))
    finally:
        ctx.set_tuning(saved)
    for k, o in zip((2, 4, 8, 16), outs[1:]):
        assert np.array_equal(o, outs[0]), k
    assert len(set(counts)) == 1 and counts[0] > 0, counts


@pytest.mark.parametrize("shading", [CKD_SHADE_FP16, CKD_SHADE_FP32])
def test_launch_shape_changes_nothing(ctx, shading):
    """16x16 or 32x32 chunks and 1 to 8 persistent blocks per CU change which
    wave traces which pixel, not the pixel: frames and the step count of a
    batch of ragged sizes (a chunk row narrower than the chunk, the largest
    height 2160, the largest width 4096) are identical"""
    sc = wl.sdf_scene(2, 8, 64)
    path = wl.landscape_path(0)
    sizes = [(8, 2160, 3.0), (4096, 9, 5.0), (1000, 563, 7.0), (333, 17, 11.0), (640, 360, 29.0),
             (31, 33, 41.0)]
    total = sum(W * H for W, H, _ in sizes)
    out = ctx.empty(total)
    offs, o = [], 0
    for W, H, _ in sizes:
        offs.append(o)
        o += W * H
    reqs = ctx.request_array([wl.make_request(wl.BatchItem(CKD_FX_SDF, W, H, t), out.ptr + 4 * offs[k], sc, path,
                                              None) for k, (W, H, t) in enumerate(sizes)])
    saved = ctx.get_tuning()
    frames, counts = [], []
    try:
        for chunk, bpc in ((32, 3), (16, 1), (16, 8), (32, 8)):
            t = ctx.get_tuning()
            t.sdfChunk, t.sdfBlocksPerCU = chunk, bpc
            ctx.set_tuning(t)
            cnt = ctx.upload(np.zeros(1, np.uint64))
            ctx.memset32(out, 0x12345678, total)
            print("launch shape", chunk, bpc, shading, flush=True)
            ctx.render_batch(reqs, sc, shading, None, 9, None, 0, cnt)
            ctx.sync()
            frames.append(out.numpy().copy())
            counts.append(int(cnt.numpy()[0]))
    finally:
        ctx.set_tuning(saved)
    assert not (frames[0] == 0x12345678).all()
    for k, f in enumerate(frames[1:], 1):
        assert np.array_equal(f, frames[0]), k
    assert len(set(counts)) == 1 and counts[0] > 0, counts
