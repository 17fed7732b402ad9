"""GPU parity of the voxel ball (code/ball.cpp), the torus twister
(code/torus-twister.cpp) and Polar_BlitA (code/polar.cpp:157-198) through
the C ABI: bit-exact.  Frames are hashed against the compiled reference's
(SHA-256 fixtures, written through its Rocket player) and compared pixel
for pixel with the oracle on random parameters and other resolutions."""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from oracles import P, aligned_empty
from cookiedough_amd import _lib
from cookiedough_amd import workloads as wl

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "voxel2.json"
W, H = 1280, 720


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def polar_maps(orc, w, h):
    m0, m1 = np.zeros(2 * w * h, np.int32), np.zeros(2 * w * h, np.int32)
    orc.orc_polar_maps(P(m0), P(m1), w, h, w, h)
    return m0, m1


@pytest.fixture(scope="module")
def gold():
    return json.loads(GOLD.read_text())


@pytest.fixture(scope="module")
def maps(ctx, orc):
    bm, tm = wl.ball_maps(W, H), wl.twister_maps(W, H)
    ctx.ball_create(bm, W, H)
    ctx.twister_create(tm["heightMap"], tm["colorMap"], tm["background"], W, H)
    return bm, tm, polar_maps(orc, W, H)


def orc_ball(orc, bm, m0, prm, w, h, target):
    d = aligned_empty(w * h)
    assert orc.orc_ball_draw(P(d), P(target), w, h, P(bm["heightMaps"]), P(bm["colorMaps"]),
                             P(bm["beamMaps"]), P(bm["envMap"]), P(bm["backgrounds"]),
                             P(bm["halo"]), P(m0), P(prm.as_array()), 16) == 0
    return d


def orc_twister(orc, tm, m0, prm, w, h):
    d, tgt = aligned_empty(w * h), aligned_empty(w * h)
    assert orc.orc_twister_draw(P(d), P(tgt), w, h, P(tm["heightMap"]), P(tm["colorMap"]),
                                P(tm["background"]), P(m0), P(prm.as_array()), 16) == 0
    return d


def zeroed(ctx, n):
    t = ctx.empty(n)
    ctx.memset32(t, 0, n)
    return t


def test_ball_frames_match_reference_hashes(ctx, maps, gold):
    for c in gold["ball"]:
        print("ball", c["time"], c["params"], flush=True)
        prm = _lib.BallParams(time=c["time"], **c["params"])
        d, t = ctx.empty(W * H), zeroed(ctx, W * H)
        ctx.ball_draw(d, t, W, H, prm)
        ctx.sync()
        assert sha(d.numpy()) == c["draw_sha256"], c["params"]


def test_twister_frames_match_reference_hashes(ctx, maps, gold):
    for c in gold["twister"]:
        print("twister", c["time"], c["params"], flush=True)
        prm = _lib.TwisterParams(time=c["time"], **c["params"])
        d, t = ctx.empty(W * H), ctx.empty(W * H)
        ctx.twister_draw(d, t, W, H, prm)
        ctx.sync()
        assert sha(d.numpy()) == c["draw_sha256"], c["params"]


def test_polar_blit_a_matches_reference_hash_and_oracle(ctx, orc, maps, gold):
    m0, m1 = maps[2]
    src, dst = wl._argb_noise(50, W * H), wl._argb_noise(51, W * H)
    for inverse, m in [(True, m1), (False, m0)]:
        print("polar_blit_a inverse", inverse, flush=True)
        dd, ds = ctx.upload(dst), ctx.upload(src)
        ctx.polar_blit_a(dd, ds, W, H, inverse)
        want = dst.copy()
        orc.orc_polar_blit_a(P(want), P(src), P(m), W, H, W, 8)
        got = dd.numpy()
        assert np.array_equal(got, want), inverse
        if inverse:
            assert sha(got) == gold["polar_blit_a_inverse"]["sha256"]


def random_ball(rng, t, radius_max=1920.0):
    return _lib.BallParams(
        time=t, blur=float(rng.choice([0.0, 0.5, 7.0, 120.0])),
        radius=float(rng.uniform(1.0, radius_max)), rayLength=float(rng.uniform(-5.0, 1100.0)),
        spikes=float(rng.integers(0, 256)), hasBeams=float(rng.integers(0, 2)),
        baseShapeIndex=float(rng.uniform(-1.0, 6.0)), speed=float(rng.uniform(-4.0, 4.0)),
        beamAttenuation=float(rng.uniform(-10.0, 300.0)),
        beamAlphaMin=float(rng.uniform(-10.0, 300.0)), rotateOffsX=float(rng.uniform(-600, 600)),
        rotateOffsY=float(rng.uniform(-600, 600)), beams1=float(rng.uniform(-0.5, 1.5)),
        beams2=float(rng.uniform(-0.5, 1.5)), beams3=float(rng.uniform(-0.5, 1.5)),
        lowBeams=float(rng.uniform(-10.0, 300.0)))


def test_ball_random_frames_match_oracle(ctx, orc, maps):
    """random parameters (clamps, rounding and out-of-range tracks
    included); the render target carries over from frame to frame, as the
    reference's g_renderTarget[0] does"""
    bm, _, (m0, _) = maps
    rng = np.random.default_rng(5)
    t = zeroed(ctx, W * H)
    host_t = np.zeros(W * H, np.uint32)
    for k in range(8):
        prm = random_ball(rng, float(rng.uniform(0.0, 100.0)))
        print("ball random", k, list(prm.as_array()), flush=True)
        d = ctx.empty(W * H)
        ctx.ball_draw(d, t, W, H, prm)
        ctx.sync()
        want = orc_ball(orc, bm, m0, prm, W, H, host_t)
        assert np.array_equal(d.numpy(), want), k
        assert np.array_equal(t.numpy(), host_t), k


def test_twister_random_frames_match_oracle(ctx, orc, maps):
    _, tm, (m0, _) = maps
    rng = np.random.default_rng(6)
    for k in range(6):
        prm = _lib.TwisterParams(time=float(rng.uniform(-50.0, 200.0)),
                                 speed=float(rng.uniform(-200.0, 200.0)),
                                 shearSpeed=float(rng.uniform(-2.0, 2.0)),
                                 blur=float(rng.choice([0.0, 0.5, 12.0, 150.0])))
        print("twister random", k, list(prm.as_array()), flush=True)
        d, t = ctx.empty(W * H), ctx.empty(W * H)
        ctx.twister_draw(d, t, W, H, prm)
        ctx.sync()
        assert np.array_equal(d.numpy(), orc_twister(orc, tm, m0, prm, W, H)), k


def test_batched_frames_equal_single_frames(ctx, maps):
    balls = wl.ball_frames(2, 0) + wl.ball_frames(2, 1, t0=30.0)
    t = zeroed(ctx, W * H)
    d = ctx.empty(len(balls) * W * H)
    ctx.ball_draw(d, t, W, H, balls)
    t1 = zeroed(ctx, W * H)
    one = ctx.empty(W * H)
    ctx.sync()
    allf = d.numpy().reshape(len(balls), -1)
    for k, prm in enumerate(balls):
        ctx.ball_draw(one, t1, W, H, prm)
        ctx.sync()
        assert np.array_equal(allf[k], one.numpy()), ("ball", k)
    tws = wl.twister_frames(3)
    d = ctx.empty(3 * W * H)
    ctx.twister_draw(d, t, W, H, tws)
    ctx.sync()
    allf = d.numpy().reshape(3, -1)
    for k, prm in enumerate(tws):
        ctx.twister_draw(one, t1, W, H, prm)
        ctx.sync()
        assert np.array_equal(allf[k], one.numpy()), ("twister", k)


@pytest.mark.parametrize("res", [(640, 360), (1920, 1080)])
def test_other_resolutions_match_oracle(orc, res):
    from cookiedough_amd.device import Context
    w, h = res
    bm = wl.ball_maps(w, h)
    m0, _ = polar_maps(orc, w, h)
    c = Context(0)
    try:
        c.ball_create(bm, w, h)
        radius_max = (w - 1) * 256.0 / 160.0
        rng = np.random.default_rng(w)
        t = zeroed(c, w * h)
        host_t = np.zeros(w * h, np.uint32)
        for k in range(3):
            prm = random_ball(rng, 10.0 + k, min(1920.0, radius_max))
            prm.hasBeams = float(k & 1)
            print("ball", res, k, flush=True)
            d = c.empty(w * h)
            c.ball_draw(d, t, w, h, prm)
            c.sync()
            assert np.array_equal(d.numpy(), orc_ball(orc, bm, m0, prm, w, h, host_t)), (res, k)
        if w >= 1200:
            tm = wl.twister_maps(w, h)
            c.twister_create(tm["heightMap"], tm["colorMap"], tm["background"], w, h)
            prm = wl.twister_frames(1)[0]
            prm.blur = 20.0
            d, tt = c.empty(w * h), c.empty(w * h)
            c.twister_draw(d, tt, w, h, prm)
            c.sync()
            assert np.array_equal(d.numpy(), orc_twister(orc, tm, m0, prm, w, h)), res
        else:
            with pytest.raises(_lib.CkdError, match="width"):
                tm = wl.twister_maps(w, h)
                c.twister_create(tm["heightMap"], tm["colorMap"], tm["background"], w, h)
    finally:
        c.close()


def test_ball_rejects_rows_overflow(ctx, maps):
    """a radius that projects the tallest map height (160) past a 640-pixel
    row is refused before anything runs"""
    from cookiedough_amd.device import Context
    c = Context(0)
    try:
        c.ball_create(wl.ball_maps(640, 360), 640, 360)
        d, t = c.empty(640 * 360), c.empty(640 * 360)
        prm = wl.ball_frames(1)[0]
        with pytest.raises(_lib.CkdError, match="past a 640-pixel row"):
            c.ball_draw(d, t, 640, 360, prm)
        with pytest.raises(_lib.CkdError, match="no ball maps"):
            c.ball_draw(d, t, 1280, 720, prm)
    finally:
        c.close()


def test_frames_from_rocket_tracks(ctx, orc, maps, tmp_path):
    """the library's track player drives the Draw calls: interpolated keys
    (linear, smooth, ramp, step) read at fractional rows, ball and twister
    drawn alternately into one render target as the demo does, bit-exact
    against the oracle on the same parameters"""
    import oracles
    from cookiedough_amd.sync import Sync
    bm, tm, (m0, _) = maps
    rng = np.random.default_rng(12)
    ranges = {"radius": (700.0, 1900.0), "rayLength": (300.0, 1024.0), "spikes": (0.0, 255.0),
              "hasBeams": (0.0, 1.0), "baseShapeIndex": (1.0, 4.0), "speed": (-4.0, 4.0),
              "beams1": (0.0, 1.0), "beams2": (0.0, 1.0), "beams3": (0.0, 1.0),
              "shearSpeed": (-0.3, 0.3)}
    for cls in (_lib.BallParams, _lib.TwisterParams):
        for fld, tr in cls.TRACKS.items():
            lo, hi = ranges.get(fld, (0.0, 60.0))
            keys = [(r, float(rng.uniform(lo, hi)), ty) for r, ty in [(0, 1), (40, 2), (80, 3), (120, 0)]]
            oracles.write_track_keys(tmp_path, tr, keys)
    rows = [3.5, 44.25, 79.75, 150.0]
    times = [Sync.row(r) / 46.4 for r in rows]
    with Sync(tmp_path) as s:
        bp, tp = s.params("ball", rows, times), s.params("twister", rows, times)
    t = zeroed(ctx, W * H)
    host_t = np.zeros(W * H, np.uint32)
    for k in range(len(rows)):
        print("tracks row", rows[k], flush=True)
        d = ctx.empty(W * H)
        ctx.ball_draw(d, t, W, H, bp[k])
        ctx.sync()
        assert np.array_equal(d.numpy(), orc_ball(orc, bm, m0, bp[k], W, H, host_t)), ("ball", k)
        ctx.twister_draw(d, t, W, H, tp[k])
        ctx.sync()
        want, host_t = aligned_empty(W * H), aligned_empty(W * H)
        assert orc.orc_twister_draw(P(want), P(host_t), W, H, P(tm["heightMap"]), P(tm["colorMap"]),
                                    P(tm["background"]), P(m0), P(tp[k].as_array()), 16) == 0
        assert np.array_equal(d.numpy(), want), ("twister", k)
        assert np.array_equal(t.numpy(), host_t), ("target", k)
