"""GPU parity of the compositing blends (code/util.cpp) through the C ABI:
bit-exact against the oracle, flat (dwordx4) and strided paths, edge
values included."""
from __future__ import annotations

import numpy as np
import pytest

from oracles import P
from helpers import rand_img

pytestmark = pytest.mark.gpu

FLAT = ["mixover32", "sub32", "excl32", "softlight32", "softlight32a", "overlay32", "overlay32a",
        "darken32_50", "mulsrc32", "mulsrc32a", "mixsrc32"]


def images(n, seed):
    d, s = rand_img(seed, n), rand_img(seed + 1, n)
    d[:4] = [0, 0xffffffff, 0x80808080, 0x7f7f7f7f]
    s[:4] = [0xffffffff, 0, 0x7f7f7f7f, 0x80808080]
    s[4:8] = [0x00ffffff, 0xff000000, 0x01020304, 0xfefdfcfb]
    return d, s


@pytest.mark.parametrize("n", [1280 * 720, 4099])   # flat path, per-pixel path
@pytest.mark.parametrize("op", FLAT)
def test_flat_blends(ctx, orc, op, n):
    d, s = images(n, 31)
    dd, ds = ctx.upload(d), ctx.upload(s)
    ctx.blend(op, dd, ds, n)
    want = d.copy()
    getattr(orc, "orc_" + op)(P(want), P(s), n)
    assert np.array_equal(dd.numpy(), want), op


@pytest.mark.parametrize("n", [1280 * 720, 4099])
def test_alpha_blends(ctx, orc, n):
    d, s = images(n, 37)
    for alpha in [0, 1, 128, 255]:
        dd, ds = ctx.upload(d), ctx.upload(s)
        ctx.blend("mix32", dd, ds, n, alpha)
        want = d.copy()
        orc.orc_mix32(P(want), P(s), n, alpha)
        assert np.array_equal(dd.numpy(), want), ("mix32", alpha)
        for rgb in [0, 0xffffff, 0x80102030]:
            dd = ctx.upload(d)
            ctx.fade32(dd, n, rgb, alpha)
            want = d.copy()
            orc.orc_fade32(P(want), n, rgb, alpha)
            assert np.array_equal(dd.numpy(), want), ("fade32", alpha, rgb)
    for a in [-1.0, 0.0, 0.3, 1.0, 2.0, float("nan")]:
        dd, ds = ctx.upload(d), ctx.upload(s)
        ctx.blend("softlight32aa", dd, ds, n, a)
        want = d.copy()
        orc.orc_softlight32aa(P(want), P(s), n, a)
        assert np.array_equal(dd.numpy(), want), ("softlight32aa", a)


@pytest.mark.parametrize("alpha", [None, 0.0, 0.25, 1.0, 17.0, 255.0])
def test_strided_blits(ctx, orc, alpha):
    W, H, SW = 1280, 37, 1000
    d, s = rand_img(41, W * H), rand_img(42, SW * H)
    dd, ds = ctx.upload(d), ctx.upload(s)
    ctx.blitsrc32(dd, ds, W, SW, H, alpha)
    want = d.copy()
    if alpha is None:
        orc.orc_blitsrc32(P(want), P(s), W, SW, H)
    else:
        orc.orc_blitsrc32a(P(want), P(s), W, SW, H, alpha)
    assert np.array_equal(dd.numpy(), want)
    d2, s2 = rand_img(43, SW * H), rand_img(44, W * H)
    dd, ds = ctx.upload(d2), ctx.upload(s2)
    ctx.mixsrc32s(dd, ds, SW, H, W)
    want = d2.copy()
    orc.orc_mixsrc32s(P(want), P(s2), SW, H, W)
    assert np.array_equal(dd.numpy(), want)


@pytest.mark.parametrize("scale", [1.0, 0.999, 0.75, 0.5, 0.1234])
def test_zoom(ctx, orc, scale):
    W, H = 1280, 720
    s = rand_img(47, W * H)
    dd, ds = ctx.empty(W * H), ctx.upload(s)
    ctx.zoom32(dd, ds, W, H, scale)
    want = np.zeros(W * H, np.uint32)
    orc.orc_zoom32(P(want), P(s), W, H, scale)
    assert np.array_equal(dd.numpy(), want)
