"""The compositing blends of code/util.cpp: the oracle against the compiled
reference (bit for bit, random images and edge values) and against what
the arithmetic fixes (identities and bounds), so a dropped term or a swapped
operand fails."""
from __future__ import annotations

import numpy as np
import pytest

from oracles import P
from helpers import rand_img

N = 4099  # odd: no vector-width assumptions


def images(seed: int):
    d, s = rand_img(seed, N), rand_img(seed + 1, N)
    # edge values: all-zero / all-one channels, and alphas 0 and 255
    d[:4] = [0, 0xffffffff, 0x80808080, 0x7f7f7f7f]
    s[:4] = [0xffffffff, 0, 0x7f7f7f7f, 0x80808080]
    s[4:8] = [0x00ffffff, 0xff000000, 0x01020304, 0xfefdfcfb]
    return d, s


FLAT = ["mixover32", "sub32", "excl32", "softlight32", "softlight32a", "overlay32", "overlay32a",
        "darken32_50", "mulsrc32", "mulsrc32a", "mixsrc32"]


@pytest.mark.ref
@pytest.mark.parametrize("op", FLAT)
def test_flat_blends_vs_reference(orc, refl, op):
    d, s = images(11)
    a, b = d.copy(), d.copy()
    getattr(orc, "orc_" + op)(P(a), P(s), N)
    getattr(refl, "ref_" + op)(P(b), P(s), N)
    assert np.array_equal(a, b), (op, int((a != b).sum()))


@pytest.mark.ref
@pytest.mark.parametrize("alpha", [0, 1, 77, 128, 254, 255])
def test_mix_and_fade_vs_reference(orc, refl, alpha):
    d, s = images(13)
    a, b = d.copy(), d.copy()
    orc.orc_mix32(P(a), P(s), N, alpha)
    refl.ref_mix32(P(b), P(s), N, alpha)
    assert np.array_equal(a, b)
    for rgb in [0, 0xffffff, 0xff102030, 0x80808080]:
        a, b = d.copy(), d.copy()
        orc.orc_fade32(P(a), N, rgb, alpha)
        refl.ref_fade32(P(b), N, rgb, alpha)
        assert np.array_equal(a, b), rgb


@pytest.mark.ref
@pytest.mark.parametrize("alpha", [-0.5, 0.0, 0.004, 0.5, 0.999, 1.0, 3.0])
def test_softlight_aa_vs_reference(orc, refl, alpha):
    d, s = images(17)
    a, b = d.copy(), d.copy()
    orc.orc_softlight32aa(P(a), P(s), N, alpha)
    refl.ref_softlight32aa(P(b), P(s), N, alpha)
    assert np.array_equal(a, b)


@pytest.mark.ref
@pytest.mark.parametrize("alpha", [0.0, 0.25, 1.0, 1.5, 17.0])
def test_strided_blits_vs_reference(orc, refl, alpha):
    W, H, SW = 97, 31, 61
    d, s = rand_img(19, W * H), rand_img(20, SW * H)
    a, b = d.copy(), d.copy()
    orc.orc_blitsrc32a(P(a), P(s), W, SW, H, alpha)
    refl.ref_blitsrc32a(P(b), P(s), W, SW, H, alpha)
    assert np.array_equal(a, b)
    a, b = d.copy(), d.copy()
    orc.orc_blitsrc32(P(a), P(s), W, SW, H)
    refl.ref_blitsrc32(P(b), P(s), W, SW, H)
    assert np.array_equal(a, b)
    a, b = d[:SW * H].copy(), d[:SW * H].copy()
    ws = rand_img(21, W * H)
    orc.orc_mixsrc32s(P(a), P(ws), SW, H, W)
    refl.ref_mixsrc32s(P(b), P(ws), SW, H, W)
    assert np.array_equal(a, b)


@pytest.mark.ref
@pytest.mark.parametrize("scale", [1.0, 0.999, 0.75, 0.5, 0.1234, 0.01])
def test_zoom_vs_reference(orc, refl, scale):
    W, H = 160, 90
    s = rand_img(23, W * H)
    a, b = np.zeros(W * H, np.uint32), np.zeros(W * H, np.uint32)
    orc.orc_zoom32(P(a), P(s), W, H, scale)
    refl.ref_zoom32(P(b), P(s), W, H, scale)
    assert np.array_equal(a, b)


def ch(px, k):
    return (px >> (8 * k)) & 0xff


def test_blend_identities(orc):
    d, s = images(29)
    # Mix32 / Fade32 with alpha 0 keep the destination
    a = d.copy()
    orc.orc_mix32(P(a), P(s), N, 0)
    assert np.array_equal(a, d)
    a = d.copy()
    orc.orc_fade32(P(a), N, 0x12345678, 0)
    assert np.array_equal(a, d)
    # Mix32: every channel is d + a*(s-d)/256, floor, in [min, max]
    a = d.copy()
    orc.orc_mix32(P(a), P(s), N, 200)
    for k in range(4):
        dk, sk = ch(d, k).astype(np.int64), ch(s, k).astype(np.int64)
        assert np.array_equal(ch(a, k), (dk * 256 + 200 * (sk - dk)) >> 8)
    # Sub32 / MulSrc32 / Darken32_50 closed forms
    a = d.copy()
    orc.orc_sub32(P(a), P(s), N)
    for k in range(4):
        assert np.array_equal(ch(a, k), np.maximum(ch(d, k).astype(np.int64) - ch(s, k), 0))
    a = d.copy()
    orc.orc_mulsrc32(P(a), P(s), N)
    for k in range(4):
        assert np.array_equal(ch(a, k), (ch(d, k) * ch(s, k)) >> 8)
    a = d.copy()
    orc.orc_darken32_50(P(a), P(s), N)
    for k in range(3):
        assert np.array_equal(ch(a, k), (ch(d, k) + np.minimum(ch(d, k), ch(s, k))) >> 1)
    assert np.array_equal(ch(a, 3), ch(d, 3))
    # MixSrc32 with an opaque source gives the source (255/256 of it, floor)
    s2 = s | np.uint32(0xff000000)
    a = d.copy()
    orc.orc_mixsrc32(P(a), P(s2), N)
    for k in range(4):
        dk, sk = ch(d, k).astype(np.int64), ch(s2, k).astype(np.int64)
        assert np.array_equal(ch(a, k), (dk + 255 * sk) >> 8)
