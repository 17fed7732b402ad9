"""Pins the build-defined per-8x8-block UV grid law (orc_grid_remap, the
spec of k_grid_remap and of the remap fused into k_tunnel_h) by an
invariant it must satisfy: when the corner grid samples an affine UV field
C(gx, gy) = A + 8 gx a + 8 gy b (24:8 integers), the integer bilinear
interpolation reproduces the field exactly at every pixel,
UV(x, y) = A + a x + b y.  The remap of such a grid must therefore equal the
bilinear fetch through the explicit per-pixel map A + a x + b y, which is
the reference's Polar_Blit fetch (code/polar.cpp:82-133, orc_polar_blit,
pinned bit for bit against the compiled reference).  The fields stay inside
the texture, where the grid's wrapped fetch and the blit's unwrapped one
coincide.  A law with a dropped term, a swapped corner or a rounding shift
off by one breaks the equality for some (x, y)."""
from __future__ import annotations

import numpy as np
import pytest

from oracles import P
from cookiedough_amd import workloads as wl

TEX_LOG2 = 9
FIELDS = [  # (A, a, b) in 24:8 per pixel, kept inside [0, 510*256]
    # (the last one only at the smaller sizes)
    ((256, 512), (95, 13), (9, 150)),
    ((130000, 700), (-97, 3), (-5, 160)),
    ((60000, 125000), (31, -50), (-41, -80)),
    ((1000, 64000), (0, 0), (0, 0)),
    ((77, 91), (255, 1), (1, 255)),
]
SIZES = [(1280, 720), (333, 187), (64, 8), (8, 64)]


def affine(W, H, A, a, b):
    gh, gw = (H + 7) // 8 + 1, (W + 7) // 8 + 1
    gy, gx = np.mgrid[0:gh, 0:gw]
    grid = np.stack([A[0] + 8 * gx * a[0] + 8 * gy * b[0],
                     A[1] + 8 * gx * a[1] + 8 * gy * b[1]], axis=-1).astype(np.int32)
    y, x = np.mgrid[0:H, 0:W]
    pmap = np.stack([A[0] + a[0] * x + b[0] * y, A[1] + a[1] * x + b[1] * y],
                    axis=-1).astype(np.int32)
    return np.ascontiguousarray(grid), np.ascontiguousarray(pmap)


def in_texture(pmap, size):
    return pmap.min() >= 0 and pmap.max() <= (size - 2) * 256 + 255


def expected(orc, tex, pmap, W, H):
    want = np.zeros(W * H, np.uint32)
    orc.orc_polar_blit(P(want), P(tex), P(pmap), W, H, 1 << TEX_LOG2, 4)
    return want


@pytest.mark.parametrize("W,H", SIZES)
@pytest.mark.parametrize("A,a,b", FIELDS)
def test_affine_grid_reproduces_per_pixel_uvs(orc, W, H, A, a, b):
    tex = wl.noise_texture(0, TEX_LOG2)
    grid, pmap = affine(W, H, A, a, b)
    if not in_texture(pmap, 1 << TEX_LOG2):
        pytest.skip("field leaves the texture at this size")
    got = np.zeros(W * H, np.uint32)
    orc.orc_grid_remap(P(got), W, H, P(grid), P(tex), TEX_LOG2, 4)
    assert np.array_equal(got, expected(orc, tex, pmap, W, H))


def test_perturbed_corner_breaks_the_invariant(orc):
    """the check has teeth: one corner moved by a fifth of a texel changes
    the pixels of the blocks around it"""
    W, H = 64, 64
    tex = wl.noise_texture(0, TEX_LOG2)
    grid, pmap = affine(W, H, *FIELDS[0])
    grid[3, 4, 0] += 51
    got = np.zeros(W * H, np.uint32)
    orc.orc_grid_remap(P(got), W, H, P(grid), P(tex), TEX_LOG2, 4)
    assert not np.array_equal(got, expected(orc, tex, pmap, W, H))


@pytest.mark.gpu
@pytest.mark.parametrize("A,a,b", FIELDS[:-1])
def test_gpu_affine_grid_remap_and_tunnel(ctx, orc, A, a, b):
    """k_grid_remap, and the remap that k_tunnel_h fuses into the tunnel
    frame (radius 1 blurs by span 3; compared through the oracle's frame of
    the same grid), both at 1280x720 against the explicit-map fetch"""
    W, H = 1280, 720
    tex = wl.noise_texture(0, TEX_LOG2)
    grid, pmap = affine(W, H, A, a, b)
    assert in_texture(pmap, 1 << TEX_LOG2)
    want = expected(orc, tex, pmap, W, H)
    print("affine grid", A, a, b, flush=True)
    d_tex, d_grid = ctx.upload(tex), ctx.upload(grid)
    d = ctx.empty(W * H)
    ctx.grid_remap(d, W, H, d_grid, d_tex, TEX_LOG2)
    ctx.sync()
    assert np.array_equal(d.numpy(), want)
    # the fused tunnel frame: remap, blur H+V (span 3), BlitAdd32 of the
    # blur onto the remap; the oracle's frame from the explicit-map remap
    tmp = ctx.empty(W * H)
    ctx.tunnel_frame(d, tmp, W, H, d_grid, d_tex, TEX_LOG2, 1)
    ctx.sync()
    blur = np.zeros(W * H, np.uint32)
    t0 = np.zeros(W * H, np.uint32)
    orc.orc_blur07_h(P(t0), P(want), W, H, np.float32(3.5 / 255), 4)
    orc.orc_blur07_v(P(blur), P(t0), W, H, np.float32(3.5 / 255), 4)
    frame = want.copy()
    orc.orc_blit_add32(P(frame), P(blur), W, W, H)
    assert np.array_equal(d.numpy(), frame)
