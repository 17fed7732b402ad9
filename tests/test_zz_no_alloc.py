"""ckd_tunnel_frame_noalloc, the tunnel frame for callers that cannot
allocate, against the allocating entry point ckd_tunnel_frame: bit-exact,
on two calls in a row (the second smaller, with the scratch array still
holding the first call's image), and a frame larger than its scratch is
refused with CKD_EINVAL."""
from __future__ import annotations

import numpy as np
import pytest

from cookiedough_amd import workloads as wl
from cookiedough_amd._lib import CKD_EINVAL, CKD_OK, last_error

pytestmark = pytest.mark.gpu


def test_noalloc_tunnel_frame_matches_allocating_entry_point(ctx):
    lib = ctx.lib
    tex = ctx.upload(wl.noise_texture(0, 9))
    for W, H, t, r in [(1280, 720, 0.7, 17), (640, 360, 2.2, 5)]:
        print("noalloc tunnel frame", W, H, t, r, flush=True)
        grid = ctx.upload(wl.tunnel_grid(W, H, t))
        want, got, tmp = ctx.empty(W * H), ctx.empty(W * H), ctx.empty(W * H)
        ctx.memset32(got, 0x5A5A5A5A, W * H)
        ctx.tunnel_frame(want, tmp, W, H, grid, tex, 9, r)
        assert lib.ckd_tunnel_frame_noalloc(ctx.h, got.ptr, W, H, grid.ptr, tex.ptr, 9, r) == CKD_OK, \
            last_error()
        ctx.sync()
        assert np.array_equal(got.numpy(), want.numpy()), (W, H)
    big = ctx.empty(4)
    assert lib.ckd_tunnel_frame_noalloc(ctx.h, big.ptr, 2561, 1440, big.ptr, tex.ptr, 9, 3) == CKD_EINVAL
    assert "scratch" in last_error()
