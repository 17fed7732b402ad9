"""Injected device-allocation failures (ckd_debug_fail_alloc): every path
that allocates returns CKD_ENOMEM with a message and leaves the context
usable; the same call succeeds once the hook is disarmed, with the same
result as on a context that never failed."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from cookiedough_amd import _lib
from cookiedough_amd import workloads as wl
from cookiedough_amd._lib import CKD_ENOMEM, CkdError
from cookiedough_amd.device import Context
from helpers import rand_img

pytestmark = pytest.mark.gpu


@pytest.fixture
def lib():
    lib = _lib.load()
    yield lib
    lib.ckd_debug_fail_alloc(-1)


def test_create_and_malloc_fail_cleanly(lib):
    h = C.c_void_p()
    lib.ckd_debug_fail_alloc(0)
    assert lib.ckd_create(0, C.byref(h)) == CKD_ENOMEM
    assert "LUT" in _lib.last_error() and not h.value
    assert lib.ckd_create(0, C.byref(h)) == 0
    p = C.c_void_p()
    lib.ckd_debug_fail_alloc(0)
    assert lib.ckd_malloc(h, C.byref(p), 4096) == CKD_ENOMEM and not p.value
    assert lib.ckd_malloc(h, C.byref(p), 4096) == 0 and p.value
    lib.ckd_free(h, p)
    lib.ckd_destroy(h)


def _expect_enomem(fn):
    with pytest.raises(CkdError) as ei:
        fn()
    assert ei.value.status == CKD_ENOMEM, str(ei.value)


def test_lazy_scratch_paths_recover(lib):
    """the 2026 blur's ping-pong scratch, the landscape maps and their
    texel quads, the batch's job staging and a captured graph's tables"""
    W, H = 320, 180
    with Context(0) as ok, Context(0) as c:
        src = rand_img(3, W * H)
        outs = []
        for ctx, fail in ((ok, False), (c, True)):
            s, d = ctx.upload(src), ctx.empty(W * H)
            if fail:
                lib.ckd_debug_fail_alloc(0)
                _expect_enomem(lambda: ctx.blur26(d, s, W, H, 10.0, 0.0, 2))
            ctx.blur26(d, s, W, H, 10.0, 0.0, 2)
            ctx.sync()
            outs.append(d.numpy())
        assert np.array_equal(*outs)

        m = wl.landscape_maps(0)
        # the n-th allocation of the call fails: height, colour, fog, then the
        # first of the context's own buffers not yet allocated (a buffer that
        # was allocated stays): the two quad maps, then the draw's work queue
        # state
        for n in range(6):
            lib.ckd_debug_fail_alloc(n)
            _expect_enomem(lambda: c.landscape_create(m.height, m.color, m.fog))
        lib.ckd_debug_fail_alloc(-1)
        c.landscape_create(m.height, m.color, m.fog)
        ok.landscape_create(m.height, m.color, m.fog)
        path = wl.landscape_path(0)
        frames = []
        for ctx in (ok, c):
            d = ctx.empty(W * H)
            ctx.landscape_draw(d, W, H, [wl.landscape_cam(path, 0.3)])
            ctx.sync()
            frames.append(d.numpy())
        assert np.array_equal(*frames)

        sc = wl.sdf_scene(0, 8, 64)
        items = [wl.BatchItem(0, W, H, 1.0), wl.BatchItem(1, W, H, 2.0)]
        out = c.empty(2 * W * H)
        reqs = c.request_array([wl.make_request(it, out.ptr + 4 * k * W * H, sc, path, None)
                                for k, it in enumerate(items)])
        lib.ckd_debug_fail_alloc(0)
        _expect_enomem(lambda: c.render_batch(reqs, sc, 1, None, 9, None, 0))
        c.render_batch(reqs, sc, 1, None, 9, None, 0)
        c.sync()
        first = out.numpy()
        lib.ckd_debug_fail_alloc(0)
        _expect_enomem(lambda: c.batch_graph(reqs, sc, 1, None, 9, None, 0))
        g = c.batch_graph(reqs, sc, 1, None, 9, None, 0)
        c.memset32(out, 0, 2 * W * H)
        g.launch()
        c.sync()
        assert np.array_equal(out.numpy(), first)
        g.close()

        # a tunnel request: the texture's texel quads (the context's for an
        # eager batch, the graph's own for a captured one)
        tex = wl.noise_texture(0, 9)
        it = wl.BatchItem(2, W, H, 3.0, 7)
        results = []
        for ctx, fail in ((ok, False), (c, True)):
            d_tex, grid = ctx.upload(tex), ctx.upload(wl.tunnel_grid(W, H, it.t))
            o, tmp = ctx.empty(W * H), ctx.empty(W * H)
            treq = ctx.request_array([wl.make_request(it, o.ptr, sc, path, grid.ptr)])
            if fail:
                lib.ckd_debug_fail_alloc(0)
                with pytest.raises(CkdError) as ei:
                    ctx.render_batch(treq, sc, 1, d_tex, 9, tmp, W * H)
                assert ei.value.status == CKD_ENOMEM and "quad" in str(ei.value), str(ei.value)
            ctx.render_batch(treq, sc, 1, d_tex, 9, tmp, W * H)
            ctx.sync()
            results.append(o.numpy())
            if fail:
                lib.ckd_debug_fail_alloc(1)  # the graph's tables, then its quads
                with pytest.raises(CkdError) as ei:
                    ctx.batch_graph(treq, sc, 1, d_tex, 9, tmp, W * H)
                assert ei.value.status == CKD_ENOMEM and "quad" in str(ei.value), str(ei.value)
                g = ctx.batch_graph(treq, sc, 1, d_tex, 9, tmp, W * H)
                ctx.memset32(o, 0, W * H)
                g.launch()
                ctx.sync()
                results.append(o.numpy())
                g.close()
        assert np.array_equal(results[0], results[1]) and np.array_equal(results[0], results[2])
