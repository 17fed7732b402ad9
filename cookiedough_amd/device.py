"""Thin Python wrapper over the C ABI: a context, device arrays and one
method per entry point.  Used by the tests and bench.py; the C++ shim
(cookiedough_amd/shim) is the drop-in for the reference's own C++ callers.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, ptr


@dataclass
class DevArray:
    ctx: "Context"
    ptr: int
    shape: tuple
    dtype: np.dtype

    @property
    def nbytes(self) -> int:
        return int(np.prod(self.shape)) * np.dtype(self.dtype).itemsize

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(self.ctx.lib.ckd_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def view(self, offset_elems: int, shape) -> "DevArray":
        it = np.dtype(self.dtype).itemsize
        return DevArray(self.ctx, self.ptr + offset_elems * it, tuple(shape), self.dtype)


class Context:
    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        check(self.lib.ckd_create(device, C.byref(h)))
        self.h = h
        self._owned: list[int] = []

    # ---- memory ------------------------------------------------------------
    def empty(self, shape, dtype=np.uint32) -> DevArray:
        shape = tuple(shape) if isinstance(shape, (tuple, list)) else (shape,)
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = C.c_void_p()
        check(self.lib.ckd_malloc(self.h, C.byref(p), max(n, 4)))
        self._owned.append(p.value)
        return DevArray(self, p.value, shape, np.dtype(dtype))

    def upload(self, a: np.ndarray) -> DevArray:
        a = np.ascontiguousarray(a)
        d = self.empty(a.shape, a.dtype)
        check(self.lib.ckd_memcpy_h2d(self.h, d.ptr, a.ctypes.data, a.nbytes))
        self.sync()
        return d

    def write(self, d: DevArray, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=d.dtype)
        assert a.nbytes == d.nbytes
        check(self.lib.ckd_memcpy_h2d(self.h, d.ptr, a.ctypes.data, a.nbytes))
        self.sync()

    def free(self, d: DevArray) -> None:
        if d.ptr in self._owned:
            self._owned.remove(d.ptr)
            check(self.lib.ckd_free(self.h, d.ptr))

    def sync(self) -> None:
        check(self.lib.ckd_synchronize(self.h))

    def close(self) -> None:
        if self.h:
            for p in self._owned:
                self.lib.ckd_free(self.h, p)
            self._owned.clear()
            self.lib.ckd_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- timing --------------------------------------------------------------
    def record(self, slot: int) -> None:
        check(self.lib.ckd_event_record(self.h, slot))

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_float()
        check(self.lib.ckd_event_elapsed_ms(self.h, a, b, C.byref(ms)))
        return ms.value

    def kernel_timing(self, enable: bool) -> None:
        check(self.lib.ckd_kernel_timing(self.h, int(enable)))

    def kernel_timing_read(self, kernel: int) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_ulonglong()
        check(self.lib.ckd_kernel_timing_read(self.h, kernel, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # ---- entry points --------------------------------------------------------
    def memcpy_d2d(self, dst, src, nbytes):
        check(self.lib.ckd_memcpy_d2d(self.h, ptr(dst), ptr(src), nbytes))

    def memset32(self, dst, value, n):
        check(self.lib.ckd_memset32(self.h, ptr(dst), value, n))

    def fx_blit_2x2(self, dst, fxmap, W, H):
        check(self.lib.ckd_fx_blit_2x2(self.h, ptr(dst), ptr(fxmap), W, H))

    def add32(self, dst, src, n):
        check(self.lib.ckd_add32(self.h, ptr(dst), ptr(src), n))

    def blit_add32(self, dst, src, destResX, srcResX, yRes):
        check(self.lib.ckd_blit_add32(self.h, ptr(dst), ptr(src), destResX, srcResX, yRes))

    def blit_add32a(self, dst, src, destResX, srcResX, yRes, alpha):
        check(self.lib.ckd_blit_add32a(self.h, ptr(dst), ptr(src), destResX, srcResX, yRes, alpha))

    def tape_warp32(self, dst, src, W, H, strength, speed):
        check(self.lib.ckd_tape_warp32(self.h, ptr(dst), ptr(src), W, H, strength, speed))

    def blur07_h(self, dst, src, W, H, strength):
        check(self.lib.ckd_blur07_h(self.h, ptr(dst), ptr(src), W, H, strength))

    def blur07_v(self, dst, src, W, H, strength):
        check(self.lib.ckd_blur07_v(self.h, ptr(dst), ptr(src), W, H, strength))

    def blur07(self, dst, src, W, H, strength):
        check(self.lib.ckd_blur07(self.h, ptr(dst), ptr(src), W, H, strength))

    def blur07_inplace_batch(self, imgs, strengths, W, H, passes):
        """in-place blur07 of every image of `imgs` (device tensors of W*H
        pixels), passes 1 = rows, 2 = columns, 3 = both"""
        n = len(imgs)
        arr = (C.c_void_p * max(n, 1))(*[ptr(t) for t in imgs])
        st = (C.c_float * max(n, 1))(*[float(x) for x in strengths])
        check(self.lib.ckd_blur07_inplace_batch(self.h, arr, st, n, W, H, passes))

    def blur26_h(self, dst, src, W, H, strength, gain, passes):
        check(self.lib.ckd_blur26_h(self.h, ptr(dst), ptr(src), W, H, strength, gain, passes))

    def blur26_v(self, dst, src, W, H, strength, gain, passes):
        check(self.lib.ckd_blur26_v(self.h, ptr(dst), ptr(src), W, H, strength, gain, passes))

    def blur26(self, dst, src, W, H, strength, gain, passes):
        check(self.lib.ckd_blur26(self.h, ptr(dst), ptr(src), W, H, strength, gain, passes))

    def polar_create(self, W, H):
        check(self.lib.ckd_polar_create(self.h, W, H))

    def polar_blit(self, dst, src, W, H, inverse):
        check(self.lib.ckd_polar_blit(self.h, ptr(dst), ptr(src), W, H, int(inverse)))

    def grid_remap(self, dst, W, H, grid, tex, texLog2):
        check(self.lib.ckd_grid_remap(self.h, ptr(dst), W, H, ptr(grid), ptr(tex), texLog2))

    def landscape_create(self, hmap: np.ndarray, cmap: np.ndarray, fog: np.ndarray):
        hmap = np.ascontiguousarray(hmap, np.uint8)
        cmap = np.ascontiguousarray(cmap, np.uint32)
        fog = np.ascontiguousarray(fog, np.uint32)
        check(self.lib.ckd_landscape_create(self.h, hmap.ctypes.data, cmap.ctypes.data, fog.ctypes.data))

    # ---- compositing blends (code/util.cpp), in place on dst
    def blend(self, op, dst, src, n, *args):
        """op: mix32 (args: alpha), mixover32, sub32, excl32, softlight32,
        softlight32a, softlight32aa (alpha), overlay32, overlay32a,
        darken32_50, mulsrc32, mulsrc32a, mixsrc32"""
        check(getattr(self.lib, f"ckd_{op}")(self.h, ptr(dst), ptr(src), n, *args))

    def mixsrc32s(self, dst, src, resX, resY, srcStride):
        check(self.lib.ckd_mixsrc32s(self.h, ptr(dst), ptr(src), resX, resY, srcStride))

    def blitsrc32(self, dst, src, destResX, srcResX, yRes, alpha=None):
        if alpha is None:
            check(self.lib.ckd_blitsrc32(self.h, ptr(dst), ptr(src), destResX, srcResX, yRes))
        else:
            check(self.lib.ckd_blitsrc32a(self.h, ptr(dst), ptr(src), destResX, srcResX, yRes,
                                          alpha))

    def fade32(self, dst, n, rgb, alpha):
        check(self.lib.ckd_fade32(self.h, ptr(dst), n, rgb, alpha))

    def zoom32(self, dst, src, W, H, scale):
        check(self.lib.ckd_zoom32(self.h, ptr(dst), ptr(src), W, H, scale))

    # ---- the reference's raymarched effects (code/shadertoy.cpp)
    def sdf_work_counter(self, counters):
        """count the batched SDF kernel's sphere-evaluating steps and shades
        into two device u64 (None: off)"""
        check(self.lib.ckd_sdf_work_counter(self.h, ptr(counters)))

    def shadertoy_step_counter(self, counter):
        """count the FX-map kernels' march steps into a device u64 (None: off)"""
        check(self.lib.ckd_shadertoy_step_counter(self.h, ptr(counter)))

    def fdtunnel_create(self, tex, tex_highlights):
        tex = np.ascontiguousarray(tex, np.uint32)
        hi = np.ascontiguousarray(tex_highlights, np.uint32)
        if tex.size != 1024 * 1024 or hi.size != 1024 * 1024:
            raise ValueError("fdtunnel textures are 1024x1024")
        check(self.lib.ckd_fdtunnel_create(self.h, tex.ctypes.data, hi.ctypes.data))

    def spikey_blurmaps(self, maps, resX=1280, resY=720):
        """the close-up spike's two blur maps, back to back in `maps`"""
        maps = np.ascontiguousarray(maps, np.uint32).reshape(-1)
        n = (resX // 2 + 4) * (resY // 2 + 4)
        if maps.size != 2 * n:
            raise ValueError(f"need 2 x {n} pixels")
        check(self.lib.ckd_spikey_blurmaps(self.h, maps.ctypes.data, maps[n:].ctypes.data, resX,
                                           resY))

    @staticmethod
    def _frames(effect, frames):
        cls = _lib.ST_PARAMS[effect]
        if isinstance(frames, C.Array):
            return frames, len(frames)
        if isinstance(frames, cls):
            frames = [frames]
        return (cls * len(frames))(*frames), len(frames)

    def st_map(self, effect, maps, resX, resY, frames, glow=None, close=False):
        """FX maps of nFrames frames ((resX/2+4) x (resY/2+4) each)"""
        arr, n = self._frames(effect, frames)
        fn = getattr(self.lib, f"ckd_{effect}_map")
        if effect == "spikey":
            check(fn(self.h, ptr(maps), resX, resY, arr, n, int(bool(close))))
        elif effect == "fdtunnel":
            check(fn(self.h, ptr(maps), ptr(glow), resX, resY, arr, n))
        else:
            check(fn(self.h, ptr(maps), resX, resY, arr, n))

    def st_draw(self, effect, dst, resX, resY, frames, close=False):
        """the effect's Draw function into nFrames resX x resY frames"""
        arr, n = self._frames(effect, frames)
        fn = getattr(self.lib, f"ckd_{effect}_draw")
        if effect == "spikey":
            check(fn(self.h, ptr(dst), resX, resY, arr, n, int(bool(close))))
        else:
            check(fn(self.h, ptr(dst), resX, resY, arr, n))

    # ---- voxel ball, torus twister (code/ball.cpp, code/torus-twister.cpp)
    def polar_blit_a(self, dst, src, W, H, inverse=False):
        check(self.lib.ckd_polar_blit_a(self.h, ptr(dst), ptr(src), W, H, int(inverse)))

    @staticmethod
    def _host(a, dtype, n, what):
        a = np.ascontiguousarray(a, dtype).reshape(-1)
        if a.size != n:
            raise ValueError(f"{what}: need {n} elements, got {a.size}")
        return a

    def twister_create(self, hmap, cmap, background, W=1280, H=720):
        m = 1024 * 1024
        hmap = self._host(hmap, np.uint8, m, "height map")
        cmap = self._host(cmap, np.uint32, m, "colour map")
        bg = self._host(background, np.uint32, W * H, "background")
        check(self.lib.ckd_twister_create(self.h, hmap.ctypes.data, cmap.ctypes.data,
                                          bg.ctypes.data, W, H))

    def twister_draw(self, dst, target, W, H, frames):
        frames = [frames] if isinstance(frames, _lib.TwisterParams) else list(frames)
        arr = (_lib.TwisterParams * len(frames))(*frames)
        check(self.lib.ckd_twister_draw(self.h, ptr(dst), ptr(target), W, H, arr, len(frames)))

    def ball_create(self, maps: dict, W=1280, H=720):
        """maps as workloads.ball_maps: heightMaps (5 x 1024^2 bytes),
        colorMaps (2 x 1024^2), beamMaps (3 x 1024^2), envMap (1024^2),
        backgrounds (2 x W*H), halo (W*H)"""
        m = 1024 * 1024
        h = self._host(maps["heightMaps"], np.uint8, 5 * m, "height maps")
        c = self._host(maps["colorMaps"], np.uint32, 2 * m, "colour maps")
        b = self._host(maps["beamMaps"], np.uint32, 3 * m, "beam maps")
        e = self._host(maps["envMap"], np.uint32, m, "environment map")
        bg = self._host(maps["backgrounds"], np.uint32, 2 * W * H, "backgrounds")
        halo = self._host(maps["halo"], np.uint32, W * H, "halo")
        check(self.lib.ckd_ball_create(self.h, h.ctypes.data, c.ctypes.data, b.ctypes.data,
                                       e.ctypes.data, bg.ctypes.data, halo.ctypes.data, W, H))

    def ball_draw(self, dst, target, W, H, frames):
        frames = [frames] if isinstance(frames, _lib.BallParams) else list(frames)
        arr = (_lib.BallParams * len(frames))(*frames)
        check(self.lib.ckd_ball_draw(self.h, ptr(dst), ptr(target), W, H, arr, len(frames)))

    def landscape_cams(self, cams):
        """a ctypes array of cameras, built once and reusable across calls"""
        return (_lib.LandscapeCam * len(cams))(*cams)

    def landscape_draw(self, dst, W, H, cams):
        arr = cams if isinstance(cams, C.Array) else self.landscape_cams(cams)
        check(self.lib.ckd_landscape_draw(self.h, ptr(dst), W, H, arr, len(cams)))

    def tunnelscape_create(self, hmap, cmap, fog):
        hmap = np.ascontiguousarray(hmap, np.uint8)
        cmap = np.ascontiguousarray(cmap, np.uint32)
        fog = np.ascontiguousarray(fog, np.uint32)
        check(self.lib.ckd_tunnelscape_create(self.h, hmap.ctypes.data, cmap.ctypes.data, fog.ctypes.data))

    def tunnelscape_rows(self, dst, W, H, stepU, stepV, speed, time):
        check(self.lib.ckd_tunnelscape_rows(self.h, ptr(dst), W, H, stepU, stepV, speed, time))

    def sdf_render(self, dst, W, H, scene, cam12, shading, step_counter=None):
        cam = (C.c_float * 12)(*[float(v) for v in cam12])
        check(self.lib.ckd_sdf_render(self.h, ptr(dst), W, H, C.byref(scene), cam, shading,
                                      ptr(step_counter)))

    def tunnel_frame(self, dst, tmp, W, H, grid, tex, texLog2, radius):
        check(self.lib.ckd_tunnel_frame(self.h, ptr(dst), ptr(tmp), W, H, ptr(grid), ptr(tex),
                                        texLog2, radius))

    def request_array(self, requests):
        """a ctypes array of requests, built once and reusable across calls"""
        return (_lib.Request * len(requests))(*requests)

    def render_batch(self, requests, scene, shading, tex, texLog2, tmp, tmp_pixels,
                     step_counter=None):
        arr = requests if isinstance(requests, C.Array) else self.request_array(requests)
        check(self.lib.ckd_render_batch(self.h, arr, len(requests),
                                        C.byref(scene) if scene is not None else None, shading,
                                        ptr(tex), texLog2, ptr(tmp), tmp_pixels, ptr(step_counter)))

    # ---- launch shapes and graphs --------------------------------------------
    def get_tuning(self) -> "_lib.Tuning":
        t = _lib.Tuning()
        check(self.lib.ckd_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, t: "_lib.Tuning") -> None:
        check(self.lib.ckd_set_tuning(self.h, C.byref(t)))

    def autotune_batch(self, requests, scene, shading, tex, texLog2, tmp, tmp_pixels):
        arr = requests if isinstance(requests, C.Array) else self.request_array(requests)
        t = _lib.Tuning()
        check(self.lib.ckd_autotune_batch(self.h, arr, len(arr),
                                          C.byref(scene) if scene is not None else None, shading,
                                          ptr(tex), texLog2, ptr(tmp), tmp_pixels, C.byref(t)))
        return t

    def batch_graph(self, requests, scene, shading, tex, texLog2, tmp, tmp_pixels,
                    step_counter=None) -> "BatchGraph":
        arr = requests if isinstance(requests, C.Array) else self.request_array(requests)
        g = C.c_void_p()
        check(self.lib.ckd_batch_graph_create(self.h, arr, len(arr),
                                              C.byref(scene) if scene is not None else None,
                                              shading, ptr(tex), texLog2, ptr(tmp), tmp_pixels,
                                              ptr(step_counter), C.byref(g)))
        # the request array must outlive the graph
        return BatchGraph(self, g, arr, scene)


class BatchGraph:
    def __init__(self, ctx: Context, handle, keep_alive, scene):
        self.ctx, self.h, self._keep = ctx, handle, (keep_alive, scene)

    def launch(self) -> None:
        check(self.ctx.lib.ckd_graph_launch(self.h))

    def update(self, requests) -> None:
        """new cameras / times / grids / radii / outputs for the next replays"""
        arr = requests if isinstance(requests, C.Array) else self.ctx.request_array(requests)
        check(self.ctx.lib.ckd_graph_update(self.h, arr, len(arr)))
        self._keep = (arr, self._keep[1])

    def close(self) -> None:
        if self.h:
            self.ctx.lib.ckd_graph_destroy(self.h)
            self.h = None
