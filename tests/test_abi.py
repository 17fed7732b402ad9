"""The C ABI library loads, exports every entry point include/*.h declares,
and rejects bad arguments with a status and a message (no GPU needed)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from cookiedough_amd import _lib
from cookiedough_amd._lib import CKD_EDEVICE, CKD_EINVAL, CKD_OK, SIGNATURES

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        names |= set(re.findall(r"CKD_API\s+[\w\s\*]+?\b(ckd_\w+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported_and_bound():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 30
    for n in sorted(names):
        assert hasattr(lib, n), n
        assert n in SIGNATURES, f"{n} has no ctypes signature"


def test_create_without_device_fails_cleanly_or_succeeds_on_gpu():
    lib = _lib.load()
    h = C.c_void_p()
    s = lib.ckd_create(0, C.byref(h))
    assert s in (CKD_OK, CKD_EDEVICE)
    if s == CKD_OK:
        lib.ckd_destroy(h)
    else:
        assert "device" in _lib.last_error()


@pytest.mark.parametrize("call,msg", [
    (lambda l: l.ckd_blur07_h(None, 1, 1, 64, 64, 0.1), "null context"),
    (lambda l: l.ckd_blur07_h(None, None, 1, 64, 64, 0.1), "null buffer"),
    (lambda l: l.ckd_blur07_v(None, 1, 1, 64, 8, 10 / 255), "exceeds"),
    (lambda l: l.ckd_blur07(None, 1, 1, 64, 64, -0.5), "strength"),
    (lambda l: l.ckd_blur26(None, 1, 2, 4096, 64, 10.0, 0.0, 1), "size"),
    (lambda l: l.ckd_blur26(None, 1, 1, 64, 64, 10.0, 0.0, 1), "differ"),
    (lambda l: l.ckd_blur26_v(None, 1, 2, 66, 64, 10.0, 0.0, 1), "multiples of 4"),
    (lambda l: l.ckd_fx_blit_2x2(None, 8, 8, 1282, 720), "W/2"),
    (lambda l: l.ckd_blit_add32(None, 1, 1, 10, 20, 4), "srcResX"),
    (lambda l: l.ckd_blit_add32a(None, 1, 1, 20, 10, 4, -1.0), "alpha"),
    (lambda l: l.ckd_grid_remap(None, 1, 64, 64, 1, 1, 20), "texLog2"),
    (lambda l: l.ckd_tunnel_frame(None, 1, 1, 320, 180, 1, 1, 9, 100), "radius"),
    (lambda l: l.ckd_polar_create(None, 3, 3), "size"),
    (lambda l: l.ckd_event_record(None, 99), "slot"),
])
def test_argument_validation_without_gpu(call, msg):
    lib = _lib.load()
    assert call(lib) == CKD_EINVAL
    assert msg in _lib.last_error()


def test_landscape_rejects_negative_tilt():
    lib = _lib.load()
    cam = _lib.LandscapeCam(0, 0, 1, 0, -5)
    assert lib.ckd_landscape_draw(None, 8, 64, 64, C.byref(cam), 1) == CKD_EINVAL
    assert "mapTilt" in _lib.last_error()


def test_landscape_rejects_frames_too_wide_for_their_height():
    """W > 64 H would put the edge columns' fish-eye factor below the range
    in which the march's division is exact"""
    lib = _lib.load()
    cam = _lib.LandscapeCam(0, 0, 1, 0, 90)
    assert lib.ckd_landscape_draw(None, 8, 4096, 63, C.byref(cam), 1) == CKD_EINVAL
    assert "wider" in _lib.last_error()
    assert lib.ckd_landscape_draw(None, 8, 4096, 64, C.byref(cam), 1) == CKD_EINVAL
    assert "null context" in _lib.last_error()
    r = _lib.Request()
    r.effect = 1
    r.W, r.H, r.dst = 1024, 15, 8
    r.vcam = cam
    arr = (_lib.Request * 1)(r)
    assert lib.ckd_render_batch(None, arr, 1, None, 0, None, 9, None, 0, None) == CKD_EINVAL
    assert "wider" in _lib.last_error()


def test_sdf_rejects_bad_scene():
    lib = _lib.load()
    sc = _lib.SdfScene()
    sc.nSpheres = 40
    sc.maxSteps = 64
    cam = (C.c_float * 12)()
    assert lib.ckd_sdf_render(None, 8, 64, 64, C.byref(sc), cam, 0, None) == CKD_EINVAL
    assert "nSpheres" in _lib.last_error()


def test_batch_validation():
    lib = _lib.load()
    r = _lib.Request()
    r.effect = 7
    r.W, r.H, r.dst = 64, 64, 8
    arr = (_lib.Request * 1)(r)
    assert lib.ckd_render_batch(None, arr, 1, None, 0, None, 9, None, 0, None) == CKD_EINVAL
    assert "unknown effect" in _lib.last_error()
    r.effect = 2
    r.radius = 3
    r.grid = 8
    arr = (_lib.Request * 1)(r)
    assert lib.ckd_render_batch(None, arr, 1, None, 0, 8, 9, None, 0, None) == CKD_EINVAL
    assert "scratch" in _lib.last_error()


def test_tuning_validation_without_gpu():
    lib = _lib.load()
    for bad in [(3, 32, 8, 32), (2, 24, 8, 32), (2, 32, 0, 32), (2, 32, 8, 4), (2, 32, 8, 1024),
                (2, 32, 8, 32, 3), (2, 32, 8, 32, -1), (2, 32, 8, 32, 0, 3), (2, 32, 8, 32, 0, 16),
                (2, 32, 8, 32, 0, 8, 9), (2, 32, 8, 32, 0, 8, -1)]:
        t = _lib.Tuning(*bad)
        assert lib.ckd_set_tuning(None, C.byref(t)) == CKD_EINVAL
        assert "tuning" in _lib.last_error()
    t = _lib.Tuning(2, 32, 8, 32)
    assert lib.ckd_set_tuning(None, C.byref(t)) == CKD_EINVAL
    assert "null context" in _lib.last_error()


def test_shadertoy_argument_validation_without_gpu():
    lib = _lib.load()
    p = _lib.PlasmaParams(time=1.0, speed=1.0, hue=24.0, gamma=2.0)
    assert lib.ckd_plasma_map(None, 8, 1282, 720, C.byref(p), 1) == CKD_EINVAL
    assert "multiple of 8" in _lib.last_error()
    assert lib.ckd_plasma_draw(None, 8, 1280, 720, C.byref(p), 1) == CKD_EINVAL
    assert "null context" in _lib.last_error()
    p.gamma = float("nan")
    assert lib.ckd_plasma_map(None, 8, 1280, 720, C.byref(p), 1) == CKD_EINVAL
    assert "finite" in _lib.last_error()
    s = _lib.SpikeyParams(time=1.0, closeMixBlurOpacity=float("inf"))
    assert lib.ckd_spikey_draw(None, 8, 1280, 720, C.byref(s), 1, 1) == CKD_EINVAL
    assert "finite" in _lib.last_error()
    assert lib.ckd_spikey_blurmaps(None, None, 8, 1280, 720) == CKD_EINVAL
    assert "null map" in _lib.last_error()
    t = _lib.FdTunnelParams(time=1.0)
    assert lib.ckd_fdtunnel_map(None, 8, None, 1280, 720, C.byref(t), 1) == CKD_EINVAL
    assert "glow" in _lib.last_error()


def test_shadertoy_param_structs_match_the_header():
    """field order of each ctypes struct == the C struct in include/ckd.h"""
    text = (ROOT / "include" / "ckd.h").read_text()
    structs = dict(_lib.ST_PARAMS, ball=_lib.BallParams, twister=_lib.TwisterParams)
    for name, cls in structs.items():
        body = re.search(r"typedef struct ckd_%s_params \{(.*?)\} ckd_%s_params;" % (name, name),
                         text, re.S).group(1)
        fields = re.findall(r"float (\w+);", body)
        assert fields == [f for f, _ in cls._fields_], name
        tracks = re.findall(r"/\* ([\w:]+)\s*\*/", body)
        assert tracks == list(cls.TRACKS.values()), name


def test_voxel2_argument_validation_without_gpu():
    lib = _lib.load()
    b = _lib.BallParams(time=1.0, radius=float("nan"))
    assert lib.ckd_ball_draw(None, 8, 16, 1280, 720, C.byref(b), 1) == CKD_EINVAL
    assert "finite" in _lib.last_error()
    assert lib.ckd_ball_draw(None, 8, 8, 1280, 720, C.byref(b), 1) == CKD_EINVAL
    assert "differ" in _lib.last_error()
    b.radius = 1800.0
    assert lib.ckd_ball_draw(None, 8, 16, 1280, 720, C.byref(b), 1) == CKD_EINVAL
    assert "null context" in _lib.last_error()
    assert lib.ckd_ball_create(None, 8, 8, 8, 8, 8, None, 1280, 720) == CKD_EINVAL
    assert "null map" in _lib.last_error()
    t = _lib.TwisterParams(time=1.0, speed=float("inf"))
    assert lib.ckd_twister_draw(None, 8, 16, 1280, 720, C.byref(t), 1) == CKD_EINVAL
    assert "finite" in _lib.last_error()
    assert lib.ckd_twister_create(None, 8, 8, 8, 1000, 720) == CKD_EINVAL
    assert "width" in _lib.last_error()
    assert lib.ckd_polar_blit_a(None, 8, 8, 1280, 720, 0) == CKD_EINVAL
    assert "differ" in _lib.last_error()
