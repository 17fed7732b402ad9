"""ctypes bindings of the checkers: the oracle restatement
(oracle/liboracle.so) and the reference harness (oracle/_ref/libckdref.so,
built from the reference sources only where the reference checkout exists).

Test infrastructure only.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from cookiedough_amd._lib import LandscapeCam, SdfScene

ROOT = Path(__file__).resolve().parent.parent
ORACLE_PATH = ROOT / "oracle" / "liboracle.so"
REF_PATH = ROOT / "oracle" / "_ref" / "libckdref.so"

_p, _u, _i, _f = C.c_void_p, C.c_uint, C.c_int, C.c_float


class OrcCam(C.Structure):
    _fields_ = LandscapeCam._fields_


_ORC = {
    "orc_bsamp_prep": (None, [_i, _i, _u, _u, _p, _p, _p, _p, _p, _p]),
    "orc_bsamp8": (_u, [_p, _u, _u, _u, _u, _u, _u]),
    "orc_bsamp32_16": (None, [_p, _u, _u, _u, _u, _u, _u, _p]),
    "orc_bsamp32_32_packed": (C.c_uint32, [_p, _u, _u, _u, _u, _u, _u]),
    "orc_cspan16": (None, [_p, C.c_ssize_t, _u, _u, _p, _p]),
    "orc_vscape_ray": (None, [_p, _u, _u, _p, _p, _p, _i, _i, _i, _i, _i, _i, _f]),
    "orc_vscape_setup": (None, [_u, _u, C.POINTER(OrcCam), _p, _p, _p, _p, _p]),
    "orc_landscape_frame": (None, [_p, _u, _u, _p, _p, _p, C.POINTER(OrcCam), _i]),
    "orc_tscape_ray": (None, [_p, _u, _p, _p, _p, _i, _i, _i, _i]),
    "orc_tscape_frame": (None, [_p, _u, _u, _p, _p, _p, _f, _f, _f, _f, _i]),
    "orc_polar_maps": (None, [_p, _p, _u, _u, _u, _u]),
    "orc_polar_blit": (None, [_p, _p, _p, _u, _u, _u, _i]),
    "orc_grid_remap": (None, [_p, _u, _u, _p, _p, _u, _i]),
    "orc_blur07_span": (_i, [_f]),
    "orc_blur07_h": (None, [_p, _p, _u, _u, _f, _i]),
    "orc_blur07_v": (None, [_p, _p, _u, _u, _f, _i]),
    "orc_blur07": (None, [_p, _p, _u, _u, _f, _i]),
    "orc_blur26_h": (None, [_p, _p, _u, _u, _f, _f, _u, _p, _p]),
    "orc_blur26_v": (None, [_p, _p, _u, _u, _f, _f, _u, _p, _p]),
    "orc_blur26": (None, [_p, _p, _u, _u, _f, _f, _u, _p, _p]),
    "orc_fx_blit_2x2": (None, [_p, _p, _u, _u]),
    "orc_add32": (None, [_p, _p, _u]),
    "orc_blit_add32": (None, [_p, _p, _u, _u, _u]),
    "orc_blit_add32a": (None, [_p, _p, _u, _u, _u, _f]),
    "orc_memset32": (None, [_p, C.c_uint32, C.c_size_t]),
    "orc_ftofp24": (_i, [_f]),
    "orc_tunnel_frame": (None, [_p, _p, _p, _u, _u, _p, _p, _u, _i, _i]),
    "orc_sdf_frame": (None, [_p, _u, _u, C.POINTER(SdfScene), _f, _i, _p, _i]),
    "orc_sdf_rows": (None, [_p, _u, _u, C.POINTER(SdfScene), _f, _i, _u, _u, _p, _i]),
    "orc_sdf_rows_ex": (None, [_p, _u, _u, C.POINTER(SdfScene), _f, _i, _u, _u, _p, _p, _i]),
    "orc_sdf_camera": (None, [C.POINTER(SdfScene), _f, _p]),
    "orc_cos_lut": (None, [_p]),
    "orc_lutcosf": (_f, [_p, _f]),
    "orc_lutsinf": (_f, [_p, _f]),
    "orc_tape_warp32": (None, [_p, _p, _u, _u, _f, _f, _p]),
    "orc_st_map": (None, [_i, _p, _p, _u, _u, _p, _p, _p, _p, _i, _i]),
    "orc_st_draw": (_i, [_i, _p, _u, _u, _p, _p, _p, _p, _p, _i, _i]),
    "orc_log_ps": (_f, [_f]),
    "orc_exp_ps": (_f, [_f]),
    "orc_topixel": (C.c_uint32, [_p]),
    "orc_rsqrtss": (_f, [_f]),
    "orc_mix32": (None, [_p, _p, _u, C.c_uint8]),
    "orc_mixover32": (None, [_p, _p, _u]),
    "orc_sub32": (None, [_p, _p, _u]),
    "orc_excl32": (None, [_p, _p, _u]),
    "orc_softlight32": (None, [_p, _p, _u]),
    "orc_softlight32a": (None, [_p, _p, _u]),
    "orc_softlight32aa": (None, [_p, _p, _u, _f]),
    "orc_overlay32": (None, [_p, _p, _u]),
    "orc_overlay32a": (None, [_p, _p, _u]),
    "orc_darken32_50": (None, [_p, _p, _u]),
    "orc_mulsrc32": (None, [_p, _p, _u]),
    "orc_mulsrc32a": (None, [_p, _p, _u]),
    "orc_mixsrc32": (None, [_p, _p, _u]),
    "orc_mixsrc32s": (None, [_p, _p, _u, _u, _u]),
    "orc_blitsrc32": (None, [_p, _p, _u, _u, _u]),
    "orc_blitsrc32a": (None, [_p, _p, _u, _u, _u, _f]),
    "orc_fade32": (None, [_p, _u, C.c_uint32, C.c_uint8]),
    "orc_zoom32": (None, [_p, _p, _u, _u, _f]),
    "orc_polar_blit_a": (None, [_p, _p, _p, _u, _u, _u, _i]),
    "orc_twister_tables": (None, [_p, _p]),
    "orc_twister_frame": (None, [_p, _u, _u, _p, _p, _f, _f, _f, _i]),
    "orc_twister_draw": (_i, [_p, _p, _u, _u, _p, _p, _p, _p, _p, _i]),
    "orc_ball_rays": (None, [_u, _f, _f, _f, _f, _p, _p]),
    "orc_ball_draw": (_i, [_p, _p, _u, _u, _p, _p, _p, _p, _p, _p, _p, _p, _i]),
}

_REF = {
    "ref_bsamp_prep": (None, [_i, _i, _u, _u, _p]),
    "ref_bsamp8": (_u, [_p, _u, _u, _u, _u, _u, _u]),
    "ref_bsamp32_16": (C.c_uint32, [_p, _u, _u, _u, _u, _u, _u]),
    "ref_bsamp32_32": (C.c_uint32, [_p, _u, _u, _u, _u, _u, _u]),
    "ref_cspan16": (None, [_p, _i, _u, _u, C.c_uint32, C.c_uint32]),
    "ref_vscape_ray": (None, [_p, _p, _p, _p, _i, _i, _i, _i, _i, _f]),
    "ref_vscape_columns": (None, [_p, _p, _p, _p, _i, _i, _i, _p, _p, _p]),
    "ref_memset32": (None, [_p, C.c_uint32, C.c_size_t]),
    "ref_tscape_ray": (None, [_p, _p, _p, _p, _i, _i, _i, _i]),
    "ref_tunnelscape_setup": (_i, [C.c_char_p, _p, _p, _p]),
    "ref_tscape": (None, [_p, _f]),
    "ref_tunnelscape_draw": (None, [_p, _f, _f]),
    "ref_fx_test_pattern": (_i, [_p]),
    "ref_polar_maps": (None, [_p, _p, _u, _u, _u, _u]),
    "ref_polar_create": (_i, []),
    "ref_polar_blit_2x2": (None, [_p, _p, _i]),
    "ref_polar_pad_2x2": (None, []),
    "ref_sync_value": (C.c_double, [C.c_char_p, C.c_char_p, C.c_double]),
    "ref_polar_destroy": (None, []),
    "ref_polar_blit": (None, [_p, _p, _i]),
    "ref_blur07_h": (None, [_p, _p, _u, _u, _f]),
    "ref_blur07_v": (None, [_p, _p, _u, _u, _f]),
    "ref_blur07": (None, [_p, _p, _u, _u, _f]),
    "ref_blur26_h": (None, [_p, _p, _u, _u, _f, _f, _u]),
    "ref_blur26_v": (None, [_p, _p, _u, _u, _f, _f, _u]),
    "ref_blur26": (None, [_p, _p, _u, _u, _f, _f, _u]),
    "ref_fx_blit_2x2": (None, [_p, _p]),
    "ref_add32": (None, [_p, _p, _u]),
    "ref_omp_threads": (None, [_i]),
    "ref_blit_add32": (None, [_p, _p, _u, _u, _u]),
    "ref_blit_add32a": (None, [_p, _p, _u, _u, _u, _f]),
    "ref_cos_lut": (None, [_p]),
    "ref_lutcosf": (_f, [_f]),
    "ref_lutsinf": (_f, [_f]),
    "ref_tape_warp32": (None, [_p, _p, _u, _u, _f, _f]),
    "ref_st_init": (_i, []),
    "ref_st_rocket": (_i, [C.c_char_p]),
    "ref_st_row": (None, [C.c_double]),
    "ref_st_get": (C.c_double, [C.c_char_p]),
    "ref_st_textures": (None, [_p, _p]),
    "ref_st_blurmaps": (None, [_p, _p]),
    "ref_st_map": (None, [_i, _p, _p, _f]),
    "ref_st_draw": (None, [_i, _p, _f, _f]),
    "ref_st_gamma": (None, [_p, _f, _p]),
    "ref_st_topixel4": (None, [_p, _p]),
    "ref_st_expfog": (_f, [_f, _f]),
    "ref_st_vlerp4": (None, [_p, _p, _f, _p]),
    "ref_mix32": (None, [_p, _p, _u, C.c_uint8]),
    "ref_mixover32": (None, [_p, _p, _u]),
    "ref_sub32": (None, [_p, _p, _u]),
    "ref_excl32": (None, [_p, _p, _u]),
    "ref_softlight32": (None, [_p, _p, _u]),
    "ref_softlight32a": (None, [_p, _p, _u]),
    "ref_softlight32aa": (None, [_p, _p, _u, _f]),
    "ref_overlay32": (None, [_p, _p, _u]),
    "ref_overlay32a": (None, [_p, _p, _u]),
    "ref_darken32_50": (None, [_p, _p, _u]),
    "ref_mulsrc32": (None, [_p, _p, _u]),
    "ref_mulsrc32a": (None, [_p, _p, _u]),
    "ref_mixsrc32": (None, [_p, _p, _u]),
    "ref_mixsrc32s": (None, [_p, _p, _u, _u, _u]),
    "ref_blitsrc32": (None, [_p, _p, _u, _u, _u]),
    "ref_blitsrc32a": (None, [_p, _p, _u, _u, _u, _f]),
    "ref_fade32": (None, [_p, _u, C.c_uint32, C.c_uint8]),
    "ref_zoom32": (None, [_p, _p, _u, _u, _f]),
    "ref_polar_blit_a": (None, [_p, _p, _i]),
    "ref_polar_pad_forward": (None, []),
    "ref_shared_init": (_i, []),
    "ref_shared_target": (_p, [_u]),
    "ref_shared_clear": (None, [_u]),
    "ref_ball_setup": (_i, [C.c_char_p, _p, _p, _p, _p, _p, _p]),
    "ref_ball_draw": (None, [_p, _f, _f]),
    "ref_twister_setup": (_i, [C.c_char_p, _p, _p, _p]),
    "ref_twister_tables": (None, [_p, _p]),
    "ref_twister_draw": (None, [_p, _f, _f]),
}

# oracle / harness effect ids (oracle/ckd_oracle.h) of the effects' map
# renderers; the spike has two (distant, close-up)
ST_IDS = {"plasma": 0, "nautilus": 1, "spikey": 2, "spikey_close": 3, "sinuses": 4,
          "laura": 5, "fdtunnel": 6}
ST_STRUCT = {"spikey_close": "spikey"}  # id name -> parameter struct name


def track_file_name(name: str) -> str:
    """_<name>.track with every char outside [A-Za-z0-9._/] as -XX
    (3rdparty/rocket-stripped/lib/device.c:26-75)"""
    return "_" + "".join(ch if (ch.isascii() and (ch.isalnum() or ch in "._/")) else
                         "".join(f"-{b:02X}" for b in ch.encode()) for ch in name) + ".track"


def write_track_keys(root: Path, name: str, keys) -> None:
    """one track file with keys [(row, value, type)], rows ascending"""
    import struct
    d = Path(root) / "sync"
    d.mkdir(parents=True, exist_ok=True)
    blob = struct.pack("<i", len(keys)) + b"".join(struct.pack("<ifb", r, float(v), t)
                                                   for r, v, t in keys)
    (d / track_file_name(name)).write_bytes(blob)


def write_rocket_tracks(root: Path, values: dict) -> None:
    """Writes GNU Rocket player track files (root/sync/_<name>.track, the
    layout 3rdparty/rocket-stripped/lib/device.c reads): for each track
    name, one step key per row 0..K-1 holding values[name][row]."""
    for name, vals in values.items():
        write_track_keys(root, name, [(row, v, 0) for row, v in enumerate(vals)])


def _bind(path: Path, table: dict) -> C.CDLL:
    lib = C.CDLL(str(path))
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_oracle = None
_ref = None


def oracle() -> C.CDLL:
    global _oracle
    if _oracle is None:
        if not ORACLE_PATH.exists():
            raise FileNotFoundError(f"{ORACLE_PATH} not built (make -C oracle)")
        _oracle = _bind(ORACLE_PATH, _ORC)
    return _oracle


def ref_available() -> bool:
    return REF_PATH.exists()


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        _ref = _bind(REF_PATH, _REF)
    return _ref


def P(a: np.ndarray) -> C.c_void_p:
    """pointer argument for a C call; it holds a reference to the array, so a
    temporary (P(np.zeros(n))) stays alive until the call returns"""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def aligned_empty(n: int, dtype=np.uint32, align: int = 64, guard: int = 0) -> np.ndarray:
    """zeroed 1-D array of n (+guard) elements whose data is `align`-aligned"""
    it = np.dtype(dtype).itemsize
    raw = np.zeros((n + guard) * it + align, np.uint8)
    off = (-raw.ctypes.data) % align
    return raw[off:off + (n + guard) * it].view(dtype)


def orc_cam(cam: LandscapeCam) -> OrcCam:
    return OrcCam(cam.X1, cam.Y1, cam.viewCos, cam.viewSin, cam.mapTilt)
