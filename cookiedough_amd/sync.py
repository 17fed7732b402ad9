"""GNU Rocket track playback through the library (ckd_sync_*): a demo's
sync/ folder of .track files, their values at fractional rows, and effect
parameter structs filled from them."""
from __future__ import annotations

import ctypes as C

from . import _lib


class Sync:
    """the tracks under root/sync/ (3rdparty/rocket-stripped/lib/device.c)"""

    def __init__(self, root):
        self.lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self.lib.ckd_sync_open(str(root).encode(), C.byref(h)))
        self.h = h

    def close(self) -> None:
        if self.h:
            self.lib.ckd_sync_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def track(self, name: str) -> int:
        idx = C.c_int()
        _lib.check(self.lib.ckd_sync_track(self.h, name.encode(), C.byref(idx)))
        return idx.value

    def get(self, name: str, row: float) -> float:
        """sync_get_val of the track at `row`"""
        return self.lib.ckd_sync_get(self.h, self.track(name), float(row))

    @staticmethod
    def row(seconds: float) -> float:
        """the row at a music time in seconds (code/audio.cpp:18)"""
        return _lib.load().ckd_sync_row(float(seconds))

    def params(self, effect: str, rows, times):
        """len(rows) parameter structs of `effect` (ctypes array)"""
        cls = _lib.TRACK_PARAMS[effect]
        n = len(rows)
        out = (cls * n)()
        r = (C.c_double * n)(*[float(x) for x in rows])
        t = (C.c_float * n)(*[float(x) for x in times])
        _lib.check(self.lib.ckd_sync_params(self.h, effect.encode(), r, t, n, C.addressof(out)))
        return out


def effect_tracks(effect: str) -> list:
    """the Rocket track names of an effect's parameter struct, in field order"""
    lib = _lib.load()
    names = C.POINTER(C.c_char_p)()
    n = C.c_uint()
    _lib.check(lib.ckd_effect_tracks(effect.encode(), C.byref(names), C.byref(n)))
    return [names[i].decode() for i in range(n.value)]
