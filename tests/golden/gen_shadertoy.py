#!/usr/bin/env python3
"""Writes the golden vectors of the reference's raymarched effects
(tests/golden/shadertoy.json + shadertoy_arrays.npz) from the compiled
reference (oracle/_ref/libckdref.so, `make -C oracle ref`).

Parameters are the demo's own: the reference's Rocket player loads the
production tracks of target/sync/ and the harness reads every effect track
at chosen rows (Rocket::get, the reference's key interpolation).  The rows
are key rows of each effect's tracks, a few rows in, so interpolated values
are covered; rows where a feature is on (the spike's specular-only warm-up,
the tunnel's lit tiles, Nautilus' blur) are added.  `time` is row / 8.

For every case the fixture holds the parameters, the SHA-256 of the
reference's FX map (and of its Draw frame where the reference can draw it),
a 32x32 crop of the map and the rsqrtss values of this CPU for a probe set:
the reference's vNorm4 uses _mm_rsqrt_ps, whose table is CPU specific, so
the hashes are only reproducible where the probe matches.  The tunnel's two
1024x1024 textures are workloads.noise_texture(11, 10) and (12, 10), the
close-up's blur maps workloads.st_blur_maps(1280, 720).

Run from the repository root:  python tests/golden/gen_shadertoy.py
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import struct
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracles  # noqa: E402
from oracles import P, aligned_empty  # noqa: E402
from cookiedough_amd import _lib  # noqa: E402
from cookiedough_amd import workloads as wl  # noqa: E402

REF_TARGET = Path("/root/reference/target")
W, H = 1280, 720
MW, MH = W // 2 + 4, H // 2 + 4
CROP = (slice(166, 198), slice(306, 338))
PROBE = [float(x) for x in np.geomspace(1e-3, 1e3, 61, dtype=np.float32)]

# effect id -> (parameter struct, track whose keys give the rows, feature
# track whose non-zero keys add rows)
EFFECTS = {
    "plasma": ("plasma", "plasma:Speed", "plasma:Desaturation"),
    "nautilus": ("nautilus", "nautilus:Hue", "nautilus:Blur"),
    "spikey": ("spikey", "spike:Speed", "distSpike:Warmup"),
    "spikey_close": ("spikey", "closeSpike:Scale", "closeSpike:AspectMul"),
    "sinuses": ("sinuses", "sinusesTunnel:Roll", "sinusesTunnel:OffsX"),
    "laura": ("laura", "laura:Speed", "laura:Roll"),
    "fdtunnel": ("fdtunnel", "tunnel:Radius", "tunnel:LitTiles"),
}


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def key_rows(name: str) -> list:
    enc = "".join(ch if (ch.isalnum() or ch in "._/") else f"-{ord(ch):02X}" for ch in name)
    blob = (REF_TARGET / "sync" / f"_{enc}.track").read_bytes()
    n = struct.unpack_from("<i", blob, 0)[0]
    return [struct.unpack_from("<ifb", blob, 4 + 9 * i) for i in range(n)]


def pick_rows(main: str, feature: str) -> list:
    keys = sorted({r for r, _, _ in key_rows(main)})
    rows = [keys[0] + 3, keys[len(keys) // 2] + 5, keys[-1] - 7] if len(keys) > 2 else [keys[0] + 3]
    on = [r for r, v, _ in key_rows(feature) if v != 0.0]
    rows += [on[0] + 2] if on else []
    return sorted(set(r for r in rows if r >= 0))


def main() -> None:
    ref = oracles.ref()
    orc = oracles.oracle()
    assert ref.ref_st_init() == 0
    assert ref.ref_st_rocket(str(REF_TARGET).encode()) == 0, "Rocket player did not start"
    lut = np.zeros(2049, np.float32)
    ref.ref_cos_lut(P(lut))
    tex = np.ascontiguousarray(wl.noise_texture(11, 10).reshape(-1))
    tex_hi = np.ascontiguousarray(wl.noise_texture(12, 10).reshape(-1))
    ref.ref_st_textures(P(tex), P(tex_hi))
    blur = wl.st_blur_maps(W, H)
    ref.ref_st_blurmaps(P(blur), P(blur[blur.size // 2:]))

    gold = {"generator": "tests/golden/gen_shadertoy.py", "resolution": [W, H], "map": [MW, MH],
            "crop": [[CROP[0].start, CROP[0].stop], [CROP[1].start, CROP[1].stop]],
            "textures": "workloads.noise_texture(11, 10), workloads.noise_texture(12, 10)",
            "blur_maps": "workloads.st_blur_maps(1280, 720)",
            "rsqrtss_probe": {"x": PROBE, "y": [float(orc.orc_rsqrtss(x)) for x in PROBE]},
            "cases": []}
    arrays = {}
    for fid, (pname, main_track, feature) in EFFECTS.items():
        cls = _lib.ST_PARAMS[pname]
        for row in pick_rows(main_track, feature):
            ref.ref_st_row(float(row))
            t = row / 8.0
            vals = {f: ref.ref_st_get(tr.encode()) for f, tr in cls.TRACKS.items()}
            prm = cls(time=t, **vals)
            arr = prm.as_array()
            m, g = aligned_empty(MW * MH), aligned_empty(MW * MH)
            ref.ref_st_map(oracles.ST_IDS[fid], P(m), P(g), t)
            om, og = aligned_empty(MW * MH), aligned_empty(MW * MH)
            orc.orc_st_map(oracles.ST_IDS[fid], P(om), P(og), W, H, P(arr), P(lut), P(tex),
                           P(tex_hi), 1, 8)
            case = {"effect": fid, "row": row, "params": [float(x) for x in arr],
                    "fields": ["time"] + list(cls.TRACKS), "map_sha256": sha(m),
                    "oracle_rsqrtss_equal": bool((om == m).all())}
            k = len(gold["cases"])
            arrays[f"map_crop_{k}"] = m.reshape(MH, MW)[CROP].copy()
            if fid == "fdtunnel":
                case["glow_sha256"] = sha(g)
                case["oracle_rsqrtss_equal"] &= bool((og == g).all())
                arrays[f"glow_crop_{k}"] = g.reshape(MH, MW)[CROP].copy()
            d = aligned_empty(W * H)
            ref.ref_st_draw(oracles.ST_IDS[fid], P(d), t, 1.0)
            case["draw_sha256"] = sha(d)
            gold["cases"].append(case)
            print(fid, row, "oracle(rsqrtss) == reference:", case["oracle_rsqrtss_equal"])
    (HERE / "shadertoy.json").write_text(json.dumps(gold, indent=1) + "\n")
    np.savez_compressed(HERE / "shadertoy_arrays.npz", **arrays)
    print("wrote", len(gold["cases"]), "cases")


if __name__ == "__main__":
    main()
