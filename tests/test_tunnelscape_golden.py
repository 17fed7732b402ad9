"""Tunnelscape frames and the FX blitter's test pattern against the hashes
of the compiled reference (tests/golden/tunnelscape.json, written by
tests/golden/gen_tunnelscape.py through the reference's Rocket player):
  - the oracle's tscape rows and Tunnelscape_Draw composition (CPU), which
    pins the per-frame set-up of tscape() (code/tunnelscape.cpp:103-134),
    not only its rays;
  - ckd_tunnelscape_rows, the composition through the C ABI (rows, inverse
    polar blit, two in-place 2007 blurs), the shim's Tunnelscape_Draw and
    FxBlitter_DrawTestPattern (GPU)."""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracles import P
from cookiedough_amd import workloads as wl

GOLD = json.loads((Path(__file__).resolve().parent / "golden" / "tunnelscape.json").read_text())
EXE = Path(__file__).resolve().parent / "shim" / "shim_test"
W, H = 1280, 720


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def box_blur_scale(blur: float) -> np.float32:
    """BoxBlurScale, code/deprecated/boxblur.h:13-19"""
    b = np.float32(blur)
    return b if b == 0 else np.float32(min(max(b, np.float32(1)), np.float32(100)) * np.float32(0.01))


@pytest.fixture(scope="module")
def maps():
    return wl.tunnelscape_maps(0)


def oracle_rows(orc, m, c):
    d = np.zeros(W * H, np.uint32)
    orc.orc_tscape_frame(P(d), W, H, P(m.height), P(m.color), P(m.fog), np.float32(c["stepU"]),
                         np.float32(c["stepV"]), np.float32(c["speed"]), np.float32(c["time"]), 8)
    return d


def test_oracle_tscape_rows_match_reference(orc, maps):
    for c in GOLD["cases"]:
        assert sha(oracle_rows(orc, maps, c)) == c["rows_sha256"], c


def test_oracle_tunnelscape_draw_matches_reference(orc, maps):
    m0, m1 = np.zeros(2 * W * H, np.int32), np.zeros(2 * W * H, np.int32)
    orc.orc_polar_maps(P(m0), P(m1), W, H, W, H)
    for c in GOLD["cases"]:
        rows = oracle_rows(orc, maps, c)
        d = np.zeros(W * H, np.uint32)
        orc.orc_polar_blit(P(d), P(rows), P(m1), W, H, W, 8)
        s = box_blur_scale(c["blur"])
        if s != 0:
            orc.orc_blur07(P(d), P(d), W, H, s, 8)
            orc.orc_blur07(P(d), P(d), W, H, s, 8)
        assert sha(d) == c["draw_sha256"], c


def test_oracle_test_pattern_matches_reference(orc):
    fw, fh = W // 2 + 4, H // 2 + 4
    y, x = np.mgrid[0:fh, 0:fw]
    fx = np.where(np.where(y < fh // 2, y, x) & 1, 0xFFFFFFFF, 0).astype(np.uint32).reshape(-1)
    d = np.zeros(W * H, np.uint32)
    orc.orc_fx_blit_2x2(P(d), P(np.ascontiguousarray(fx)), W, H)
    assert sha(d) == GOLD["test_pattern_sha256"]


@pytest.mark.gpu
def test_gpu_tunnelscape_rows_and_draw(ctx, maps):
    ctx.tunnelscape_create(maps.height, maps.color, maps.fog)
    ctx.polar_create(W, H)
    rows, d = ctx.empty(W * H), ctx.empty(W * H)
    for c in GOLD["cases"]:
        print("tunnelscape", c, flush=True)
        ctx.tunnelscape_rows(rows, W, H, c["stepU"], c["stepV"], c["speed"], c["time"])
        ctx.sync()
        assert sha(rows.numpy()) == c["rows_sha256"], c
        ctx.polar_blit(d, rows, W, H, 1)
        s = box_blur_scale(c["blur"])
        if s != 0:
            ctx.blur07(d, d, W, H, float(s))
            ctx.blur07(d, d, W, H, float(s))
        ctx.sync()
        assert sha(d.numpy()) == c["draw_sha256"], c


def _shim(args, env, tmp):
    e = dict(os.environ)
    e.update(env)
    r = subprocess.run([str(EXE), *map(str, args)], capture_output=True, text=True, env=e,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return r


@pytest.mark.gpu
def test_gpu_shim_tunnelscape_draw(maps, tmp_path):
    (tmp_path / "tscape-D7-edit.png.bin").write_bytes(maps.height.tobytes())
    (tmp_path / "tscape-C7W-edit.png.bin").write_bytes(maps.color.tobytes())
    (tmp_path / "foggradient.jpg.bin").write_bytes(maps.fog.tobytes())
    for c in GOLD["cases"]:
        print("shim tunnelscape", c, flush=True)
        tracks = (f"starsTunnel:stepU={c['stepU']};starsTunnel:stepV={c['stepV']};"
                  f"starsTunnel:Speed={c['speed']};starsTunnel:Blur={c['blur']}")
        out = tmp_path / "out.bin"
        _shim(["tunnelscape", c["time"], 2, out],
              {"CKD_SHIM_ASSETS": str(tmp_path), "CKD_SHIM_TRACKS": tracks}, tmp_path)
        got = np.fromfile(out, np.uint32).reshape(2, W * H)
        assert sha(got[0]) == c["draw_sha256"], c
        assert sha(got[1]) == c["draw_sha256"], c


@pytest.mark.gpu
def test_gpu_shim_test_pattern(tmp_path):
    out = tmp_path / "tp.bin"
    _shim(["testpattern", out], {}, tmp_path)
    assert sha(np.fromfile(out, np.uint32)) == GOLD["test_pattern_sha256"]


@pytest.mark.gpu
def test_gpu_shim_lifecycle(tmp_path):
    """Create calls that fail (every asset missing) give back their context
    reference; Create/Destroy pairs, repeated Creates and repeated Destroys
    leave none behind; free functions hold exactly one"""
    r = _shim(["lifecycle"], {"CKD_SHIM_ASSETS": str(tmp_path)}, tmp_path)
    assert r.stdout.strip().endswith("ok")
