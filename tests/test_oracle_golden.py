"""The oracle against the committed golden vectors (tests/golden/, written
by gen_golden.py from the compiled reference).  Needs no reference
checkout."""
import ctypes as C
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from cookiedough_amd import workloads as wl
from oracles import P, aligned_empty, orc_cam

GOLD_DIR = Path(__file__).resolve().parent / "golden"
GOLD = json.loads((GOLD_DIR / "golden.json").read_text())
ARR = np.load(GOLD_DIR / "golden_arrays.npz", allow_pickle=False)


def img(seed, n):
    return (wl.splitmix64(seed, n) & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_bilinear(orc):
    q = img(1, 4 * 4096).reshape(4096, 4)
    fr = (wl.splitmix64(2, 2 * 4096) & np.uint64(0xFF)).astype(np.uint32).reshape(4096, 2)
    for k in range(0, 4096, 3):
        tex = q[k].copy()
        fu, fv = int(fr[k, 0]) * 0x01010101, int(fr[k, 1]) * 0x01010101
        assert orc.orc_bsamp32_32_packed(P(tex), 0, 0, 1, 2, fu, fv) == ARR["bilinear_b32"][k]
        assert ARR["bilinear_b16"][k] == ARR["bilinear_b32"][k]
        tex8 = (tex & 0xFF).astype(np.uint8)
        assert orc.orc_bsamp8(P(tex8), 0, 0, 1, 2, fu, fv) == ARR["bilinear_b8"][k]


def test_cspan(orc):
    out = []
    for a, b in GOLD["cspan_pairs"]:
        A = (C.c_uint16 * 4)(*[(a >> (8 * c)) & 0xff for c in range(4)])
        B = (C.c_uint16 * 4)(*[(b >> (8 * c)) & 0xff for c in range(4)])
        for length in range(1, 25):
            for d in range(1, length + 1):
                dst = np.zeros(d, np.uint32)
                orc.orc_cspan16(P(dst), 1, length, d, A, B)
                out.append(dst)
    assert np.array_equal(np.concatenate(out), ARR["cspan"])


def test_blur07(orc):
    g = GOLD["blur07"]
    W, H = g["W"], g["H"]
    src = img(g["seed"], W * H)
    for key, digest in g["sha256"].items():
        span, mode, inplace = key.split("/")
        s = (int(span) + 0.5) / 255.0
        f = {"h": orc.orc_blur07_h, "v": orc.orc_blur07_v, "2d": orc.orc_blur07}[mode]
        dst = src.copy() if inplace == "1" else np.zeros_like(src)
        f(P(dst), P(dst if inplace == "1" else src), W, H, s, 4)
        assert sha(dst) == digest, key


def test_blur26(orc):
    for key, digest in GOLD["blur26"]["sha256"].items():
        size, strength, gain, passes, mode = key.split("/")
        w, h = map(int, size.split("x"))
        src = aligned_empty(w * h, guard=4)
        src[:w * h] = img(4 + w + h, w * h)
        dst, s0, s1 = (aligned_empty(w * h, guard=4) for _ in range(3))
        f = {"h": orc.orc_blur26_h, "v": orc.orc_blur26_v, "2d": orc.orc_blur26}[mode]
        f(P(dst), P(src), w, h, float(strength), float(gain), int(passes), P(s0), P(s1))
        assert sha(dst[:w * h]) == digest, key


@pytest.fixture(scope="module")
def lmaps():
    m = wl.landscape_maps(0)
    assert [sha(m.height), sha(m.color), sha(m.fog)] == GOLD["landscape_maps_sha256"]
    return m


def test_landscape_frames(orc, lmaps):
    m = lmaps
    for t in GOLD["landscape"]["times"]:
        cam = orc_cam(wl.landscape_cam(wl.landscape_path(0), t))
        out = np.zeros(1280 * 720, np.uint32)
        orc.orc_landscape_frame(P(out), 1280, 720, P(m.height), P(m.color), P(m.fog), C.byref(cam), 8)
        assert sha(out) == GOLD["landscape"]["sha256"][str(t)], t
        assert np.array_equal(out.reshape(720, 1280)[400:464, 600:664], ARR[f"landscape_crop_{t}"])


def test_polar(orc):
    n = 1280 * 720
    m0, m1 = np.zeros(2 * n, np.int32), np.zeros(2 * n, np.int32)
    orc.orc_polar_maps(P(m0), P(m1), 1280, 720, 1280, 720)
    assert sha(m0) == GOLD["polar"]["map_sha256"] and sha(m1) == GOLD["polar"]["inv_map_sha256"]
    src = img(GOLD["polar"]["src_seed"], n)
    dst = np.zeros(n, np.uint32)
    orc.orc_polar_blit(P(dst), P(src), P(m1), 1280, 720, 1280, 8)
    assert sha(dst) == GOLD["polar"]["inverse_blit_sha256"]


def test_blits(orc):
    src = img(GOLD["fx_blit_2x2"]["src_seed"], 644 * 364)
    dst = np.zeros(1280 * 720, np.uint32)
    orc.orc_fx_blit_2x2(P(dst), P(src), 1280, 720)
    assert sha(dst) == GOLD["fx_blit_2x2"]["sha256"]
    g = GOLD["blits"]
    d0, s0 = img(g["dst_seed"], 97 * 31), img(g["src_seed"], 97 * 31)
    for key, digest in g["sha256"].items():
        d = d0.copy()
        if key == "add32":
            orc.orc_add32(P(d), P(s0), d.size)
        elif key == "None":
            orc.orc_blit_add32(P(d), P(s0), 97, 60, 31)
        else:
            orc.orc_blit_add32a(P(d), P(s0), 97, 60, 31, float(key))
        assert sha(d) == digest, key


def test_lut_and_tape_warp(orc):
    lut = np.zeros(2049, np.float32)
    orc.orc_cos_lut(P(lut))
    assert np.array_equal(lut.view(np.uint32), ARR["cos_lut"].view(np.uint32))
    for x, c, s in zip(ARR["lut_x"], ARR["lutcos"], ARR["lutsin"]):
        assert np.float32(orc.orc_lutcosf(P(lut), float(x))) == c
        assert np.float32(orc.orc_lutsinf(P(lut), float(x))) == s
    n = 1280 * 720
    src = img(GOLD["tape_warp"]["src_seed"], n)
    for key, digest in GOLD["tape_warp"]["sha256"].items():
        strength, speed = map(float, key.split("/"))
        dst = np.zeros(n, np.uint32)
        orc.orc_tape_warp32(P(dst), P(src), 1280, 720, strength, speed, P(lut))
        assert sha(dst) == digest, key
