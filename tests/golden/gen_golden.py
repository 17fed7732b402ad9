#!/usr/bin/env python3
"""Writes the golden vectors of tests/golden/ from the compiled reference
(oracle/_ref/libckdref.so, built from the reference sources by
`make -C oracle ref`).  Inputs are regenerated from splitmix64 seeds; the
fixtures hold the seeds, the reference outputs (small arrays) and SHA-256
digests of the large ones.  Run from the repository root:

    python tests/golden/gen_golden.py
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracles  # noqa: E402
from oracles import P, aligned_empty  # noqa: E402
from cookiedough_amd import workloads as wl  # noqa: E402

BLUR07_SPANS = [1, 2, 3, 4, 5, 8, 9, 16, 17, 32, 33, 63, 64, 100]
BLUR26_CASES = [(64, 48, 10.0, 0.0, 1), (64, 48, 37.5, 0.25, 2), (128, 96, 100.0, 1.0, 3),
                (200, 120, 3.3, 0.0, 3), (96, 200, 55.0, 0.5, 4)]
LS_TIMES = [0.0, 0.7, 1.9]


def img(seed: int, n: int) -> np.ndarray:
    return (wl.splitmix64(seed, n) & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main() -> None:
    ref = oracles.ref()
    gold: dict = {"generator": "tests/golden/gen_golden.py", "inputs": "splitmix64 seeds"}
    arrays: dict = {}

    # bilinear: 4096 random quads and fractions
    q = img(1, 4 * 4096).reshape(4096, 4)
    fr = (wl.splitmix64(2, 2 * 4096) & np.uint64(0xFF)).astype(np.uint32).reshape(4096, 2)
    b32, b16, b8 = [], [], []
    for k in range(4096):
        tex = q[k].copy()
        tex8 = (tex & 0xFF).astype(np.uint8)
        fu, fv = int(fr[k, 0]) * 0x01010101, int(fr[k, 1]) * 0x01010101
        b32.append(ref.ref_bsamp32_32(P(tex), 0, 0, 1, 2, fu, fv))
        b16.append(ref.ref_bsamp32_16(P(tex), 0, 0, 1, 2, fu, fv))
        b8.append(ref.ref_bsamp8(P(tex8), 0, 0, 1, 2, fu, fv))
    arrays["bilinear_b32"] = np.array(b32, np.uint32)
    arrays["bilinear_b16"] = np.array(b16, np.uint32)
    arrays["bilinear_b8"] = np.array(b8, np.uint32)

    # cspan: every (length <= 24, drawLength) for 4 colour pairs, concatenated
    pairs = [(0x00000000, 0xFFFFFFFF), (0xFFFFFFFF, 0x00000000), (0x12345678, 0x87654321),
             (0xFF010203, 0x00FFFEFD)]
    out = []
    for a, b in pairs:
        for length in range(1, 25):
            for d in range(1, length + 1):
                dst = np.zeros(d, np.uint32)
                ref.ref_cspan16(P(dst), 1, length, d, a, b)
                out.append(dst)
    arrays["cspan"] = np.concatenate(out)
    gold["cspan_pairs"] = [[a, b] for a, b in pairs]

    # blur07 on a 128x112 image: every mode, in and out of place (span <= both sides)
    W, H = 128, 112
    src = img(3, W * H)
    b07 = {}
    for span in BLUR07_SPANS:
        s = (span + 0.5) / 255.0
        for mode, fn in [("h", ref.ref_blur07_h), ("v", ref.ref_blur07_v), ("2d", ref.ref_blur07)]:
            for inplace in (0, 1):
                dst = src.copy() if inplace else np.zeros_like(src)
                fn(P(dst), P(dst if inplace else src), W, H, s)
                b07[f"{span}/{mode}/{inplace}"] = sha(dst)
    gold["blur07"] = {"seed": 3, "W": W, "H": H, "sha256": b07}

    # blur26 (zero guard pixels past the image)
    b26 = {}
    for (w, h, strength, gain, passes) in BLUR26_CASES:
        s26 = aligned_empty(w * h, guard=4)
        s26[:w * h] = img(4 + w + h, w * h)
        for mode, fn in [("h", ref.ref_blur26_h), ("v", ref.ref_blur26_v), ("2d", ref.ref_blur26)]:
            dst = aligned_empty(w * h, guard=4)
            fn(P(dst), P(s26), w, h, strength, gain, passes)
            b26[f"{w}x{h}/{strength}/{gain}/{passes}/{mode}"] = sha(dst[:w * h])
    gold["blur26"] = {"seed": "4+W+H", "sha256": b26}

    # voxel landscape frames: memset32(fog[0]) + reference vscape_ray per column
    m = wl.landscape_maps(0)
    gold["landscape_maps_sha256"] = [sha(m.height), sha(m.color), sha(m.fog)]
    orc = oracles.oracle()
    frames = {}
    for t in LS_TIMES:
        cam = wl.landscape_cam(wl.landscape_path(0), t)
        oc = oracles.orc_cam(cam)
        fpX1, fpY1 = C.c_int(), C.c_int()
        dX = np.zeros(1280, np.int32)
        dY = np.zeros(1280, np.int32)
        fish = np.zeros(1280, np.float32)
        orc.orc_vscape_setup(1280, 720, C.byref(oc), C.byref(fpX1), C.byref(fpY1), P(dX), P(dY),
                             P(fish))
        frame = aligned_empty(1280 * 720)
        ref.ref_memset32(P(frame), int(m.fog[0]), 1280 * 720)
        for col in range(1280):
            ref.ref_vscape_ray(P(frame[col:]), P(m.height), P(m.color), P(m.fog), cam.mapTilt,
                               fpX1.value, fpY1.value, int(dX[col]), int(dY[col]), float(fish[col]))
        frames[str(t)] = sha(frame)
        arrays[f"landscape_crop_{t}"] = frame.reshape(720, 1280)[400:464, 600:664].copy()
    gold["landscape"] = {"times": LS_TIMES, "sha256": frames,
                         "camera": "workloads.landscape_cam(landscape_path(0), t)"}

    # polar maps and the inverse blit at 1280x720
    n = 1280 * 720
    m0, m1 = np.zeros(2 * n, np.int32), np.zeros(2 * n, np.int32)
    ref.ref_polar_maps(P(m0), P(m1), 1280, 720, 1280, 720)
    ref.ref_polar_create()
    psrc = aligned_empty(n)
    psrc[:] = img(5, n)
    pdst = aligned_empty(n)
    ref.ref_polar_blit(P(pdst), P(psrc), 1)
    ref.ref_polar_destroy()
    gold["polar"] = {"map_sha256": sha(m0), "inv_map_sha256": sha(m1), "src_seed": 5,
                     "inverse_blit_sha256": sha(pdst)}

    # Fx_Blit_2x2, blits, TapeWarp32, LUT
    fx = aligned_empty(644 * 364)
    fx[:] = img(6, 644 * 364)
    fo = aligned_empty(n)
    ref.ref_fx_blit_2x2(P(fo), P(fx))
    gold["fx_blit_2x2"] = {"src_seed": 6, "sha256": sha(fo)}
    d0, s0 = img(7, 97 * 31), img(8, 97 * 31)
    blits = {}
    for alpha in [None, 0.0, 0.3, 1.0, 1.7]:
        d = d0.copy()
        if alpha is None:
            ref.ref_blit_add32(P(d), P(s0), 97, 60, 31)
        else:
            ref.ref_blit_add32a(P(d), P(s0), 97, 60, 31, alpha)
        blits[str(alpha)] = sha(d)
    d = d0.copy()
    ref.ref_add32(P(d), P(s0), d.size)
    blits["add32"] = sha(d)
    gold["blits"] = {"dst_seed": 7, "src_seed": 8, "sha256": blits}
    lut = np.zeros(2049, np.float32)
    ref.ref_cos_lut(P(lut))
    arrays["cos_lut"] = lut
    xs = (wl.Rng(9).uniform(256, -100, 100)).astype(np.float32)
    arrays["lut_x"] = xs
    arrays["lutcos"] = np.array([ref.ref_lutcosf(float(x)) for x in xs], np.float32)
    arrays["lutsin"] = np.array([ref.ref_lutsinf(float(x)) for x in xs], np.float32)
    tw = {}
    for strength, speed in [(3.0, 0.05), (12.5, 0.013)]:
        tdst = aligned_empty(n)
        ref.ref_tape_warp32(P(tdst), P(psrc), 1280, 720, strength, speed)
        tw[f"{strength}/{speed}"] = sha(tdst)
    gold["tape_warp"] = {"src_seed": 5, "sha256": tw}

    (HERE / "golden.json").write_text(json.dumps(gold, indent=1, sort_keys=True) + "\n")
    np.savez_compressed(HERE / "golden_arrays.npz", **arrays)
    print("wrote", HERE / "golden.json", HERE / "golden_arrays.npz")


if __name__ == "__main__":
    main()
