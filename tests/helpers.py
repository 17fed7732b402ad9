"""Shared test helpers: seeded images, error statistics."""
from __future__ import annotations

import numpy as np


def rand_img(seed: int, n: int) -> np.ndarray:
    return np.random.default_rng(seed).integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)


def channel_errors(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """|a-b| per channel, shape (n, 4)"""
    ca = a.reshape(-1).view(np.uint8).reshape(-1, 4).astype(np.int16)
    cb = b.reshape(-1).view(np.uint8).reshape(-1, 4).astype(np.int16)
    return np.abs(ca - cb)


def error_stats(a: np.ndarray, b: np.ndarray, tol: int = 2) -> dict:
    e = channel_errors(a, b)
    per_px = e.max(axis=1)
    return {
        "max": int(e.max()),
        "mean": float(e.mean()),
        "frac_diff": float((per_px > 0).mean()),
        "frac_over_tol": float((per_px > tol).mean()),
        "p9999": float(np.quantile(per_px, 0.9999)),
        "pixels": int(per_px.size),
    }


# ---- the SDF scene against its FP64 render -------------------------------
#
# A pixel is "rounding-decided" when its FP64 colour moves by more than the
# error bar under a 1% change of the march's stopping threshold or a one-step
# change of its step budget (orc_sdf_rows_ex's conditioning): a ray that
# passes a surface at about eps, or is still moving when its steps run out.
# FP32 distances (errors of order 1e-6 against eps = 1e-3) can stop such a
# march elsewhere, and the colour follows the rounding rather than the
# scene.  Everywhere else the GPU must be within the error bar of FP64.


def sdf_fp64(orc, sc, W, H, t, row_start=0, row_step=1, nthreads=8):
    """FP64 render of rows row_start, row_start+row_step, ...: pixels,
    march steps and per-pixel conditioning"""
    import ctypes as C
    from oracles import P
    rows = len(range(row_start, H, row_step))
    ref = np.zeros(rows * W, np.uint32)
    st = np.zeros(rows * W, np.uint16)
    cond = np.zeros(rows * W, np.uint8)
    orc.orc_sdf_rows_ex(P(ref), W, H, C.byref(sc), t, 64, row_start, row_step, P(st), P(cond),
                        nthreads)
    return ref, st, cond


def sdf_report(got: np.ndarray, ref: np.ndarray, cond: np.ndarray, tol: int = 2) -> dict:
    e = channel_errors(got, ref)
    per_px = e.max(axis=1)
    cls = cond > tol
    out = ~cls
    return {
        "max": int(per_px.max()) if per_px.size else 0,
        "mean": float(e.mean()) if e.size else 0.0,
        "frac_diff": float((per_px > 0).mean()) if per_px.size else 0.0,
        "frac_over_tol": float((per_px > tol).mean()) if per_px.size else 0.0,
        "pixels": int(per_px.size),
        "rounding_decided_frac": float(cls.mean()) if per_px.size else 0.0,
        "max_outside_rounding_decided": int(per_px[out].max()) if out.any() else 0,
        "over_tol_inside_rounding_decided": int((per_px[cls] > tol).sum()),
    }


def merge_reports(reports: list[dict]) -> dict:
    """pixel-weighted merge of sdf_report / error_stats dictionaries"""
    n = sum(r["pixels"] for r in reports)
    out = {"pixels": n}
    for k in ("max", "max_outside_rounding_decided"):
        if k in reports[0]:
            out[k] = max(r[k] for r in reports)
    for k in ("mean", "frac_diff", "frac_over_tol", "rounding_decided_frac"):
        if k in reports[0]:
            out[k] = sum(r[k] * r["pixels"] for r in reports) / max(1, n)
    if "over_tol_inside_rounding_decided" in reports[0]:
        out["over_tol_inside_rounding_decided"] = sum(r["over_tol_inside_rounding_decided"]
                                                      for r in reports)
    return out
