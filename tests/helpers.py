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
    }
