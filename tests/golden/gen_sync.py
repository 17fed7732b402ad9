#!/usr/bin/env python3
"""Writes tests/golden/sync.json: synthetic GNU Rocket tracks (every key
type, negative rows, repeated values, large rows) and the values the
reference's own player (code/rocket.cpp over
3rdparty/rocket-stripped/lib/*.c, built as oracle/_ref/libckdref.so with
SYNC_PLAYER) returns for them at integer and fractional rows before, between,
on and after the keys.

Run from the repository root:  python tests/golden/gen_sync.py
"""
from __future__ import annotations

import json
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracles  # noqa: E402


def tracks() -> dict:
    rng = np.random.default_rng(2026)
    out = {
        "gen:step": [(0, 1.5, 0), (10, -2.25, 0), (11, 7.0, 0)],
        "gen:linear": [(-8, 3.0, 1), (4, -1.0, 1), (100, 1e6, 1), (101, 0.1, 1)],
        "gen:smooth": [(2, 0.0, 2), (18, 1.0, 2), (50, -300.5, 2), (51, 12.0, 0)],
        "gen:ramp": [(0, 10.0, 3), (64, 74.0, 3), (65, 0.333, 3)],
        "gen:single": [(5, 42.0, 1)],
        "gen:empty": [],
        "gen:mixed:Deep-Name ok": [(0, 0.1, 1), (3, 0.2, 2), (9, 0.3, 3), (20, 0.4, 0),
                                   (21, 0.5, 1), (1 << 20, -0.5, 1)],
    }
    rows = sorted(set(int(r) for r in rng.integers(0, 4000, 24)))
    out["gen:random"] = [(r, float(np.float32(rng.normal() * 100)), int(rng.integers(0, 4)))
                         for r in rows]
    return out


def sample_rows(keys) -> list:
    rows = [-100.0, -8.5, -0.25, 0.0, 0.5, 1.0, 1e7]
    for r, _, _ in keys:
        rows += [r - 1.0, r - 0.5, float(r), r + 0.25, r + 0.999, r + 1.0]
    return sorted(set(rows))


def main() -> None:
    ref = oracles.ref()
    gold = {"generator": "tests/golden/gen_sync.py", "tracks": []}
    with tempfile.TemporaryDirectory() as td:
        tr = tracks()
        for name, keys in tr.items():
            oracles.write_track_keys(Path(td), name, keys)
        assert ref.ref_st_rocket(td.encode()) == 0
        for name, keys in tr.items():
            rows = sample_rows(keys)
            vals = [ref.ref_sync_value(td.encode(), name.encode(), r) for r in rows]
            gold["tracks"].append({"name": name, "keys": keys, "rows": rows, "values": vals})
    (HERE / "sync.json").write_text(json.dumps(gold, indent=1) + "\n")
    print("wrote", HERE / "sync.json")


if __name__ == "__main__":
    main()
