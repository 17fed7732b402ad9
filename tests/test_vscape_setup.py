"""The voxel landscape's per-frame camera set-up (orc_vscape_setup, which
restates code/landscape.cpp:164-190 and which the GPU's ray set-up matches
bit for bit) against closed forms.  The reference's vscape() reads the
gamepad and Rocket and cannot be called here, so its set-up is pinned by
what the mathematics fixes:
  - column i casts towards (0.25 (i - 640), rayY) with
    rayY = 1024 * (720/1280) * (pi/10) (code/landscape.cpp:32,169,176),
    rotated by the view angle (voxel::vrot2D, code/voxel-shared.h:8-14);
  - its 24:8 direction is that vector normalised and truncated, and
    fishMul = rayY / |ray| (code/landscape.cpp:188), independent of the
    angle;
  - the origin is the camera position truncated to 24:8.
A swapped sine, a wrong sign in the rotation or a fishMul taken after the
rotation fails these."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from oracles import OrcCam, P

W, H = 1280, 720
RAY_Y = 1024.0 * (720.0 / 1280.0) * (math.pi * 0.1)


def setup(orc, X1, Y1, angle):
    oc = OrcCam(np.float32(X1), np.float32(Y1), np.float32(math.cos(angle)),
                np.float32(math.sin(angle)), 90)
    fpX1, fpY1 = C.c_int(), C.c_int()
    dX, dY = np.zeros(W, np.int32), np.zeros(W, np.int32)
    fish = np.zeros(W, np.float32)
    orc.orc_vscape_setup(W, H, C.byref(oc), C.byref(fpX1), C.byref(fpY1), P(dX), P(dY), P(fish))
    return fpX1.value, fpY1.value, dX, dY, fish


@pytest.mark.parametrize("angle", [0.0, 0.3, -1.2, 2.9, math.pi / 2])
def test_vscape_rays_are_the_rotated_column_vectors(orc, angle):
    X1, Y1 = 123.456, -77.7
    fpX1, fpY1, dX, dY, fish = setup(orc, X1, Y1, angle)
    assert fpX1 == int(np.float32(X1) * np.float32(256)) and fpY1 == int(np.float32(Y1) * np.float32(256))
    rx = 0.25 * (np.arange(W) - 640.0)
    ry = np.full(W, RAY_Y)
    n = np.hypot(rx, ry)
    c, s = math.cos(angle), math.sin(angle)
    ex = (c * rx - s * ry) / n
    ey = (s * rx + c * ry) / n
    # 24:8 truncation of FP32 values: within one unit of the exact direction
    assert np.abs(dX - ex * 256).max() <= 1.0
    assert np.abs(dY - ey * 256).max() <= 1.0
    np.testing.assert_allclose(fish, RAY_Y / n, rtol=3e-6)


def test_vscape_fishmul_symmetric_and_bounded(orc):
    _, _, dX, dY, fish = setup(orc, 0.0, 0.0, 0.0)
    i = np.arange(1, 640)
    assert np.array_equal(fish[i], fish[1280 - i])
    assert np.array_equal(dX[i], -dX[1280 - i]) and np.array_equal(dY[i], dY[1280 - i])
    assert fish.max() == np.float32(1.0) and fish.argmax() == 640
    assert fish.min() >= np.float32(RAY_Y / math.hypot(160.0, RAY_Y)) - 1e-6
