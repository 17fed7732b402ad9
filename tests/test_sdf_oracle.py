"""Pins the renderer of the synthetic SDF scene (oracle/orc_sdf.c, the FP64
yardstick of the sphere tracer) against closed forms it does not use.

The oracle sphere-traces the scene.  The checks here ray-cast it
analytically instead, in numpy FP64: the nearest root of
|o + t d - c|^2 = r^2 over the spheres and of o_y + t d_y = groundY for
the plane, the analytic normal (p - c)/r or (0, 1, 0), and the shading of
that hit.  A miss pixel must be exactly the fog colour raised to gamma
with alpha 255.  The shading of a hit is computed two ways: from the
scene's shading law in FP64, and through the reference's own FP32 helpers
(ExpFog, vLerp4, GammaAdj and ToPixel4 of code/shadertoy-util.h:62,136-146,
186-190,237-241, called in the compiled reference) that the scene's
shading mirrors.

A sphere tracer and a ray caster disagree, legitimately, where the march
is ill-conditioned; those pixels are excluded and counted:
  - the ray passes a sphere it does not hit within 1e-2 of its surface
    (the tracer slows down there, and stops if it comes within eps);
  - the hit is grazing, |n.d| < 0.15;
  - another surface lies within 2e-2 of the hit point (the central
    difference normal and the nearest-object id mix two surfaces);
  - the tracer used all of its steps.
The exclusions must stay a small fraction of the frame, so that the
comparison keeps its teeth.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracles import P
from cookiedough_amd import workloads as wl
from cookiedough_amd._lib import SdfScene

SIL, GRAZE, CREASE = 1e-2, 0.15, 2e-2


def _rays(orc, sc, W, H, t):
    cam = np.zeros(12, np.float32)
    orc.orc_sdf_camera(C.byref(sc), t, P(cam))
    c = cam.astype(np.float64)
    x, y = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    u = (2 * x - W) / H
    v = (H - 2 * y) / H
    d = c[3:6] + u[..., None] * c[6:9] + v[..., None] * c[9:12]
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return c[0:3], d.reshape(-1, 3)


def analytic(orc, sc, W, H, t):
    """per pixel: t_hit (inf = miss), object id (nSpheres = plane, -1 miss),
    unit normal, and the ill-conditioning mask"""
    o, d = _rays(orc, sc, W, H, t)
    n = sc.nSpheres
    cen = np.array([[sc.sphere[i][k] for k in range(3)] for i in range(n)], np.float64)
    rad = np.array([sc.sphere[i][3] for i in range(n)], np.float64)
    oc = o[None, :] - cen                       # (n, 3)
    b = d @ oc.T                                # (N, n)  d.(o-c)
    cc = (oc * oc).sum(1) - rad * rad           # (n,)
    disc = b * b - cc[None, :]
    root = np.where(disc >= 0, -b - np.sqrt(np.maximum(disc, 0)), np.inf)
    root = np.where(root > 0, root, np.inf)
    with np.errstate(divide="ignore", invalid="ignore"):
        tp = np.where(d[:, 1] < 0, (sc.groundY - o[1]) / d[:, 1], np.inf)
    allt = np.concatenate([root, tp[:, None]], axis=1)   # (N, n+1)
    obj = allt.argmin(1)
    th = allt.min(1)
    miss = ~np.isfinite(th) | (th >= sc.tMax)
    p = o[None, :] + d * np.where(miss, 0, th)[:, None]
    nrm = np.zeros_like(p)
    isp = (obj < n) & ~miss
    nrm[isp] = (p[isp] - cen[obj[isp]]) / rad[obj[isp]][:, None]
    nrm[(obj == n) & ~miss] = [0.0, 1.0, 0.0]
    # closest approach of the segment [0, t_end] to every sphere it does not hit
    tend = np.where(miss, sc.tMax, th)
    tc = np.clip(-b, 0, tend[:, None])
    gap = np.linalg.norm(oc[None, :, :] + tc[..., None] * d[:, None, :], axis=-1) - rad[None, :]
    hit_sphere = np.zeros_like(gap, bool)
    hit_sphere[isp, obj[isp]] = True
    bad = ((gap < SIL) & ~hit_sphere).any(1)
    inc = np.abs((nrm * d).sum(1))
    bad |= ~miss & (inc < GRAZE)
    # distance of the hit point to every other surface
    ds = np.linalg.norm(p[:, None, :] - cen[None], axis=-1) - rad[None, :]
    dsurf = np.concatenate([ds, (p[:, 1] - sc.groundY)[:, None]], axis=1)
    other = np.ones_like(dsurf, bool)
    other[np.arange(len(obj)), obj] = False
    bad |= ~miss & ((np.abs(dsurf) < CREASE) & other).any(1)
    return th, np.where(miss, -1, obj), nrm, miss, bad


def to_chan(c):
    return np.clip(np.rint(255.0 * np.nan_to_num(c, nan=0.0)), 0, 255).astype(np.uint32)


def shade_fp64(sc, th, obj, nrm, miss):
    """the scene's shading law: Lambert + ambient, ExpFog, fog blend incl.
    alpha, gamma per lane, ToPixel4 packing"""
    L = np.array([sc.light[k] for k in range(3)])
    diff = np.maximum(nrm @ L, 0.0)
    lit = sc.ambient + (1 - sc.ambient) * diff
    fog = np.where(miss, 1.0, 1 - np.exp(-sc.fogDensity * np.where(miss, 0, th) ** 3))
    alb = np.array([[sc.albedo[i][k] for k in range(3)] for i in range(sc.nSpheres + 1)])
    base = alb[np.maximum(obj, 0)] * lit[:, None]
    fc = np.array([sc.fogColor[k] for k in range(3)])
    col = np.concatenate([base + fog[:, None] * (fc - base), fog[:, None]], axis=1)
    with np.errstate(divide="ignore"):
        g = np.where(col > 0, np.exp(sc.gamma * np.log(np.maximum(col, 1e-300))), 0.0)
    ch = to_chan(g)
    return ch[:, 0] | (ch[:, 1] << 8) | (ch[:, 2] << 16) | (ch[:, 3] << 24)


def one_sphere_scene() -> SdfScene:
    sc = wl.sdf_scene(0, 8, 1024)
    sc.nSpheres = 1
    sc.sphere[0][0], sc.sphere[0][1], sc.sphere[0][2], sc.sphere[0][3] = 0.5, -1.0, 0.25, 2.0
    for k, a in enumerate([0.2, 0.55, 0.9]):
        sc.albedo[0][k] = a
    sc.albedo[1][0], sc.albedo[1][1], sc.albedo[1][2] = 0.45, 0.5, 0.55
    return sc


CASES = [("one sphere + plane", one_sphere_scene, 0.0),
         ("one sphere + plane", one_sphere_scene, 2.7),
         ("8 spheres", lambda: wl.sdf_scene(0, 8, 1024), 0.4),
         ("32 spheres", lambda: wl.sdf_scene(0, 32, 1024), 1.1)]


def render(orc, sc, W, H, t, prec):
    img = np.zeros(W * H, np.uint32)
    st = np.zeros(W * H, np.uint16)
    orc.orc_sdf_frame(P(img), W, H, C.byref(sc), t, prec, P(st), 8)
    return img, st


def chan_err(a, b):
    ca = a.view(np.uint8).reshape(-1, 4).astype(int)
    cb = b.view(np.uint8).reshape(-1, 4).astype(int)
    return np.abs(ca - cb).max(1)


@pytest.mark.parametrize("name,make,t", CASES)
@pytest.mark.parametrize("prec", [64, 32])
def test_sdf_render_matches_analytic_ray_cast(orc, name, make, t, prec):
    sc = make()
    W, H = 192, 108
    img, st = render(orc, sc, W, H, t, prec)
    th, obj, nrm, miss, bad = analytic(orc, sc, W, H, t)
    bad |= st >= sc.maxSteps
    want = shade_fp64(sc, th, obj, nrm, miss)
    ok = ~bad
    err = chan_err(img, want)
    print(name, t, prec, "hits", int((~miss).sum()), "excluded", float(bad.mean()),
          "max", int(err[ok].max()), "mean", float(err[ok].mean()))
    assert bad.mean() < 0.05, "too many ill-conditioned pixels for the check to mean much"
    assert (~miss & ok).sum() > 0.3 * W * H
    # the miss pixels are exactly the fog colour with alpha 255
    fogpx = to_chan(np.array([sc.fogColor[k] for k in range(3)]) ** sc.gamma)
    fogpx = int(fogpx[0] | (fogpx[1] << 8) | (fogpx[2] << 16) | (255 << 24))
    assert (img[miss & ok] == fogpx).all()
    # every other pixel within 1/255 of the analytic render (FP32: 2/255)
    assert err[ok].max() <= (1 if prec == 64 else 2)


def test_sdf_wrong_law_is_caught(orc):
    """the check has teeth: the analytic render with the light reversed, or
    with the albedo's B and R swapped, differs by far more than 1/255"""
    sc = wl.sdf_scene(0, 8, 1024)
    W, H = 96, 54
    img, st = render(orc, sc, W, H, 0.4, 64)
    th, obj, nrm, miss, bad = analytic(orc, sc, W, H, 0.4)
    ok = ~bad & (st < sc.maxSteps) & ~miss
    flipped = wl.sdf_scene(0, 8, 1024)
    for k in range(3):
        flipped.light[k] = -sc.light[k]
    assert chan_err(img, shade_fp64(flipped, th, obj, nrm, miss))[ok].max() > 20
    swapped = wl.sdf_scene(0, 8, 1024)
    for i in range(9):
        swapped.albedo[i][0], swapped.albedo[i][2] = sc.albedo[i][2], sc.albedo[i][0]
    assert chan_err(img, shade_fp64(swapped, th, obj, nrm, miss))[ok].max() > 20


@pytest.mark.ref
def test_sdf_shading_matches_reference_helpers(orc, refl):
    """hits shaded through the reference's FP32 ExpFog, vLerp4, GammaAdj and
    ToPixel4 on the analytic geometry equal the oracle's FP64 pixels within
    1/255 (the reference's Cephes exp/log against exact ones)"""
    sc = wl.sdf_scene(0, 8, 1024)
    W, H = 160, 90
    t = 0.9
    img, st = render(orc, sc, W, H, t, 64)
    th, obj, nrm, miss, bad = analytic(orc, sc, W, H, t)
    idx = np.nonzero(~bad & (st < sc.maxSteps) & ~miss)[0][::7]
    assert len(idx) > 500
    L = np.array([sc.light[k] for k in range(3)], np.float32)
    f32 = np.float32
    a4, b4, o4 = (np.zeros(4, np.float32) for _ in range(3))
    c16 = np.zeros(16, np.float32)
    out4 = np.zeros(4, np.uint32)
    for i in idx:
        n = nrm[i].astype(np.float32)
        diff = max(f32(n[0] * L[0] + n[1] * L[1] + n[2] * L[2]), f32(0))
        lit = f32(sc.ambient) + (f32(1) - f32(sc.ambient)) * diff
        fog = refl.ref_st_expfog(float(th[i]), sc.fogDensity)
        for k in range(3):
            a4[k] = f32(sc.albedo[obj[i]][k]) * lit
            b4[k] = sc.fogColor[k]
        a4[3], b4[3] = 0.0, 1.0
        refl.ref_st_vlerp4(P(a4), P(b4), fog, P(o4))
        refl.ref_st_gamma(P(o4), sc.gamma, P(c16))
        c16[4:] = np.tile(c16[0:4], 3)
        refl.ref_st_topixel4(P(c16), P(out4))
        e = chan_err(np.array([out4[0]], np.uint32), img[i:i + 1])[0]
        assert e <= 1, (i, hex(int(out4[0])), hex(int(img[i])))
