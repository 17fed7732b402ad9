"""Seeded synthetic inputs of the benchmark configurations.

The reference renders from demo assets and GNU Rocket tracks; the benchmark
replaces both with seeded synthetic data of the same shapes:

  sdf        8 (or 32) spheres with centres U[-4,4]^3, radii U[0.3,1.2] and a
             ground plane, an orbiting camera, frames at t = k/60 s
  landscape  1024x1024 u8 heightmap of 6-octave value noise (persistence
             0.5), a 1024x1024 ARGB colormap through a 256-entry seeded
             gradient palette, a 256-entry fog ramp, a Lissajous camera path
  tunnel     a 512x512 ARGB value-noise texture sampled through per-8x8-block
             24:8 UV grids of a tunnel law, blur radius U{1..32} per frame
  batch      512 requests mixing the three in a 40/30/30 ratio at resolutions
             U{320x180, 640x360, 1280x720, 1920x1080, 2560x1440}, t U[0,60]

Random numbers come from splitmix64 (Steele, Lea, Flood 2014), counter
based so that every array is generated in one vectorised call.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import LandscapeCam, Request, SdfScene, CKD_FX_SDF, CKD_FX_TUNNEL, CKD_FX_VOXEL

_M64 = (1 << 64) - 1
_GOLDEN = 0x9E3779B97F4A7C15


def splitmix64(state: int, n: int) -> np.ndarray:
    """The first n outputs of splitmix64 started from `state`."""
    with np.errstate(over="ignore"):
        z = np.uint64(state & _M64) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(_GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


class Rng:
    """A splitmix64 stream; (seed, stream) pairs give independent sequences."""

    def __init__(self, seed: int, stream: int = 0):
        self.state = (seed + stream * 0x632BE59BD9B4E019) & _M64

    def u64(self, n: int) -> np.ndarray:
        out = splitmix64(self.state, n)
        self.state = (self.state + n * _GOLDEN) & _M64
        return out

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
        return lo + (hi - lo) * u

    def integers(self, n: int, lo: int, hi: int) -> np.ndarray:
        """uniform integers in [lo, hi] (inclusive)"""
        return lo + np.floor(self.uniform(n) * (hi - lo + 1)).astype(np.int64)


# ---------------------------------------------------------------- SDF scene

def sdf_scene(seed: int = 0, n_spheres: int = 8, max_steps: int = 64) -> SdfScene:
    """Spheres (centres U[-4,4]^3, radii U[0.3,1.2]) over a ground plane.
    Geometry and albedo come from separate streams, so the first 8 spheres
    of the 32-sphere scene are the 8-sphere scene."""
    sc = SdfScene()
    sc.nSpheres = n_spheres
    geo = Rng(seed, 1).uniform(4 * n_spheres).reshape(n_spheres, 4)
    alb = Rng(seed, 2).uniform(3 * n_spheres, 0.25, 0.95).reshape(n_spheres, 3)
    for i in range(n_spheres):
        sc.sphere[i][0] = -4.0 + 8.0 * geo[i, 0]
        sc.sphere[i][1] = -4.0 + 8.0 * geo[i, 1]
        sc.sphere[i][2] = -4.0 + 8.0 * geo[i, 2]
        sc.sphere[i][3] = 0.3 + 0.9 * geo[i, 3]
        for c in range(3):
            sc.albedo[i][c] = alb[i, c]
    sc.albedo[n_spheres][0], sc.albedo[n_spheres][1], sc.albedo[n_spheres][2] = 0.45, 0.5, 0.55
    sc.groundY = -4.5
    lx, ly, lz = -0.4, 0.8, -0.45
    ln = math.sqrt(lx * lx + ly * ly + lz * lz)
    sc.light[0], sc.light[1], sc.light[2] = lx / ln, ly / ln, lz / ln
    sc.fogColor[0], sc.fogColor[1], sc.fogColor[2] = 0.85, 0.72, 0.62
    sc.fogDensity = 2e-4
    sc.gamma = 0.8
    sc.ambient = 0.18
    sc.eps = 1e-3
    sc.tMax = 60.0
    sc.maxSteps = max_steps
    sc.camRadius = 12.0
    sc.camHeight = 3.0
    sc.camSpeed = 0.25
    sc.focal = 1.5
    return sc


def _f32(x: float) -> float:
    return float(np.float32(x))


def sdf_camera(sc: SdfScene, t: float) -> list[float]:
    """Float camera basis of frame time t: origin, forward*focal, right, up.
    Same double-precision expression order as the CPU renderer, libm sin/cos,
    rounded to float at the end."""
    t = _f32(t)
    a = float(np.float32(sc.camSpeed)) * t
    ox = float(np.float32(sc.camRadius)) * math.sin(a)
    oy = float(np.float32(sc.camHeight))
    oz = -float(np.float32(sc.camRadius)) * math.cos(a)
    fx, fy, fz = 0.0 - ox, -1.0 - oy, 0.0 - oz
    l = math.sqrt(fx * fx + fy * fy + fz * fz)
    fx, fy, fz = fx / l, fy / l, fz / l
    rx, ry, rz = fz, 0.0, -fx
    l = math.sqrt(rx * rx + rz * rz)
    rx, rz = rx / l, rz / l
    ux = fy * rz - fz * ry
    uy = fz * rx - fx * rz
    uz = fx * ry - fy * rx
    foc = float(np.float32(sc.focal))
    vals = [ox, oy, oz, fx * foc, fy * foc, fz * foc, rx, ry, rz, ux, uy, uz]
    return [_f32(v) for v in vals]


# ----------------------------------------------------------------- textures

def value_noise(size: int, octaves: int, persistence: float, base_cells: int,
                rng: Rng) -> np.ndarray:
    """Tileable value noise in [0,1]: smoothstep-interpolated lattice values,
    octave o has base_cells << o cells per side and weight persistence^o."""
    total = np.zeros((size, size))
    amp, norm = 1.0, 0.0
    for o in range(octaves):
        cells = base_cells << o
        lat = rng.uniform(cells * cells).reshape(cells, cells)
        x = np.arange(size) * (cells / size)
        i0 = np.floor(x).astype(np.int64)
        f = x - i0
        f = f * f * (3.0 - 2.0 * f)
        i1 = (i0 + 1) % cells
        i0 = i0 % cells
        top = lat[i0][:, i0] * (1 - f)[None, :] + lat[i0][:, i1] * f[None, :]
        bot = lat[i1][:, i0] * (1 - f)[None, :] + lat[i1][:, i1] * f[None, :]
        total += amp * (top * (1 - f)[:, None] + bot * f[:, None])
        norm += amp
        amp *= persistence
    return total / norm


def _to_u8(x: np.ndarray) -> np.ndarray:
    lo, hi = x.min(), x.max()
    return np.floor((x - lo) / max(hi - lo, 1e-12) * 255.999).astype(np.uint8)


def _argb(b, g, r, a) -> np.ndarray:
    return (np.asarray(b, np.uint32) | (np.asarray(g, np.uint32) << 8) |
            (np.asarray(r, np.uint32) << 16) | (np.asarray(a, np.uint32) << 24)).astype(np.uint32)


def gradient_palette(rng: Rng, stops: int = 6) -> np.ndarray:
    """256-entry ARGB palette, linear between `stops` random colours."""
    pos = np.linspace(0, 255, stops)
    cols = rng.uniform(stops * 3, 0, 255).reshape(stops, 3)
    idx = np.arange(256)
    ch = [np.floor(np.interp(idx, pos, cols[:, c])).astype(np.uint32) for c in range(3)]
    return _argb(ch[0], ch[1], ch[2], 255)


def fog_ramp(rng: Rng) -> np.ndarray:
    """256-entry fog LUT subtracted by the voxel casters: base + i*slope/256."""
    base = rng.integers(4, 0, 24)
    slope = rng.integers(4, 60, 220)
    i = np.arange(256)[:, None]
    v = np.minimum(255, base[None, :] + (i * slope[None, :]) // 256)
    return _argb(v[:, 0], v[:, 1], v[:, 2], v[:, 3])


@dataclass
class LandscapeMaps:
    height: np.ndarray  # (1024, 1024) u8
    color: np.ndarray   # (1024, 1024) u32 ARGB
    fog: np.ndarray     # (256,) u32 ARGB


def landscape_maps(seed: int = 0, size: int = 1024) -> LandscapeMaps:
    n = value_noise(size, 6, 0.5, 4, Rng(seed, 10))
    h = _to_u8(n)
    pal = gradient_palette(Rng(seed, 11))
    return LandscapeMaps(h, pal[h], fog_ramp(Rng(seed, 12)))


def tunnelscape_maps(seed: int = 0) -> LandscapeMaps:
    """2048^2 maps for the tunnelscape rows; heights stay <= 159."""
    n = value_noise(2048, 6, 0.5, 4, Rng(seed, 20))
    h = (_to_u8(n).astype(np.uint32) * 159 // 255).astype(np.uint8)
    pal = gradient_palette(Rng(seed, 21))
    return LandscapeMaps(h, pal[h], fog_ramp(Rng(seed, 22)))


@dataclass
class LissajousPath:
    ax: float
    ay: float
    px: float
    py: float
    av: float
    pv: float
    at: float
    pt: float


def landscape_path(seed: int = 0) -> LissajousPath:
    u = Rng(seed, 13).uniform(8)
    return LissajousPath(ax=0.05 + 0.1 * u[0], ay=0.05 + 0.1 * u[1], px=2 * math.pi * u[2],
                         py=2 * math.pi * u[3], av=0.2 + 0.3 * u[4], pv=2 * math.pi * u[5],
                         at=0.3 + 0.5 * u[6], pt=2 * math.pi * u[7])


def landscape_cam(path: LissajousPath, t: float) -> LandscapeCam:
    ang = 0.6 * math.sin(path.av * t + path.pv)
    tilt = 40.0 * math.sin(path.at * t + path.pt)
    return LandscapeCam(
        X1=_f32(512.0 + 400.0 * math.sin(2 * math.pi * path.ax * t + path.px)),
        Y1=_f32(512.0 + 400.0 * math.sin(2 * math.pi * path.ay * t + path.py)),
        viewCos=_f32(math.cos(ang)), viewSin=_f32(math.sin(ang)),
        mapTilt=90 + int(max(-90.0, min(90.0, tilt))))


def noise_texture(seed: int = 0, log2: int = 9) -> np.ndarray:
    size = 1 << log2
    ch = [_to_u8(value_noise(size, 4, 0.55, 8, Rng(seed, 30 + c))) for c in range(3)]
    return _argb(ch[0], ch[1], ch[2], 255).reshape(size, size)


def grid_shape(W: int, H: int) -> tuple[int, int]:
    """(rows, cols) of the per-8x8-block corner grid of a W x H frame."""
    return (H + 7) // 8 + 1, (W + 7) // 8 + 1


def tunnel_grid(W: int, H: int, t: float, tex_log2: int = 9) -> np.ndarray:
    """24:8 (U,V) corner grid of the tunnel law at time t, int32
    (rows, cols, 2).  U follows 1/r (depth), V the angle; the per-frame
    scroll is reduced modulo the texture so the values stay small."""
    gh, gw = grid_shape(W, H)
    size = float(1 << tex_log2) * 256.0
    gy, gx = np.mgrid[0:gh, 0:gw].astype(np.float64)
    s = H / 2.0
    dx = (8.0 * gx - W / 2.0) / s
    dy = (8.0 * gy - H / 2.0) / s
    r = np.maximum(np.hypot(dx, dy), 0.05)
    ang = np.arctan2(dy, dx)
    su = math.fmod(0.35 * t, 1.0)
    sv = math.fmod(0.05 * t, 1.0)
    U = size * (0.6 / r + su)
    V = size * (ang / (2 * math.pi) + sv)
    return np.stack([np.rint(U), np.rint(V)], axis=-1).astype(np.int32)


def tunnel_radii(seed: int, n: int) -> np.ndarray:
    return Rng(seed, 31).integers(n, 1, 32)


# ------------------------------------------------------------------- batch

RESOLUTIONS = [(320, 180), (640, 360), (1280, 720), (1920, 1080), (2560, 1440)]


@dataclass
class BatchItem:
    effect: int
    W: int
    H: int
    t: float
    radius: int = 0


def batch_requests(seed: int = 0, n: int = 512) -> list[BatchItem]:
    rng = Rng(seed, 40)
    u = rng.uniform(3 * n).reshape(n, 3)
    radii = tunnel_radii(seed + 1, n)
    out = []
    for i in range(n):
        eff = CKD_FX_SDF if u[i, 0] < 0.4 else (CKD_FX_VOXEL if u[i, 0] < 0.7 else CKD_FX_TUNNEL)
        W, H = RESOLUTIONS[min(4, int(u[i, 1] * 5))]
        out.append(BatchItem(eff, W, H, 60.0 * u[i, 2], int(radii[i])))
    return out


def make_request(item: BatchItem, dst_ptr: int, scene: SdfScene, path: LissajousPath,
                 grid_ptr: int | None) -> Request:
    r = Request()
    r.effect = item.effect
    r.W, r.H = item.W, item.H
    r.dst = dst_ptr
    if item.effect == CKD_FX_SDF:
        for k, v in enumerate(sdf_camera(scene, item.t)):
            r.cam[k] = v
    elif item.effect == CKD_FX_VOXEL:
        r.vcam = landscape_cam(path, item.t)
    else:
        r.grid = grid_ptr
        r.radius = item.radius
    return r


# ---- the reference's raymarched effects ----------------------------------

# One parameter set per effect: the demo's own Rocket track values
# (target/sync of the reference) at a row where the effect is on screen.
ST_DEMO = {
    "plasma": {"speed": 1.33, "hue": 24.0, "gamma": 2.0, "desaturation": 0.0},
    "nautilus": {"roll": 0.0, "hue": 72.0, "speed": 3.14, "desaturation": 0.63,
                 "blur": 32.2194},
    "spikey": {"speed": 12.0, "roll": 1.1, "specular": 22.0, "desaturation": -1.0, "hue": 66.0,
               "gamma": 0.3125, "distZOffs": 0.66, "closeXOffs": 0.314, "closeZOffs": 1.0,
               "closeZOffsScale": 2.0, "closeNormalGrain": 1.0, "closeScale": 1.0,
               "closeRim": 1.0},
    "spikey_close": {"speed": 10.0, "roll": 4.0, "specular": 8.0, "desaturation": 0.0,
                     "hue": 33.0, "gamma": 1.66, "distXOffs": 0.5, "distYOffs": -1.0,
                     "distZOffs": 0.8, "closeXOffs": 0.314, "closeZOffs": 0.999548,
                     "closeZOffsScale": 2.0, "closeNormalGrain": 0.999729, "closeScale": 0.88,
                     "closeRim": 1.0, "closeAspectMul": 1.0},
    "sinuses": {"specular": 30.0, "roll": -2.9997, "speed": -5.0, "offsX": 0.0,
                "gamma": 2.2422, "hue": 50.3438, "desaturation": -7.8517},
    "laura": {"speed": 3.14, "yaw": 0.5991, "pitch": 0.14, "roll": 0.01, "hue": 24.0,
              "saturate": 3.9746},
    "fdtunnel": {"boxy": 0.9999, "flowerScale": 0.2, "flowerFreq": 4.0, "flowerPhase": 0.4999,
                 "speed": 5.28, "roll": 1.0, "pitch": 1.4, "radius": 0.8375, "mulU": 1024.0,
                 "mulV": 64.0, "litTiles": 1.0, "litBlur": 3.0, "fog1": 0.0, "fog2": 64.0},
}


def st_frames(effect: str, n: int, t0: float = 10.0):
    """n frames of an effect (`spikey_close` is the spike's close-up) with the
    demo's parameters at t = t0 + k/60, as ctypes parameter structs"""
    from . import _lib
    cls = _lib.ST_PARAMS["spikey" if effect == "spikey_close" else effect]
    return [cls(time=t0 + k / 60.0, **ST_DEMO[effect]) for k in range(n)]


def st_blur_maps(resX: int = 1280, resY: int = 720) -> np.ndarray:
    """the close-up spike's two blur maps for an output of resX x resY: two
    smooth noise images of the FX map's size, back to back (uint32)"""
    mw, mh = resX // 2 + 4, resY // 2 + 4
    log2 = max(10, int(np.ceil(np.log2(max(mw, mh)))))
    return np.concatenate([np.ascontiguousarray(noise_texture(13 + k, log2)[:mh, :mw]).reshape(-1)
                           for k in range(2)])


# ---- the voxel ball and the torus twister ----------------------------------

def _height_map(seed: int, top: int) -> np.ndarray:
    t = noise_texture(seed, 10)
    return ((t & 0xff).astype(np.uint32) * top // 255).astype(np.uint8).reshape(-1)


def _argb_noise(seed: int, n: int) -> np.ndarray:
    return (splitmix64(seed, n) & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def ball_maps(W: int = 1280, H: int = 720) -> dict:
    """the voxel ball's inputs: five 1024^2 height maps (heights <= 160, so
    no projected height leaves a 1280-pixel row at the largest radius), two
    colour maps, three beam maps, an environment map (1024^2 ARGB), two
    W x H backgrounds and a W x H halo"""
    return {
        "heightMaps": np.concatenate([_height_map(20 + i, 160) for i in range(5)]),
        "colorMaps": np.concatenate([noise_texture(30 + i, 10).reshape(-1) for i in range(2)]),
        "beamMaps": np.concatenate([noise_texture(32 + i, 10).reshape(-1) for i in range(3)]),
        "envMap": np.ascontiguousarray(noise_texture(35, 10).reshape(-1)),
        "backgrounds": np.concatenate([_argb_noise(36 + i, W * H) for i in range(2)]),
        "halo": _argb_noise(38, W * H),
    }


def twister_maps(W: int = 1280, H: int = 720) -> dict:
    """the torus twister's inputs: a 1024^2 height map, a 1024^2 colour map
    and a W x H background"""
    return {"heightMap": _height_map(40, 255),
            "colorMap": np.ascontiguousarray(noise_texture(41, 10).reshape(-1)),
            "background": _argb_noise(42, W * H)}


# the demo's track values (target/sync of the reference) at rows where the
# effects are on screen
BALL_DEMO = [
    {"blur": 10.0, "radius": 1800.0, "rayLength": 512.0, "spikes": 0.0, "hasBeams": 1.0,
     "baseShapeIndex": 1.0, "speed": 3.14, "beamAttenuation": 32.0, "beamAlphaMin": 196.0,
     "beams1": 0.5, "beams2": 0.66, "beams3": 1.0, "lowBeams": 196.0},
    {"blur": 30.0, "radius": 1750.0, "rayLength": 512.0, "spikes": 200.0, "hasBeams": 0.0,
     "baseShapeIndex": 4.0, "speed": 4.0, "beamAttenuation": 32.0, "beamAlphaMin": 196.0,
     "rotateOffsX": 256.0, "rotateOffsY": -256.0, "lowBeams": 128.0},
]
TWISTER_DEMO = {"speed": 8.0, "shearSpeed": 0.19, "blur": 0.0}


def ball_frames(n: int, demo: int = 0, t0: float = 10.0):
    from . import _lib
    return [_lib.BallParams(time=t0 + k / 60.0, **BALL_DEMO[demo]) for k in range(n)]


def twister_frames(n: int, t0: float = 10.0):
    from . import _lib
    return [_lib.TwisterParams(time=t0 + k / 60.0, **TWISTER_DEMO) for k in range(n)]
