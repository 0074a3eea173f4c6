"""Deterministic synthetic YUV420P8 clips for the motion-path workloads.

Content model (integer arithmetic only, so the bytes are reproducible):
  * background: uniform u8 noise from a seeded PCG64 stream, blurred by a separable
    integer Gaussian (sigma 2, radius 6, round-half-up), on a periodic 2x canvas;
  * global motion in half-pel steps: frame t samples the 2x canvas at offset
    2*v*t and decimates with a 2x2 box, so half-pel shifts are exact;
  * textured rectangles (own seeds, own velocities) painted in order, optionally
    rotated by a per-frame angle with nearest-neighbour sampling in Q16;
  * chroma: an independent blurred texture at half resolution moved by floor(v/2);
  * luma is mapped to 16..127 (16 + 7*t/16): the scene detector's luma weight
    wraps to a tiny value once the two frames' block means sum past 255, so content
    centred on mid-gray would be classified as a scene change in every frame;
  * scene cuts: every `cut_every` frames the seed advances and everything is redrawn.
"""
from __future__ import annotations

import math

import numpy as np

_TEX = 1024  # texture period (1x pixels)


def _gauss_kernel(sigma: float = 2.0, radius: int = 6) -> np.ndarray:
    k = [round(math.exp(-(i * i) / (2 * sigma * sigma)) * 4096) for i in range(-radius, radius + 1)]
    return np.array(k, np.int64)


_K = _gauss_kernel()
_KS = int(_K.sum())


def _blur_periodic(a: np.ndarray) -> np.ndarray:
    r = len(_K) // 2
    x = a.astype(np.int64)
    acc = np.zeros_like(x)
    for i, w in enumerate(_K):
        acc += w * np.roll(x, i - r, axis=1)
    x = (acc + _KS // 2) // _KS
    acc = np.zeros_like(x)
    for i, w in enumerate(_K):
        acc += w * np.roll(x, i - r, axis=0)
    return ((acc + _KS // 2) // _KS).clip(0, 255).astype(np.uint8)


def texture(seed: int, h: int, w: int) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return _blur_periodic(rng.integers(0, 256, size=(h, w), dtype=np.uint8))


def _sample_box2(canvas: np.ndarray, oy2: int, ox2: int, h: int, w: int) -> np.ndarray:
    """Frame of h x w from a periodic 2x canvas at 2x offset (oy2, ox2), 2x2 box."""
    H, W = canvas.shape
    ys = (oy2 + 2 * np.arange(h)) % H
    xs = (ox2 + 2 * np.arange(w)) % W
    ys1 = (ys + 1) % H
    xs1 = (xs + 1) % W
    c = canvas.astype(np.uint16)
    s = c[ys][:, xs] + c[ys][:, xs1] + c[ys1][:, xs] + c[ys1][:, xs1]
    return ((s + 2) >> 2).astype(np.uint8)


def _sample_int(tex: np.ndarray, oy: int, ox: int, h: int, w: int) -> np.ndarray:
    H, W = tex.shape
    return tex[(oy + np.arange(h)) % H][:, (ox + np.arange(w)) % W]


class _Scene:
    def __init__(self, seed: int, width: int, height: int, vmax: float, nobj: tuple, rot_deg: float,
                 half_pel: bool):
        rng = np.random.Generator(np.random.PCG64(seed + 7))
        self.canvas = texture(seed, 2 * _TEX, 2 * _TEX)
        self.ctex = [texture(seed + 1, _TEX, _TEX), texture(seed + 2, _TEX, _TEX)]
        step = 0.5 if half_pel else 1.0
        n = int(2 * vmax / step)
        self.v = (rng.integers(-n // 2, n // 2 + 1) * step, rng.integers(-n // 2, n // 2 + 1) * step)
        self.objs = []
        k = int(rng.integers(nobj[0], nobj[1] + 1))
        for j in range(k):
            ow = int(rng.integers(max(8, width // 16), max(9, width // 4)))
            oh = int(rng.integers(max(8, height // 16), max(9, height // 4)))
            o = {
                "w": ow, "h": oh,
                "x": float(rng.integers(0, width)), "y": float(rng.integers(0, height)),
                "vx": float(rng.integers(-int(vmax), int(vmax) + 1)),
                "vy": float(rng.integers(-int(vmax), int(vmax) + 1)),
                "rot": float(rng.uniform(-rot_deg, rot_deg)) if rot_deg > 0 else 0.0,
                "tex": texture(seed + 100 + j, 256, 256),
                "cu": int(rng.integers(40, 216)), "cv": int(rng.integers(40, 216)),
            }
            self.objs.append(o)


def _paint(y: np.ndarray, u: np.ndarray, v: np.ndarray, o: dict, t: int) -> None:
    H, W = y.shape
    cx = o["x"] + o["vx"] * t
    cy = o["y"] + o["vy"] * t
    ang = math.radians(o["rot"] * t)
    c16 = int(round(math.cos(ang) * 65536))
    s16 = int(round(math.sin(ang) * 65536))
    half = int(math.ceil(math.hypot(o["w"], o["h"]) / 2)) + 1
    icx, icy = int(math.floor(cx)), int(math.floor(cy))
    x0, x1 = max(0, icx - half), min(W, icx + half)
    y0, y1 = max(0, icy - half), min(H, icy + half)
    if x0 >= x1 or y0 >= y1:
        return
    yy, xx = np.mgrid[y0:y1, x0:x1]
    dx = (xx - icx).astype(np.int64)
    dy = (yy - icy).astype(np.int64)
    # inverse rotation in Q16, nearest neighbour
    lx = (c16 * dx + s16 * dy + 32768) >> 16
    ly = (-s16 * dx + c16 * dy + 32768) >> 16
    inside = (np.abs(lx) * 2 < o["w"]) & (np.abs(ly) * 2 < o["h"])
    if not inside.any():
        return
    tex = o["tex"]
    ty = (ly + o["h"] // 2) % tex.shape[0]
    tx = (lx + o["w"] // 2) % tex.shape[1]
    y[y0:y1, x0:x1][inside] = tex[ty[inside], tx[inside]]
    # chroma: the object's flat colour at half resolution
    cy0, cy1, cx0, cx1 = y0 // 2, (y1 + 1) // 2, x0 // 2, (x1 + 1) // 2
    cin = inside[:: 2, :: 2][: cy1 - cy0, : cx1 - cx0]
    u[cy0:cy0 + cin.shape[0], cx0:cx0 + cin.shape[1]][cin] = o["cu"]
    v[cy0:cy0 + cin.shape[0], cx0:cx0 + cin.shape[1]][cin] = o["cv"]


def clip(width: int, height: int, frames: int, seed: int, vmax: float = 6.0, nobj=(2, 2), rot_deg: float = 0.0,
         cut_every: int = 0, half_pel: bool = True):
    """Returns a list of (Y, U, V) uint8 planes (H x W, H/2 x W/2, H/2 x W/2)."""
    out = []
    scene = None
    cut = -1
    for t in range(frames):
        c = t // cut_every if cut_every else 0
        if c != cut:
            cut = c
            scene = _Scene(seed + c, width, height, vmax, nobj, rot_deg, half_pel)
            t0 = t
        lt = t - t0
        vx, vy = scene.v
        ox2, oy2 = int(round(2 * vx * lt)), int(round(2 * vy * lt))
        Y = _sample_box2(scene.canvas, oy2, ox2, height, width)
        cox, coy = int(math.floor(vx * lt / 2)), int(math.floor(vy * lt / 2))
        U = _sample_int(scene.ctex[0], coy, cox, height // 2, width // 2).copy()
        V = _sample_int(scene.ctex[1], coy, cox, height // 2, width // 2).copy()
        for o in scene.objs:
            _paint(Y, U, V, o, lt)
        Y = (16 + ((Y.astype(np.uint16) * 7) >> 4)).astype(np.uint8)
        out.append((Y, U, V))
    return out


# Workloads of the benchmark configurations (sizes and motion model per config).
CONFIGS = {
    1: dict(width=640, height=360, frames=32, seed=1000, vmax=6.0, nobj=(2, 2), rot_deg=0.0, cut_every=0),
    2: dict(width=1920, height=1080, frames=120, seed=2000, vmax=12.0, nobj=(6, 12), rot_deg=2.0, cut_every=60),
    3: dict(width=3840, height=2160, frames=48, seed=3000, vmax=24.0, nobj=(6, 12), rot_deg=2.0, cut_every=60),
}


def config_clip(cfg: int, frames: int | None = None):
    kw = dict(CONFIGS[cfg])
    if frames is not None:
        kw["frames"] = frames
    return clip(**kw)
