"""Seeded synthetic inputs for the parity tests and the benchmark.

Piecewise-smooth frames (a linear-gradient background, random axis-aligned
rectangles with constant or gradient fill and 2-4 sinusoidal textures,
values clamped to [0, 1]) plus additive white Gaussian noise given in 8-bit
units, and a moving-object video for V-BM3D.  numpy's PCG64 makes every
sequence reproducible on any host.
"""
from __future__ import annotations

import numpy as np


def piecewise_smooth(h: int, w: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = rng.uniform(0.2, 0.8) + rng.uniform(-0.3, 0.3) * xx / w + rng.uniform(-0.3, 0.3) * yy / h
    for _ in range(int(rng.integers(8, 33))):
        rh = int(rng.integers(max(2, h // 20), max(3, h // 3)))
        rw = int(rng.integers(max(2, w // 20), max(3, w // 3)))
        y0 = int(rng.integers(0, max(1, h - rh)))
        x0 = int(rng.integers(0, max(1, w - rw)))
        if rng.random() < 0.5:
            img[y0:y0 + rh, x0:x0 + rw] = rng.uniform(0, 1)
        else:
            a, bx, by = rng.uniform(0, 1), rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5)
            img[y0:y0 + rh, x0:x0 + rw] = (a + bx * (xx[y0:y0 + rh, x0:x0 + rw] - x0) / rw
                                           + by * (yy[y0:y0 + rh, x0:x0 + rw] - y0) / rh)
    for _ in range(int(rng.integers(2, 5))):
        f = rng.uniform(0.02, 0.3)
        th = rng.uniform(0, np.pi)
        amp = rng.uniform(0.02, 0.15)
        ph = rng.uniform(0, 2 * np.pi)
        img += amp * np.sin(2 * np.pi * f * (np.cos(th) * xx + np.sin(th) * yy) + ph)
    return np.clip(img, 0.0, 1.0)


def awgn(img: np.ndarray, sigma8: float, seed: int) -> np.ndarray:
    """img in [0, 1]; sigma in 8-bit units; result float64, not clipped."""
    rng = np.random.default_rng(seed)
    return img + rng.standard_normal(img.shape) * (sigma8 / 255.0)


def gray_float(h: int, w: int, sigma8: float, seed: int) -> np.ndarray:
    clean = piecewise_smooth(h, w, seed)
    return awgn(clean, sigma8, seed + 7919).astype(np.float32), clean.astype(np.float32)


def rgb16(h: int, w: int, sigma8: float, seed: int):
    """RGB48 frame (3, h, w) uint16 with per-channel textures; AWGN sigma*257 LSB."""
    chans, clean = [], []
    for c in range(3):
        cl = piecewise_smooth(h, w, seed + 100000 * c)
        noisy = awgn(cl, sigma8, seed + 100000 * c + 7919)
        chans.append(np.clip(np.rint(noisy * 65535.0), 0, 65535).astype(np.uint16))
        clean.append(np.clip(np.rint(cl * 65535.0), 0, 65535).astype(np.uint16))
    return np.stack(chans), np.stack(clean)


def moving_video(nframes: int, h: int, w: int, sigma8: float, seed: int, planes: int = 3):
    """Textured objects under global translation plus independent per-object
    integer / sub-pixel motion; returns float32 (T, planes, h, w) noisy and clean."""
    rng = np.random.default_rng(seed)
    pad = 64
    bg = [piecewise_smooth(h + 2 * pad, w + 2 * pad, seed + 17 * p) for p in range(planes)]
    nobj = int(rng.integers(4, 9))
    objs = []
    for _ in range(nobj):
        oh, ow = int(rng.integers(h // 10, h // 4)), int(rng.integers(w // 10, w // 4))
        tex = [piecewise_smooth(oh, ow, int(rng.integers(1 << 30))) for _ in range(planes)]
        objs.append(dict(tex=tex, y=rng.uniform(0, h - oh), x=rng.uniform(0, w - ow),
                         vy=rng.uniform(-2, 2), vx=rng.uniform(-2, 2), h=oh, w=ow))
    gvy, gvx = rng.integers(-2, 3), rng.integers(-2, 3)
    clean = np.zeros((nframes, planes, h, w), np.float64)
    for t in range(nframes):
        oy = int(np.clip(pad + gvy * t, 0, 2 * pad))
        ox = int(np.clip(pad + gvx * t, 0, 2 * pad))
        for p in range(planes):
            clean[t, p] = bg[p][oy:oy + h, ox:ox + w]
        for o in objs:
            y = o["y"] + o["vy"] * t
            x = o["x"] + o["vx"] * t
            iy, ix = int(np.floor(y)), int(np.floor(x))
            fy, fx = y - iy, x - ix
            y0, x0 = max(iy, 0), max(ix, 0)
            y1, x1 = min(iy + o["h"] - 1, h), min(ix + o["w"] - 1, w)
            if y1 <= y0 or x1 <= x0:
                continue
            for p in range(planes):
                tx = o["tex"][p]
                a = tx[y0 - iy:y1 - iy, x0 - ix:x1 - ix]
                b = tx[y0 - iy + 1:y1 - iy + 1, x0 - ix:x1 - ix]
                c = tx[y0 - iy:y1 - iy, x0 - ix + 1:x1 - ix + 1]
                d = tx[y0 - iy + 1:y1 - iy + 1, x0 - ix + 1:x1 - ix + 1]
                clean[t, p, y0:y1, x0:x1] = ((1 - fy) * ((1 - fx) * a + fx * c) + fy * ((1 - fx) * b + fx * d))
    noisy = clean + rng.standard_normal(clean.shape) * (sigma8 / 255.0)
    return noisy.astype(np.float32), clean.astype(np.float32)
