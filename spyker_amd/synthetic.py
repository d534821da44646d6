"""Seeded synthetic inputs in the shapes the layer workloads are quoted on.

strokes(n, size, seed)  -- uint8 images of 3-6 anti-aliased line strokes
                           (width 2-3 px, intensity 0-255) on zero background
bars(n, size, seed)     -- float32 images in [0, 1]: sums of 8-20 oriented
                           Gaussian bars (sigma 1-4 px), normalised by the max
events(n, bins, size, seed)
                        -- uint8 DVS-like event frames [n, bins, 2, size, size]:
                           Bernoulli events at 2 % mean density, modulated by
                           3-5 Gaussian blobs moving 0.5-2 px per bin

Both draw from numpy's PCG64 with the given seed, so every run and every
machine sees the same images.
"""

import numpy as np


def _segment_distance(py, px, y0, x0, y1, x1):
    dy, dx = y1 - y0, x1 - x0
    ll = dy * dy + dx * dx
    t = np.where(ll > 0, ((py - y0) * dy + (px - x0) * dx) / np.maximum(ll, 1e-12), 0.0)
    t = np.clip(t, 0.0, 1.0)
    cy, cx = y0 + t * dy, x0 + t * dx
    return np.sqrt((py - cy) ** 2 + (px - cx) ** 2)


def strokes(n: int, size: int = 28, seed: int = 1) -> np.ndarray:
    """[n, size, size] uint8."""
    rng = np.random.Generator(np.random.PCG64(seed))
    yy, xx = np.mgrid[0:size, 0:size].astype(np.float64)
    py, px = yy + 0.5, xx + 0.5
    out = np.zeros((n, size, size), np.float64)
    for i in range(n):
        img = out[i]
        for _ in range(int(rng.integers(3, 7))):
            y0, x0, y1, x1 = rng.uniform(0, size, 4)
            width = rng.uniform(2.0, 3.0)
            level = float(rng.integers(0, 256))
            d = _segment_distance(py, px, y0, x0, y1, x1)
            cover = np.clip(width / 2 + 0.5 - d, 0.0, 1.0)
            np.maximum(img, level * cover, out=img)
    return np.rint(out).astype(np.uint8)


def bars(n: int, size: int = 128, seed: int = 3) -> np.ndarray:
    """[n, size, size] float32 in [0, 1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    yy, xx = np.mgrid[0:size, 0:size].astype(np.float32)
    out = np.zeros((n, size, size), np.float32)
    for i in range(n):
        img = out[i]
        for _ in range(int(rng.integers(8, 21))):
            cy, cx = rng.uniform(0, size, 2)
            ang = rng.uniform(0, np.pi)
            sigma = rng.uniform(1.0, 4.0)
            length = rng.uniform(0.1, 0.5) * size
            u = (xx - cx) * np.cos(ang) + (yy - cy) * np.sin(ang)
            v = -(xx - cx) * np.sin(ang) + (yy - cy) * np.cos(ang)
            along = np.clip(np.abs(u) - length / 2, 0, None)
            img += np.exp(-(v * v + along * along) / (2 * sigma * sigma)).astype(np.float32)
        m = img.max()
        if m > 0:
            img /= m
    return out


def events(n: int, bins: int = 20, size: int = 128, seed: int = 5, density: float = 0.02) -> np.ndarray:
    """[n, bins, 2, size, size] uint8 in {0, 1}: per bin and polarity, a pixel
    fires with probability density * (1 + sum of moving blobs), the
    modulation renormalised so the mean probability is `density`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    yy, xx = np.mgrid[0:size, 0:size].astype(np.float32)
    out = np.zeros((n, bins, 2, size, size), np.uint8)
    for i in range(n):
        k = int(rng.integers(3, 6))
        cy, cx = rng.uniform(0, size, k), rng.uniform(0, size, k)
        speed = rng.uniform(0.5, 2.0, k)
        ang = rng.uniform(0, 2 * np.pi, k)
        sigma = rng.uniform(4.0, 12.0, k)
        for b in range(bins):
            mod = np.ones((size, size), np.float32)
            for j in range(k):
                y = cy[j] + b * speed[j] * np.sin(ang[j])
                x = cx[j] + b * speed[j] * np.cos(ang[j])
                mod += 4.0 * np.exp(-((yy - y) ** 2 + (xx - x) ** 2) / (2 * sigma[j] ** 2))
            prob = density * mod / mod.mean()
            out[i, b] = rng.random((2, size, size)) < prob[None]
    return out
