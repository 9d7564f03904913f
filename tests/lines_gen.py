"""Seeded synthetic segment generators shared by tests and the benchmark.

Conventions: numpy ``default_rng(seed)``, drawn in float64 and cast to the
target dtype; rows ``[x0, y0, x1, y1, w]`` (x = column); random-angle segments
start uniformly in the frame, so end points may leave it.
"""
from __future__ import annotations

import numpy as np


def random_segments(rng, n, H, W, lmin, lmax, wmin, wmax, pad=0.0):
    x0 = rng.uniform(-pad, W + pad, n)
    y0 = rng.uniform(-pad, H + pad, n)
    th = rng.uniform(0.0, 2.0 * np.pi, n)
    L = rng.uniform(lmin, lmax, n)
    w = rng.uniform(wmin, wmax, n)
    return np.stack([x0, y0, x0 + L * np.cos(th), y0 + L * np.sin(th), w], 1)


def config1(dtype=np.float32):
    """One 512x512 frame, 1000 segments with end points uniform in the frame,
    widths uniform in [1, 4], seed 0."""
    rng = np.random.default_rng(0)
    pts = rng.uniform(0, 512, (1000, 4))
    w = rng.uniform(1, 4, (1000, 1))
    return np.concatenate([pts, w], 1).astype(dtype)


def config2(frames=64, per_frame=2000, H=2048, W=2048, dtype=np.float32, seed=1):
    """Frames of HxW, per_frame segments each, lengths U[20, 400], random
    angles, widths U[2, 6]; idxs = repeat(arange(frames), per_frame)."""
    rng = np.random.default_rng(seed)
    lines = random_segments(rng, frames * per_frame, H, W, 20.0, 400.0, 2.0, 6.0)
    idxs = np.repeat(np.arange(frames), per_frame)
    it = np.int64 if np.dtype(dtype) == np.float64 else np.int32
    return lines.astype(dtype), idxs.astype(it)


def config3(frames=256, H=1024, W=1024, dtype=np.float32, seed=2):
    """Frames of HxW with per-frame counts log-uniform in [100, 20000], C2
    segment distributions, profile per segment triangular or gaussian (p=1/2)."""
    rng = np.random.default_rng(seed)
    counts = np.round(np.exp(rng.uniform(np.log(100), np.log(20000), frames))).astype(np.int64)
    lines = random_segments(rng, int(counts.sum()), H, W, 20.0, 400.0, 2.0, 6.0)
    idxs = np.repeat(np.arange(frames), counts)
    profile = rng.integers(0, 2, int(counts.sum()))  # 0 = triangular, 1 = gaussian
    it = np.int64 if np.dtype(dtype) == np.float64 else np.int32
    return lines.astype(dtype), idxs.astype(it), profile, counts


def segments3(rng, n, shape, length=10.0, wmin=2.0, wmax=2.0, pad=0.0):
    """3D lines as rows ``[x0, y0, z0, x1, y1, z1, w]`` in a volume ``shape``
    (..., D, H, W): start uniform in the volume (x along W), random direction,
    length U[0, length] (the generator of the reference's own draw_lines tests
    for ndim = 3), widths U[wmin, wmax]."""
    D, H, W = shape[-3:]
    lo, hi = -pad, np.array([W, H, D], np.float64) + pad
    p0 = lo + (hi - lo) * rng.random((n, 3))
    vec = rng.normal(size=(n, 3))
    lengths = length * rng.random(n)
    p1 = p0 + vec * (lengths / np.sqrt(np.sum(vec ** 2, axis=-1)))[:, None]
    w = rng.uniform(wmin, wmax, (n, 1))
    return np.concatenate([p0, p1, w], 1)
