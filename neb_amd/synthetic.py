"""Synthetic inputs of the benchmark configurations (host-side NumPy).

These only build inputs (points, quadrature weights, target fields); every
contraction on them runs on the GPU.  Conventions follow the reference's
drivers: the periodic grid of experiments/2d/_backend/periodic_core/
vortex_merger/canonical_gauge/scripts/evaluate_large_comparison.py:26-28
(meshgrid indexing='xy'), the analytic-curl targets of
experiments/metaball/large_eddy.py:8-30 (here with Gaussian-weighted random
wavevectors |k| <= kmax, as the 3D projection workload specifies).
"""
from __future__ import annotations

import math

import numpy as np


def periodic_grid(n: int) -> np.ndarray:
    """n x n grid on [0, 2pi)^2, x fastest (meshgrid indexing='xy')."""
    axis = np.arange(n) * 2 * np.pi / n
    x, y = np.meshgrid(axis, axis, indexing="xy")
    return np.column_stack((x.ravel(), y.ravel()))


def periodic_uniform(n: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).uniform(0, 2 * np.pi, (n, 2))


def blob_domain(count: int, seed: int, blobs: int = 6, sigma: float = 0.35, level: float = 0.5,
                batch: int = 1 << 20):
    """``count`` points drawn uniformly from {x in [-1,1]^3 :
    sum_b exp(-|x - c_b|^2 / (2 sigma^2)) > level}, centres c_b uniform in
    [-0.6, 0.6]^3, by rejection sampling from one seeded stream.
    Returns (points [count, 3], centres [blobs, 3], volume estimate)."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-0.6, 0.6, (blobs, 3))
    out, have, drawn = [], 0, 0
    while have < count:
        p = rng.uniform(-1.0, 1.0, (batch, 3))
        s = np.zeros(batch)
        for c in centres:
            s += np.exp(-np.sum((p - c) ** 2, 1) / (2 * sigma * sigma))
        keep = p[s > level]
        out.append(keep)
        have += len(keep)
        drawn += batch
    pts = np.concatenate(out)[:count]
    return pts, centres, 8.0 * have / drawn


def curl_targets(points: np.ndarray, count: int, seed: int, kmax: int = 6, waves: int = 16) -> np.ndarray:
    """Divergence-free targets u = curl A, A = sum_w a_w sin(k_w . x + phi_w)
    with integer wavevectors (times pi/2) of |k| <= kmax and Gaussian
    amplitude decay; returns [n, 3, count] (analytic curl)."""
    rng = np.random.default_rng(seed)
    n = len(points)
    out = np.zeros((n, 3, count))
    for t in range(count):
        ks = []
        while len(ks) < waves:
            k = rng.integers(-kmax, kmax + 1, 3)
            if 0 < k @ k <= kmax * kmax:
                ks.append(k)
        ks = np.array(ks, dtype=np.float64) * (np.pi / 2)
        amp = rng.normal(size=(waves, 3)) * np.exp(-np.sum(ks * ks, 1) / (2 * (np.pi * kmax / 4) ** 2))[:, None]
        ph = rng.uniform(0, 2 * np.pi, waves)
        c = np.cos(points @ ks.T + ph)
        J = np.einsum("nw,wc,wj->ncj", c, amp, ks)
        out[:, :, t] = np.stack((J[:, 2, 1] - J[:, 1, 2], J[:, 0, 2] - J[:, 2, 0], J[:, 1, 0] - J[:, 0, 1]), 1)
    return out


def grid_weight(n: int) -> float:
    return (2 * math.pi / n) ** 2
