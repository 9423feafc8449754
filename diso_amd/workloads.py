"""Synthetic SDF grids for the benchmark and parity workloads.

Conventions: sample points p lie on a uniform grid over [-0.5, 0.5]^3 built with
linspace(-0.5, 0.5, N) per axis and "ij" indexing (the same layout the reference example uses,
test/example.py:33-42); values are computed in float64 and cast at the end; randomness comes from
numpy.random.default_rng(seed), drawn in the order sdf noise -> deform -> upstream gradient.
"""
import numpy as np


def points(n):
    a = np.linspace(-0.5, 0.5, n)
    return np.stack(np.meshgrid(a, a, a, indexing="ij"), -1)


def sphere(n, radius=0.4, center=(0.0, 0.0, 0.0)):
    p = points(n) - np.asarray(center)
    return np.linalg.norm(p, axis=-1) - radius


def round_box(n, half_extent=0.35, radius=0.1):
    """Rounded box: q = |p| - b + r; |max(q,0)| + min(max(q), 0) - r."""
    b = np.broadcast_to(np.asarray(half_extent, dtype=np.float64), (3,))
    q = np.abs(points(n)) - b + radius
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inside = np.minimum(np.max(q, axis=-1), 0.0)
    return outside + inside - radius


def blobs(n, rng, count=8):
    """Sum of `count` Gaussian blobs (centres U(-0.3,0.3)^3, sigma U(0.05,0.15)) minus 0.5."""
    p = points(n)
    c = rng.uniform(-0.3, 0.3, size=(count, 3))
    s = rng.uniform(0.05, 0.15, size=count)
    f = np.zeros(p.shape[:3])
    for k in range(count):
        d2 = ((p - c[k]) ** 2).sum(-1)
        f += np.exp(-d2 / (2 * s[k] ** 2))
    return f - 0.5


def noisy_round_box(n, seed=0, amp=0.05, deform_amp=0.3, dtype=np.float32):
    """The N=256 forward+backward workload: rounded box (half-extent 0.35, radius 0.1) plus
    i.i.d. U(-amp, amp) noise, and deform U(-deform_amp, deform_amp) of shape [N,N,N,3].
    Returns (sdf, deform, rng) - the rng continues with the upstream-gradient draw."""
    rng = np.random.default_rng(seed)
    sdf = round_box(n) + rng.uniform(-amp, amp, size=(n, n, n))
    deform = rng.uniform(-deform_amp, deform_amp, size=(n, n, n, 3))
    return sdf.astype(dtype), deform.astype(dtype), rng


def upstream(rng, nverts, dtype=np.float32):
    """Upstream dL/dverts ~ N(0,1) of shape [V,3], drawn after V is known."""
    return rng.standard_normal((nverts, 3)).astype(dtype)


def family_grid(kind, n, rng):
    """One grid of the mixed-family batch: 'sphere', 'round_box', 'blobs' or 'noise'."""
    if kind == "sphere":
        return sphere(n, radius=rng.uniform(0.2, 0.45))
    if kind == "round_box":
        return round_box(n, half_extent=rng.uniform(0.2, 0.4, size=3), radius=0.1)
    if kind == "blobs":
        return blobs(n, rng)
    if kind == "noise":
        return rng.uniform(-1.0, 1.0, size=(n, n, n))
    raise ValueError(kind)
