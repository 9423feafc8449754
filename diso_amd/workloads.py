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


def noisy_round_box_torch(n, seed=0, amp=0.05, deform_amp=0.3, device="cuda", dtype=None):
    """noisy_round_box's distribution generated on the device with torch (for sizes where numpy
    is slow; the values differ from noisy_round_box's). Returns (sdf [N,N,N], deform [N,N,N,3])."""
    import torch

    dtype = dtype or torch.float32
    g = torch.Generator(device=device).manual_seed(seed)
    a = torch.linspace(-0.5, 0.5, n, dtype=torch.float64, device=device)
    p = torch.stack(torch.meshgrid(a, a, a, indexing="ij"), -1)
    q = p.abs() - 0.35 + 0.1
    f = q.clamp_min(0).norm(dim=-1) + q.max(-1)[0].clamp_max(0) - 0.1
    del p, q
    f += (torch.rand(f.shape, generator=g, dtype=torch.float64, device=device) * 2 - 1) * amp
    d = (torch.rand((n, n, n, 3), generator=g, dtype=torch.float64, device=device) * 2 - 1) * deform_amp
    return f.to(dtype).contiguous(), d.to(dtype).contiguous()


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


FAMILIES = ("sphere", "round_box", "blobs", "noise")


def _family_torch(kind, n, u, ud, device):
    """One float64 field of `kind` on an n^3 grid (torch, on `device`); `u` draws the shape
    parameters (CPU generator), `ud` the noise field (device generator)."""
    import torch

    a = torch.linspace(-0.5, 0.5, n, dtype=torch.float64, device=device)
    p = torch.stack(torch.meshgrid(a, a, a, indexing="ij"), -1)
    if kind == "sphere":
        return p.norm(dim=-1) - (0.2 + 0.25 * float(u(1)))
    if kind == "round_box":
        b = (0.2 + 0.2 * u(3)).to(device)
        q = p.abs() - b + 0.1
        return q.clamp_min(0).norm(dim=-1) + q.max(-1)[0].clamp_max(0) - 0.1
    if kind == "blobs":
        c = (u(8, 3) * 0.6 - 0.3).to(device)
        sg = (0.05 + 0.1 * u(8)).to(device)
        f = torch.zeros(p.shape[:3], dtype=torch.float64, device=device)
        for k in range(8):
            f += torch.exp(-((p - c[k]) ** 2).sum(-1) / (2 * sg[k] ** 2))
        return f - 0.5
    return ud(n, n, n) * 2 - 1


def uneven_batch_torch(count=64, seed=0, sizes=(32, 48, 64, 96, 128, 192, 256), deform_amp=0.3, device="cuda",
                       dtype=None):
    """The uneven batch workload, generated on the device with torch (numpy would take minutes
    for the 256^3 members): `count` grids with N drawn uniformly from `sizes`, family uniform over
    sphere (radius U(0.2,0.45)), rounded box (half-extents U(0.2,0.4), radius 0.1), 8 Gaussian
    blobs (centres U(-0.3,0.3), sigma U(0.05,0.15)) minus 0.5, and i.i.d. U(-1,1); deform
    U(-deform_amp, deform_amp); algorithms alternate MC, DMC. Values in float64, then cast.
    Returns [(sdf, deform, algo, family, n)]."""
    import torch

    dtype = dtype or torch.float32
    g = torch.Generator(device="cpu").manual_seed(seed)  # sizes, families and shape parameters
    gd = torch.Generator(device=device).manual_seed(seed + 1)  # the large noise and deform fields
    u = lambda *shape: torch.rand(*shape, generator=g, dtype=torch.float64)
    ud = lambda *shape: torch.rand(*shape, generator=gd, dtype=torch.float64, device=device)
    out = []
    for i in range(count):
        n = int(sizes[int(torch.randint(len(sizes), (1,), generator=g))])
        kind = FAMILIES[int(torch.randint(len(FAMILIES), (1,), generator=g))]
        f = _family_torch(kind, n, u, ud, device)
        d = (ud(n, n, n, 3) * 2 - 1) * deform_amp
        out.append((f.to(dtype).contiguous(), d.to(dtype).contiguous(), "mc" if i % 2 == 0 else "dmc", kind, n))
        del f
    return out


def request_stream(count, rate, seed=0, pool=48, lo=32, hi=384, deform_amp=0.3, backward_frac=0.3, device="cuda",
                   dtype=None):
    """The serving workload: `count` requests arriving as a Poisson process of `rate` per second.
    Each request takes a grid from a pool of `pool` grids with N log-uniform over [lo, hi] and
    the family uniform as in the uneven batch (with deform U(-deform_amp, deform_amp)); the
    algorithm is MC or DMC with probability 1/2 and a backward (upstream N(0,1)) is requested
    with probability `backward_frac`. Returns (pool [(sdf, deform, family, n)], requests
    [(arrival_s, pool_index, algo, backward)])."""
    import math

    import torch

    dtype = dtype or torch.float32
    g = torch.Generator(device="cpu").manual_seed(seed)
    gd = torch.Generator(device=device).manual_seed(seed + 1)
    u = lambda *shape: torch.rand(*shape, generator=g, dtype=torch.float64)
    ud = lambda *shape: torch.rand(*shape, generator=gd, dtype=torch.float64, device=device)
    grids = []
    for _ in range(pool):
        n = int(round(math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * float(u(1)))))
        kind = FAMILIES[int(torch.randint(len(FAMILIES), (1,), generator=g))]
        f = _family_torch(kind, n, u, ud, device)
        d = (ud(n, n, n, 3) * 2 - 1) * deform_amp
        grids.append((f.to(dtype).contiguous(), d.to(dtype).contiguous(), kind, n))
        del f
    gaps = -torch.log1p(-u(count)) / rate
    arrivals = torch.cumsum(gaps, 0).tolist()
    pick = torch.randint(pool, (count,), generator=g).tolist()
    algo = ["mc" if float(x) < 0.5 else "dmc" for x in u(count)]
    bwd = [float(x) < backward_frac for x in u(count)]
    return grids, list(zip(arrivals, pick, algo, bwd))
