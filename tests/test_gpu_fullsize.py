"""Full-size checks (needs a GPU): at the benchmark sizes the oracle is too slow to run, so the HIP
path is tied to size-independent properties instead.

- 256^3 benchmark workload (rounded box + noise, deform): vertex, face, quad and patch counts equal
  independent counts computed with torch from the case tables (crossing edges; triangles per
  Bourke code; patches per DMC code after the face-neighbour rule); every face index is in range.
- 256^3 sphere: the MC and DMC meshes are closed, manifold and consistently oriented, with
  Euler characteristic 2 and the sphere's volume.
- 512^3: the recompute-classification backward equals the stored one bit for bit, and a
  fixed-capacity extraction equals the plain one.
"""
import numpy as np
import pytest
import torch

import diso_amd
from diso_amd import FixedExtractor, mesh, workloads
from oracle import oracle as O
from tests.torch_counts import expected_counts

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def test_counts_at_256_match_independent_counts():
    print("case fullsize counts 256", flush=True)
    tables = {k: torch.tensor(v, device=DEV, dtype=torch.int64) for k, v in O.expand_tables().items()}
    sdf_np, def_np, _ = workloads.noisy_round_box(256, seed=0)
    sdf, deform = torch.tensor(sdf_np, device=DEV), torch.tensor(def_np, device=DEV)
    nx, ntri, npatch = expected_counts(sdf, tables)
    v, f = diso_amd.DiffMC()(sdf, deform)
    assert v.shape[0] == nx and f.shape[0] == ntri
    assert int(f.min()) >= 0 and int(f.max()) < nx
    v, q = diso_amd.DiffDMC()(sdf, deform, return_quads=True)
    assert q.shape[0] == nx and v.shape[0] == npatch
    assert int(q.min()) >= 0 and int(q.max()) < npatch
    assert bool(torch.isfinite(v).all())


def test_sphere_meshes_at_256_are_closed_and_oriented():
    print("case fullsize sphere 256", flush=True)
    sdf = torch.tensor(workloads.sphere(256, radius=0.4).astype(np.float32), device=DEV)
    vol = 4.0 / 3.0 * np.pi * 0.4 ** 3
    for mod, kw in ((diso_amd.DiffMC(), {}), (diso_amd.DiffDMC(), {"return_quads": True})):
        v, f = mod(sdf, **kw)
        fv = f.cpu().numpy()
        r = mesh.check(fv, n_verts=v.shape[0])
        assert r["oriented"] and r["watertight"] and r["euler"] == 2 and r["components"] == 1, r
        # normalised vertices span [0, 1]^3 over the grid's [-0.5, 0.5]^3
        assert abs(abs(mesh.signed_volume(v.cpu().numpy().astype(np.float64), fv)) - vol) < 0.01 * vol


def test_512_recompute_and_fixed_capacity_match_stored():
    print("case fullsize 512", flush=True)
    n = 512
    g = torch.Generator(device=DEV).manual_seed(3)
    a = torch.linspace(-0.5, 0.5, n, device=DEV)
    x, y, z = torch.meshgrid(a, a, a, indexing="ij")
    sdf0 = (torch.sqrt(x * x + y * y + z * z) - 0.4 + 0.01 * torch.randn((n, n, n), generator=g, device=DEV)).contiguous()
    del x, y, z
    def0 = ((torch.rand((n, n, n, 3), generator=g, device=DEV) * 2 - 1) * 0.1).contiguous()
    res = []
    for recompute in (False, True):
        sdf, deform = sdf0.clone().requires_grad_(True), def0.clone().requires_grad_(True)
        v, q = diso_amd.DiffDMC(recompute_backward=recompute)(sdf, deform, return_quads=True)
        up = torch.randn(v.shape, generator=torch.Generator(device=DEV).manual_seed(4), device=DEV)
        torch.autograd.backward(v, grad_tensors=up)
        res.append((v.detach(), q, sdf.grad, deform.grad, up))
        del sdf, deform
    (v0, q0, gs0, gd0, up), (v1, q1, gs1, gd1, _) = res
    assert torch.equal(v0, v1) and torch.equal(q0, q1)
    assert torch.equal(gs0, gs1) and torch.equal(gd0, gd1)
    ex = FixedExtractor("dmc", torch.float32, vert_capacity=v0.shape[0] + 1000, face_capacity=q0.shape[0] + 1000,
                        recompute_backward=True)
    sdf, deform = sdf0.clone().requires_grad_(True), def0.clone().requires_grad_(True)
    v2, q2, c = ex(sdf, deform)
    assert ex.check(c) == (v0.shape[0], q0.shape[0])
    assert torch.equal(v2[: v0.shape[0]], v0) and torch.equal(q2[: q0.shape[0]], q0)
    upp = torch.zeros_like(v2)
    upp[: up.shape[0]] = up
    torch.autograd.backward(v2, grad_tensors=upp)
    assert torch.equal(sdf.grad, gs0) and torch.equal(deform.grad, gd0)
