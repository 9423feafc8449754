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
import torch.nn.functional as F

import diso_amd
from diso_amd import FixedExtractor, mesh, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"

MC_CORNERS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def _padded_inside(sdf, iso=0.0):
    return F.pad(sdf, (1, 1, 1, 1, 1, 1), value=iso + 1.0) >= iso


def _crossings(ins):
    return sum(int((ins.narrow(a, 0, ins.shape[a] - 1) != ins.narrow(a, 1, ins.shape[a] - 1)).sum()) for a in range(3))


def _codes(ins, corners):
    X, Y, Z = (d - 1 for d in ins.shape)
    code = torch.zeros((X, Y, Z), dtype=torch.int64, device=ins.device)
    for i, (dx, dy, dz) in enumerate(corners):
        code |= ins[dx:dx + X, dy:dy + Y, dz:dz + Z].long() << i
    return code


def _good_codes(code, pc):
    """The face-neighbour rule (src/cudualmc.cu:815-839) on a whole code grid."""
    d = pc[code]
    out = code.clone()
    dims = code.shape
    for comp in range(3):
        for step in (1, -1):
            sel = d == (2 * comp + (1 if step > 0 else 0))
            if not bool(sel.any()):
                continue
            idx = sel.nonzero(as_tuple=True)
            n = [t.clone() for t in idx]
            n[comp] = n[comp] + step
            ok = (n[comp] >= 0) & (n[comp] < dims[comp])
            nb = torch.full_like(idx[0], 255)
            nb[ok] = pc[code[n[0][ok], n[1][ok], n[2][ok]]]
            flip = ok & (nb != 255)
            out[idx[0][flip], idx[1][flip], idx[2][flip]] ^= 255
    return out


def test_counts_at_256_match_independent_counts():
    print("case fullsize counts 256", flush=True)
    tables = {k: torch.tensor(v, device=DEV, dtype=torch.int64) for k, v in O.expand_tables().items()}
    sdf_np, def_np, _ = workloads.noisy_round_box(256, seed=0)
    sdf, deform = torch.tensor(sdf_np, device=DEV), torch.tensor(def_np, device=DEV)
    ins = _padded_inside(sdf)
    nx = _crossings(ins)
    mc_code = _codes(ins, MC_CORNERS)
    ntri = int(((tables["fmi"][1:] - tables["fmi"][:-1]) // 3)[mc_code].sum())
    dmc_code = _good_codes(_codes(ins, [(i & 1, (i >> 1) & 1, (i >> 2) & 1) for i in range(8)]), tables["pc"])
    npatch = int((tables["fpi"][1:] - tables["fpi"][:-1])[dmc_code].sum())
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
