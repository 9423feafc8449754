"""The benchmark configurations' own inputs on the HIP path (needs a GPU), tied to the oracle.

- N=64 sphere ||p|| - 0.4, MC forward, no deform, normalised: bit-exact against the oracle, with
  the expected counts (V = 11,952, F = 23,900; DMC Vd = 11,954, Q = 11,952).
- N=32 fp64 rounded box + noise with deform: central finite differences (eps 1e-6) of
  L = sum(g * verts) through DiffMC / DiffDMC on the GPU against the GPU backward, on samples whose
  +-eps step changes no topology and whose incident crossing edges are not nearly flat.
- N=256 benchmark workload: two exact identities of the default backward in fp64 accumulation:
  sum_i sdf_i * dL/dsdf_i = 0 at iso 0 (t is invariant under scaling the sdf) and
  sum_i dL/ddeform_i = sum_v g_v / (dims - 1) per axis (endpoint weights sum to one, DMC means keep
  that), given that no crossing touches the padding.
- The uneven batch (64 grids, N in 32..256, four families, MC / DMC alternating, quads):
  one BatchExtractor call, forward and backward; grids with N <= 64 against the oracle; every
  grid against independent torch counts, index ranges, and the per-grid path bit for bit.
- The serving stream (N log-uniform 32..384, MC / DMC, 30 % with a backward): a 200-request slice
  through Server; every result equals the request run alone, and requests with N <= 64 equal the
  oracle.
"""
import threading

import numpy as np
import pytest
import torch

import diso_amd
from diso_amd import workloads
from diso_amd.serve import Request, Server
from oracle import oracle as O
from tests.helpers import check_against_oracle
from tests.torch_counts import expected_counts

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _tables():
    return {k: torch.tensor(v, device=DEV, dtype=torch.int64) for k, v in O.expand_tables().items()}


def test_sphere64_mc_forward_is_bit_exact():
    print("case sphere 64 r=0.4", flush=True)
    sdf_np = workloads.sphere(64, radius=0.4).astype(np.float32)
    sdf = torch.tensor(sdf_np, device=DEV)
    v, f = diso_amd.DiffMC(torch.float32)(sdf, None, 0.0, normalize=True)
    assert (v.shape[0], f.shape[0]) == (11952, 23900)
    assert f.dtype == torch.int64
    check_against_oracle("mc", sdf_np, None, v, f)
    v, q = diso_amd.DiffDMC(torch.float32)(sdf, None, 0.0, return_quads=True, normalize=True)
    assert (v.shape[0], q.shape[0]) == (11954, 11952)
    check_against_oracle("dmc", sdf_np, None, v, q)


def _fd_stable(sdf):
    """Samples whose +-1e-6 step flips no inside bit (|sdf| > 1e-5) and whose crossing edges all
    have |d1 - d0| >= 1e-3 (so the central difference's third-order error stays far below the
    tolerance), drawn at random among those with a non-zero gradient."""
    P = np.pad(sdf, 1, constant_values=1.0)
    ok = np.abs(sdf) > 1e-5
    c = P[1:-1, 1:-1, 1:-1]
    for ax in range(3):
        for s in (-1, 1):
            nb = np.roll(P, -s, axis=ax)[1:-1, 1:-1, 1:-1]
            cross = (c >= 0) != (nb >= 0)
            ok &= ~(cross & (np.abs(nb - c) < 1e-3))
    return ok


@pytest.mark.parametrize("algo", ["mc", "dmc"])
def test_fp64_central_differences_n32(algo):
    print(f"case fd n=32 fp64 {algo}", flush=True)
    n, eps = 32, 1e-6
    sdf_np, def_np, rng = workloads.noisy_round_box(n, seed=0, dtype=np.float64)
    mod = diso_amd.DiffMC(torch.float64) if algo == "mc" else diso_amd.DiffDMC(torch.float64)
    kw = {"return_quads": True} if algo == "dmc" else {}

    def loss(s, d, g):
        with torch.no_grad():
            v, _ = mod(torch.tensor(s, device=DEV), torch.tensor(d, device=DEV), **kw)
            return float((v * g).sum())

    sdf = torch.tensor(sdf_np, device=DEV, requires_grad=True)
    deform = torch.tensor(def_np, device=DEV, requires_grad=True)
    v, _ = mod(sdf, deform, **kw)
    g = torch.tensor(rng.standard_normal(tuple(v.shape)), device=DEV)
    (v * g).sum().backward()
    gs, gd = sdf.grad.cpu().numpy(), deform.grad.cpu().numpy()
    ok = _fd_stable(sdf_np)
    cand = np.argwhere(ok & (np.abs(gs) > 0))
    assert len(cand) > 100
    for i, j, k in cand[rng.choice(len(cand), 40, replace=False)]:
        sp, sm = sdf_np.copy(), sdf_np.copy()
        sp[i, j, k] += eps
        sm[i, j, k] -= eps
        fd = (loss(sp, def_np, g) - loss(sm, def_np, g)) / (2 * eps)
        assert abs(fd - gs[i, j, k]) <= 1e-6 * max(1.0, abs(fd)), (i, j, k, fd, gs[i, j, k])
    cand = np.argwhere(np.abs(gd) > 0)
    for i, j, k, c in cand[rng.choice(len(cand), 40, replace=False)]:
        dp, dm = def_np.copy(), def_np.copy()
        dp[i, j, k, c] += eps
        dm[i, j, k, c] -= eps
        fd = (loss(sdf_np, dp, g) - loss(sdf_np, dm, g)) / (2 * eps)
        assert abs(fd - gd[i, j, k, c]) <= 1e-6 * max(1.0, abs(fd)), (i, j, k, c, fd, gd[i, j, k, c])


@pytest.mark.parametrize("algo", ["mc", "dmc"])
@pytest.mark.parametrize("recompute", [False, True])
def test_gradient_identities_at_256(algo, recompute):
    print(f"case gradient identities 256 {algo} recompute={recompute}", flush=True)
    n = 256
    sdf_np, def_np, rng = workloads.noisy_round_box(n, seed=0)
    # the rounded box stays >= 0.1 from the grid's faces, so no crossing touches the padding
    assert min(sdf_np[0].min(), sdf_np[-1].min(), sdf_np[:, 0].min(), sdf_np[:, -1].min(),
               sdf_np[:, :, 0].min(), sdf_np[:, :, -1].min()) > 0
    sdf = torch.tensor(sdf_np, device=DEV, requires_grad=True)
    deform = torch.tensor(def_np, device=DEV, requires_grad=True)
    mod = (diso_amd.DiffMC(recompute_backward=recompute) if algo == "mc"
           else diso_amd.DiffDMC(recompute_backward=recompute))
    kw = {"return_quads": True} if algo == "dmc" else {}
    v, _ = mod(sdf, deform, **kw)
    g = torch.tensor(workloads.upstream(rng, v.shape[0]), device=DEV)
    torch.autograd.backward(v, grad_tensors=g)
    s64, gs64 = sdf.detach().double(), sdf.grad.double()
    terms = s64 * gs64
    lhs, scale = float(terms.sum()), float(terms.abs().sum())
    assert scale > 0 and abs(lhs) <= 1e-5 * scale, (lhs, scale)
    want = g.double().sum(0) / (n - 1)
    got = deform.grad.double().reshape(-1, 3).sum(0)
    scale = g.double().abs().sum(0) / (n - 1)
    assert torch.all((got - want).abs() <= 1e-5 * scale), (got, want, scale)


def _upstream(v, seed):
    gen = torch.Generator(device=DEV).manual_seed(seed)
    return torch.randn(tuple(v.shape), generator=gen, dtype=v.dtype, device=DEV)


def test_uneven_batch_against_oracle_counts_and_per_grid():
    print("case uneven batch 64 grids", flush=True)
    data = workloads.uneven_batch_torch(64, seed=0, device=DEV)
    tables = _tables()
    sdfs = [s.clone().requires_grad_(True) for s, _, _, _, _ in data]
    defs = [d.clone().requires_grad_(True) for _, d, _, _, _ in data]
    algos = [a for _, _, a, _, _ in data]
    outs = diso_amd.BatchExtractor(torch.float32)(sdfs, algos, defs, return_quads=True)
    ups = [_upstream(v, 100 + i) for i, (v, _) in enumerate(outs)]
    live = [i for i, (v, _) in enumerate(outs) if v.shape[0]]
    torch.autograd.backward([outs[i][0] for i in live], grad_tensors=[ups[i] for i in live])
    n_oracle = 0
    for i, ((v, f), (s0, d0, algo, kind, n)) in enumerate(zip(outs, data)):
        print(f"  grid {i} n={n} {algo} {kind}", flush=True)
        nx, ntri, npatch = expected_counts(s0, tables)
        if v.shape[0]:
            want = (nx, ntri) if algo == "mc" else (npatch, nx)
            assert (v.shape[0], f.shape[0]) == want, (i, want)
            assert int(f.min()) >= 0 and int(f.max()) < v.shape[0]
        if n <= 64:
            n_oracle += 1
            check_against_oracle(algo, s0.cpu().numpy(), d0.cpu().numpy(), v, f, ups[i], sdfs[i].grad, defs[i].grad)
        # the same grid alone through the per-grid path
        s1, d1 = s0.clone().requires_grad_(True), d0.clone().requires_grad_(True)
        mod = diso_amd.DiffMC() if algo == "mc" else diso_amd.DiffDMC()
        pv, pf = mod(s1, d1, return_quads=True) if algo == "dmc" else mod(s1, d1)
        assert torch.equal(pv.detach(), v.detach()) and torch.equal(pf, f), i
        if v.shape[0]:
            torch.autograd.backward(pv, grad_tensors=ups[i])
            assert torch.equal(s1.grad, sdfs[i].grad) and torch.equal(d1.grad, defs[i].grad), i
    assert n_oracle >= 10


def test_serving_stream_slice():
    print("case serving stream 200 requests", flush=True)
    pool, reqs = workloads.request_stream(200, rate=1000.0, seed=0, pool=24, device=DEV)
    made = []
    for k, (_, pi, algo, bwd) in enumerate(reqs):
        sdf, deform, _, n = pool[pi]
        up = (lambda nv, s=1000 + k: torch.randn((nv, 3), generator=torch.Generator(device=DEV).manual_seed(s),
                                                 device=DEV)) if bwd else None
        made.append((Request(sdf, algo, deform, upstream=up), pi, algo, bwd, 1000 + k))
    with Server(budget_bytes=200 << 30, workers=4) as srv:
        threads = [threading.Thread(target=lambda ms=made[j::2]: [srv.submit(m[0]) for m in ms]) for j in range(2)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        results = [m[0].future.result(timeout=300) for m in made]
    assert len(srv.latencies) == len(made)
    n_oracle = 0
    for res, (_, pi, algo, bwd, seed) in zip(results, made):
        sdf, deform, _, n = pool[pi]
        s = sdf.detach().requires_grad_(bwd)
        d = deform.detach().requires_grad_(bwd)
        mod = diso_amd.DiffMC() if algo == "mc" else diso_amd.DiffDMC()
        v, f = mod(s, d, return_quads=True) if algo == "dmc" else mod(s, d)
        assert torch.equal(res[0], v.detach()) and torch.equal(res[1], f)
        up = gs = gd = None
        if bwd and v.shape[0]:
            up = torch.randn((v.shape[0], 3), generator=torch.Generator(device=DEV).manual_seed(seed), device=DEV)
            gs, gd = torch.autograd.grad([v], [s, d], grad_outputs=[up])
            assert torch.equal(res[2], gs) and torch.equal(res[3], gd)
        if n <= 64:
            n_oracle += 1
            check_against_oracle(algo, sdf.cpu().numpy(), deform.cpu().numpy(), res[0], res[1], up,
                                 None if up is None else res[2], None if up is None else res[3])
    assert n_oracle >= 10
