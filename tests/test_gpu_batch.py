"""Batched extraction against per-grid extraction and the oracle (needs a GPU): every grid of a
mixed batch (MC and DMC, uneven and odd dims, with and without deform, an empty grid) gives
bit-identical vertices, faces and gradients to DiffMC / DiffDMC on that grid alone, and equals the
oracle (faces and vertices bit-exact, gradients within the parity tolerance)."""
import numpy as np
import pytest
import torch

import diso_amd
from tests.helpers import check_against_oracle, inputs

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TDT = {np.float32: torch.float32, np.float64: torch.float64}

BATCH = [  # dims, family, algo, deform
    ((24, 24, 24), "noisy_box", "mc", True),
    ((17, 33, 70), "noise", "dmc", True),
    ((40, 9, 130), "blobs", "mc", False),
    ((64, 64, 64), "sphere", "dmc", False),
    ((5, 3, 66), "noise", "dmc", True),
    ((2, 2, 2), "noise", "mc", False),
    ((31, 47, 29), "sphere", "mc", True),
    ((8, 8, 8), "empty", "dmc", False),  # all samples above iso: the early-out
    ((48, 40, 72), "noise", "dmc", True),
]


def _grid(dims, kind, seed, dtype, with_deform):
    if kind == "empty":
        return np.full(dims, 0.5, dtype), None
    return inputs(dims, kind, seed, dtype, with_deform)


def _plain(algo, dtype, sdf_np, def_np, ups, compat, recompute, return_quads):
    sdf = torch.tensor(sdf_np, device=DEV, requires_grad=True)
    deform = None if def_np is None else torch.tensor(def_np, device=DEV, requires_grad=True)
    if algo == "mc":
        mod = diso_amd.DiffMC(TDT[dtype], recompute_backward=recompute)
        v, f = mod(sdf, deform)
    else:
        mod = diso_amd.DiffDMC(TDT[dtype], reference_compat=compat, recompute_backward=recompute)
        v, f = mod(sdf, deform, return_quads=return_quads)
    if v.shape[0]:
        torch.autograd.backward(v, grad_tensors=ups)
    return v.detach(), f, sdf.grad, None if deform is None else deform.grad


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("compat,recompute,return_quads", [(False, False, True), (True, True, False)])
def test_batch_matches_per_grid(dtype, compat, recompute, return_quads):
    print(f"case batch dtype={np.dtype(dtype).name} compat={compat} recompute={recompute} quads={return_quads}",
          flush=True)
    data = [_grid(d, k, i, dtype, w) for i, (d, k, _, w) in enumerate(BATCH)]
    algos = [a for _, _, a, _ in BATCH]
    sdfs = [torch.tensor(s, device=DEV, requires_grad=True) for s, _ in data]
    defs = [None if d is None else torch.tensor(d, device=DEV, requires_grad=True) for _, d in data]
    ex = diso_amd.BatchExtractor(TDT[dtype], reference_compat=compat, recompute_backward=recompute)
    outs = ex(sdfs, algos, defs, return_quads=return_quads)
    rng = np.random.default_rng(5)
    ups = [torch.tensor(rng.standard_normal(tuple(v.shape)), dtype=TDT[dtype], device=DEV) for v, _ in outs]
    live = [(v, u) for (v, _), u in zip(outs, ups) if v.shape[0]]
    torch.autograd.backward([v for v, _ in live], grad_tensors=[u for _, u in live])
    for i, ((v, f), (s_np, d_np)) in enumerate(zip(outs, data)):
        pv, pf, pgs, pgd = _plain(algos[i], dtype, s_np, d_np, ups[i], compat, recompute, return_quads)
        assert torch.equal(v.detach(), pv), f"grid {i} vertices"
        assert f.dtype == pf.dtype and torch.equal(f, pf), f"grid {i} faces"
        if v.shape[0]:
            assert torch.equal(sdfs[i].grad, pgs), f"grid {i} sdf gradient"
            if defs[i] is not None:
                assert torch.equal(defs[i].grad, pgd), f"grid {i} deform gradient"
        else:
            assert sdfs[i].grad is None or not sdfs[i].grad.any()
        if algos[i] == "mc" or return_quads:  # (the split's cut is tie-tolerant, tested elsewhere)
            check_against_oracle(algos[i], s_np, d_np, v, f, ups[i], sdfs[i].grad,
                                 None if defs[i] is None else defs[i].grad, compat=compat)


def test_batch_twice_and_nan():
    """The control block is reused across batches of different sizes; NaN is reported per grid."""
    ex = diso_amd.BatchExtractor()
    g1 = torch.tensor(inputs((20, 21, 22), "sphere", 0, np.float32, False)[0], device=DEV)
    g2 = torch.tensor(inputs((33, 12, 40), "noise", 1, np.float32, False)[0], device=DEV)
    for grids in ([g1, g2], [g2], [g1, g2, g1]):
        outs = ex(grids, ["mc", "dmc", "mc"][: len(grids)], return_quads=True)
        for (v, f), g, a in zip(outs, grids, ["mc", "dmc", "mc"]):
            mod = diso_amd.DiffMC() if a == "mc" else diso_amd.DiffDMC()
            pv, pf = mod(g, return_quads=True) if a == "dmc" else mod(g)
            assert torch.equal(v, pv) and torch.equal(f, pf)
    bad = g2.clone()
    bad[3, 4, 5] = float("nan")
    with pytest.raises(RuntimeError, match="grid 1"):
        ex([g1, bad], ["mc", "mc"])


def test_batch_reuses_a_counting_pass():
    """count() gives the sizes the extraction produces; a call with its handle skips the pass and
    gives the same results; a handle from other grids or settings is refused."""
    ex = diso_amd.BatchExtractor()
    grids = [torch.tensor(inputs(d, k, i, np.float32, False)[0], device=DEV)
             for i, (d, k) in enumerate([((20, 21, 22), "sphere"), ((33, 12, 40), "noise"), ((8, 8, 8), "blobs")])]
    algos = ["mc", "dmc", "dmc"]
    ref = ex(grids, algos, return_quads=True)
    c = ex.count(grids, algos)
    assert [(n_v, n_f) for n_v, n_f, _ in c.sizes] == [(v.shape[0], f.shape[0]) for v, f in ref]
    outs = ex(grids, algos, return_quads=True, counted=c)
    for (v, f), (pv, pf) in zip(outs, ref):
        assert torch.equal(v, pv) and torch.equal(f, pf)
    c = ex.count(grids, algos)
    with pytest.raises(ValueError, match="counted"):
        ex(grids, ["mc", "mc", "dmc"], counted=c)
    with pytest.raises(ValueError, match="counted"):
        ex(grids[:2], algos[:2], counted=c)
