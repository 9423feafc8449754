"""Fixed-capacity extraction and captured steps (needs a GPU).

FixedExtractor against DiffMC / DiffDMC on the same grid (bit-identical rows, counts and
gradients, which the per-grid parity tests tie to the oracle); the capacity overflow; a captured
forward+backward replayed on changing inputs with eager work and synchronisations between the
replays, every replay checked against an eager run; and the captured sdf-fitting step against
the same steps run eagerly."""
import numpy as np
import pytest
import torch

import diso_amd
from diso_amd import FixedExtractor
from diso_amd.fit import SdfFit
from tests.helpers import check_against_oracle, inputs

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TDT = {np.float32: torch.float32, np.float64: torch.float64}


def _plain(algo, dtype, sdf_np, def_np, ups_seed, recompute):
    sdf = torch.tensor(sdf_np, device=DEV, requires_grad=True)
    deform = None if def_np is None else torch.tensor(def_np, device=DEV, requires_grad=True)
    if algo == "mc":
        v, f = diso_amd.DiffMC(TDT[dtype], recompute_backward=recompute)(sdf, deform)
    else:
        v, f = diso_amd.DiffDMC(TDT[dtype], recompute_backward=recompute)(sdf, deform, return_quads=True)
    g = torch.randn(v.shape, generator=torch.Generator(device=DEV).manual_seed(ups_seed), dtype=v.dtype, device=DEV)
    torch.autograd.backward(v, grad_tensors=g)
    return v.detach(), f, g, sdf.grad, None if deform is None else deform.grad


def _fixed(ex, sdf_np, def_np, g):
    sdf = torch.tensor(sdf_np, device=DEV, requires_grad=True)
    deform = None if def_np is None else torch.tensor(def_np, device=DEV, requires_grad=True)
    v, f, c = ex(sdf, deform)
    gp = torch.zeros_like(v)
    gp[: g.shape[0]] = g
    torch.autograd.backward(v, grad_tensors=gp)
    return v.detach(), f, c, sdf.grad, None if deform is None else deform.grad


@pytest.mark.parametrize("algo", ["mc", "dmc"])
@pytest.mark.parametrize("dtype,with_deform", [(np.float32, True), (np.float64, False)])
@pytest.mark.parametrize("recompute", [False, True])
def test_fixed_matches_plain(algo, dtype, with_deform, recompute):
    print(f"case fixed {algo} dtype={np.dtype(dtype).name} deform={with_deform} recompute={recompute}", flush=True)
    sdf_np, def_np = inputs((37, 29, 70), "noise", 3, dtype, with_deform)
    v, f, g, gs, gd = _plain(algo, dtype, sdf_np, def_np, 7, recompute)
    nv, nf = v.shape[0], f.shape[0]
    ex = FixedExtractor(algo, TDT[dtype], vert_capacity=nv + 37, face_capacity=nf + 11, recompute_backward=recompute)
    v2, f2, c, gs2, gd2 = _fixed(ex, sdf_np, def_np, g)
    counts = c.tolist()
    assert counts[:2] == [nv, nf]
    both = diso_amd._lib.TOTAL_HAS_BELOW | diso_amd._lib.TOTAL_HAS_ABOVE
    assert counts[2] & both == both
    assert ex.check(c) == (nv, nf)
    assert v2.shape == (nv + 37, 3) and f2.shape == (nf + 11, 4 if algo == "dmc" else 3)
    assert torch.equal(v2[:nv], v)
    assert not v2[nv:].any()
    assert torch.equal(f2[:nf], f)
    assert torch.equal(gs2, gs)
    if with_deform:
        assert torch.equal(gd2, gd)


def test_fixed_overflow_writes_nothing():
    sdf_np, _ = inputs((24, 24, 24), "sphere", 0, np.float32, False)
    sdf = torch.tensor(sdf_np, device=DEV)
    nv = diso_amd.DiffDMC()(sdf, return_quads=True)[0].shape[0]
    ex = FixedExtractor("dmc", torch.float32, vert_capacity=nv - 1, face_capacity=10 * nv)
    v, _, c = ex(sdf)
    assert int(c[0]) == nv
    assert not v.any()
    with pytest.raises(RuntimeError, match="exceed"):
        ex.check(c)


@pytest.mark.parametrize("algo", ["mc", "dmc"])
@pytest.mark.parametrize("recount_spins", [None, 0])
def test_captured_step_replays_match_eager(algo, recount_spins):
    """recount_spins 0: every count block counts the ranges of predecessors that have not yet
    published instead of waiting; replays must stay exact (no host step between them resets
    anything)."""
    print(f"case graph replay {algo} recount_spins={recount_spins}", flush=True)
    L = diso_amd._lib.lib()
    spins = L.diso_get_param(diso_amd._lib.PARAM_RECOUNT_SPINS)
    if recount_spins is not None:
        assert L.diso_set_param(diso_amd._lib.PARAM_RECOUNT_SPINS, recount_spins) == 0
    try:
        _replays(algo)
    finally:
        L.diso_set_param(diso_amd._lib.PARAM_RECOUNT_SPINS, spins)


def _replays(algo):
    dims = (40, 36, 44)
    fields = [inputs(dims, kind, seed, np.float32, True) for seed, kind in enumerate(["noise", "sphere", "blobs"])]
    cap_v = 4 * 42 * 38 * 46  # more than any grid of these dims can produce (padded positions x 4)
    ex = FixedExtractor(algo, torch.float32, vert_capacity=cap_v, face_capacity=2 * cap_v)
    sdf = torch.tensor(fields[0][0], device=DEV, requires_grad=True)
    deform = torch.tensor(fields[0][1], device=DEV, requires_grad=True)
    w = torch.randn((cap_v, 3), generator=torch.Generator(device=DEV).manual_seed(1), device=DEV)

    def step():
        v, f, c = ex(sdf, deform)
        torch.autograd.backward(v, grad_tensors=w)
        return v, f, c

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(2):
            sdf.grad = deform.grad = None
            step()
    torch.cuda.current_stream().wait_stream(s)
    sdf.grad = deform.grad = None
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        v_s, f_s, c_s = step()
    for k in range(6):
        sdf_np, def_np = fields[k % 3]
        with torch.no_grad():
            sdf.copy_(torch.tensor(sdf_np, device=DEV))
            deform.copy_(torch.tensor(def_np, device=DEV))
        graph.replay()
        # eager work between replays, then the reference for this input
        ex_e = FixedExtractor(algo, torch.float32, vert_capacity=cap_v, face_capacity=2 * cap_v)
        v_e, f_e, c_e, gs_e, gd_e = _fixed(ex_e, sdf_np, def_np, w)
        torch.cuda.synchronize()
        nv, nf = ex.check(c_s)
        assert torch.equal(c_s, c_e), k
        assert torch.equal(v_s, v_e), k
        assert torch.equal(f_s[:nf], f_e[:nf]), k
        assert torch.equal(sdf.grad, gs_e), k
        assert torch.equal(deform.grad, gd_e), k
        # and the replay against the oracle
        check_against_oracle(algo, sdf_np, def_np, v_s[:nv], f_s[:nf], w[:nv], sdf.grad, deform.grad)


def test_captured_fit_matches_eager_fit():
    print("case fit n=48", flush=True)
    a = SdfFit(48, device=DEV, seed=0)
    b = SdfFit(48, device=DEV, seed=0)
    assert a.ex.vert_capacity == b.ex.vert_capacity
    a.capture(warmup=3)
    for _ in range(3):
        b.step()
    losses = []
    for k in range(8):
        a.replay()
        b.step()
        torch.cuda.synchronize()
        a.check()
        assert torch.equal(a.counts, b.counts), k
        assert torch.equal(a.loss, b.loss), k
        assert torch.equal(a.sdf, b.sdf), k
        assert torch.equal(a.deform, b.deform), k
        losses.append(float(a.loss))
    assert losses[-1] < losses[0]
