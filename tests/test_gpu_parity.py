"""HIP engine against the CPU oracle on the same seeded inputs (needs a GPU).

Topology (faces / quads, their order and the vertex count) and fp32/fp64 vertex positions must
be bit-identical to the oracle, which restates the reference engine operation by operation.
Gradients are compared against the oracle's sequential scatter within a tolerance relative to
the gradient's scale: 2e-5 (fp32) / 1e-12 (fp64), since the gather sums the same terms in a
different order.
"""
import numpy as np
import pytest
import torch

import diso_amd
from diso_amd import _C
from oracle import oracle as O
from tests.helpers import grad_close, inputs

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TOL = {np.float32: 2e-5, np.float64: 1e-12}
TDT = {np.float32: torch.float32, np.float64: torch.float64}

CASES = [
    ((24, 24, 24), "noisy_box"),
    ((17, 33, 70), "noise"),
    ((5, 3, 66), "noise"),
    ((40, 9, 130), "blobs"),
    ((64, 64, 64), "sphere"),
    ((2, 2, 2), "noise"),
    ((1, 7, 3), "noise"),  # X=1: normalisation divides by X-1 = 0, so run unnormalised
]


def _module(algo, dtype, compat=False):
    return diso_amd.DiffMC(TDT[dtype]) if algo == O.MC else diso_amd.DiffDMC(TDT[dtype], reference_compat=compat)


def _run_both(algo, dtype, dims, kind, with_deform, seed=0, normalize=True, iso=0.0, compat=False):
    sdf_np, def_np = inputs(dims, kind, seed, dtype, with_deform)
    print(f"case algo={algo} dtype={np.dtype(dtype).name} dims={dims} kind={kind} deform={with_deform}", flush=True)
    ref = O.Extractor(algo, dtype)
    rv, rf = ref.forward(sdf_np, def_np, iso, normalize, return_quads=True)
    sdf = torch.tensor(sdf_np, device=DEV, requires_grad=True)
    deform = None if def_np is None else torch.tensor(def_np, device=DEV, requires_grad=True)
    kw = {"return_quads": True} if algo == O.DMC else {}
    verts, faces = _module(algo, dtype, compat)(sdf, deform, iso, normalize=normalize, **kw)
    return ref, rv, rf, sdf, deform, verts, faces


@pytest.mark.parametrize("algo", [O.MC, O.DMC])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("with_deform", [False, True])
@pytest.mark.parametrize("dims,kind", CASES)
def test_forward_backward_match_oracle(algo, dtype, with_deform, dims, kind):
    ref, rv, rf, sdf, deform, verts, faces = _run_both(algo, dtype, dims, kind, with_deform,
                                                       normalize=min(dims) > 1)
    if ref.empty:
        assert verts.shape == (0, 3) and faces.shape[0] == 0
        return
    assert faces.dtype == torch.int64
    assert tuple(faces.shape) == rf.shape
    assert np.array_equal(faces.cpu().numpy(), rf)
    assert np.array_equal(verts.detach().cpu().numpy(), rv)
    rng = np.random.default_rng(7)
    g = rng.standard_normal(rv.shape).astype(dtype)
    (verts * torch.tensor(g, device=DEV)).sum().backward()
    gs, gd = ref.backward(g)
    grad_close(sdf.grad.cpu().numpy(), gs, TOL[dtype])
    if deform is not None:
        grad_close(deform.grad.cpu().numpy(), gd, TOL[dtype])


@pytest.mark.parametrize("algo", [O.MC, O.DMC])
def test_unnormalized_and_isovalue(algo):
    ref, rv, rf, sdf, deform, verts, faces = _run_both(algo, np.float32, (20, 31, 67), "blobs", True,
                                                       normalize=False, iso=0.1)
    assert np.array_equal(faces.cpu().numpy(), rf)
    assert np.array_equal(verts.detach().cpu().numpy(), rv)


def test_dmc_reference_compat_backward():
    ref, rv, rf, sdf, deform, verts, faces = _run_both(O.DMC, np.float64, (19, 22, 70), "noise", True,
                                                       compat=True)
    g = np.random.default_rng(3).standard_normal(rv.shape)
    (verts * torch.tensor(g, device=DEV)).sum().backward()
    gs, gd = ref.backward(g, reference_compat=True)
    grad_close(sdf.grad.cpu().numpy(), gs, 1e-12)
    grad_close(deform.grad.cpu().numpy(), gd, 1e-12)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_dmc_triangles_match_oracle_split(dtype):
    sdf_np, def_np = inputs((33, 30, 41), "noisy_box", 5, dtype, True)
    ref = O.Extractor(O.DMC, dtype)
    rv, rt = ref.forward(sdf_np, def_np, 0.0, True, return_quads=False)
    _, rq = O.Extractor(O.DMC, dtype).forward(sdf_np, def_np, 0.0, True, return_quads=True)
    _, s1, s2 = O.split_quads(rv, rq)
    verts, tris = diso_amd.DiffDMC(TDT[dtype])(torch.tensor(sdf_np, device=DEV), torch.tensor(def_np, device=DEV))
    tris = tris.cpu().numpy()
    assert np.array_equal(verts.cpu().numpy(), rv)
    assert tris.shape == rt.shape
    # The cut choice compares float scores; quads whose two scores tie within rounding may be cut
    # either way (and then move between the two halves of the output), so compare per-quad cuts.
    first = s1 < s2
    tie = np.abs(s1 - s2) <= 4 * np.finfo(dtype).eps
    n1 = int(first.sum())
    got_first = np.zeros(len(rq), bool)
    got = {tuple(q) for q in tris[: 2 * n1].reshape(-1, 6)} if n1 else set()
    for i, q in enumerate(rq):
        got_first[i] = tuple(q[[0, 1, 3, 1, 2, 3]]) in got
    assert np.all((got_first == first) | tie)
    if not tie.any():
        assert np.array_equal(tris, rt)


@pytest.mark.parametrize("cls,algo,dtype", [(_C.CUMCFloat, O.MC, np.float32), (_C.CUMCDouble, O.MC, np.float64),
                                            (_C.CUDMCFloat, O.DMC, np.float32), (_C.CUDMCDouble, O.DMC, np.float64)])
def test_binding_classes_match_engine(cls, algo, dtype):
    sdf_np, def_np = inputs((21, 18, 69), "noise", 11, dtype, True)
    g_np, d_np = O.pad(sdf_np, def_np, 0.0)
    eng = O.Engine(algo, dtype)
    rv, rf, st = eng.forward(g_np, d_np, 0.0)
    assert st == 0
    b = cls()
    grid = torch.tensor(g_np, device=DEV)
    deform = torch.tensor(d_np, device=DEV)
    verts, faces = b.forward(grid, deform, 0.0)
    assert faces.dtype == torch.int32
    assert np.array_equal(faces.cpu().numpy().astype(np.int64), rf)
    assert np.array_equal(verts.cpu().numpy(), rv)
    adj = np.random.default_rng(2).standard_normal(rv.shape).astype(dtype)
    ag = torch.zeros_like(grid)
    ad = torch.zeros_like(deform)
    b.backward(grid, deform, 0.0, torch.tensor(adj, device=DEV), ag, ad)
    gs, gd = eng.backward(g_np, d_np, 0.0, adj, reference_compat=(algo == O.DMC))
    grad_close(ag.cpu().numpy(), gs, TOL[dtype])
    grad_close(ad.cpu().numpy(), gd, TOL[dtype])
    # no-deform overloads
    verts2, faces2 = b.forward(grid, 0.0)
    rv2, rf2, _ = O.Engine(algo, dtype).forward(g_np, None, 0.0)
    assert np.array_equal(faces2.cpu().numpy().astype(np.int64), rf2)
    assert np.array_equal(verts2.cpu().numpy(), rv2)


def test_early_out_and_nan():
    mc = diso_amd.DiffMC()
    g = torch.full((8, 8, 8), 0.5, device=DEV)
    v, f = mc(g)
    assert v.shape == (0, 3) and f.shape == (0, 3) and f.dtype == torch.int32
    g[3, 3, 3] = 0.0  # max == iso... min == iso: min >= iso holds -> empty
    v, f = mc(g, isovalue=0.0)
    assert v.shape == (0, 3)
    v, f = diso_amd.DiffDMC()(torch.full((8, 8, 8), -1.0, device=DEV))
    assert v.shape == (0, 3) and f.shape == (0, 4) and f.dtype == torch.int32
    g = torch.randn(8, 8, 8, device=DEV)
    g[1, 2, 3] = float("nan")
    with pytest.raises(RuntimeError):
        mc(g)


def test_deterministic_and_repeatable():
    sdf_np, def_np = inputs((48, 40, 72), "noise", 4, np.float32, True)
    outs = []
    for _ in range(2):
        for mod in (diso_amd.DiffMC(), diso_amd.DiffDMC()):
            sdf = torch.tensor(sdf_np, device=DEV, requires_grad=True)
            deform = torch.tensor(def_np, device=DEV, requires_grad=True)
            v, f = mod(sdf, deform)
            (v * v).sum().backward()
            outs.append((v.detach().cpu().numpy(), f.cpu().numpy(), sdf.grad.cpu().numpy(), deform.grad.cpu().numpy()))
    for a, b in zip(outs[:2], outs[2:]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("algo", [O.MC, O.DMC])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dims,kind", [((37, 29, 131), "noise"), ((24, 24, 24), "noisy_box")])
def test_recompute_backward_is_identical(algo, dtype, dims, kind):
    """The recompute-classification backward (only the segment bases kept from the forward)
    gives bit-identical gradients to the stored-classification backward, and both match the
    oracle."""
    sdf_np, def_np = inputs(dims, kind, 7, dtype, True)
    print(f"case recompute algo={algo} dtype={np.dtype(dtype).name} dims={dims} kind={kind}", flush=True)
    ups = None
    grads = []
    for rec in (False, True):
        mod = (diso_amd.DiffMC(TDT[dtype], recompute_backward=rec) if algo == O.MC
               else diso_amd.DiffDMC(TDT[dtype], recompute_backward=rec))
        sdf = torch.tensor(sdf_np, device=DEV, requires_grad=True)
        deform = torch.tensor(def_np, device=DEV, requires_grad=True)
        kw = {"return_quads": True} if algo == O.DMC else {}
        v, f = mod(sdf, deform, **kw)
        if ups is None:
            ups = torch.tensor(np.random.default_rng(3).standard_normal(tuple(v.shape)), dtype=v.dtype, device=DEV)
        torch.autograd.backward(v, grad_tensors=ups)
        grads.append((sdf.grad.cpu().numpy(), deform.grad.cpu().numpy()))
    for a, b in zip(grads[0], grads[1]):
        assert np.array_equal(a, b)
    ref = O.Extractor(algo, dtype)
    ref.forward(sdf_np, def_np, 0.0, True, return_quads=True)
    gs, gd = ref.backward(ups.cpu().numpy())
    grad_close(grads[1][0], gs, TOL[dtype])
    grad_close(grads[1][1], gd, TOL[dtype])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nq,idx", [(0, torch.int64), (1, torch.int64), (3000, torch.int32), (200_003, torch.int64)])
def test_split_kernel_matches_oracle(dtype, nq, idx):
    """The HIP split on random quads (including degenerate ones with repeated corners) against
    the oracle's restatement of diso/__init__.py:117-147."""
    print(f"case split dtype={np.dtype(dtype).name} nq={nq} idx={idx}", flush=True)
    rng = np.random.default_rng(nq)
    nv = max(4, nq // 2)
    verts = rng.standard_normal((nv, 3)).astype(dtype)
    quads = rng.integers(0, nv, (nq, 4))
    quads[: nq // 50, 2] = quads[: nq // 50, 0]  # zero-length diagonals / sides
    want, s1, s2 = O.split_quads(verts, quads)
    got = diso_amd.split_quads(torch.tensor(verts, device=DEV), torch.tensor(quads, dtype=idx, device=DEV))
    assert got.dtype == torch.int64 and tuple(got.shape) == (2 * nq, 3)
    assert np.array_equal(got.cpu().numpy(), want)


def test_extracted_meshes_are_watertight_and_export(tmp_path):
    """End to end: HIP meshes of a closed shape pass the topology checks and survive export."""
    from diso_amd import mesh

    sdf = torch.tensor(inputs((48, 40, 56), "sphere", 0, np.float32, False)[0], device=DEV)
    for mod, k in ((diso_amd.DiffMC(), 3), (diso_amd.DiffDMC(), 3), (diso_amd.DiffDMC(), 4)):
        v, f = mod(sdf, return_quads=True) if k == 4 else mod(sdf)
        r = mesh.check(f, v.shape[0])
        assert f.shape[1] == k and r["oriented"] and r["euler"] == 2 and r["components"] == 1
        assert mesh.signed_volume(v, f) > 0
        mesh.write_ply(tmp_path / "m.ply", v, f)
        v2, f2 = mesh.read_ply(tmp_path / "m.ply")
        assert np.array_equal(v2, v.cpu().numpy()) and np.array_equal(f2, f.cpu().numpy())


def test_results_do_not_depend_on_launch_parameters():
    from diso_amd import autotune

    sdf_np, def_np = inputs((37, 50, 150), "noisy_box", 9, np.float32, True)
    before = autotune.current()
    L = diso_amd._lib.lib()
    spins = L.diso_get_param(diso_amd._lib.PARAM_RECOUNT_SPINS)
    outs = []
    try:
        # recount spins 0: every count block counts the ranges of the predecessors that have not
        # published when it first looks (the path that keeps the scan live when launches share
        # the device)
        for run, cpc, rs in ((1, 1, spins), (7, 3, 0), (32, 4, spins), (64, 2, 0), (16, 4, 0), (3, 1, spins)):
            print(f"case launch run={run} count_blocks_per_cu={cpc} recount_spins={rs}", flush=True)
            if run == 3:  # only the bucket these grids and the batch fall in (64^3-class) differs
                autotune.apply({"run": 48, "count_blocks_per_cu": 4})
                autotune.apply({"64": {"run": run, "count_blocks_per_cu": cpc}})
            else:
                autotune.apply({"run": run, "count_blocks_per_cu": cpc})
            assert L.diso_set_param(diso_amd._lib.PARAM_RECOUNT_SPINS, rs) == 0
            res = []
            for mod in (diso_amd.DiffMC(), diso_amd.DiffDMC()):
                sdf = torch.tensor(sdf_np, device=DEV, requires_grad=True)
                deform = torch.tensor(def_np, device=DEV, requires_grad=True)
                v, f = mod(sdf, deform)
                (v * v).sum().backward()
                res += [v.detach().cpu().numpy(), f.cpu().numpy(), sdf.grad.cpu().numpy(), deform.grad.cpu().numpy()]
            # the batched count pass (one scan over the segments of several grids)
            grids = [torch.tensor(sdf_np, device=DEV), torch.tensor(sdf_np[:20, :33, :70].copy(), device=DEV)]
            for v, f in diso_amd.BatchExtractor()(grids, ["mc", "dmc"], return_quads=True):
                res += [v.cpu().numpy(), f.cpu().numpy()]
            outs.append(res)
    finally:
        autotune.apply(before)
        L.diso_set_param(diso_amd._lib.PARAM_RECOUNT_SPINS, spins)
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("algo", [O.MC, O.DMC])
@pytest.mark.parametrize("recompute", [False, True])
def test_reused_module_matches_fresh_module(algo, recompute):
    """A module reused on grids of the same dims emits speculatively right behind its counting
    pass (into buffers sized from the previous call): counts that fit (views of those buffers),
    counts that outgrow them (emitted again), an empty grid and the triangle split all give the
    results of a fresh module, gradients included."""
    dims = (30, 26, 70)
    mk = (lambda: diso_amd.DiffMC(recompute_backward=recompute)) if algo == O.MC else \
        (lambda: diso_amd.DiffDMC(recompute_backward=recompute))
    reused = mk()
    seq = [("sphere", 0), ("sphere", 1), ("noise", 2), ("noisy_box", 3), ("empty", 4), ("sphere", 5), ("noise", 6)]
    for i, (kind, seed) in enumerate(seq):
        print(f"case reuse algo={algo} recompute={recompute} step={i} kind={kind}", flush=True)
        if kind == "empty":
            sdf_np, def_np = np.full(dims, 0.5, np.float32), np.zeros(dims + (3,), np.float32)
        else:
            sdf_np, def_np = inputs(dims, kind, seed, np.float32, True)
        res = []
        for mod in (reused, mk()):
            sdf = torch.tensor(sdf_np, device=DEV, requires_grad=True)
            deform = torch.tensor(def_np, device=DEV, requires_grad=True)
            kw = {"return_quads": i % 2 == 0} if algo == O.DMC else {}
            v, f = mod(sdf, deform, **kw)
            if v.shape[0]:
                (v * v).sum().backward()
            res.append((v.detach(), f, sdf.grad, deform.grad))
        (v1, f1, gs1, gd1), (v2, f2, gs2, gd2) = res
        assert v1.shape == v2.shape and f1.dtype == f2.dtype and torch.equal(v1, v2) and torch.equal(f1, f2), i
        if v2.shape[0]:
            assert torch.equal(gs1, gs2) and torch.equal(gd1, gd2), i


@pytest.mark.parametrize("algo", [O.MC, O.DMC])
@pytest.mark.parametrize("normalize", [True, False])
def test_fp32_vertices_within_1e6_of_fp64_oracle(algo, normalize):
    """fp32 extraction against the fp64 oracle on the same (fp32-representable) samples: identical
    topology (the classification compares the same values) and vertices within 1e-6 of the
    coordinate scale (max |v|)."""
    print(f"case fp32 vs fp64 oracle algo={algo} normalize={normalize}", flush=True)
    sdf_np, def_np = inputs((64, 57, 70), "noisy_box", 9, np.float32, True)
    ref = O.Extractor(algo, np.float64)
    rv, rf = ref.forward(sdf_np.astype(np.float64), def_np.astype(np.float64), 0.0, normalize, return_quads=True)
    kw = {"return_quads": True} if algo == O.DMC else {}
    v, f = _module(algo, np.float32)(torch.tensor(sdf_np, device=DEV), torch.tensor(def_np, device=DEV),
                                     normalize=normalize, **kw)
    assert np.array_equal(f.cpu().numpy(), rf)
    err = np.abs(v.cpu().numpy().astype(np.float64) - rv).max()
    assert err <= 1e-6 * np.abs(rv).max(), err


@pytest.mark.parametrize("algo", [O.MC, O.DMC])
def test_backward_makes_no_face_gradient(algo):
    """The faces are integer outputs: the backward allocates the two input gradients (and the DMC
    edge scratch) but no zero-filled gradient of the faces."""
    print(f"case no face gradient algo={algo}", flush=True)
    sdf_np, def_np = inputs((80, 80, 80), "noise", 4, np.float32, True)
    sdf = torch.tensor(sdf_np, device=DEV, requires_grad=True)
    deform = torch.tensor(def_np, device=DEV, requires_grad=True)
    kw = {"return_quads": True} if algo == O.DMC else {}
    mod = _module(algo, np.float32)
    v, f = mod(sdf, deform, **kw)
    up = torch.randn(v.shape, device=DEV, generator=torch.Generator(device=DEV).manual_seed(1))
    face_bytes = f.numel() * f.element_size()
    scratch = 0 if algo == O.MC else 12 * f.shape[0]
    assert face_bytes > 2 << 20  # large enough to show above the allocator's rounding
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    torch.autograd.backward(v, grad_tensors=up)
    torch.cuda.synchronize()
    peak = torch.cuda.max_memory_allocated() - base
    grads = (sdf.numel() + deform.numel()) * 4
    assert peak < grads + scratch + face_bytes // 2, (peak, grads, scratch, face_bytes)


@pytest.mark.parametrize("algo,quads", [(O.MC, False), (O.DMC, False), (O.DMC, True)])
def test_reused_module_outputs_allow_inplace_edits(algo, quads):
    """A module's second call on grids of the same dims returns the speculatively emitted rows;
    those are tensors of their own, so in-place edits of the vertices (and of MC / split DMC
    triangles) under autograd work as on the first call, and the gradient flows through the
    edited vertices. DiffDMC's quads with return_quads=True are what its backward reads: an
    in-place edit of them makes the backward raise autograd's saved-tensor error (never a
    silently wrong gradient)."""
    print(f"case inplace edit algo={algo} quads={quads}", flush=True)
    sdf_np, def_np = inputs((24, 26, 30), "noisy_box", 2, np.float32, True)
    mod = _module(algo, np.float32)
    kw = {"return_quads": quads} if algo == O.DMC else {}
    grads = []
    for call in range(3):
        sdf = torch.tensor(sdf_np, device=DEV, requires_grad=True)
        deform = torch.tensor(def_np, device=DEV, requires_grad=True)
        v, f = mod(sdf, deform, **kw)
        v[:, 0] += 1.0
        v.mul_(2.0)
        if quads:
            (v * v).sum().backward(retain_graph=True)
            f += 0
            with pytest.raises(RuntimeError, match="inplace"):
                (v * v).sum().backward()
        else:
            f += 0
            (v * v).sum().backward()
        grads.append((v.detach().clone(), sdf.grad, deform.grad))
    for v, gs, gd in grads[1:]:
        assert torch.equal(v, grads[0][0]) and torch.equal(gs, grads[0][1]) and torch.equal(gd, grads[0][2])


# z extents whose padded rows end in a short last chunk (1, 17 and 32 positions: counted several
# rows per wave iteration) and one just past it (33 positions: one row per iteration)
TAIL_DIMS = [(6, 5, 63), (5, 7, 79), (7, 4, 94), (4, 6, 95)]


@pytest.mark.parametrize("dims", TAIL_DIMS)
@pytest.mark.parametrize("algo", [O.MC, O.DMC])
def test_short_row_tails_match_oracle(algo, dims):
    ref, rv, rf, sdf, deform, verts, faces = _run_both(algo, np.float32, dims, "noise", True)
    assert np.array_equal(faces.cpu().numpy(), rf)
    assert np.array_equal(verts.detach().cpu().numpy(), rv)
    g = np.random.default_rng(5).standard_normal(rv.shape).astype(np.float32)
    (verts * torch.tensor(g, device=DEV)).sum().backward()
    gs, gd = ref.backward(g)
    grad_close(sdf.grad.cpu().numpy(), gs, TOL[np.float32])
    grad_close(deform.grad.cpu().numpy(), gd, TOL[np.float32])


def test_short_row_tails_in_one_batch():
    """The same grids in one batched call (their tails counted inside the batch's ranges, with
    the other grids' segments around them): every grid as the oracle has it."""
    from tests.helpers import check_against_oracle

    data = [inputs(d, "noise", 20 + i, np.float32, True) for i, d in enumerate(TAIL_DIMS)]
    algos = ["mc", "dmc", "mc", "dmc"]
    print(f"case batch dims={TAIL_DIMS}", flush=True)
    sdfs = [torch.tensor(s, device=DEV, requires_grad=True) for s, _ in data]
    defs = [torch.tensor(d, device=DEV, requires_grad=True) for _, d in data]
    outs = diso_amd.BatchExtractor()(sdfs, algos, defs, return_quads=True)
    ups = [torch.randn(v.shape, device=DEV) for v, _ in outs]
    torch.autograd.backward([v for v, _ in outs], grad_tensors=ups)
    for i, (v, f) in enumerate(outs):
        check_against_oracle(algos[i], data[i][0], data[i][1], v, f, ups[i], sdfs[i].grad, defs[i].grad)
