"""Pins the CPU oracle (the checker) with facts that do not come from its own formulas.

* Counts: vertices = crossing edges of the padded grid, MC triangles = sum of the table's
  triangle counts, DMC vertices = sum of patch counts of the resolved codes - all recomputed
  here with independent vectorised numpy.
* Geometry: on a linear field every MC vertex lies exactly on the zero plane, and every DMC
  vertex too (a mean of points on a plane); every triangle's three vertices lie on edges of one
  cell; DMC vertices lie inside their cell.
* Topology: on smooth closed surfaces the MC mesh and the DMC quad mesh are closed, manifold
  and consistently oriented with Euler characteristic 2, and the enclosed volume matches the
  sphere's.
* Gradients: central finite differences in fp64.
* Known published counts: the 64^3 sphere gives V = 11952, F = 23900 (MC) and Vd = 11954,
  Q = 11952 (DMC).
"""
import numpy as np
import pytest

from diso_amd import workloads
from oracle import oracle as O
from tests.helpers import inputs

MC_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])


@pytest.fixture(scope="module")
def tables():
    return O.expand_tables()


def crossing_edges_np(g, iso):
    """Boolean masks of crossing edges per axis (edge from v to v+e_axis), independent numpy."""
    inside = g >= iso
    out = []
    for ax in range(3):
        a = inside.take(range(g.shape[ax] - 1), axis=ax)
        b = inside.take(range(1, g.shape[ax]), axis=ax)
        out.append(a != b)
    return out


def codes_np(g, iso, order):
    inside = (g >= iso).astype(np.int64)
    X, Y, Z = g.shape
    code = np.zeros((X - 1, Y - 1, Z - 1), np.int64)
    for i in range(8):
        dx, dy, dz = MC_CORNERS[i] if order == "mc" else (i & 1, (i >> 1) & 1, (i >> 2) & 1)
        code |= inside[dx:X - 1 + dx, dy:Y - 1 + dy, dz:Z - 1 + dz] << i
    return code


def good_codes_np(code, tables):
    pc = tables["pc"]
    dims = code.shape
    out = code.copy()
    amb = np.argwhere(pc[code] != 255)
    for x, y, z in amb:
        d = pc[code[x, y, z]]
        n = [x, y, z]
        n[d >> 1] += 1 if d & 1 else -1
        if 0 <= n[d >> 1] < dims[d >> 1] and pc[code[tuple(n)]] != 255:
            out[x, y, z] ^= 255
    return out


@pytest.mark.parametrize("kind,dims,seed", [("noise", (9, 11, 13), 0), ("blobs", (20, 17, 15), 1),
                                            ("noisy_box", (16, 16, 16), 2), ("sphere", (12, 7, 19), 3)])
def test_counts_match_numpy(kind, dims, seed, tables):
    sdf, _ = inputs(dims, kind, seed, np.float64, False)
    g, _ = O.pad(sdf, None, 0.0)
    nx = sum(int(m.sum()) for m in crossing_edges_np(g, 0.0))
    mc_code = codes_np(g, 0.0, "mc")
    used = (mc_code != 0) & (mc_code != 255)
    ntri = (np.diff(tables["fmi"]) // 3)[mc_code].sum()
    dmc_code = good_codes_np(codes_np(g, 0.0, "dmc"), tables)
    npatch = np.diff(tables["fpi"])[dmc_code].sum()
    mc = O.Engine(O.MC, np.float64)
    v, f, st = mc.forward(g, None, 0.0)
    assert st == 0 and mc.n_used() == used.sum()
    assert len(v) == nx and len(f) == ntri
    dmc = O.Engine(O.DMC, np.float64)
    v, q, st = dmc.forward(g, None, 0.0)
    assert st == 0 and len(q) == nx and len(v) == npatch
    codes, idx = dmc.codes()
    G1, G2 = g.shape[1], g.shape[2]
    assert np.array_equal(codes, dmc_code[idx // (G1 * G2), (idx // G2) % G1, idx % G2])


def test_sphere64_known_counts():
    sdf = workloads.sphere(64).astype(np.float32)
    v, f = O.Extractor(O.MC, np.float32).forward(sdf, None, 0.0, True)
    assert v.shape == (11952, 3) and f.shape == (23900, 3)
    v, q = O.Extractor(O.DMC, np.float32).forward(sdf, None, 0.0, True, return_quads=True)
    assert v.shape == (11954, 3) and q.shape == (11952, 4)


@pytest.mark.parametrize("algo", [O.MC, O.DMC])
def test_linear_field_vertices_on_plane(algo):
    """Linear interpolation is exact on a linear field: vertices on edges (MC) and means of such
    vertices (DMC) lie on the zero plane. Vertices on edges into the iso+1 padding are excluded."""
    n = 13
    p = workloads.points(n)
    a = np.array([0.3, -0.7, 0.5])
    sdf = (p @ a + 0.05).astype(np.float64)
    ex = O.Extractor(algo, np.float64)
    # unnormalised output is in grid units: sample (i,j,k) sits at p = -0.5 + (i,j,k)/(n-1)
    v, _ = ex.forward(sdf, None, 0.0, normalize=False, return_quads=True)
    if algo == O.MC:
        keep = np.all((v >= 0) & (v <= n - 1), axis=1)
    else:
        codes, idx = ex.engine.codes()
        G = n + 2
        cell = np.stack([idx // (G * G), (idx // G) % G, idx % G], 1)
        cell = np.repeat(cell, np.diff(O.expand_tables()["fpi"])[codes], axis=0)
        keep = np.all((cell >= 1) & (cell <= n - 1), axis=1)  # all eight corners are data
    assert keep.sum() > 50
    pts = -0.5 + v[keep] / (n - 1)
    assert np.abs(pts @ a + 0.05).max() < 1e-12


def _mesh_checks(faces, nverts):
    """Closed, manifold, consistently oriented polygon mesh: every directed edge exactly once and
    its reverse exactly once."""
    k = faces.shape[1]
    e = np.stack([faces, np.roll(faces, -1, axis=1)], -1).reshape(-1, 2)
    directed = {tuple(x) for x in e.tolist()}
    assert len(directed) == len(e), "a directed edge repeats (non-manifold or flipped face)"
    assert all((b, a) in directed for a, b in directed), "boundary edge (hole)"
    n_edges = len(e) // 2
    used = np.unique(faces)
    assert len(used) == nverts, "unreferenced vertices"
    return nverts - n_edges + len(faces)


@pytest.mark.parametrize("shape", ["sphere", "round_box"])
def test_mc_mesh_closed_manifold_oriented(shape):
    n = 40
    sdf = workloads.sphere(n, 0.33) if shape == "sphere" else workloads.round_box(n, 0.3, 0.1)
    v, f = O.Extractor(O.MC, np.float64).forward(sdf, None, 0.0, normalize=False)
    assert _mesh_checks(f, len(v)) == 2
    # signed volume (divergence theorem) in grid units: sign fixes the winding convention
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0
    h = 1.0 / (n - 1)
    if shape == "sphere":
        expect = 4.0 / 3.0 * np.pi * (0.33 / h) ** 3
        assert abs(abs(vol) / expect - 1) < 0.02
    assert vol > 0  # the tables wind triangles counter-clockwise seen from the sdf >= iso side


def test_dmc_quads_closed_manifold_oriented_and_in_cells():
    n = 40
    sdf = workloads.sphere(n, 0.33)
    ex = O.Extractor(O.DMC, np.float64)
    v, q = ex.forward(sdf, None, 0.0, normalize=False, return_quads=True)
    assert _mesh_checks(q, len(v)) == 2
    # each DMC vertex lies inside (the closure of) the cell that produced it, in padded coords
    codes, idx = ex.engine.codes()
    G = n + 2
    cell = np.stack([idx // (G * G), (idx // G) % G, idx % G], 1)
    npatch = np.diff(O.expand_tables()["fpi"])[codes]
    owner = np.repeat(cell, npatch, axis=0) - 1  # shift like the output
    assert np.all(v >= owner - 1e-12) and np.all(v <= owner + 1 + 1e-12)


def test_mc_triangle_vertices_lie_on_one_cell():
    sdf, _ = inputs((14, 12, 16), "noise", 5, np.float64, False)
    ex = O.Extractor(O.MC, np.float64)
    v, f = ex.forward(sdf, None, 0.0, normalize=False)
    tri = v[f]  # [F,3,3]
    lo = np.floor(tri.min(1) + 1e-9)
    assert np.all(tri.max(1) - lo <= 1 + 1e-9), "triangle spans more than one cell"
    # every vertex lies on a grid edge: two of its three coordinates are integers
    frac = np.abs(v - np.round(v)) < 1e-12
    assert np.all(frac.sum(1) >= 2)


def _loss(ex, sdf, deform, g, algo):
    v, _ = ex.forward(sdf, deform, 0.0, normalize=True, return_quads=True)
    return float((v * g).sum())


@pytest.mark.parametrize("algo", [O.MC, O.DMC])
@pytest.mark.parametrize("with_deform", [False, True])
def test_gradients_match_central_differences(algo, with_deform):
    n = 12
    rng = np.random.default_rng(3)
    # a smooth field keeps |d1 - d0| large on every crossing edge and no sample near iso, so a
    # +-1e-6 step changes no topology and the third-derivative error of the FD stays tiny
    sdf = workloads.sphere(n, 0.31) + 0.02 * np.sin(7 * workloads.points(n)).sum(-1)
    sdf = sdf + np.where(np.abs(sdf) < 1e-3, 2e-3, 0.0)
    deform = rng.uniform(-0.2, 0.2, (n, n, n, 3)) if with_deform else None
    ex = O.Extractor(algo, np.float64)
    v, _ = ex.forward(sdf, deform, 0.0, True, return_quads=True)
    g = rng.standard_normal(v.shape)
    gs, gd = ex.backward(g)
    eps = 1e-6
    touched = np.argwhere(np.abs(gs) > 0)
    pick = touched[rng.choice(len(touched), size=min(40, len(touched)), replace=False)]
    for i, j, k in pick:
        sp, sm = sdf.copy(), sdf.copy()
        sp[i, j, k] += eps
        sm[i, j, k] -= eps
        fd = (_loss(O.Extractor(algo, np.float64), sp, deform, g, algo) -
              _loss(O.Extractor(algo, np.float64), sm, deform, g, algo)) / (2 * eps)
        assert abs(fd - gs[i, j, k]) <= 1e-6 * max(1.0, abs(fd)), (i, j, k, fd, gs[i, j, k])
    if with_deform:
        tdef = np.argwhere(np.abs(gd) > 0)
        for i, j, k, c in tdef[rng.choice(len(tdef), size=min(40, len(tdef)), replace=False)]:
            dp, dm = deform.copy(), deform.copy()
            dp[i, j, k, c] += eps
            dm[i, j, k, c] -= eps
            fd = (_loss(O.Extractor(algo, np.float64), sdf, dp, g, algo) -
                  _loss(O.Extractor(algo, np.float64), sdf, dm, g, algo)) / (2 * eps)
            assert abs(fd - gd[i, j, k, c]) <= 1e-6 * max(1.0, abs(fd))


def test_dmc_reference_compat_differs_only_in_multi_patch_cells():
    # on a smooth sphere every used cell has one patch: both backward modes coincide
    sdf = workloads.sphere(24, 0.35)
    ex = O.Extractor(O.DMC, np.float64)
    v, _ = ex.forward(sdf, None, 0.0, True, return_quads=True)
    codes, _ = ex.engine.codes()
    assert np.diff(O.expand_tables()["fpi"])[codes].max() == 1
    g = np.random.default_rng(0).standard_normal(v.shape)
    a, _ = ex.backward(g)
    b, _ = ex.backward(g, reference_compat=True)
    assert np.array_equal(a, b)
    # on noise many cells have several patches and the modes differ
    sdf, _ = inputs((10, 10, 10), "noise", 1, np.float64, False)
    v, _ = ex.forward(sdf, None, 0.0, True, return_quads=True)
    g = np.random.default_rng(0).standard_normal(v.shape)
    a, _ = ex.backward(g)
    b, _ = ex.backward(g, reference_compat=True)
    assert not np.allclose(a, b)


@pytest.mark.parametrize("with_deform", [False, True])
def test_dmc_reference_compat_is_true_backward_of_first_patch_gradients(with_deform):
    """The reference's DMC backward reads each cell's first patch-vertex gradient for every patch
    of the cell (src/cudualmc.cu:990). So compat(g) equals the true backward of g', where g' copies
    the gradient of each cell's first patch vertex to all of that cell's patch vertices. Pins
    which patch compat reads (the first, not the last or its own) independently of the
    compat code path; the vertex ranges come from the cells' codes and the patch-count table."""
    sdf, deform = inputs((11, 10, 12), "noise", 5, np.float64, with_deform)
    ex = O.Extractor(O.DMC, np.float64)
    v, _ = ex.forward(sdf, deform, 0.0, True, return_quads=True)
    codes, _ = ex.engine.codes()
    npatch = np.diff(O.expand_tables()["fpi"])[codes]
    assert npatch.min() >= 1 and npatch.sum() == v.shape[0]
    assert npatch.max() >= 2  # multi-patch cells present
    first = np.concatenate([[0], np.cumsum(npatch)[:-1]])
    cell_of_vertex = np.repeat(np.arange(len(codes)), npatch)
    g = np.random.default_rng(3).standard_normal(v.shape)
    g1 = g[first[cell_of_vertex]]
    a_s, a_d = ex.backward(g, reference_compat=True)
    b_s, b_d = ex.backward(g1)
    np.testing.assert_allclose(a_s, b_s, rtol=1e-12, atol=1e-12 * np.abs(b_s).max())
    if with_deform:
        np.testing.assert_allclose(a_d, b_d, rtol=1e-12, atol=1e-12 * np.abs(b_d).max())
    # and g' differs from g, so the identity is not the trivial one
    c_s, _ = ex.backward(g)
    assert not np.allclose(a_s, c_s)


def test_dmc_correct_mode_matches_fd_on_multi_patch_cells():
    """Finite differences on a field with multi-patch cells agree with the default backward."""
    rng = np.random.default_rng(11)
    n = 8
    sdf = rng.uniform(-1, 1, (n, n, n))
    sdf = np.where(np.abs(sdf) < 0.05, np.sign(sdf) * 0.05, sdf)  # no sample near iso
    ex = O.Extractor(O.DMC, np.float64)
    v, _ = ex.forward(sdf, None, 0.0, True, return_quads=True)
    codes, _ = ex.engine.codes()
    assert np.diff(O.expand_tables()["fpi"])[codes].max() >= 2
    g = rng.standard_normal(v.shape)
    gs, _ = ex.backward(g)
    eps = 1e-7
    for i, j, k in np.argwhere(np.abs(gs) > 0)[:60]:
        sp, sm = sdf.copy(), sdf.copy()
        sp[i, j, k] += eps
        sm[i, j, k] -= eps
        fd = (_loss(O.Extractor(O.DMC, np.float64), sp, None, g, O.DMC) -
              _loss(O.Extractor(O.DMC, np.float64), sm, None, g, O.DMC)) / (2 * eps)
        assert abs(fd - gs[i, j, k]) <= 1e-5 * max(1.0, abs(fd))


def test_wrapper_semantics():
    # early-out uses min >= iso or max <= iso, in the grid dtype
    assert O.is_empty(np.full((3, 3, 3), 0.5, np.float32), 0.0)
    g = np.full((3, 3, 3), 0.5, np.float32)
    g[1, 1, 1] = 0.0
    assert O.is_empty(g, 0.0)  # min == iso
    g[1, 1, 1] = -0.1
    assert not O.is_empty(g, 0.0)
    # pad value is (dtype)(iso + 1)
    p, _ = O.pad(np.zeros((2, 2, 2), np.float32), None, 0.1)
    assert p[0, 0, 0] == np.float32(1.1)
    # empty results: (0,3) vertices and int32 faces, (0,4) for DMC even for triangles
    v, f = O.Extractor(O.DMC, np.float32).forward(np.ones((4, 4, 4), np.float32), None, 0.0)
    assert v.shape == (0, 3) and f.shape == (0, 4) and f.dtype == np.int32


def test_split_known_answer():
    """A planar kite whose short diagonal (1-3) gives two fat triangles and whose long diagonal
    (0-2) gives two thin ones: cut (0,1,3)+(1,2,3) must be chosen, listed first."""
    verts = np.array([[0.0, 0, 0], [1.0, 0.3, 0], [4.0, 0, 0], [1.0, -0.3, 0]])
    quads = np.array([[0, 1, 2, 3], [0, 3, 2, 1]])
    faces, s1, s2 = O.split_quads(verts, quads)
    # quad 0: cut along 1-3 => triangles (0,1,3),(1,2,3)
    assert s1[0] < s2[0]
    assert faces[:2].tolist() == [[0, 1, 3], [1, 2, 3]]
    # a square ties: the second cut is taken
    sq = np.array([[0.0, 0, 0], [1.0, 0, 0], [1.0, 1, 0], [0.0, 1, 0]])
    faces, s1, s2 = O.split_quads(sq, np.array([[0, 1, 2, 3]]))
    assert faces.tolist() == [[0, 1, 2], [0, 2, 3]]
