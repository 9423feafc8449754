"""The packed case tables: provenance against the reference text and geometric invariants.

The invariants are derived from geometry alone (corner positions, edge endpoints), so a table
entry transcribed wrongly (a swapped edge, a wrong patch, a wrong ambiguity direction) fails
here even without the reference present.
"""
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

REF = "/root/reference/src"
MC_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
DMC_CORNERS = np.array([[i & 1, (i >> 1) & 1, (i >> 2) & 1] for i in range(8)])
EDGE_OFS = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1], [0, 0, 0], [0, 1, 0], [1, 1, 0],
                     [0, 1, 1], [0, 1, 0], [0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]])
EDGE_AXIS = np.array([0, 2, 0, 2, 0, 2, 0, 2, 1, 1, 1, 1])
QUAD = np.array([
    [[0, 0, 0, 0], [0, -1, 0, 4], [0, -1, -1, 6], [0, 0, -1, 2]],
    [[0, 0, 0, 8], [0, 0, -1, 11], [-1, 0, -1, 10], [-1, 0, 0, 9]],
    [[0, 0, 0, 3], [-1, 0, 0, 1], [-1, -1, 0, 5], [0, -1, 0, 7]],
    [[0, 0, 0, 0], [0, 0, -1, 2], [0, -1, -1, 6], [0, -1, 0, 4]],
    [[0, 0, 0, 8], [-1, 0, 0, 9], [-1, 0, -1, 10], [0, 0, -1, 11]],
    [[0, 0, 0, 3], [0, -1, 0, 7], [-1, -1, 0, 5], [-1, 0, 0, 1]]])


@pytest.fixture(scope="module")
def tables():
    return O.expand_tables()


def edge_corners(corners):
    out = []
    for e in range(12):
        a = EDGE_OFS[e]
        b = a.copy()
        b[EDGE_AXIS[e]] += 1
        out.append((int(np.where((corners == a).all(1))[0][0]), int(np.where((corners == b).all(1))[0][0])))
    return out


def crossing(code, ec):
    return {e for e, (a, b) in enumerate(ec) if ((code >> a) & 1) != ((code >> b) & 1)}


def _grab(text, name):
    i = text.rindex(name) if name.startswith("mcCorners") else text.index(name)
    j = text.index("{", i)
    depth, k = 0, j
    while True:
        depth += text[k] == "{"
        depth -= text[k] == "}"
        if depth == 0:
            break
        k += 1
    return np.array([int(x) for x in re.findall(r"-?\d+", text[j:k + 1])])


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not mounted")
def test_tables_equal_reference_text(tables):
    mc = open(os.path.join(REF, "cumc.cu")).read()
    dmc = open(os.path.join(REF, "cudualmc.cu")).read()
    assert np.array_equal(_grab(mc, "firstMarchingCubesId[257]"), tables["fmi"])
    assert np.array_equal(_grab(mc, "marchingCubesIds[2460]"), tables["ids"])
    assert np.array_equal(_grab(dmc, "mcFirstPatchIndex[257]"), tables["fpi"])
    assert np.array_equal(_grab(dmc, "mcFirstEdgeIndex[359]"), tables["fei"])
    assert np.array_equal(_grab(dmc, "mcEdgeIndex[1536]"), tables["ei"])
    assert np.array_equal(_grab(dmc, "dmcEdgeOffset[256][12]").reshape(256, 12), tables["deo"])
    assert np.array_equal(_grab(dmc, "problematicConfigs[256]"), tables["pc"])
    assert np.array_equal(_grab(mc, "problematicConfigs[256]"), tables["pc"])
    assert np.array_equal(_grab(dmc, "dmcQuad[6][4][4]").reshape(6, 4, 4), QUAD)
    assert np.array_equal(_grab(mc, "mcEdgeLocations[12][4]").reshape(12, 4)[:, :3], EDGE_OFS)
    assert np.array_equal(_grab(mc, "mcEdgeLocations[12][4]").reshape(12, 4)[:, 3], EDGE_AXIS)
    assert np.array_equal(_grab(mc, "mcCorners[8][3] =").reshape(8, 3), MC_CORNERS)
    assert np.array_equal(_grab(dmc, "mcCorners[8][3] =").reshape(8, 3), DMC_CORNERS)


def test_mc_triangles_use_exactly_the_crossing_edges(tables):
    ec = edge_corners(MC_CORNERS)
    fmi, ids = tables["fmi"], tables["ids"]
    ntri = np.diff(fmi) // 3
    assert fmi[-1] == 2460 and ntri.sum() == 820 and ntri.max() == 5
    for code in range(256):
        tri = ids[fmi[code]:fmi[code + 1]]
        assert len(tri) % 3 == 0
        assert set(tri.tolist()) == crossing(code, ec), code
        # no degenerate triangle: three distinct edges each
        for t in tri.reshape(-1, 3):
            assert len(set(t.tolist())) == 3


def test_mc_triangle_orientation_is_consistent(tables):
    """Winding is fixed by the geometry: with edge midpoints as vertices, each triangle's normal
    has the same sign against the inside->outside direction of its own three edges in every
    case (all triangles face outwards, or all inwards)."""
    fmi, ids = tables["fmi"], tables["ids"]
    ec = edge_corners(MC_CORNERS)
    mid = EDGE_OFS + 0.5 * np.eye(3)[EDGE_AXIS]
    signs = []
    for code in range(1, 255):
        for tri in ids[fmi[code]:fmi[code + 1]].reshape(-1, 3):
            a, b, c = tri
            n = np.cross(mid[b] - mid[a], mid[c] - mid[a])
            flux = 0.0
            for e in tri:
                i0, i1 = ec[e]
                inside, outside = (i0, i1) if (code >> i0) & 1 else (i1, i0)
                flux += np.dot(n, MC_CORNERS[outside] - MC_CORNERS[inside])
            signs.append(np.sign(flux))
    assert set(signs) == {signs[0]} and signs[0] != 0


def test_dmc_patches_partition_the_crossing_edges(tables):
    ec = edge_corners(DMC_CORNERS)
    fpi, fei, ei, deo = tables["fpi"], tables["fei"], tables["ei"], tables["deo"]
    assert fpi[-1] == 358
    assert np.bincount(np.diff(fpi), minlength=5).tolist() == [2, 162, 82, 8, 2]
    for code in range(256):
        seen = []
        for k, p in enumerate(range(fpi[code], fpi[code + 1])):
            es = ei[fei[p]:fei[p + 1]].tolist()
            assert es == sorted(es) and 3 <= len(es) <= 7
            assert all(deo[code, e] == k for e in es)
            seen += es
        assert sorted(seen) == sorted(crossing(code, ec)), code
        assert all(deo[code, e] == -1 for e in range(12) if e not in seen)


def test_dmc_patches_are_the_mc_surface_components(tables):
    """For every corner configuration, the DMC patches are exactly the connected components (by
    shared edges) of the MC triangulation of the same inside corners: the two tables agree on
    which crossing edges belong to one surface sheet, including on ambiguous faces."""
    fmi, ids, fpi, fei, ei = tables["fmi"], tables["ids"], tables["fpi"], tables["fei"], tables["ei"]
    for code in range(256):
        inside = {tuple(DMC_CORNERS[i]) for i in range(8) if (code >> i) & 1}
        mcode = sum(1 << j for j in range(8) if tuple(MC_CORNERS[j]) in inside)
        parent = {}

        def root(x):
            while parent.setdefault(x, x) != x:
                x = parent[x]
            return x

        tris = ids[fmi[mcode]:fmi[mcode + 1]].reshape(-1, 3)
        for a, b, c in tris:
            parent[root(b)] = root(a)
            parent[root(c)] = root(a)
        comps = {}
        for e in set(tris.flatten().tolist()):
            comps.setdefault(root(e), []).append(e)
        comps = sorted(sorted(c) for c in comps.values())
        patches = sorted(sorted(ei[fei[p]:fei[p + 1]].tolist()) for p in range(fpi[code], fpi[code + 1]))
        assert comps == patches, code


def test_ambiguity_table(tables):
    pc = tables["pc"]
    amb = np.where(pc != 255)[0]
    assert len(amb) == 36 and pc[amb].max() <= 5
    fpi = tables["fpi"]
    for code in amb:
        # ambiguous cases have one patch and their complement two
        assert fpi[code + 1] - fpi[code] == 1
        assert fpi[(code ^ 255) + 1] - fpi[code ^ 255] == 2
        # the tabulated face neighbour shares the face on which the case is ambiguous: that face's
        # four corners alternate in sign (a saddle)
        comp, side = pc[code] >> 1, pc[code] & 1
        face = [i for i in range(8) if DMC_CORNERS[i][comp] == side]
        bits = {i: (code >> i) & 1 for i in face}
        diag = [(a, b) for a in face for b in face if a < b and np.abs(DMC_CORNERS[a] - DMC_CORNERS[b]).sum() == 2]
        assert all(bits[a] == bits[b] for a, b in diag)
        assert len(set(bits.values())) == 2


def test_quad_stencil_cells_share_the_edge():
    for ty in range(6):
        axis = ty % 3
        for dx, dy, dz, eid in QUAD[ty]:
            # the cell at (dx,dy,dz) holds the owner's axis edge as its edge `eid`
            assert EDGE_AXIS[eid] == axis
            assert np.array_equal(np.array([dx, dy, dz]) + EDGE_OFS[eid], [0, 0, 0])
        # exiting quads list the same cells in the opposite cyclic order
        if ty >= 3:
            assert np.array_equal(QUAD[ty][[0, 3, 2, 1]], QUAD[ty - 3])
