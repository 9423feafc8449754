"""Mesh export (OBJ / PLY round trips) and topology validation on hand-built meshes with known
answers, and on oracle iso-surfaces."""
import numpy as np
import pytest

from diso_amd import mesh
from oracle import oracle as O
from tests.helpers import field

CUBE_V = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], np.float64)
# outward-facing quads of the unit cube (vertex id = 4x + 2y + z)
CUBE_Q = np.array([[0, 1, 3, 2], [4, 6, 7, 5], [0, 4, 5, 1], [2, 3, 7, 6], [0, 2, 6, 4], [1, 5, 7, 3]])
TET_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)
TET_F = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]])


def test_cube_and_tetrahedron_reports():
    r = mesh.check(CUBE_Q)
    assert r["oriented"] and r["watertight"] and r["euler"] == 2 and r["components"] == 1
    assert (r["n_faces"], r["n_verts"], r["n_edges"]) == (6, 8, 12)
    assert mesh.signed_volume(CUBE_V, CUBE_Q) == pytest.approx(1.0)
    r = mesh.check(TET_F)
    assert r["oriented"] and r["euler"] == 2
    assert mesh.signed_volume(TET_V, TET_F) == pytest.approx(1 / 6)


def test_defects_are_reported():
    r = mesh.check(TET_F[:3])  # one face removed
    assert not r["closed"] and r["boundary_edges"] == 3 and r["manifold"]
    flipped = TET_F.copy()
    flipped[0] = flipped[0][::-1]
    r = mesh.check(flipped)
    assert r["watertight"] and not r["oriented"] and r["misoriented_edges"] == 3
    fan = np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]])  # three triangles on edge (0, 1)
    assert mesh.check(fan)["nonmanifold_edges"] == 1
    two = np.concatenate([TET_F, TET_F + 4])
    assert mesh.check(two)["components"] == 2 and mesh.check(two)["euler"] == 4
    assert mesh.check(np.array([[0, 0, 1]]))["degenerate_faces"] == 1
    with pytest.raises(ValueError):
        mesh.check(TET_F, n_verts=3)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("faces", [TET_F, CUBE_Q, np.zeros((0, 3), np.int64)])
def test_obj_and_ply_round_trips(tmp_path, dtype, faces):
    rng = np.random.default_rng(0)
    v = rng.standard_normal((8, 3)).astype(dtype)
    mesh.write_obj(tmp_path / "m.obj", v, faces)
    v2, f2 = mesh.read_obj(tmp_path / "m.obj")
    assert np.array_equal(v2.astype(dtype), v) and np.array_equal(f2.reshape(faces.shape), faces)
    for binary in (True, False):
        mesh.write_ply(tmp_path / "m.ply", v, faces, binary=binary)
        v3, f3 = mesh.read_ply(tmp_path / "m.ply")
        assert v3.dtype == dtype and np.array_equal(v3, v)
        assert np.array_equal(f3.reshape(faces.shape), faces)


def test_oracle_iso_surfaces_are_watertight(tmp_path):
    sdf = field((40, 36, 44), "sphere", 0)
    for algo, k in ((O.MC, 3), (O.DMC, 4)):
        v, f = O.Extractor(algo, np.float64).forward(sdf, None, 0.0, True, return_quads=True)
        r = mesh.check(f, len(v))
        assert f.shape[1] == k and r["oriented"] and r["euler"] == 2 and r["components"] == 1
        assert r["n_verts"] == len(v)
        mesh.write_ply(tmp_path / "s.ply", v, f)
        v2, f2 = mesh.read_ply(tmp_path / "s.ply")
        assert np.array_equal(v2, v) and np.array_equal(f2, f)
