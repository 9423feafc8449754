"""Python face of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module. It binds
oracle/liboracle.so (the C restatement of the reference engine, src/cumc.cu and src/cudualmc.cu)
and restates, in numpy, what the reference's Python wrapper does around that engine
(diso/__init__.py:48-147): the empty-grid early-out, the iso+1 / zero padding, the `-1` shift,
the per-axis normalisation, the int64 widening, the DiffDMC quad->triangle split, and the chain
rule that carries an upstream vertex gradient back to the unpadded sdf and deform.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.environ.get("DISO_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")  # override: a sanitizer build
_lib = None

MC, DMC = 0, 1


def build():
    """Compiles liboracle.so from oracle.c (gcc, OpenMP)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(_LIB_PATH)
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_destroy.argtypes = [vp]
        L.oracle_forward.restype = ctypes.c_int
        L.oracle_forward.argtypes = [vp, vp, vp, i64, i64, i64, ctypes.c_double]
        for f in ("oracle_n_verts", "oracle_n_faces", "oracle_n_used"):
            getattr(L, f).restype = i64
            getattr(L, f).argtypes = [vp]
        L.oracle_copy_out.argtypes = [vp, vp, vp]
        L.oracle_copy_codes.argtypes = [vp, vp, vp]
        L.oracle_backward.restype = ctypes.c_int
        L.oracle_backward.argtypes = [vp, vp, vp, ctypes.c_double, vp, vp, vp, ctypes.c_int]
        L.oracle_expand_tables.argtypes = [vp] * 7
        L.oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def expand_tables():
    """The packed case tables expanded to the reference's layouts (see oracle.c)."""
    arrs = dict(
        fmi=np.zeros(257, np.int32), ids=np.zeros(2460, np.int32), fpi=np.zeros(257, np.int32),
        fei=np.zeros(359, np.int32), ei=np.zeros(1536, np.int32),
        deo=np.zeros(256 * 12, np.int32), pc=np.zeros(256, np.int32))
    lib().oracle_expand_tables(*[_ptr(arrs[k]) for k in ("fmi", "ids", "fpi", "fei", "ei", "deo", "pc")])
    arrs["deo"] = arrs["deo"].reshape(256, 12)
    return arrs


class Engine:
    """The reference's C++ extractor (CuMC / CUDualMC) on an already padded grid.

    forward() returns (verts [V,3] in padded grid coordinates, faces [F,3|4] int64, status);
    backward() must follow a forward on the same inputs, as in the reference."""

    def __init__(self, algo, dtype, threads=1):
        self.algo = algo
        self.dtype = np.dtype(dtype)
        code = 0 if self.dtype == np.float32 else 1
        self.h = lib().oracle_create(algo, code, threads)
        if not self.h:
            raise ValueError("bad oracle configuration")

    def __del__(self):
        h, self.h = getattr(self, "h", None), None
        if h and _lib is not None:
            _lib.oracle_destroy(h)

    def forward(self, grid, deform, iso):
        grid = np.ascontiguousarray(grid, self.dtype)
        deform = None if deform is None else np.ascontiguousarray(deform, self.dtype)
        self._keep = (grid, deform)
        X, Y, Z = grid.shape
        st = lib().oracle_forward(self.h, _ptr(grid), _ptr(deform), X, Y, Z, float(iso))
        if st < 0:
            raise MemoryError("oracle allocation failed")
        nv, nf = lib().oracle_n_verts(self.h), lib().oracle_n_faces(self.h)
        verts = np.zeros((nv, 3), self.dtype)
        faces = np.zeros((nf, 3 if self.algo == MC else 4), np.int64)
        lib().oracle_copy_out(self.h, _ptr(verts), _ptr(faces))
        return verts, faces, st

    def n_used(self):
        return lib().oracle_n_used(self.h)

    def codes(self):
        n = self.n_used()
        codes = np.zeros(n, np.uint8)
        idx = np.zeros(n, np.int64)
        lib().oracle_copy_codes(self.h, _ptr(codes), _ptr(idx))
        return codes, idx

    def backward(self, grid, deform, iso, adj_verts, reference_compat=False):
        grid, deform = self._keep
        adj_verts = np.ascontiguousarray(adj_verts, self.dtype)
        adj_grid = np.zeros(grid.shape, self.dtype)
        adj_deform = None if deform is None else np.zeros(deform.shape, self.dtype)
        lib().oracle_backward(self.h, _ptr(grid), _ptr(deform), float(iso), _ptr(adj_verts),
                              _ptr(adj_grid), _ptr(adj_deform), int(bool(reference_compat)))
        return adj_grid, adj_deform


def pad(sdf, deform, iso):
    """F.pad(grid, 1, value=iso+1) and F.pad(deform, 0) (diso/__init__.py:52-54)."""
    s = sdf.dtype.type
    g = np.pad(sdf, 1, mode="constant", constant_values=s(float(iso) + 1.0))
    d = None if deform is None else np.pad(deform, ((1, 1), (1, 1), (1, 1), (0, 0)))
    return g, d


def is_empty(sdf, iso):
    """The early-out of diso/__init__.py:49,103, compared in the grid's dtype."""
    s = sdf.dtype.type
    return bool(sdf.min() >= s(iso) or sdf.max() <= s(iso))


def _epilogue(verts, shape, normalize):
    s = verts.dtype.type
    verts = verts - s(1)
    if normalize:
        verts = verts / (np.array(shape, dtype=verts.dtype) - s(1))
    return verts


def split_quads(verts, quads):
    """The DiffDMC quad->triangle split (diso/__init__.py:117-147): each quad is cut along the
    diagonal whose worse triangle has the smaller maximum corner cosine; all quads that take cut
    (0,1,3)+(1,2,3) come first, in quad order, then all quads that take (0,1,2)+(0,2,3).
    Returns (faces [2Q,3] int64, score1, score2)."""
    s = verts.dtype.type

    def normalize(v):
        n = np.sqrt((v * v).sum(-1, keepdims=True))
        return v / np.maximum(n, s(1e-12))

    def score(cfg):
        worst = []
        for a, b, c in cfg:
            v0, v1, v2 = verts[quads[:, a]], verts[quads[:, b]], verts[quads[:, c]]
            c1 = (normalize(v1 - v0) * normalize(v2 - v0)).sum(-1)
            c2 = (normalize(v2 - v1) * normalize(v0 - v1)).sum(-1)
            c3 = (normalize(v0 - v2) * normalize(v1 - v2)).sum(-1)
            worst.append(np.maximum(np.maximum(c1, c2), c3))
        return np.maximum(worst[0], worst[1])

    s1 = score([(0, 1, 3), (1, 2, 3)])
    s2 = score([(0, 1, 2), (0, 2, 3)])
    m = s1 < s2
    f1 = quads[m][:, [0, 1, 3, 1, 2, 3]].reshape(-1, 3)
    f2 = quads[~m][:, [0, 1, 2, 0, 2, 3]].reshape(-1, 3)
    return np.concatenate([f1, f2], 0).astype(np.int64), s1, s2


class Extractor:
    """DiffMC / DiffDMC forward and backward as the reference computes them, on the CPU.

    forward(sdf [X,Y,Z], deform [X,Y,Z,3] | None, iso, normalize, return_quads)
      -> (verts, faces) with the wrapper's exact dtype semantics.
    backward(g) -> (grad_sdf [X,Y,Z], grad_deform [X,Y,Z,3] | None) for upstream dL/dverts = g.
    """

    def __init__(self, algo, dtype=np.float32, threads=1):
        self.algo = algo
        self.dtype = np.dtype(dtype)
        self.engine = Engine(algo, dtype, threads)

    def forward(self, sdf, deform=None, iso=0.0, normalize=True, return_quads=False):
        sdf = np.ascontiguousarray(sdf, self.dtype)
        deform = None if deform is None else np.ascontiguousarray(deform, self.dtype)
        self.shape, self.iso, self.normalize = sdf.shape, iso, normalize
        self.has_deform = deform is not None
        k = 3 if (self.algo == MC or not return_quads) else 4
        if is_empty(sdf, iso):
            self.empty = True
            return np.zeros((0, 3), self.dtype), np.zeros((0, 4 if self.algo == DMC else 3), np.int32)
        self.empty = False
        g, d = pad(sdf, deform, iso)
        self.padded = (g, d)
        verts, faces, self.status = self.engine.forward(g, d, iso)
        verts = _epilogue(verts, sdf.shape, normalize)
        if self.algo == DMC and not return_quads:
            faces, _, _ = split_quads(verts, faces)
        assert faces.shape[1] == k
        return verts, faces

    def backward(self, g, reference_compat=False):
        if self.empty:
            return np.zeros(self.shape, self.dtype), (np.zeros(self.shape + (3,), self.dtype)
                                                       if self.has_deform else None)
        g = np.ascontiguousarray(g, self.dtype)
        if self.normalize:
            g = g / (np.array(self.shape, dtype=self.dtype) - self.dtype.type(1))
        pg, pdef = self.padded
        ag, ad = self.engine.backward(pg, pdef, self.iso, g, reference_compat)
        ag = ag[1:-1, 1:-1, 1:-1]
        ad = None if ad is None else ad[1:-1, 1:-1, 1:-1]
        return np.ascontiguousarray(ag), (None if ad is None else np.ascontiguousarray(ad))
