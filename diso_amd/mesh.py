"""Host-side mesh export and validation for extracted surfaces.

The reference's example writes its meshes with trimesh (test/example.py:72-73,86-87,98-99); here
OBJ and PLY are written directly from the tensors, and `check` reports the topological properties
an iso-surface of a closed shape must have. Faces may be triangles [F,3] or DiffDMC quads [Q,4].
Inputs are torch tensors or numpy arrays; nothing here touches the GPU besides the one copy to
the host.
"""
import numpy as np


def _np(a):
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


def write_obj(path, verts, faces):
    """Wavefront OBJ: `v x y z` lines, then `f a b c [d]` lines with 1-based indices."""
    v = _np(verts)
    g = "%.17g" if v.dtype == np.float64 else "%.9g"  # enough digits to read the same value back
    v = v.astype(np.float64)
    f = _np(faces).astype(np.int64) + 1
    with open(path, "w") as fh:
        fh.write(f"# {len(v)} vertices, {len(f)} faces\n")
        if len(v):
            np.savetxt(fh, v, fmt=f"v {g} {g} {g}")
        if len(f):
            np.savetxt(fh, f, fmt="f" + " %d" * f.shape[1])


def read_obj(path):
    """The vertices and faces of an OBJ written by write_obj (or any OBJ with plain `v`/`f` lines;
    `f a/b/c` forms keep the vertex index)."""
    verts, faces = [], []
    with open(path) as fh:
        for ln in fh:
            p = ln.split()
            if not p:
                continue
            if p[0] == "v":
                verts.append([float(x) for x in p[1:4]])
            elif p[0] == "f":
                faces.append([int(x.split("/")[0]) - 1 for x in p[1:]])
    k = len(faces[0]) if faces else 3
    return np.array(verts, np.float64).reshape(-1, 3), np.array(faces, np.int64).reshape(-1, k)


def write_ply(path, verts, faces, binary=True):
    """PLY with float/double vertex coordinates (as given) and int32 face lists."""
    v = _np(verts)
    v = v.astype(np.float64 if v.dtype == np.float64 else np.float32)
    f = _np(faces).astype(np.int64)
    if len(f) and (f.min() < 0 or f.max() >= 2 ** 31):
        raise ValueError("face indices do not fit PLY int32")
    k = f.shape[1] if f.ndim == 2 else 3
    vt = "double" if v.dtype == np.float64 else "float"
    head = ["ply", "format " + ("binary_little_endian" if binary else "ascii") + " 1.0",
            f"element vertex {len(v)}", f"property {vt} x", f"property {vt} y", f"property {vt} z",
            f"element face {len(f)}", "property list uchar int vertex_indices", "end_header"]
    with open(path, "wb") as fh:
        fh.write(("\n".join(head) + "\n").encode())
        if binary:
            fh.write(np.ascontiguousarray(v, v.dtype.newbyteorder("<")).tobytes())
            rec = np.zeros(len(f), dtype=[("n", "u1"), ("i", "<i4", (k,))])
            rec["n"] = k
            rec["i"] = f
            fh.write(rec.tobytes())
        else:
            for p in v:
                fh.write((" ".join(repr(float(x)) for x in p) + "\n").encode())
            for q in f:
                fh.write((f"{k} " + " ".join(str(int(i)) for i in q) + "\n").encode())


def read_ply(path):
    """The vertices and faces of a PLY written by write_ply."""
    with open(path, "rb") as fh:
        head = []
        while True:
            ln = fh.readline().decode().strip()
            head.append(ln)
            if ln == "end_header":
                break
        nv = int(next(h for h in head if h.startswith("element vertex")).split()[-1])
        nf = int(next(h for h in head if h.startswith("element face")).split()[-1])
        vt = np.float64 if any(h == "property double x" for h in head) else np.float32
        if any("binary_little_endian" in h for h in head):
            v = np.frombuffer(fh.read(nv * 3 * np.dtype(vt).itemsize), dtype=np.dtype(vt).newbyteorder("<"))
            v = v.reshape(nv, 3).astype(vt)
            rest = fh.read()
            if nf == 0:
                return v, np.zeros((0, 3), np.int64)
            k = rest[0]
            rec = np.frombuffer(rest, dtype=[("n", "u1"), ("i", "<i4", (k,))], count=nf)
            return v, rec["i"].astype(np.int64)
        lines = fh.read().decode().split("\n")
        v = np.array([[float(x) for x in lines[i].split()] for i in range(nv)], vt).reshape(-1, 3)
        f = [[int(x) for x in lines[nv + i].split()[1:]] for i in range(nf)]
        return v, np.array(f, np.int64).reshape(nf, -1) if nf else np.zeros((0, 3), np.int64)


def check(faces, n_verts=None):
    """Topology report of a polygon mesh (triangles or quads):
      n_faces, n_verts (referenced), n_edges, boundary_edges (used by one face),
      nonmanifold_edges (used by more than two faces), misoriented_edges (a directed edge used
      twice, i.e. two neighbours with opposite winding), euler (V - E + F over referenced
      vertices), components (face-connected), degenerate_faces (a repeated corner);
      closed = no boundary edge, manifold = no non-manifold edge,
      oriented = closed and manifold and no misoriented edge,
      watertight = closed and manifold."""
    f = _np(faces).astype(np.int64)
    if f.ndim != 2 or (len(f) and f.shape[1] < 3):
        raise ValueError("faces must be [F, k] with k >= 3")
    nf, k = f.shape if len(f) else (0, 3)
    used = np.unique(f) if nf else np.zeros(0, np.int64)
    if n_verts is not None and nf and used.max() >= n_verts:
        raise ValueError("a face index is out of range")
    a = f.reshape(-1)
    b = np.roll(f, -1, axis=1).reshape(-1)
    fid = np.repeat(np.arange(nf), k)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    und, inv, cnt = np.unique(np.stack([lo, hi], 1), axis=0, return_inverse=True, return_counts=True)
    inv = inv.reshape(-1)
    _, dcnt = np.unique(np.stack([a, b], 1), axis=0, return_counts=True)
    # face components via union-find over shared undirected edges
    parent = np.arange(nf)

    def find(x):
        r = x
        while parent[r] != r:
            r = parent[r]
        while parent[x] != r:
            parent[x], x = r, parent[x]
        return r

    order = np.argsort(inv, kind="stable")
    for i in range(1, len(order)):
        if inv[order[i]] == inv[order[i - 1]]:
            ra, rb = find(fid[order[i]]), find(fid[order[i - 1]])
            if ra != rb:
                parent[ra] = rb
    comps = len({find(i) for i in range(nf)}) if nf else 0
    degenerate = int(sum(len(set(row)) < k for row in f.tolist())) if nf else 0
    rep = {
        "n_faces": int(nf), "n_verts": int(len(used)), "n_edges": int(len(und)),
        "boundary_edges": int((cnt == 1).sum()), "nonmanifold_edges": int((cnt > 2).sum()),
        "misoriented_edges": int((dcnt > 1).sum()), "euler": int(len(used) - len(und) + nf),
        "components": int(comps), "degenerate_faces": degenerate,
    }
    rep["closed"] = rep["boundary_edges"] == 0
    rep["manifold"] = rep["nonmanifold_edges"] == 0
    rep["watertight"] = rep["closed"] and rep["manifold"]
    rep["oriented"] = rep["watertight"] and rep["misoriented_edges"] == 0
    return rep


def signed_volume(verts, faces):
    """Volume enclosed by a closed, oriented mesh (quads fanned into triangles); positive when
    faces wind counter-clockwise seen from outside."""
    v = _np(verts).astype(np.float64)
    f = _np(faces).astype(np.int64)
    vol = 0.0
    for j in range(1, f.shape[1] - 1):
        p0, p1, p2 = v[f[:, 0]], v[f[:, j]], v[f[:, j + 1]]
        vol += float(np.einsum("ij,ij->i", p0, np.cross(p1, p2)).sum()) / 6.0
    return vol
