"""Fixed-capacity extraction: no host synchronisation and no data-dependent shapes.

A plain DiffMC / DiffDMC forward waits once for the counting pass to size its outputs. A
`FixedExtractor` writes into outputs of a fixed capacity instead and leaves the counts on the
device, so a training step built on it enqueues work only: it can be captured once in a HIP graph
(torch.cuda.graph) and replayed.

    ex = FixedExtractor("dmc", torch.float32, vert_capacity=V, face_capacity=Q)
    verts, faces, counts = ex(sdf, deform)     # verts [V,3]; counts = (vertices, faces, flags)
    live = torch.arange(V, device=dev) < counts[0]
    loss = torch.where(live, f(verts), 0).sum() / counts[0]
    loss.backward()

The first counts[0] rows of `verts` and counts[1] rows of `faces` are exactly what DiffMC /
DiffDMC(return_quads=True) return for the same input; vertex rows past the count are zero, face
rows past it are undefined. The backward is the same gather (with `recompute_backward` the
classification is re-derived from the samples). Not applied here, unlike the plain modules: the
reference wrapper's empty-grid early-out (diso/__init__.py:49-50,103-104) and the NaN check;
counts[2] holds the flags (diso_amd._lib.TOTAL_*) that decide them. If a count exceeds its
capacity, nothing is written: `check(counts)` raises in that case (it synchronises, so call it
outside the captured region).
"""
import ctypes

import torch

from . import _lib
from .engine import _DTYPES, Extractor, _ptr, _stream

__all__ = ["FixedExtractor"]

_ALGOS = {"mc": _lib.MC, "dmc": _lib.DMC}


class _FixedExtract(torch.autograd.Function):
    @staticmethod
    def forward(ctx, grid, deform, fx, iso, flags):
        L = _lib.lib()
        ex = fx.ex
        dev = grid.device
        dims = tuple(int(d) for d in grid.shape)
        dmc = ex.algo == _lib.DMC
        nseg = L.diso_seg_bytes(*dims, flags & ~_lib.FLAG_RECOMPUTE)
        off = L.diso_totals_offset(*dims, flags)
        if nseg < 0 or off < 0:
            _lib.check(min(nseg, off), "workspace sizing")
        seg = torch.empty(nseg, dtype=torch.uint8, device=dev)
        ctl = ex.ctl_for(dev, dims, flags)
        st = _stream(dev)
        rc = L.diso_count(ex.algo, ex.code, _ptr(grid), *dims, iso, flags, _ptr(seg), nseg, _ptr(ctl), ctl.numel(),
                          None, st)
        _lib.check(rc, "count pass")
        vcap, fcap = fx.vert_capacity, fx.face_capacity
        verts = torch.zeros((vcap, 3), dtype=ex.dtype, device=dev)
        faces = torch.empty((fcap, 4 if dmc else 3), dtype=torch.int64, device=dev)
        vinfo = torch.empty(vcap if dmc else 0, dtype=torch.uint8, device=dev)
        rc = L.diso_emit(ex.algo, ex.code, _ptr(grid), _ptr(deform), *dims, iso, flags, _ptr(seg), nseg, _ptr(verts),
                         vcap, _ptr(faces), fcap, _ptr(vinfo), st)
        _lib.check(rc, "emit pass")
        tot = seg[off:off + 32].view(torch.int64)  # (crossing edges, MC triangles | DMC vertices, flags, algo)
        counts = torch.stack([tot[1], tot[0], tot[2]]) if dmc else tot[:3].clone()
        if flags & _lib.FLAG_RECOMPUTE:
            seg = seg[:L.diso_seg_bytes(*dims, flags)].clone()  # the bases (and totals) only
        ctx.fx, ctx.iso, ctx.flags, ctx.dims, ctx.seg = fx, iso, flags, dims, seg
        ctx.save_for_backward(grid, deform, faces if dmc else None, vinfo if dmc else None)
        ctx.mark_non_differentiable(faces, counts)
        ctx.set_materialize_grads(False)  # no zero-filled gradients for faces, counts or unused vertices
        return verts, faces, counts

    @staticmethod
    def backward(ctx, adj_verts, adj_faces, adj_counts):
        grid, deform, quads, vinfo = ctx.saved_tensors
        if adj_verts is None:
            return None, None, None, None, None
        fx = ctx.fx
        ex = fx.ex
        L = _lib.lib()
        dev = grid.device
        adj_verts = adj_verts.contiguous()
        dmc = ex.algo == _lib.DMC
        nsc = L.diso_scratch_bytes(ex.algo, ex.code, fx.face_capacity if dmc else 0)
        if nsc < 0:
            _lib.check(nsc, "scratch sizing")
        scratch = torch.empty(max(nsc, 1), dtype=torch.uint8, device=dev)
        grad_grid = torch.empty_like(grid)
        grad_deform = None if deform is None else torch.empty_like(deform)
        with torch.cuda.device(dev):
            rc = L.diso_backward(ex.algo, ex.code, _ptr(grid), _ptr(deform), *ctx.dims, ctx.iso, ctx.flags,
                                 _ptr(ctx.seg), ctx.seg.numel(), _ptr(quads), fx.face_capacity if dmc else 0,
                                 _ptr(vinfo), _ptr(adj_verts), fx.vert_capacity, _ptr(grad_grid),
                                 _ptr(grad_deform), _ptr(scratch), nsc, _stream(dev))
        _lib.check(rc, "backward pass")
        return grad_grid, grad_deform, None, None, None


class FixedExtractor:
    """Marching Cubes ("mc") or Dual Marching Cubes ("dmc", quads) into outputs of fixed
    capacity, counts on the device. `reference_compat` / `recompute_backward` as for DiffDMC /
    DiffMC. Call it once outside a capture before capturing (that call allocates the extractor's
    scan control block)."""

    def __init__(self, algo, dtype=torch.float32, vert_capacity=0, face_capacity=0, reference_compat=False,
                 recompute_backward=False):
        if algo not in _ALGOS:
            raise ValueError(f"algorithm {algo!r}: use 'mc' or 'dmc'")
        if dtype not in _DTYPES:
            raise ValueError(f"diso_amd: dtype must be torch.float32 or torch.float64, got {dtype}")
        if vert_capacity < 1 or face_capacity < 1:
            raise ValueError("capacities must be positive")
        self.ex = Extractor(_ALGOS[algo], dtype)
        self.algo = algo
        self.dtype = dtype
        self.vert_capacity = int(vert_capacity)
        self.face_capacity = int(face_capacity)
        self.reference_compat = reference_compat
        self.recompute_backward = recompute_backward

    def __call__(self, grid, deform=None, isovalue=0.0, normalize=True):
        dims = tuple(grid.shape)
        if len(dims) != 3:
            raise ValueError(f"grid must have shape [X, Y, Z], got {list(dims)}")
        self.ex.check_grid(grid, "grid")
        if deform is not None:
            self.ex.check_grid(deform, "deformation")
            if tuple(deform.shape) != dims + (3,) or deform.device != grid.device:
                raise ValueError(f"deform must be a {list(dims) + [3]} tensor on the grid's device")
        grid = grid.contiguous()
        deform = None if deform is None else deform.contiguous()
        flags = _lib.FLAG_VIRTUAL_PAD | _lib.FLAG_SHIFT | _lib.FLAG_INDEX64
        if normalize:
            flags |= _lib.FLAG_NORMALIZE
        if self.reference_compat and self.algo == "dmc":
            flags |= _lib.FLAG_REFERENCE_COMPAT
        if self.recompute_backward:
            flags |= _lib.FLAG_RECOMPUTE
        with torch.cuda.device(grid.device):
            return _FixedExtract.apply(grid, deform, self, float(isovalue), flags)

    def check(self, counts):
        """Raises if the extraction behind `counts` overflowed a capacity, saw a NaN or timed out
        (host synchronisation)."""
        nv, nf, fl = (int(x) for x in counts.tolist())
        if fl & _lib.TOTAL_HAS_NAN:
            raise RuntimeError("diso_amd: sdf grid contains NaN values")
        if fl & _lib.TOTAL_SCAN_TIMEOUT:
            raise RuntimeError("diso_amd: single-pass scan did not complete")
        if nv > self.vert_capacity or nf > self.face_capacity:
            raise RuntimeError(f"diso_amd: {nv} vertices / {nf} faces exceed the capacities "
                               f"{self.vert_capacity} / {self.face_capacity}; nothing was written")
        return nv, nf
