"""diso_amd - differentiable iso-surface extraction (Marching Cubes and Dual Marching Cubes) on
AMD Instinct MI355X, a drop-in for the reference package `diso` (diso/__init__.py:1-147).

    from diso_amd import DiffMC, DiffDMC
    verts, faces = DiffMC(dtype=torch.float32)(sdf, deform=None, isovalue=0.0, normalize=True)
    verts, faces = DiffDMC(dtype=torch.float32)(sdf, deform=None, isovalue=0.0,
                                                return_quads=False, normalize=True)

Same arguments, outputs, dtypes and ordering as the reference; backward flows to `sdf` and
`deform` through torch.autograd. What the reference does with several ATen calls around its
engine (early-out test, F.pad, `-1`, normalisation, `.long()`) is done inside the HIP passes
here: padding is virtual, the epilogue is written by the emit kernel and the early-out test is a
by-product of the counting pass, so a forward costs one host synchronisation (to size the
outputs) and a backward none.
"""
import torch
from torch import nn

from . import _lib
from .batch import BatchExtractor
from .engine import Extractor, split_quads
from .fixed import FixedExtractor

__all__ = ["DiffMC", "DiffDMC", "BatchExtractor", "FixedExtractor", "split_quads"]
__version__ = "0.1.0"


def _check_inputs(ex, grid, deform):
    dims = tuple(grid.shape)
    if len(dims) != 3:
        raise ValueError(f"grid must have shape [X, Y, Z], got {list(dims)}")
    ex.check_grid(grid, "grid")
    if deform is not None:
        ex.check_grid(deform, "deformation")
        if tuple(deform.shape) != dims + (3,):
            raise ValueError(f"deform must have shape {list(dims) + [3]}, got {list(deform.shape)}")
        if deform.device != grid.device:
            raise RuntimeError("grid and deform must be on the same device")


class _Extract(torch.autograd.Function):
    """Emission of an already counted grid, with the gather backward of the HIP engine.

    The reference recomputes its whole forward inside backward (diso/__init__.py:32,86) because
    its engine keeps per-call state; here the per-call state (segment bases) is kept by autograd
    in `counted`, so the backward is a single kernel."""

    @staticmethod
    def forward(ctx, grid, deform, ex, counted, flags):
        verts, faces = ex.emit(grid, deform, counted, flags)
        ctx.ex, ctx.counted, ctx.flags = ex, counted, flags
        # DMC backward reads the quads; keep them through autograd, not through `counted`
        counted.quads = None
        ctx.save_for_backward(grid, deform, faces if ex.algo == _lib.DMC else None)
        ctx.mark_non_differentiable(faces)
        # no zero-filled gradient for the (integer) faces, nor for unused vertices
        ctx.set_materialize_grads(False)
        return verts, faces

    @staticmethod
    def backward(ctx, adj_verts, adj_faces):
        if adj_verts is None:  # the vertices do not reach the loss: zero gradients
            return None, None, None, None, None
        grid, deform, quads = ctx.saved_tensors
        ctx.counted.quads = quads
        grad_grid = torch.empty_like(grid)
        grad_deform = None if deform is None else torch.empty_like(deform)
        with torch.cuda.device(grid.device):
            ctx.ex.backward(grid, deform, ctx.counted, ctx.flags, adj_verts, grad_grid, grad_deform)
        return grad_grid, grad_deform, None, None, None


def _run(ex, grid, deform, isovalue, normalize, compat, recompute, empty_k):
    _check_inputs(ex, grid, deform)
    grid = grid.contiguous()
    deform = None if deform is None else deform.contiguous()
    iso = float(isovalue)
    flags = _lib.FLAG_VIRTUAL_PAD | _lib.FLAG_SHIFT | _lib.FLAG_INDEX64
    if normalize:
        flags |= _lib.FLAG_NORMALIZE
    if compat:
        flags |= _lib.FLAG_REFERENCE_COMPAT
    if recompute:
        flags |= _lib.FLAG_RECOMPUTE
    with torch.cuda.device(grid.device):
        counted = ex.count_speculative(grid, deform, iso, flags)
        if counted.has_nan:
            raise RuntimeError("diso_amd: sdf grid contains NaN values")
        if counted.empty_grid:
            # diso/__init__.py:49-50,103-104: fresh tensors, int32 faces, no autograd graph.
            return (torch.zeros((0, 3), dtype=ex.dtype, device=grid.device),
                    torch.zeros((0, empty_k), dtype=torch.int32, device=grid.device), True)
        verts, faces = _Extract.apply(grid, deform, ex, counted, flags)
        return verts, faces, False


class DiffMC(nn.Module):
    """Differentiable Marching Cubes (diso/__init__.py:9-61).

    recompute_backward=True keeps only the segment bases between forward and backward (16 bytes
    per 64 grid positions instead of 2 bytes per position) and re-derives the crossing masks from
    the samples in the backward; results are identical."""

    def __init__(self, dtype=torch.float32, recompute_backward=False):
        super().__init__()
        self.dtype = dtype
        self.recompute_backward = recompute_backward
        self._ex = Extractor(_lib.MC, dtype)

    def forward(self, grid, deform=None, isovalue=0.0, normalize=True):
        verts, faces, _ = _run(self._ex, grid, deform, isovalue, normalize, False, self.recompute_backward, 3)
        return verts, faces


class DiffDMC(nn.Module):
    """Differentiable Dual Marching Cubes (diso/__init__.py:63-147).

    reference_compat=True reproduces the reference backward's read of each cell's first patch
    gradient for all of the cell's patches (src/cudualmc.cu:990); the default computes the true
    gradient of the forward (the two agree whenever every cell has at most one patch).
    recompute_backward as for DiffMC."""

    def __init__(self, dtype=torch.float32, reference_compat=False, recompute_backward=False):
        super().__init__()
        self.dtype = dtype
        self.reference_compat = reference_compat
        self.recompute_backward = recompute_backward
        self._ex = Extractor(_lib.DMC, dtype)

    def forward(self, grid, deform=None, isovalue=0.0, return_quads=False, normalize=True):
        verts, quads, empty = _run(self._ex, grid, deform, isovalue, normalize, self.reference_compat,
                                   self.recompute_backward, 4)
        if return_quads or empty:
            return verts, quads
        return verts, split_quads(verts, quads)
