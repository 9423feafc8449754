"""Drop-in for the reference's binding module `diso._C` (src/pybind.cpp:419-447).

CUMCFloat / CUMCDouble / CUDMCFloat / CUDMCDouble keep the binding's method shapes:

  forward(grid, deform, iso) / forward(grid, iso) -> (verts [V,3], faces int32 [F,3] | [Q,4])
      grid is the *padded* grid as the reference's Python wrapper hands it over
      (diso/__init__.py:52-55); vertices are in padded grid coordinates.
  backward(grid, deform, iso, adj_verts, adj_grid, adj_deform) / backward(grid, iso, adj_verts,
      adj_grid) -> None
      accumulates into the caller-zeroed adj_grid / adj_deform, using the state left by the last
      forward on the same object (src/cumc.cu:734-743, src/cudualmc.cu:1130-1138).

Input checks raise RuntimeError like the binding's AT_ASSERTM / TORCH_INTERNAL_ASSERT
(src/pybind.cpp:5-11,45-60,142-160). Results equal the reference's whenever its precondition
holds (a ring of samples >= iso around the grid, which the wrapper's padding guarantees).
"""
import torch

from . import _lib
from .engine import Extractor, check_backward_shapes


class _Binding:
    _algo = None
    _dtype = None

    def __init__(self):
        self._ex = Extractor(self._algo, self._dtype)
        self._last = None
        # The binding reproduces the reference engine, so its DMC backward keeps the reference's
        # per-cell gradient read (src/cudualmc.cu:990) unless this is switched off.
        self.reference_compat = self._algo == _lib.DMC

    def _check_contiguous(self, t, name):
        self._ex.check_grid(t, name)
        if not t.is_contiguous():
            raise RuntimeError(f"{name} must be contiguous")

    def forward(self, grid, *args):
        if len(args) == 2:
            deform, iso = args
        elif len(args) == 1:
            deform, iso = None, args[0]
        else:
            raise TypeError("forward(grid, deform, iso) or forward(grid, iso)")
        self._check_contiguous(grid, "grid")
        if grid.dim() != 3:
            raise RuntimeError("grid must be a 3-D tensor [X, Y, Z]")
        if deform is not None:
            self._check_contiguous(deform, "deform")
            if tuple(deform.shape) != tuple(grid.shape) + (3,):
                raise RuntimeError("deform must have shape [X, Y, Z, 3]")
        with torch.cuda.device(grid.device):
            counted = self._ex.count(grid, iso, 0)
            verts, faces = self._ex.emit(grid, deform, counted, 0)
        self._last = counted
        return verts, faces

    def backward(self, grid, *args):
        if len(args) == 5:
            deform, iso, adj_verts, adj_grid, adj_deform = args
        elif len(args) == 3:
            deform, adj_deform = None, None
            iso, adj_verts, adj_grid = args
        else:
            raise TypeError("backward(grid, deform, iso, adj_verts, adj_grid, adj_deform) or "
                            "backward(grid, iso, adj_verts, adj_grid)")
        self._check_contiguous(adj_verts, "adj_verts")
        self._check_contiguous(adj_grid, "adj_grid")
        if deform is not None:
            self._check_contiguous(adj_deform, "adj_deform")
        if self._last is None:
            raise RuntimeError("backward called before forward")
        check_backward_shapes(self._last.dims, grid, deform, adj_grid, adj_deform)
        flags = _lib.FLAG_ACCUMULATE | (_lib.FLAG_REFERENCE_COMPAT if self.reference_compat else 0)
        with torch.cuda.device(grid.device):
            self._ex.backward(grid.contiguous(), None if deform is None else deform.contiguous(), self._last,
                              flags, adj_verts, adj_grid, adj_deform)


class CUMCFloat(_Binding):
    _algo, _dtype = _lib.MC, torch.float32


class CUMCDouble(_Binding):
    _algo, _dtype = _lib.MC, torch.float64


class CUDMCFloat(_Binding):
    _algo, _dtype = _lib.DMC, torch.float32


class CUDMCDouble(_Binding):
    _algo, _dtype = _lib.DMC, torch.float64
