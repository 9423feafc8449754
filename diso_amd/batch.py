"""Batched extraction: many grids of uneven size, Marching Cubes and Dual Marching Cubes mixed,
forward and backward with one launch per pass (the C ABI's diso_batch_*).

    ex = BatchExtractor(torch.float32)
    outs = ex([sdf0, sdf1, ...], ["mc", "dmc", ...], deforms=[d0, None, ...], return_quads=True)
    loss = sum(f(v) for v, _ in outs); loss.backward()

Each (verts, faces) pair equals what DiffMC / DiffDMC (diso_amd/__init__.py, the reference's
diso/__init__.py:9-147) return for that grid alone, bit for bit, gradients included: the same
kernels run, only the launches are shared. A batch costs one host synchronisation in the forward
(to size the outputs) plus one short descriptor upload per pass.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .engine import _DTYPES, _ptr, _stream, split_quads

__all__ = ["BatchExtractor"]


class GridT(ctypes.Structure):
    """diso_grid_t (include/diso_hip.h)."""

    _fields_ = [("algo", ctypes.c_int32), ("flags", ctypes.c_int32), ("X", ctypes.c_int64), ("Y", ctypes.c_int64),
                ("Z", ctypes.c_int64), ("iso", ctypes.c_double), ("sdf", ctypes.c_void_p),
                ("deform", ctypes.c_void_p), ("verts", ctypes.c_void_p), ("faces", ctypes.c_void_p),
                ("vinfo", ctypes.c_void_p), ("vert_capacity", ctypes.c_int64), ("face_capacity", ctypes.c_int64),
                ("adj_verts", ctypes.c_void_p), ("adj_sdf", ctypes.c_void_p), ("adj_deform", ctypes.c_void_p),
                ("totals", _lib.Totals)]


_ALGOS = {"mc": _lib.MC, "dmc": _lib.DMC, _lib.MC: _lib.MC, _lib.DMC: _lib.DMC}


def _grid_np_dtype():
    """A numpy record type laid out exactly as GridT (totals flattened to t_*), so that the fields
    of all descriptors are read and written as columns instead of one ctypes access each."""
    conv = {ctypes.c_int32: np.int32, ctypes.c_int64: np.int64, ctypes.c_double: np.float64, ctypes.c_void_p: np.uint64}
    names, formats, offsets = [], [], []
    for name, ct in GridT._fields_:
        off = getattr(GridT, name).offset
        if ct is _lib.Totals:
            for tn, tct in _lib.Totals._fields_:
                names.append("t_" + tn)
                formats.append(conv[tct])
                offsets.append(off + getattr(_lib.Totals, tn).offset)
        else:
            names.append(name)
            formats.append(conv[ct])
            offsets.append(off)
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": ctypes.sizeof(GridT)})


_GRID_NP = _grid_np_dtype()


def _table(desc):
    """Column view of a GridT array: writes go to the descriptors themselves."""
    return np.frombuffer(desc, dtype=_GRID_NP)


class _State:
    """Per-call state shared by forward and backward: descriptors, workspace, saved outputs."""

    def __init__(self, ex, n):
        self.ex = ex
        self.n = n
        self.desc = (GridT * n)()
        self.table = _table(self.desc)
        self.seg = None
        self.quads = [None] * n
        self.vinfo = [None] * n
        self.empty = [False] * n
        self.has_deform = [False] * n


class _BatchExtract(torch.autograd.Function):
    @staticmethod
    def forward(ctx, st, *tensors):
        n = st.n
        sdfs, deforms = tensors[:n], tensors[n:]
        L = _lib.lib()
        code = st.ex.code
        dev = sdfs[0].device
        stream = _stream(dev)
        d = st.desc
        if st.seg is None:  # else counted already (BatchExtractor.count)
            nbytes = L.diso_batch_seg_bytes(ctypes.addressof(d), n)
            _lib.check(0 if nbytes >= 0 else nbytes, "batch workspace sizing")
            st.seg = torch.empty(nbytes, dtype=torch.uint8, device=dev)
            ctl = st.ex.ctl_for(dev, n)
            rc = L.diso_batch_count(code, ctypes.addressof(d), n, _ptr(st.seg), nbytes, _ptr(ctl), ctl.numel(),
                                    stream)
            _lib.check(rc, "batch count pass")
        nbytes = st.seg.numel()
        both = _lib.TOTAL_HAS_BELOW | _lib.TOTAL_HAS_ABOVE
        a = st.table
        fl = a["t_flags"]
        nan = np.flatnonzero(fl & _lib.TOTAL_HAS_NAN)
        if nan.size:
            raise RuntimeError(f"diso_amd: sdf grid {int(nan[0])} of the batch contains NaN values")
        empty = (fl & both) != both
        # diso/__init__.py:49-50,103-104: a grid's early-out
        a["t_n_verts"][empty] = 0
        a["t_n_faces"][empty] = 0
        st.empty = empty.tolist()
        is_dmc = (a["algo"] == _lib.DMC).tolist()
        nvs, nfs = a["t_n_verts"].tolist(), a["t_n_faces"].tolist()
        verts, faces = [], []
        vp, fp, vip = [0] * n, [0] * n, [0] * n
        for i in range(n):
            dmc = is_dmc[i]
            if st.empty[i]:
                verts.append(torch.zeros((0, 3), dtype=st.ex.dtype, device=dev))
                faces.append(torch.zeros((0, 4 if dmc else 3), dtype=torch.int32, device=dev))
                continue
            v = torch.empty((nvs[i], 3), dtype=st.ex.dtype, device=dev)
            f = torch.empty((nfs[i], 4 if dmc else 3), dtype=torch.int64, device=dev)
            vp[i], fp[i] = v.data_ptr(), f.data_ptr()
            if dmc:
                vi = torch.empty(nvs[i], dtype=torch.uint8, device=dev)
                vip[i] = vi.data_ptr()
                st.quads[i], st.vinfo[i] = f, vi
            verts.append(v)
            faces.append(f)
        a["verts"], a["faces"], a["vinfo"] = vp, fp, vip
        a["vert_capacity"], a["face_capacity"] = a["t_n_verts"], a["t_n_faces"]
        rc = L.diso_batch_emit(code, ctypes.addressof(d), n, _ptr(st.seg), nbytes, stream)
        _lib.check(rc, "batch emit pass")
        ctx.st = st
        ctx.save_for_backward(*sdfs, *[x if x is not None else None for x in deforms])
        ctx.mark_non_differentiable(*faces)
        ctx.set_materialize_grads(False)  # no zero-filled gradients for faces or unused vertices
        return (*verts, *faces)

    @staticmethod
    def backward(ctx, *grads):
        st = ctx.st
        n = st.n
        saved = ctx.saved_tensors
        sdfs, deforms = saved[:n], saved[n:]
        L = _lib.lib()
        d = st.desc
        g_sdf, g_def = [None] * n, [None] * n
        keep = []
        ap, sp, dp = [0] * n, [0] * n, [0] * n
        for i in range(n):
            gv = grads[i]
            if gv is None or st.empty[i]:
                continue
            gv = gv.contiguous()
            if gv.dtype != st.ex.dtype:
                raise RuntimeError("adj_verts type must match the extractor's dtype")
            g_sdf[i] = torch.empty_like(sdfs[i])
            if deforms[i] is not None:
                g_def[i] = torch.empty_like(deforms[i])
                dp[i] = g_def[i].data_ptr()
            keep.append(gv)
            ap[i], sp[i] = gv.data_ptr(), g_sdf[i].data_ptr()
        a = st.table
        a["adj_verts"], a["adj_sdf"], a["adj_deform"] = ap, sp, dp
        if keep:
            dev = sdfs[0].device
            nsc = L.diso_batch_scratch_bytes(st.ex.code, ctypes.addressof(d), n)
            _lib.check(0 if nsc >= 0 else nsc, "batch scratch sizing")
            scratch = torch.empty(max(nsc, 1), dtype=torch.uint8, device=dev)
            with torch.cuda.device(dev):
                rc = L.diso_batch_backward(st.ex.code, ctypes.addressof(d), n, _ptr(st.seg), st.seg.numel(), _ptr(scratch), nsc,
                                           _stream(dev))
            _lib.check(rc, "batch backward pass")
        return (None, *g_sdf, *g_def)


class _Counted:
    """Handle of a counting pass (BatchExtractor.count)."""

    def __init__(self, st, sizes, key):
        self.st = st
        self.sizes = sizes
        self.key = key


class BatchExtractor:
    """Extracts a list of grids in one batch. `algos` holds "mc" / "dmc" per grid; `deforms` is
    None or a list with a [X,Y,Z,3] tensor or None per grid. All grids share the dtype, the
    device, `isovalue` and `normalize`; DMC grids return quads when `return_quads`, else the
    reference's triangle split."""

    def __init__(self, dtype=torch.float32, reference_compat=False, recompute_backward=False):
        if dtype not in _DTYPES:
            raise ValueError(f"diso_amd: dtype must be torch.float32 or torch.float64, got {dtype}")
        self.dtype = dtype
        self.code = _DTYPES[dtype]
        self.reference_compat = reference_compat
        self.recompute_backward = recompute_backward
        self._ctl = {}
        _lib.lib()

    def ctl_for(self, device, n):
        need = _lib.lib().diso_batch_ctl_bytes(n)
        _lib.check(0 if need >= 0 else need, "batch control sizing")
        ctl = self._ctl.get(device.index)
        if ctl is None or ctl.numel() < need:
            ctl = torch.zeros(need * 5 // 4 + 256, dtype=torch.uint8, device=device)  # zero once, then reused
            self._ctl[device.index] = ctl
        return ctl

    def _describe(self, grids, algos, deforms, isovalue, normalize):
        n = len(grids)
        if len(algos) != n:
            raise ValueError("one algorithm per grid")
        deforms = [None] * n if deforms is None else list(deforms)
        if len(deforms) != n:
            raise ValueError("one deform (or None) per grid")
        dev = grids[0].device
        st = _State(self, n)
        sdfs, defs = [], []
        base = _lib.FLAG_VIRTUAL_PAD | _lib.FLAG_SHIFT | _lib.FLAG_INDEX64
        if normalize:
            base |= _lib.FLAG_NORMALIZE
        if self.recompute_backward:
            base |= _lib.FLAG_RECOMPUTE
        col = {k: [0] * n for k in ("algo", "flags", "X", "Y", "Z", "sdf", "deform")}
        for i in range(n):
            g, df = grids[i], deforms[i]
            if not isinstance(g, torch.Tensor) or g.dim() != 3:
                raise ValueError(f"grid {i} must be a [X, Y, Z] tensor")
            if g.device.type != "cuda" or g.device != dev:
                raise RuntimeError(f"grid {i} must be a CUDA tensor on {dev}")
            if g.dtype != self.dtype:
                raise RuntimeError(f"grid {i} type must match the extractor's dtype")
            if df is not None and (tuple(df.shape) != tuple(g.shape) + (3,) or df.dtype != self.dtype
                                   or df.device != dev):
                raise ValueError(f"deform {i} must be a {list(g.shape) + [3]} tensor of the grid's dtype and device")
            algo = _ALGOS.get(algos[i])
            if algo is None:
                raise ValueError(f"algorithm {algos[i]!r}: use 'mc' or 'dmc'")
            g = g.contiguous()
            df = None if df is None else df.contiguous()
            sdfs.append(g)
            defs.append(df)
            col["algo"][i] = algo
            col["flags"][i] = base | (_lib.FLAG_REFERENCE_COMPAT if self.reference_compat and algo == _lib.DMC else 0)
            col["X"][i], col["Y"][i], col["Z"][i] = g.shape
            col["sdf"][i] = g.data_ptr()
            col["deform"][i] = 0 if df is None else df.data_ptr()
        a = st.table
        for k, v in col.items():
            a[k] = v
        a["iso"] = float(isovalue)
        return st, sdfs, defs, dev

    def count(self, grids, algos, isovalue=0.0, normalize=True):
        """The counting pass alone, before any output is allocated: returns a handle whose
        `sizes` holds per grid (n_verts, n_faces, flags) as the extraction will produce them
        (DMC faces are quads). Passing the handle as `counted` to a call on the same grids,
        algorithms, isovalue and normalisation reuses this pass. Costs the workspace and one host
        synchronisation."""
        if not grids:
            return _Counted(None, [], [])
        st, sdfs, _, dev = self._describe(grids, algos, None, isovalue, normalize)
        L = _lib.lib()
        d = st.desc
        with torch.cuda.device(dev):
            nbytes = L.diso_batch_seg_bytes(ctypes.addressof(d), st.n)
            _lib.check(0 if nbytes >= 0 else nbytes, "batch workspace sizing")
            st.seg = torch.empty(nbytes, dtype=torch.uint8, device=dev)
            ctl = self.ctl_for(dev, st.n)
            rc = L.diso_batch_count(self.code, ctypes.addressof(d), st.n, _ptr(st.seg), nbytes, _ptr(ctl),
                                    ctl.numel(), _stream(dev))
            _lib.check(rc, "batch count pass")
        a = st.table
        sizes = list(zip(a["t_n_verts"].tolist(), a["t_n_faces"].tolist(), a["t_flags"].tolist()))
        algo = a["algo"].tolist()
        return _Counted(st, sizes, [(g.data_ptr(), algo[i], tuple(g.shape)) for i, g in enumerate(sdfs)])

    def __call__(self, grids, algos, deforms=None, isovalue=0.0, normalize=True, return_quads=False, counted=None):
        n = len(grids)
        if n == 0:
            return []
        st, sdfs, defs, dev = self._describe(grids, algos, deforms, isovalue, normalize)
        if counted is not None:
            # reuse the counting pass: same grids (storage, dims, algorithm) and settings
            a, algo = st.table, st.table["algo"].tolist()
            same = counted.st is not None and counted.st.n == n and all(
                (g.data_ptr(), algo[i], tuple(g.shape)) == counted.key[i] for i, g in enumerate(sdfs)) and \
                np.array_equal(counted.st.table["flags"], a["flags"]) and np.array_equal(counted.st.table["iso"], a["iso"])
            if not same:
                raise ValueError("counted: the handle's counting pass was not run on these grids and settings")
            counted.st.table["deform"] = a["deform"]
            st = counted.st
            counted.st = None  # a handle serves one call
        with torch.cuda.device(dev):
            outs = _BatchExtract.apply(st, *sdfs, *defs)
            verts, faces = list(outs[:n]), list(outs[n:])
            if not return_quads:
                algo = st.table["algo"].tolist()
                for i in range(n):
                    if algo[i] == _lib.DMC and not st.empty[i]:
                        faces[i] = split_quads(verts[i], faces[i])
        return list(zip(verts, faces))
