"""Thin host-side driver of libdiso_hip.so for torch tensors.

`Extractor` is the counterpart of the reference's engine objects (CuMC, src/cumc.h:70-139;
CUDualMC, src/cudualmc.h:70-150): it owns the scan control workspace of each device it has run
on (zero-filled once, grown on demand) and runs the count / emit / backward passes on the
current torch stream. Per-call state (the segment bases the backward needs) is returned to the
caller instead of being kept inside the object, so one extractor can serve many grids whose
backward runs later (autograd, activation checkpointing).
"""
import ctypes

import torch

from . import _lib

_DTYPES = {torch.float32: _lib.F32, torch.float64: _lib.F64}


class Counted:
    """Result of the counting pass: segment bases on the device and the totals on the host.
    After emission it also holds what a DMC backward needs from the forward: the quads and one
    byte of patch information per DMC vertex."""

    __slots__ = ("seg", "n_verts", "n_faces", "n_crossing", "flags", "dims", "iso", "flags_in", "quads", "vinfo",
                 "spec")

    def __init__(self, seg, totals, dims, iso, flags_in):
        self.seg = seg
        self.n_verts = int(totals.n_verts)
        self.n_faces = int(totals.n_faces)
        self.n_crossing = int(totals.n_crossing)
        self.flags = int(totals.flags)
        self.dims = dims
        self.iso = iso
        self.flags_in = flags_in
        self.quads = None
        self.vinfo = None
        self.spec = None  # (verts, faces, vinfo) already emitted into buffers these counts fit

    @property
    def empty_grid(self):
        """The reference wrapper's early-out: min(sdf) >= iso or max(sdf) <= iso
        (diso/__init__.py:49,103)."""
        both = _lib.TOTAL_HAS_BELOW | _lib.TOTAL_HAS_ABOVE
        return (self.flags & both) != both

    @property
    def has_nan(self):
        return bool(self.flags & _lib.TOTAL_HAS_NAN)


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _prefix(buf, n):
    """The first n rows of `buf` as a tensor of their own over the same memory. Not a view (as
    buf[:n] would be): autograd forbids in-place edits of views made inside a custom Function's
    forward, and callers may edit the returned vertices in place."""
    size = (n,) + tuple(buf.shape[1:])
    return torch.empty(0, dtype=buf.dtype, device=buf.device).set_(buf.untyped_storage(), buf.storage_offset(), size,
                                                                    buf.stride())


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def _stream(device):
    """The current HIP stream of `device` as a raw handle (the direct query when torch has it:
    the public one costs some ten microseconds per call)."""
    if _raw_stream is not None:
        return ctypes.c_void_p(_raw_stream(device.index if device.index is not None else torch.cuda.current_device()))
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


_hiprt = None


def _hip():
    """The HIP runtime torch runs on (same soname, so the already loaded one), for the few
    stream and event calls of the totals read-back."""
    global _hiprt
    if _hiprt is None:
        h = ctypes.CDLL("libamdhip64.so.7")
        vp, i32 = ctypes.c_void_p, ctypes.c_int
        h.hipEventCreateWithFlags.argtypes = [ctypes.POINTER(vp), ctypes.c_uint]
        h.hipEventRecord.argtypes = [vp, vp]
        h.hipStreamWaitEvent.argtypes = [vp, vp, ctypes.c_uint]
        h.hipMemcpyAsync.argtypes = [vp, vp, ctypes.c_size_t, i32, vp]
        h.hipStreamSynchronize.argtypes = [vp]
        for f in (h.hipEventCreateWithFlags, h.hipEventRecord, h.hipStreamWaitEvent, h.hipMemcpyAsync,
                  h.hipStreamSynchronize):
            f.restype = i32
        _hiprt = h
    return _hiprt


def _hip_check(rc, what):
    if rc != 0:
        raise RuntimeError(f"diso_amd: {what} failed (hipError {rc})")


class Extractor:
    def __init__(self, algo, dtype):
        if dtype not in _DTYPES:
            raise ValueError(f"diso_amd: dtype must be torch.float32 or torch.float64, got {dtype}")
        self.algo = algo
        self.dtype = dtype
        self.code = _DTYPES[dtype]
        self._ctl = {}
        self._guess = {}  # (device, dims, flags) -> (vertices, faces) of the last call
        self._side = {}  # device -> (side stream handle, pinned totals, event, side stream)
        _lib.lib()  # fail loudly now if the HIP library is not there

    def check_grid(self, grid, name="grid"):
        if not isinstance(grid, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor")
        if grid.device.type != "cuda":
            raise RuntimeError(f"{name} must be a CUDA tensor")
        if grid.dtype != self.dtype:
            raise RuntimeError(f"{name} type must match the {'mc' if self.algo == _lib.MC else 'dmc'} class")

    def ctl_for(self, device, dims, flags):
        need = _lib.lib().diso_ctl_bytes(*dims, flags)
        if need < 0:
            _lib.check(need, "workspace sizing")
        key = device.index
        ctl = self._ctl.get(key)
        if ctl is None or ctl.numel() < need:
            # Grown buffers start zero-filled; the scan keeps them consistent from then on (its
            # words carry their call's epoch, so not even a timed-out call needs a reset).
            ctl = torch.zeros(max(need, 1 << 16) * 5 // 4, dtype=torch.uint8, device=device)
            self._ctl[key] = ctl
        return ctl

    def count(self, grid, iso, flags):
        dims = tuple(int(d) for d in grid.shape)
        L = _lib.lib()
        nseg = L.diso_seg_bytes(*dims, flags & ~_lib.FLAG_RECOMPUTE)  # the full workspace
        if nseg < 0:
            _lib.check(nseg, "workspace sizing")
        seg = torch.empty(nseg, dtype=torch.uint8, device=grid.device)
        ctl = self.ctl_for(grid.device, dims, flags)
        tot = _lib.Totals()
        rc = L.diso_count(self.algo, self.code, _ptr(grid), *dims, float(iso), flags, _ptr(seg), nseg,
                          _ptr(ctl), ctl.numel(), ctypes.byref(tot), _stream(grid.device))
        _lib.check(rc, "count pass")
        return Counted(seg, tot, dims, float(iso), flags)

    def count_speculative(self, grid, deform, iso, flags):
        """The counting pass, and, when an earlier call on a grid of these dims gives a size
        estimate, the emission queued right behind it into buffers of that size plus a margin:
        the host then waits only for the counts (copied on a side stream as soon as the counting
        kernel is done), so its work until the next launch overlaps the emission instead of
        leaving the device idle. If the counts fit, `counted.spec` holds the emitted buffers
        (emit() returns views of them); otherwise emit() runs the emission again at the exact
        size (the queued one writes nothing when the counts exceed its capacities)."""
        dims = tuple(int(d) for d in grid.shape)
        dev = grid.device
        key = (dev.index, dims, flags)
        guess = self._guess.get(key)
        if guess is None:
            counted = self.count(grid, iso, flags)
            self._remember(key, counted.n_verts, counted.n_faces)
            return counted
        L = _lib.lib()
        nseg = L.diso_seg_bytes(*dims, flags & ~_lib.FLAG_RECOMPUTE)
        off = L.diso_totals_offset(*dims, flags)
        if nseg < 0 or off < 0:
            _lib.check(min(nseg, off), "workspace sizing")
        seg = torch.empty(nseg, dtype=torch.uint8, device=dev)
        ctl = self.ctl_for(dev, dims, flags)
        main = _stream(dev)
        rc = L.diso_count(self.algo, self.code, _ptr(grid), *dims, float(iso), flags, _ptr(seg), nseg, _ptr(ctl),
                          ctl.numel(), None, main)
        _lib.check(rc, "count pass")
        # the side stream copies the totals as soon as the count kernel is done
        side, host, ev = self._side_for(dev)
        hip = _hip()
        _hip_check(hip.hipEventRecord(ev, main), "event record")
        _hip_check(hip.hipStreamWaitEvent(side, ev, 0), "stream wait")
        _hip_check(hip.hipMemcpyAsync(ctypes.c_void_p(host.data_ptr()), ctypes.c_void_p(seg.data_ptr() + off), 32,
                                      2, side), "totals copy")  # 2: device to host
        dmc = self.algo == _lib.DMC
        vcap, fcap = guess[0] + guess[0] // 4 + 1024, guess[1] + guess[1] // 4 + 1024
        idx = torch.int64 if flags & _lib.FLAG_INDEX64 else torch.int32
        verts = torch.empty((vcap, 3), dtype=self.dtype, device=dev)
        faces = torch.empty((fcap, 4 if dmc else 3), dtype=idx, device=dev)
        vinfo = torch.empty(vcap if dmc else 0, dtype=torch.uint8, device=dev)
        rc = L.diso_emit(self.algo, self.code, _ptr(grid), _ptr(deform), *dims, float(iso), flags, _ptr(seg), nseg,
                         _ptr(verts), vcap, _ptr(faces), fcap, _ptr(vinfo), main)
        _lib.check(rc, "emit pass")
        _hip_check(hip.hipStreamSynchronize(side), "totals wait")
        crossing, other, tflags, _ = host.tolist()
        tot = _lib.Totals()
        tot.n_verts, tot.n_faces = (other, crossing) if dmc else (crossing, other)
        tot.flags, tot.n_crossing = tflags, crossing
        if tflags & _lib.TOTAL_SCAN_TIMEOUT:
            raise RuntimeError("diso_amd: count pass failed: single-pass scan did not complete")
        counted = Counted(seg, tot, dims, float(iso), flags)
        if tot.n_verts <= vcap and tot.n_faces <= fcap:
            counted.spec = (verts, faces, vinfo)
        self._remember(key, tot.n_verts, tot.n_faces)
        return counted

    def _side_for(self, dev):
        """Per device: a side stream (default priority), a pinned 32-byte landing buffer and an
        event without timing, created once."""
        got = self._side.get(dev.index)
        if got is None:
            side = torch.cuda.Stream(device=dev)
            host = torch.empty(4, dtype=torch.int64, pin_memory=True)
            ev = ctypes.c_void_p()
            with torch.cuda.device(dev):
                _hip_check(_hip().hipEventCreateWithFlags(ctypes.byref(ev), 2), "event create")  # 2: no timing
            got = (ctypes.c_void_p(side.cuda_stream), host, ev, side)
            self._side[dev.index] = got
        return got[:3]

    def _remember(self, key, n_verts, n_faces):
        if len(self._guess) >= 64 and key not in self._guess:
            self._guess.clear()  # many different grids: start over
        self._guess[key] = (n_verts, n_faces)

    def _scratch(self, counted, device):
        n = _lib.lib().diso_scratch_bytes(self.algo, self.code, counted.n_crossing)
        if n < 0:
            _lib.check(n, "scratch sizing")
        return torch.empty(max(n, 1), dtype=torch.uint8, device=device), n

    def emit(self, grid, deform, counted, flags):
        """Vertices [V,3] and faces ([F,3] MC triangles / [Q,4] DMC quads) of a counted grid."""
        L = _lib.lib()
        dmc = self.algo == _lib.DMC
        k = 4 if dmc else 3
        idx = torch.int64 if flags & _lib.FLAG_INDEX64 else torch.int32
        if counted.spec is not None:  # emitted already (count_speculative): exact-size prefixes
            v, f, vi = counted.spec
            counted.spec = None
            verts, faces = _prefix(v, counted.n_verts), _prefix(f, counted.n_faces)
            vinfo = _prefix(vi, counted.n_verts) if dmc else vi
        else:
            verts = torch.empty((counted.n_verts, 3), dtype=self.dtype, device=grid.device)
            faces = torch.empty((counted.n_faces, k), dtype=idx, device=grid.device)
            vinfo = torch.empty(counted.n_verts if dmc else 0, dtype=torch.uint8, device=grid.device)
            rc = L.diso_emit(self.algo, self.code, _ptr(grid), _ptr(deform), *counted.dims, counted.iso, flags,
                             _ptr(counted.seg), counted.seg.numel(), _ptr(verts), counted.n_verts, _ptr(faces),
                             counted.n_faces, _ptr(vinfo), _stream(grid.device))
            _lib.check(rc, "emit pass")
        if dmc:
            counted.quads, counted.vinfo = faces, vinfo
        if flags & _lib.FLAG_RECOMPUTE:
            # The backward will re-derive the classification from the samples: keep only the
            # segment bases and release the classification words.
            nb = L.diso_seg_bytes(*counted.dims, flags)
            counted.seg = counted.seg[:nb].clone()
        return verts, faces

    def backward(self, grid, deform, counted, flags, adj_verts, adj_grid, adj_deform):
        """Writes (or, with FLAG_ACCUMULATE, adds) d(sum(adj_verts*verts))/d(grid, deform)."""
        check_backward_shapes(counted.dims, grid, deform, adj_grid, adj_deform)
        L = _lib.lib()
        adj_verts = adj_verts.contiguous()
        if adj_verts.dtype != self.dtype:
            raise RuntimeError("adj_verts type must match the extractor's dtype")
        if adj_verts.shape != (counted.n_verts, 3):
            raise ValueError(f"adj_verts must have shape ({counted.n_verts}, 3), got {tuple(adj_verts.shape)}")
        dmc = self.algo == _lib.DMC
        quads = counted.quads if dmc else None
        if dmc and quads is None:
            raise RuntimeError("DMC backward needs the quads of its forward")
        if dmc:
            # the quads' index width decides how the backward reads them
            flags = (flags & ~_lib.FLAG_INDEX64) | (_lib.FLAG_INDEX64 if quads.dtype == torch.int64 else 0)
        scratch, nscratch = self._scratch(counted, grid.device)
        rc = L.diso_backward(self.algo, self.code, _ptr(grid), _ptr(deform), *counted.dims, counted.iso, flags,
                             _ptr(counted.seg), counted.seg.numel(), _ptr(quads), counted.n_faces if dmc else 0,
                             _ptr(counted.vinfo if dmc else None), _ptr(adj_verts), counted.n_verts,
                             _ptr(adj_grid), _ptr(adj_deform), _ptr(scratch), nscratch, _stream(grid.device))
        _lib.check(rc, "backward pass")


def check_backward_shapes(dims, grid, deform, adj_grid, adj_deform):
    """The backward writes dims worth of gradients: the grid, the gradient buffers and (with
    deform) the deform must have the forward's dims, or the kernels would address memory outside
    them. Raises RuntimeError (the reference binding's error type) otherwise."""
    dims = tuple(int(d) for d in dims)
    if tuple(grid.shape) != dims:
        raise RuntimeError(f"grid has shape {tuple(grid.shape)}, the forward's was {dims}")
    if adj_grid is None or tuple(adj_grid.shape) != dims:
        raise RuntimeError(f"adj_grid must have shape {dims}, got {None if adj_grid is None else tuple(adj_grid.shape)}")
    if deform is not None:
        want = dims + (3,)
        if tuple(deform.shape) != want:
            raise RuntimeError(f"deform must have shape {want}, got {tuple(deform.shape)}")
        if adj_deform is None or tuple(adj_deform.shape) != want:
            raise RuntimeError(f"adj_deform must have shape {want}, got "
                               f"{None if adj_deform is None else tuple(adj_deform.shape)}")


def split_quads(verts, quads):
    """DiffDMC quad -> triangle split on the GPU (diso_split; diso/__init__.py:117-147):
    [Q,4] quads -> [2Q,3] int64 triangles, first-cut quads first."""
    if verts.device.type != "cuda" or quads.device.type != "cuda":
        raise RuntimeError("split_quads: verts and quads must be CUDA tensors")
    if verts.dtype not in _DTYPES:
        raise RuntimeError(f"split_quads: verts must be float32 or float64, got {verts.dtype}")
    if quads.dtype not in (torch.int32, torch.int64) or quads.dim() != 2 or quads.shape[1] != 4:
        raise ValueError("split_quads: quads must be an int32/int64 [Q, 4] tensor")
    L = _lib.lib()
    nq = int(quads.shape[0])
    tris = torch.empty((2 * nq, 3), dtype=torch.int64, device=verts.device)
    if nq == 0:
        return tris
    verts = verts.detach().contiguous()
    quads = quads.contiguous()
    nbytes = L.diso_split_scratch_bytes(nq)
    if nbytes < 0:
        _lib.check(nbytes, "split sizing")
    with torch.cuda.device(verts.device):
        scratch = torch.empty(nbytes, dtype=torch.uint8, device=verts.device)
        flags = _lib.FLAG_INDEX64 if quads.dtype == torch.int64 else 0
        rc = L.diso_split(_DTYPES[verts.dtype], _ptr(verts), int(verts.shape[0]), _ptr(quads), nq, flags, _ptr(tris),
                          2 * nq, _ptr(scratch), nbytes, _stream(verts.device))
    _lib.check(rc, "split")
    return tris
