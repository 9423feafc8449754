"""diso_count_noalloc (the count pass with the library's own control block, no allocation)
against diso_count, through the C ABI: the whole segment workspace (bases, totals, classification
words) and the host totals must be identical, on two calls in a row whose second grid is smaller,
for MC and DMC in fp32 and fp64."""
import ctypes

import numpy as np
import pytest
import torch

from diso_amd import _lib
from tests.helpers import inputs

DEV = torch.device("cuda:0")
FLAGS = _lib.FLAG_VIRTUAL_PAD | _lib.FLAG_SHIFT | _lib.FLAG_NORMALIZE | _lib.FLAG_INDEX64


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _seg(L, dims):
    n = L.diso_seg_bytes(*dims, FLAGS)
    assert n > 0
    return n


def _count(L, algo, code, sdf, dims, noalloc):
    nseg = _seg(L, dims)
    seg = torch.full((nseg,), 0xA5, dtype=torch.uint8, device=DEV)  # no stale zeros to hide a missed write
    tot = _lib.Totals()
    st = ctypes.c_void_p(torch.cuda.current_stream(DEV).cuda_stream)
    if noalloc:
        rc = L.diso_count_noalloc(algo, code, _ptr(sdf), *dims, 0.0, FLAGS, _ptr(seg), nseg, ctypes.byref(tot), st)
    else:
        need = L.diso_ctl_bytes(*dims, FLAGS)
        ctl = torch.zeros(need, dtype=torch.uint8, device=DEV)
        rc = L.diso_count(algo, code, _ptr(sdf), *dims, 0.0, FLAGS, _ptr(seg), nseg, _ptr(ctl), need,
                          ctypes.byref(tot), st)
    _lib.check(rc, "count")
    torch.cuda.synchronize()
    return seg.cpu().numpy(), (tot.n_verts, tot.n_faces, tot.flags, tot.n_crossing)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("algo", [_lib.MC, _lib.DMC])
def test_noalloc_count_equals_count(algo, dtype):
    L = _lib.lib()
    code = _lib.F32 if dtype == np.float32 else _lib.F64
    for dims, kind, seed in (((72, 66, 130), "noise", 3), ((20, 33, 17), "noisy_box", 4)):
        print(f"case noalloc algo={algo} dtype={dtype.__name__} dims={dims} kind={kind}", flush=True)
        sdf_np, _ = inputs(dims, kind, seed, dtype, False)
        sdf = torch.tensor(sdf_np, device=DEV)
        want_seg, want_tot = _count(L, algo, code, sdf, dims, noalloc=False)
        got_seg, got_tot = _count(L, algo, code, sdf, dims, noalloc=True)
        assert got_tot == want_tot
        assert want_tot[0] > 0 and want_tot[1] > 0
        assert np.array_equal(got_seg, want_seg)


@pytest.mark.gpu
def test_noalloc_count_rejects_a_small_workspace():
    L = _lib.lib()
    dims = (16, 16, 16)
    sdf = torch.zeros(dims, dtype=torch.float32, device=DEV)
    nseg = _seg(L, dims)
    seg = torch.empty(nseg - 1, dtype=torch.uint8, device=DEV)
    rc = L.diso_count_noalloc(_lib.MC, _lib.F32, _ptr(sdf), *dims, 0.0, FLAGS, _ptr(seg), nseg - 1, None,
                              ctypes.c_void_p(torch.cuda.current_stream(DEV).cuda_stream))
    assert rc != 0 and b"too small" in L.diso_last_error()
