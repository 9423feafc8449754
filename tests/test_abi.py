"""C ABI and host logic, without a GPU: the library loads, exports every declared symbol, sizes
its workspaces and rejects bad arguments before touching a device; the Python layer raises the
reference's exception types."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

import diso_amd
from diso_amd import _C, _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "diso_hip.h")


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(diso_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert set(names) >= {"diso_count", "diso_emit", "diso_backward", "diso_seg_bytes", "diso_ctl_bytes"}
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.EXPORTS)
    assert b"gfx950" in _lib.lib().diso_version()


def test_workspace_sizes():
    L = _lib.lib()
    n = 256
    seg = L.diso_seg_bytes(n, n, n, _lib.FLAG_VIRTUAL_PAD)
    # one 16-byte base per (row, 64-wide z chunk) of the padded grid, plus two totals entries, then
    # one 16-bit classification word per lane of every segment
    nseg = (n + 2) * (n + 2) * ((n + 2 + 63) // 64)
    assert seg == (nseg + 2) * 16 + nseg * 64 * 2
    # a recompute-classification backward keeps only the bases
    assert L.diso_seg_bytes(n, n, n, _lib.FLAG_VIRTUAL_PAD | _lib.FLAG_RECOMPUTE) == (nseg + 2) * 16
    # the totals follow the bases, inside the recompute prefix
    assert L.diso_totals_offset(n, n, n, _lib.FLAG_VIRTUAL_PAD) == nseg * 16
    assert L.diso_totals_offset(0, 4, 4, 0) < 0
    nseg = n * n * ((n + 63) // 64)
    assert L.diso_seg_bytes(n, n, n, 0) == (nseg + 2) * 16 + nseg * 64 * 2
    assert L.diso_ctl_bytes(n, n, n, 0) > 0
    assert L.diso_seg_bytes(0, 4, 4, 0) < 0
    assert L.diso_seg_bytes(1 << 30, 1 << 30, 1 << 30, 0) < 0


def test_invalid_arguments_fail_before_any_device_work():
    L = _lib.lib()
    tot = _lib.Totals()
    null = ctypes.c_void_p(0)
    rc = L.diso_count(7, 0, null, 4, 4, 4, 0.0, 0, null, 0, null, 0, ctypes.byref(tot), null)
    assert rc == -1 and b"algo" in L.diso_last_error()
    rc = L.diso_count(0, 5, null, 4, 4, 4, 0.0, 0, null, 0, null, 0, ctypes.byref(tot), null)
    assert rc == -1 and b"dtype" in L.diso_last_error()
    rc = L.diso_count(0, 0, null, 4, 4, 4, 0.0, 0, null, 0, null, 0, ctypes.byref(tot), null)
    assert rc == -1 and b"null" in L.diso_last_error()
    fake = ctypes.c_void_p(0x1000)
    rc = L.diso_count(0, 0, fake, 4, 4, 4, 0.0, 0, fake, 16, fake, 1 << 20, ctypes.byref(tot), null)
    assert rc == -3 and b"seg" in L.diso_last_error()
    rc = L.diso_emit(0, 0, fake, null, 4, 4, 4, 0.0, 0, fake, 1 << 20, null, 5, null, 0, null, null)
    assert rc == -1
    rc = L.diso_emit(1, 0, fake, null, 4, 4, 4, 0.0, 0, fake, 1 << 20, fake, 5, fake, 7, null, null)
    assert rc == -1 and b"vinfo" in L.diso_last_error()
    rc = L.diso_emit(0, 0, fake, null, 4, 4, 4, 0.0, 0, fake, 16, fake, 5, fake, 7, null, null)
    assert rc == -3 and b"seg" in L.diso_last_error()
    rc = L.diso_backward(0, 0, fake, fake, 4, 4, 4, 0.0, 0, fake, 1 << 20, null, 0, null, fake, 3, fake, null,
                         null, 0, null)
    assert rc == -1 and b"adj_deform" in L.diso_last_error()
    rc = L.diso_backward(1, 0, fake, null, 4, 4, 4, 0.0, 0, fake, 1 << 20, null, 9, null, fake, 3, fake, null,
                         null, 0, null)
    assert rc == -1 and b"quads" in L.diso_last_error()
    rc = L.diso_backward(1, 0, fake, null, 4, 4, 4, 0.0, 0, fake, 1 << 20, fake, 9, fake, fake, 3, fake, null,
                         fake, 16, null)
    assert rc == -3 and b"scratch" in L.diso_last_error()
    assert L.diso_scratch_bytes(0, 0, 100) == 0 and L.diso_scratch_bytes(1, 1, 100) == 2400
    with pytest.raises(RuntimeError, match="count pass failed"):
        _lib.check(-1, "count pass")


def test_python_api_errors_match_the_reference():
    with pytest.raises(ValueError):
        diso_amd.DiffMC(torch.float16)
    mc = diso_amd.DiffMC()
    with pytest.raises(RuntimeError, match="CUDA"):
        mc(torch.zeros(4, 4, 4))  # the reference's binding rejects non-CUDA tensors
    with pytest.raises(ValueError):
        mc(torch.zeros(4, 4))
    b = _C.CUDMCDouble()
    with pytest.raises(RuntimeError, match="CUDA"):
        b.forward(torch.zeros(3, 3, 3, dtype=torch.float64), 0.0)
    with pytest.raises(TypeError):
        b.forward(torch.zeros(3, 3, 3, dtype=torch.float64))
    assert _C.CUDMCFloat().reference_compat and not _C.CUMCFloat().reference_compat


def test_split_arguments_checked_before_any_device_work():
    L = _lib.lib()
    null = ctypes.c_void_p(0)
    fake = ctypes.c_void_p(0x1000)
    assert L.diso_split(0, null, 0, null, 0, 0, null, 0, null, 0, null) == 0  # nothing to do
    assert L.diso_split(3, fake, 4, fake, 1, 0, fake, 2, fake, 1 << 20, null) == -1
    assert L.diso_split(0, fake, 4, null, 1, 0, fake, 2, fake, 1 << 20, null) == -1
    rc = L.diso_split(0, fake, 4, fake, 5, 0, fake, 9, fake, 1 << 20, null)
    assert rc == -3 and b"triangle buffer" in L.diso_last_error()
    rc = L.diso_split(0, fake, 4, fake, 5, 0, fake, 10, fake, 8, null)
    assert rc == -3 and b"scratch" in L.diso_last_error()
    assert L.diso_split_scratch_bytes(1000) >= 1000 and L.diso_split_scratch_bytes(-1) < 0
    with pytest.raises(RuntimeError, match="CUDA"):
        diso_amd.split_quads(torch.zeros(4, 3), torch.zeros(1, 4, dtype=torch.int64))


def test_batch_arguments_checked_before_any_device_work():
    from diso_amd.batch import GridT

    L = _lib.lib()
    null = ctypes.c_void_p(0)
    fake = ctypes.c_void_p(0x1000)
    d = (GridT * 2)()
    for i, dims in enumerate([(8, 9, 10), (70, 3, 5)]):
        d[i].algo, d[i].flags = i, _lib.FLAG_VIRTUAL_PAD
        d[i].X, d[i].Y, d[i].Z = dims
        d[i].sdf = 0x1000
    addr = ctypes.addressof(d)
    nseg = sum((x + 2) * (y + 2) * ((z + 2 + 63) // 64) for x, y, z in [(8, 9, 10), (70, 3, 5)])
    assert L.diso_batch_seg_bytes(addr, 2) >= (nseg + 2) * 16 + nseg * 128
    assert L.diso_batch_seg_bytes(addr, 0) < 0 and L.diso_batch_ctl_bytes(0) < 0
    ctlb = L.diso_batch_ctl_bytes(2)
    assert ctlb > 0
    assert L.diso_batch_count(0, addr, 2, fake, 16, fake, ctlb, null) == -3
    assert L.diso_batch_count(0, addr, 2, fake, 1 << 30, fake, 8, null) == -3
    assert L.diso_batch_count(4, addr, 2, fake, 1 << 30, fake, ctlb, null) == -1
    d[1].algo = 9
    assert L.diso_batch_count(0, addr, 2, fake, 1 << 30, fake, ctlb, null) == -1
    assert b"grid 1" in L.diso_last_error()
    d[1].algo = 1
    d[0].sdf = None
    assert L.diso_batch_emit(0, addr, 2, fake, 1 << 30, null) == -1
    d[0].sdf = 0x1000
    d[1].totals.n_faces = 10
    assert L.diso_batch_scratch_bytes(0, addr, 2) == 120 and L.diso_batch_scratch_bytes(1, addr, 2) == 240
    rc = L.diso_batch_backward(0, addr, 2, fake, 1 << 30, null, 0, null)
    assert rc == -3 and b"scratch" in L.diso_last_error()


def test_launch_parameters():
    from diso_amd import autotune

    L = _lib.lib()
    before = autotune.current()
    try:
        assert L.diso_set_param(_lib.PARAM_RUN, 0) == -1 and L.diso_set_param(_lib.PARAM_RUN, 65) == -1
        assert L.diso_set_param(_lib.PARAM_COUNT_BLOCKS_PER_CU, 5) == -1 and L.diso_set_param(9, 1) == -1
        autotune.apply({"run": 16, "count_blocks_per_cu": 2})
        assert autotune.current() == {str(n): {"run": 16, "count_blocks_per_cu": 2} for n in autotune.BUCKET_SIZES}
        assert L.diso_get_param(_lib.PARAM_RUN) == 16 and L.diso_get_param(_lib.PARAM_COUNT_BLOCKS_PER_CU) == 2
        with pytest.raises(RuntimeError):
            autotune.apply({"run": 100, "count_blocks_per_cu": 2})
        with pytest.raises(ValueError):
            autotune.apply({"100": {"run": 8}})
        spins = L.diso_get_param(_lib.PARAM_RECOUNT_SPINS)
        assert spins == 8192
        assert L.diso_set_param(_lib.PARAM_RECOUNT_SPINS, -1) == -1
        assert L.diso_set_param(_lib.PARAM_RECOUNT_SPINS, (1 << 20) + 1) == -1
        assert L.diso_set_param(_lib.PARAM_RECOUNT_SPINS, 0) == 0 and L.diso_get_param(_lib.PARAM_RECOUNT_SPINS) == 0
        assert L.diso_set_param(_lib.PARAM_RECOUNT_SPINS, spins) == 0
    finally:
        autotune.apply(before)


def test_bucket_parameters():
    """Per-bucket values are independent; diso_get_param reads bucket 3; invalid buckets, values
    and parameters are rejected with the message set."""
    from diso_amd import autotune

    L = _lib.lib()
    nb = _lib.PARAM_BUCKETS
    before = autotune.current()
    try:
        for b in range(nb):
            assert L.diso_set_bucket_param(_lib.PARAM_RUN, b, 1 + b) == 0
            assert L.diso_set_bucket_param(_lib.PARAM_COUNT_BLOCKS_PER_CU, b, 1 + b % 4) == 0
        assert [L.diso_get_bucket_param(_lib.PARAM_RUN, b) for b in range(nb)] == [1, 2, 3, 4, 5]
        assert [L.diso_get_bucket_param(_lib.PARAM_COUNT_BLOCKS_PER_CU, b) for b in range(nb)] == [1, 2, 3, 4, 1]
        assert L.diso_get_param(_lib.PARAM_RUN) == 4
        assert L.diso_set_bucket_param(_lib.PARAM_RUN, nb, 8) == -1 and b"bucket" in L.diso_last_error()
        assert L.diso_set_bucket_param(_lib.PARAM_RUN, -1, 8) == -1
        assert L.diso_set_bucket_param(_lib.PARAM_RUN, 0, 65) == -1 and b"run length" in L.diso_last_error()
        assert L.diso_set_bucket_param(_lib.PARAM_RECOUNT_SPINS, 0, 8) == -1
        assert b"no size buckets" in L.diso_last_error()
        assert L.diso_get_bucket_param(_lib.PARAM_RUN, nb) == -1
        autotune.apply({"64": {"run": 8}, "512": {"count_blocks_per_cu": 3}})
        cur = autotune.current()
        assert cur["64"]["run"] == 8 and cur["512"]["count_blocks_per_cu"] == 3 and cur["32"]["run"] == 1
    finally:
        autotune.apply(before)


def test_param_bucket_of_grids():
    """Cubic grids of the representative sizes land in their own bucket (virtual pad), and the
    limits follow the segment count ((X+2)(Y+2) rows of ceil((Z+2)/64) segments)."""
    L = _lib.lib()
    pad = _lib.FLAG_VIRTUAL_PAD
    assert [L.diso_param_bucket(n, n, n, pad) for n in (32, 64, 128, 256, 512)] == [0, 1, 2, 3, 4]
    assert [L.diso_param_bucket(n, n, n, pad) for n in (16, 48, 96, 192, 384, 1024)] == [0, 0, 1, 3, 4, 4]
    assert L.diso_param_bucket(54, 54, 62, 0) == 0 and L.diso_param_bucket(57, 57, 62, 0) == 1  # 2,916 / 3,249
    assert L.diso_param_bucket(0, 4, 4, 0) == -1 and b"dims" in L.diso_last_error()


def test_batch_descriptor_columns_alias_the_structs():
    """The batch path reads and writes descriptor fields as numpy columns over the ctypes array:
    the record layout must be GridT's, field for field (totals flattened)."""
    from diso_amd.batch import _GRID_NP, _table, GridT

    assert _GRID_NP.itemsize == ctypes.sizeof(GridT)
    d = (GridT * 3)()
    a = _table(d)
    a["algo"] = [0, 1, 0]
    a["flags"] = [1, 3, 7]
    a["X"], a["Y"], a["Z"] = [5, 6, 7], [8, 9, 10], [11, 12, 13]
    a["iso"] = 0.25
    a["sdf"] = [0x1000, 0x2000, 1 << 47]
    a["deform"] = [0, 0x3000, 0]
    a["vert_capacity"] = [100, 200, 300]
    a["adj_deform"] = [0, 0, 0x4000]
    assert (d[1].algo, d[2].flags, d[0].X, d[1].Y, d[2].Z) == (1, 7, 5, 9, 13)
    assert d[0].iso == 0.25 and d[2].sdf == 1 << 47 and d[0].deform is None and d[1].deform == 0x3000
    assert d[2].vert_capacity == 300 and d[2].adj_deform == 0x4000
    d[1].totals.n_verts, d[1].totals.n_faces, d[1].totals.flags, d[1].totals.n_crossing = 11, 22, 3, 44
    assert (int(a["t_n_verts"][1]), int(a["t_n_faces"][1]), int(a["t_flags"][1]), int(a["t_n_crossing"][1])) == \
        (11, 22, 3, 44)


def test_backward_shape_checks():
    """A backward whose grid, deform or gradient buffers do not have the forward's dims is refused
    before any kernel runs (the kernels write the forward's dims worth of gradients)."""
    from diso_amd.engine import check_backward_shapes

    dims = (4, 5, 6)
    g, d = torch.zeros(dims), torch.zeros(dims + (3,))
    check_backward_shapes(dims, g, d, torch.zeros(dims), torch.zeros(dims + (3,)))
    check_backward_shapes(dims, g, None, torch.zeros(dims), None)
    bad = [
        (torch.zeros(4, 5, 7), d, torch.zeros(dims), torch.zeros(dims + (3,))),  # grid of other dims
        (g, d, torch.zeros(4, 5, 5), torch.zeros(dims + (3,))),  # smaller adj_grid
        (g, d, torch.zeros(dims), torch.zeros(4, 5, 6, 2)),  # adj_deform of the wrong shape
        (g, d, torch.zeros(dims), None),  # no adj_deform with a deform
        (g, torch.zeros(4, 5, 6, 2), torch.zeros(dims), torch.zeros(dims + (3,))),  # deform shape
        (g, None, None, None),  # no adj_grid
    ]
    for args in bad:
        with pytest.raises(RuntimeError):
            check_backward_shapes(dims, *args)


def test_reference_import_names_resolve():
    """`from diso import DiffMC, DiffDMC` and `diso._C.CUMCFloat` (the reference's import names)
    resolve to this package's classes."""
    import diso
    import diso._C
    import diso_amd
    import diso_amd._C

    assert diso.DiffMC is diso_amd.DiffMC and diso.DiffDMC is diso_amd.DiffDMC
    for name in ("CUMCFloat", "CUMCDouble", "CUDMCFloat", "CUDMCDouble"):
        assert getattr(diso._C, name) is getattr(diso_amd._C, name)
