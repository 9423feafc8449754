"""Loader for libdiso_hip.so (the C ABI declared in include/diso_hip.h) and its build recipe.

There is no fallback: if the library is missing or cannot be loaded, every extraction raises.
"""
import ctypes
import os
import subprocess

import torch  # noqa: F401  (loads torch's HIP runtime first, so the library binds to the same one)

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
# DISO_AMD_LIB selects another build of the same library (e.g. the host-sanitizer build the
# tests make); by default the in-tree one.
LIB_PATH = os.environ.get("DISO_AMD_LIB") or os.path.join(HERE, "libdiso_hip.so")
SOURCES = [os.path.join(HERE, "csrc", "diso_hip.hip")]
HEADERS = [os.path.join(ROOT, "include", "diso_hip.h"), os.path.join(HERE, "csrc", "diso_tables.h"),
           os.path.join(HERE, "csrc", "diso_scan_word.h")]

# Flags: gfx950 only; no FP contraction and correctly rounded division so results match an IEEE
# evaluation of the reference's expressions; packed-fp32 instructions disabled.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

MC, DMC = 0, 1
F32, F64 = 0, 1
FLAG_VIRTUAL_PAD = 1 << 0
FLAG_SHIFT = 1 << 1
FLAG_NORMALIZE = 1 << 2
FLAG_INDEX64 = 1 << 3
FLAG_REFERENCE_COMPAT = 1 << 4
FLAG_ACCUMULATE = 1 << 5
FLAG_RECOMPUTE = 1 << 6
TOTAL_HAS_BELOW = 1 << 0
TOTAL_HAS_ABOVE = 1 << 1
TOTAL_HAS_NAN = 1 << 2
TOTAL_SCAN_TIMEOUT = 1 << 3
E_TIMEOUT = -4  # DISO_E_TIMEOUT: the single-pass scan did not complete (never expected)
PARAM_RUN = 0
PARAM_COUNT_BLOCKS_PER_CU = 1
PARAM_RECOUNT_SPINS = 2
PARAM_BUCKETS = 5  # size buckets of the run length and count blocks per CU (~32^3 .. ~512^3 grids)

EXPORTS = [
    "diso_seg_bytes", "diso_ctl_bytes", "diso_scratch_bytes", "diso_count", "diso_emit", "diso_backward",
    "diso_read_totals", "diso_last_error", "diso_version", "diso_split_scratch_bytes", "diso_split",
    "diso_batch_seg_bytes", "diso_batch_ctl_bytes", "diso_batch_count", "diso_batch_emit", "diso_batch_scratch_bytes",
    "diso_batch_backward", "diso_set_param", "diso_get_param", "diso_totals_offset", "diso_count_noalloc",
    "diso_set_bucket_param", "diso_get_bucket_param", "diso_param_bucket",
]


def build(verbose=False):
    """Compiles libdiso_hip.so in-tree with hipcc for gfx950."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "csrc"),
           "-o", tmp, *SOURCES]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-4000:])
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


class Totals(ctypes.Structure):
    _fields_ = [("n_verts", ctypes.c_int64), ("n_faces", ctypes.c_int64), ("flags", ctypes.c_int64),
                ("n_crossing", ctypes.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"diso_amd: HIP library not built ({LIB_PATH}); run `python -c \"import "
                f"__graft_entry__ as g; g.build()\"` from the repository root")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
        L.diso_seg_bytes.restype = i64
        L.diso_seg_bytes.argtypes = [i64, i64, i64, i32]
        L.diso_ctl_bytes.restype = i64
        L.diso_ctl_bytes.argtypes = [i64, i64, i64, i32]
        L.diso_totals_offset.restype = i64
        L.diso_totals_offset.argtypes = [i64, i64, i64, i32]
        L.diso_count.restype = i32
        L.diso_count.argtypes = [i32, i32, vp, i64, i64, i64, dbl, i32, vp, i64, vp, i64,
                                 ctypes.POINTER(Totals), vp]
        L.diso_count_noalloc.restype = i32
        L.diso_count_noalloc.argtypes = [i32, i32, vp, i64, i64, i64, dbl, i32, vp, i64, ctypes.POINTER(Totals), vp]
        L.diso_scratch_bytes.restype = i64
        L.diso_scratch_bytes.argtypes = [i32, i32, i64]
        L.diso_emit.restype = i32
        L.diso_emit.argtypes = [i32, i32, vp, vp, i64, i64, i64, dbl, i32, vp, i64, vp, i64, vp, i64, vp, vp]
        L.diso_backward.restype = i32
        L.diso_backward.argtypes = [i32, i32, vp, vp, i64, i64, i64, dbl, i32, vp, i64, vp, i64, vp, vp,
                                    i64, vp, vp, vp, i64, vp]
        L.diso_split_scratch_bytes.restype = i64
        L.diso_split_scratch_bytes.argtypes = [i64]
        L.diso_split.restype = i32
        L.diso_split.argtypes = [i32, vp, i64, vp, i64, i32, vp, i64, vp, i64, vp]
        # batch entry points take the diso_grid_t array by address (diso_amd.batch.GridT)
        L.diso_batch_seg_bytes.restype = i64
        L.diso_batch_seg_bytes.argtypes = [vp, i32]
        L.diso_batch_ctl_bytes.restype = i64
        L.diso_batch_ctl_bytes.argtypes = [i32]
        L.diso_batch_count.restype = i32
        L.diso_batch_count.argtypes = [i32, vp, i32, vp, i64, vp, i64, vp]
        L.diso_batch_emit.restype = i32
        L.diso_batch_emit.argtypes = [i32, vp, i32, vp, i64, vp]
        L.diso_batch_scratch_bytes.restype = i64
        L.diso_batch_scratch_bytes.argtypes = [i32, vp, i32]
        L.diso_batch_backward.restype = i32
        L.diso_batch_backward.argtypes = [i32, vp, i32, vp, i64, vp, i64, vp]
        L.diso_set_param.restype = i32
        L.diso_set_param.argtypes = [i32, i32]
        L.diso_get_param.restype = i32
        L.diso_get_param.argtypes = [i32]
        L.diso_set_bucket_param.restype = i32
        L.diso_set_bucket_param.argtypes = [i32, i32, i32]
        L.diso_get_bucket_param.restype = i32
        L.diso_get_bucket_param.argtypes = [i32, i32]
        L.diso_param_bucket.restype = i32
        L.diso_param_bucket.argtypes = [i64, i64, i64, i32]
        L.diso_read_totals.restype = i32
        L.diso_read_totals.argtypes = [vp, i64, i64, i64, i64, i32, ctypes.POINTER(Totals), vp]
        L.diso_last_error.restype = ctypes.c_char_p
        L.diso_last_error.argtypes = []
        L.diso_version.restype = ctypes.c_char_p
        L.diso_version.argtypes = []
        _lib = L
        from . import autotune

        autotune.load()  # launch parameters saved for this device, if any
    return _lib


def check(rc, what):
    if rc != 0:
        msg = lib().diso_last_error().decode(errors="replace")
        raise RuntimeError(f"diso_amd: {what} failed ({rc}): {msg}")
