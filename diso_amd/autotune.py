"""Launch-parameter autotuner.

The engine has two launch parameters per size bucket (include/diso_hip.h, diso_set_bucket_param):
the number of segments each emission / backward block takes (the "run", 1..64) and the number of
resident count blocks per CU (1..4). A launch takes the bucket of its segment count; buckets hold
grids of about 32^3, 64^3, 128^3, 256^3 and 512^3 samples (a batched count: all its grids
together; batched emission and backward: each grid's own run length). Results never depend on
the parameters; speed does, by device and grid size.

`tune()` times, per bucket, the count launches of DiffMC + DiffDMC for every count-blocks choice
and their emission + backward launches for every run length (the two parameters act on different
kernels, so they are searched independently), on a rounded box with noise and a deformation at
the bucket's size. It applies the fastest and can save them to tuning.json under the device name;
`load()` (called when the library is first loaded) applies the saved choice.

    python -m diso_amd.autotune --save
"""
import argparse
import ctypes
import json
import os
import time

from . import _lib

TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
BUCKET_SIZES = (32, 64, 128, 256, 512)  # representative cubic grid per bucket (bucket b: index b)
RUNS = (1, 2, 4, 8, 16, 32, 64)
COUNT_BLOCKS_PER_CU = (1, 2, 3, 4)
_applied = {"source": "library defaults"}


def _set(which, bucket, value, what):
    _lib.check(_lib.lib().diso_set_bucket_param(which, bucket, int(value)), what)


def apply(params):
    """Applies launch parameters: either one choice for every bucket ({"run": r,
    "count_blocks_per_cu": c}, either key optional) or a table keyed by bucket size ({"32": {...},
    ..., "512": {...}}, as current() returns)."""
    if any(k in params for k in ("run", "count_blocks_per_cu")):
        params = {str(n): params for n in BUCKET_SIZES}
    for key, p in params.items():
        if key not in {str(n) for n in BUCKET_SIZES}:
            raise ValueError(f"unknown size bucket {key!r}")
        b = BUCKET_SIZES.index(int(key))
        if "run" in p:
            _set(_lib.PARAM_RUN, b, p["run"], "set run length")
        if "count_blocks_per_cu" in p:
            _set(_lib.PARAM_COUNT_BLOCKS_PER_CU, b, p["count_blocks_per_cu"], "set count blocks per CU")


def current():
    """The parameters in force, per bucket: {"32": {"run": r, "count_blocks_per_cu": c}, ...}."""
    L = _lib.lib()
    return {str(n): {"run": L.diso_get_bucket_param(_lib.PARAM_RUN, b),
                     "count_blocks_per_cu": L.diso_get_bucket_param(_lib.PARAM_COUNT_BLOCKS_PER_CU, b)}
            for b, n in enumerate(BUCKET_SIZES)}


def applied():
    """Where the parameters in force came from (tuning file and device, or library defaults)."""
    return dict(_applied)


def _device_name():
    """Key of this device in the tuning file: its name and architecture (e.g. "AMD Radeon
    Graphics (gfx950)"; ROCm reports MI355X under that name)."""
    import torch

    if not torch.cuda.is_available():
        return None
    arch = getattr(torch.cuda.get_device_properties(0), "gcnArchName", "").split(":")[0]
    return f"{torch.cuda.get_device_name(0)} ({arch})" if arch else torch.cuda.get_device_name(0)


def load(path=TUNING_FILE):
    """Applies the saved parameters for this device, if any; returns them (or None)."""
    if not os.path.exists(path):
        return None
    try:
        name = _device_name()
        with open(path) as f:
            saved = json.load(f)
        entry = saved.get(name) if name else None
    except (OSError, ValueError, RuntimeError):
        return None
    if not entry:
        return None
    params = entry.get("buckets", entry)
    apply(params)
    _applied.clear()
    _applied.update({"source": os.path.relpath(path, os.path.dirname(os.path.dirname(path))), "device": name})
    return params


def _inputs(n, dev, seed=0):
    """Rounded box (half-extent 0.35, radius 0.1) + U(-0.05, 0.05) noise, deform U(-0.3, 0.3),
    generated on the device."""
    from . import workloads

    return workloads.noisy_round_box_torch(n, seed=seed, device=dev)


def _gpu_ms(fn, reps, inner):
    """GPU time (ms) of one fn(): `inner` calls captured in a HIP graph (so host overhead does not
    hide the kernels of small grids), replayed `reps` times between events."""
    import torch

    fn()  # eager once: first-call set-up stays out of the graph
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(inner):
            fn()
    g.replay()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        g.replay()
    e1.record(st)
    torch.cuda.synchronize()
    del g
    return e0.elapsed_time(e1) / (reps * inner)


def _bucket_launches(n, dev):
    """(count, emit+backward) launch closures of DiffMC + DiffDMC on the bucket's grid, through
    the C ABI into preallocated buffers (no allocation, no host synchronisation)."""
    import torch

    import diso_amd

    sdf, deform = _inputs(n, dev)
    flags = _lib.FLAG_VIRTUAL_PAD | _lib.FLAG_SHIFT | _lib.FLAG_NORMALIZE | _lib.FLAG_INDEX64
    L = _lib.lib()
    ptr = lambda t: ctypes.c_void_p(0 if t is None else t.data_ptr())
    stream = lambda: ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    state = []
    for mod in (diso_amd.DiffMC(torch.float32), diso_amd.DiffDMC(torch.float32)):
        ex = mod._ex
        c = ex.count(sdf, 0.0, flags)
        verts, faces = ex.emit(sdf, deform, c, flags)
        vinfo = c.vinfo if ex.algo == _lib.DMC else None
        up = torch.randn_like(verts)
        gs, gd = torch.empty_like(sdf), torch.empty_like(deform)
        nsc = L.diso_scratch_bytes(ex.algo, ex.code, c.n_crossing)
        scratch = torch.empty(max(nsc, 1), dtype=torch.uint8, device=dev)
        ctl = ex.ctl_for(sdf.device, c.dims, flags)
        state.append((ex, c, verts, faces, vinfo, up, gs, gd, scratch, nsc, ctl))

    def count():
        for ex, c, *_, ctl in state:
            _lib.check(L.diso_count(ex.algo, ex.code, ptr(sdf), *c.dims, 0.0, flags, ptr(c.seg), c.seg.numel(),
                                    ptr(ctl), ctl.numel(), None, stream()), "count")

    def emit_backward():
        for ex, c, verts, faces, vinfo, up, gs, gd, scratch, nsc, _ in state:
            st = stream()
            _lib.check(L.diso_emit(ex.algo, ex.code, ptr(sdf), ptr(deform), *c.dims, 0.0, flags, ptr(c.seg),
                                   c.seg.numel(), ptr(verts), c.n_verts, ptr(faces), c.n_faces, ptr(vinfo), st),
                       "emit")
            dmc = ex.algo == _lib.DMC
            _lib.check(L.diso_backward(ex.algo, ex.code, ptr(sdf), ptr(deform), *c.dims, 0.0, flags, ptr(c.seg),
                                       c.seg.numel(), ptr(faces if dmc else None), c.n_faces if dmc else 0,
                                       ptr(vinfo), ptr(up), c.n_verts, ptr(gs), ptr(gd), ptr(scratch), nsc, st),
                       "backward")

    return count, emit_backward


def tune(sizes=BUCKET_SIZES, reps=10, save=False, path=TUNING_FILE, verbose=True):
    """Per bucket, the count-blocks choice with the fastest MC + DMC count launches and the run
    length with the fastest MC + DMC emission + backward launches (GPU time of graph replays,
    events on the launching stream). Applies them; with save=True writes them to `path` under the
    device name."""
    import torch

    dev = torch.device("cuda:0")
    before = current()
    table, report = dict(before), {}
    for n in sizes:
        key, b = str(n), BUCKET_SIZES.index(n)
        count, emit_backward = _bucket_launches(n, dev)
        inner = 20 if n <= 64 else 4 if n <= 256 else 1
        t_cpc, t_run = {}, {}
        for c in COUNT_BLOCKS_PER_CU:
            _set(_lib.PARAM_COUNT_BLOCKS_PER_CU, b, c, "set count blocks per CU")
            t_cpc[c] = _gpu_ms(count, reps, inner)
        for r in RUNS:
            _set(_lib.PARAM_RUN, b, r, "set run length")
            t_run[r] = _gpu_ms(emit_backward, reps, inner)
        best_c, best_r = min(t_cpc, key=t_cpc.get), min(t_run, key=t_run.get)
        _set(_lib.PARAM_COUNT_BLOCKS_PER_CU, b, best_c, "set count blocks per CU")
        _set(_lib.PARAM_RUN, b, best_r, "set run length")
        table[key] = {"run": best_r, "count_blocks_per_cu": best_c}
        default = t_cpc[4] + t_run[32]
        best = t_cpc[best_c] + t_run[best_r]
        report[key] = {"count_ms": {str(k): round(v, 4) for k, v in t_cpc.items()},
                       "emit_backward_ms": {str(k): round(v, 4) for k, v in t_run.items()},
                       "step_ms_default": round(default, 4), "step_ms_tuned": round(best, 4)}
        if verbose:
            print(f"N={n}: count blocks/CU {best_c} (count {t_cpc[best_c]:.4f} ms), run {best_r} "
                  f"(emit+backward {t_run[best_r]:.4f} ms); MC+DMC fwd+bwd launches {default:.4f} -> {best:.4f} ms",
                  flush=True)
        del count, emit_backward
        torch.cuda.empty_cache()
    apply(table)
    if save:
        name = _device_name()
        saved = {}
        if os.path.exists(path):
            with open(path) as f:
                saved = json.load(f)
        saved[name] = {"buckets": table, "measured": report, "reps": reps,
                       "workload": "DiffMC + DiffDMC(quads) count / emit + backward launches, rounded box + "
                                   "U(-0.05,0.05) noise, deform U(-0.3,0.3), fp32, N per bucket",
                       "date": time.strftime("%Y-%m-%d")}
        with open(path, "w") as f:
            json.dump(saved, f, indent=1, sort_keys=True)
            f.write("\n")
        _applied.clear()
        _applied.update({"source": "diso_amd/tuning.json", "device": name})
    return table


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", type=int, nargs="+", default=list(BUCKET_SIZES))
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--save", action="store_true")
    ap.add_argument("--path", default=TUNING_FILE)
    a = ap.parse_args()
    print(json.dumps(tune(tuple(a.sizes), a.reps, a.save, a.path)))
