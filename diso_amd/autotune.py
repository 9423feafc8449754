"""Launch-parameter autotuner.

The engine has two process-wide launch parameters (include/diso_hip.h, diso_set_param): the
number of segments each emission / backward block takes (the "run", 1..64) and the number of
resident count blocks per CU (1..4). Results never depend on them; speed does, by device and
workload. `tune()` times DiffMC + DiffDMC forward+backward on a synthetic grid for every
candidate, applies the fastest and can save it; `load()` (called when the library is first
loaded) applies a saved choice.

    python -m diso_amd.autotune --n 256 --save
"""
import argparse
import json
import os
import time

from . import _lib

TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
RUNS = (8, 16, 32, 64)
COUNT_BLOCKS_PER_CU = (2, 3, 4)


def apply(params):
    L = _lib.lib()
    if "run" in params:
        _lib.check(L.diso_set_param(_lib.PARAM_RUN, int(params["run"])), "set run length")
    if "count_blocks_per_cu" in params:
        _lib.check(L.diso_set_param(_lib.PARAM_COUNT_BLOCKS_PER_CU, int(params["count_blocks_per_cu"])),
                   "set count blocks per CU")


def current():
    L = _lib.lib()
    return {"run": L.diso_get_param(_lib.PARAM_RUN),
            "count_blocks_per_cu": L.diso_get_param(_lib.PARAM_COUNT_BLOCKS_PER_CU)}


def load(path=TUNING_FILE):
    """Applies the saved parameters for this device, if any; returns them (or None)."""
    if not os.path.exists(path):
        return None
    try:
        import torch

        name = torch.cuda.get_device_name(0) if torch.cuda.is_available() else None
        saved = json.load(open(path))
        params = saved.get(name) if name else None
    except (OSError, ValueError, RuntimeError):
        return None
    if params:
        apply(params)
    return params


def tune(n=256, reps=5, save=False, path=TUNING_FILE, verbose=True):
    """Times DiffMC + DiffDMC forward+backward (rounded box + noise, deform) at n^3 for every
    candidate and applies the fastest."""
    import torch

    import diso_amd
    from diso_amd import workloads

    dev = torch.device("cuda:0")
    sdf_np, def_np, rng = workloads.noisy_round_box(n, seed=0)
    sdf0, def0 = torch.tensor(sdf_np, device=dev), torch.tensor(def_np, device=dev)
    mods = [(diso_amd.DiffMC(), {}), (diso_amd.DiffDMC(), {"return_quads": True})]
    ups = []
    with torch.no_grad():
        for mod, kw in mods:
            v, _ = mod(sdf0, def0, **kw)
            ups.append(torch.randn(v.shape, device=dev))

    def step():
        for (mod, kw), g in zip(mods, ups):
            sdf, deform = sdf0.detach().requires_grad_(True), def0.detach().requires_grad_(True)
            v, _ = mod(sdf, deform, **kw)
            torch.autograd.backward(v, grad_tensors=g)

    best = None
    for run in RUNS:
        for cpc in COUNT_BLOCKS_PER_CU:
            params = {"run": run, "count_blocks_per_cu": cpc}
            apply(params)
            step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                step()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / reps * 1e3
            if verbose:
                print(f"run={run} count_blocks_per_cu={cpc}: {ms:.3f} ms per MC+DMC fwd+bwd", flush=True)
            if best is None or ms < best[0]:
                best = (ms, params)
    apply(best[1])
    if save:
        name = torch.cuda.get_device_name(0)
        saved = json.load(open(path)) if os.path.exists(path) else {}
        saved[name] = dict(best[1], ms=round(best[0], 3), n=n)
        json.dump(saved, open(path, "w"), indent=1, sort_keys=True)
    return best[1]


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--save", action="store_true")
    a = ap.parse_args()
    print(json.dumps(tune(a.n, a.reps, a.save)))
