"""Benchmark: DiffMC + DiffDMC forward+backward on a 256^3 fp32 grid, one MI355X.

Workload (the N=256 forward+backward configuration): sdf = rounded box (half-extent 0.35,
radius 0.1) + i.i.d. U(-0.05, 0.05) noise, deform = U(-0.3, 0.3) [N,N,N,3], upstream dL/dverts
~ N(0,1) [V,3], all from numpy default_rng(seed) in that order, generated in float64 and cast to
float32. One step = DiffMC forward+backward + DiffDMC forward+backward (quads) through the public
API, each on its own copy of the inputs; two input sets alternate between steps so that the
256 MB last-level cache cannot hold a step's inputs.

Reported value: Gvoxel/s = (voxels extracted per step) / (seconds per step) / 1e9, with
2 * N^3 voxels per step (one MC and one DMC extraction). Inputs are resident in HBM before the
timed region. `roofline` times each engine kernel with HIP events on the stream it runs on and
reports the slowest one against HBM peak using its algorithmic bytes (sdf / deform / gradient
samples it must touch plus the vertices, faces and dense gradients it must write).
`cpu_baseline` runs the OpenMP CPU oracle (the reference's per-cell algorithm) on the same
workload on the host cores.

`roofline.step_frac` is the whole step's algorithmic bytes (both algorithms' count, emit and
backward) over `ms_per_step`, against HBM peak.

`config.launch_params` names the launch parameters in force (diso_amd/tuning.json for this
device when the autotuner saved one, else the library defaults), the size bucket of the headline
grid and the table of all buckets.

Further legs, keys of the same line: `batched_step` (the headline step as one batched call over
its two extractions, against the plain step); `sphere_64` (the reference example's 64^3 sphere,
DiffMC forward, one call at a time); `round_cube_64` (a 64^3 rounded box, one call at a time, for
context beside the reference's README table); `batched` (64 grids, N in 32..256) and
`batched_small` (512 grids, N in 16..64), each as one batched call per pass against a loop of
per-grid calls; `serving` (continuous batching under Poisson arrivals); `learn` (graph-captured
512^3 fit).
"""
import argparse
import ctypes
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--n", type=int, default=256)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--kernel-reps", type=int, default=10)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-reps", type=int, default=5)
    p.add_argument("--cube-reps", type=int, default=50, help="calls per module of the 64^3 round-cube leg (0: skip)")
    p.add_argument("--copy-mib", type=int, default=256)
    p.add_argument("--batch-count", type=int, default=64, help="grids of the uneven batch (0: skip it)")
    p.add_argument("--batch-small-count", type=int, default=512,
                   help="grids of the small-grid batch, N in 16..64 (0: skip it)")
    p.add_argument("--batch-reps", type=int, default=5)
    p.add_argument("--learn-n", type=int, default=512, help="learnable-sdf fit grid (0: skip the fit)")
    p.add_argument("--learn-stored-n", type=int, default=384)
    p.add_argument("--learn-steps", type=int, default=100)
    p.add_argument("--serve-requests", type=int, default=20000, help="requests of the serving stream (0: skip it)")
    p.add_argument("--serve-load", type=float, default=0.95, help="arrival rate / the plain path's service rate")
    p.add_argument("--serve-workers", type=int, default=4)
    return p.parse_args()


def make_inputs(n, seed, dev):
    import torch

    from diso_amd import workloads

    sdf, deform, rng = workloads.noisy_round_box(n, seed=seed)
    return (torch.tensor(sdf, device=dev), torch.tensor(deform, device=dev), rng, sdf, deform)


def touched_vertices(sdf, iso=0.0):
    """Grid vertices incident to at least one crossing edge of the padded grid."""
    import torch
    import torch.nn.functional as F

    g = F.pad(sdf, (1, 1, 1, 1, 1, 1), value=iso + 1.0)
    inside = g >= iso
    t = torch.zeros_like(inside)
    for ax in range(3):
        a = inside.narrow(ax, 0, inside.shape[ax] - 1)
        b = inside.narrow(ax, 1, inside.shape[ax] - 1)
        c = a != b
        t.narrow(ax, 0, inside.shape[ax] - 1).logical_or_(c)
        t.narrow(ax, 1, inside.shape[ax] - 1).logical_or_(c)
    return int(t[1:-1, 1:-1, 1:-1].sum().item())


def time_kernels(ex, sdf, deform, flags, g, reps):
    """Average duration (ms) of the count, emit and backward launches, HIP events on the
    current stream (the one the library launches on)."""
    import torch

    from diso_amd import _lib

    st = torch.cuda.current_stream()
    res = {}
    counted = ex.count(sdf, 0.0, flags)
    verts, faces = ex.emit(sdf, deform, counted, flags)
    gs, gd = torch.empty_like(sdf), torch.empty_like(deform)
    L = _lib.lib()
    ctl = ex.ctl_for(sdf.device, counted.dims, flags)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    sp = ctypes.c_void_p(st.cuda_stream)

    def run_count():
        rc = L.diso_count(ex.algo, ex.code, ptr(sdf), *counted.dims, 0.0, flags, ptr(counted.seg),
                          counted.seg.numel(), ptr(ctl), ctl.numel(), None, sp)
        _lib.check(rc, "count")

    def run_emit():
        ex.emit(sdf, deform, counted, flags)

    def run_bwd():
        ex.backward(sdf, deform, counted, flags, g, gs, gd)

    for name, fn in (("count", run_count), ("emit", run_emit), ("backward", run_bwd)):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            fn()
        e1.record(st)
        torch.cuda.synchronize()
        res[name] = e0.elapsed_time(e1) / reps
    return res, counted


def algorithmic_bytes(kind, n, s, tv, nv, nf, k):
    """Bytes a kernel must move (the path's byte model, split by kernel):
    count: every sample once; emit: samples + deform at touched vertices, vertices, int64 faces;
    backward: samples + deform at touched vertices, upstream vertex gradients, dense sdf and deform
    gradients."""
    n3 = n ** 3
    if kind == "count":
        return s * n3
    if kind == "emit":
        return s * tv + 3 * s * tv + 3 * s * nv + 8 * k * nf
    return s * tv + 3 * s * tv + 3 * s * nv + s * n3 + 3 * s * n3


def copy_calibration(dev, mib, reps):
    """A plain elementwise copy (torch.neg, fp32) of `mib` MiB in and out: its duration is the
    practical HBM streaming rate, and under rocprofv3 its FETCH_SIZE / WRITE_SIZE readings
    (known to be `mib` MiB each) calibrate the counters' units."""
    import torch

    n = (mib << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.neg(src, out=dst)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        torch.neg(src, out=dst)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return {"kernel": "torch.neg fp32", "bytes_read": n * 4, "bytes_written": n * 4, "ms": round(ms, 4),
            "GBs": round(2 * n * 4 / (ms / 1e3) / 1e9, 1)}


def profiled_traffic(phase, n):
    """HBM bytes per launch of `phase` measured with rocprofv3 FETCH_SIZE / WRITE_SIZE passes of
    this workload (calibrated against a plain copy), from the committed summary; None when there
    is no summary for this workload."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_bench_n{n}.txt")
    if not os.path.exists(path):
        return None
    for ln in open(path):
        f = ln.split()
        if len(f) > 2 and f[0] == "traffic" and f[1] == phase:
            kv = dict(x.split("=", 1) for x in f[2:] if "=" in x)
            return int(kv["total_bytes"])
    return None


def batched_vs_plain(dev, count, seed, reps, sizes=(32, 48, 64, 96, 128, 192, 256)):
    """The uneven batch: `count` grids, N uniform over `sizes`, mixed families, deform,
    MC/DMC alternating, DMC returning quads; forward + backward of all grids per pass, timed as
    one BatchExtractor call and as a Python loop of per-grid DiffMC / DiffDMC calls."""
    import torch

    import diso_amd
    from diso_amd import workloads

    data = workloads.uneven_batch_torch(count, seed=seed, sizes=sizes, device=dev)
    voxels = sum(n ** 3 for *_, n in data)
    algos = [a for _, _, a, _, _ in data]
    mods = {"mc": diso_amd.DiffMC(torch.float32), "dmc": diso_amd.DiffDMC(torch.float32)}
    bex = diso_amd.BatchExtractor(torch.float32)
    gen = torch.Generator(device=dev).manual_seed(seed + 7)
    ups = None

    def leaves():
        return ([s.detach().requires_grad_(True) for s, *_ in data],
                [d.detach().requires_grad_(True) for _, d, *_ in data])

    def plain():
        sdfs, defs = leaves()
        outs = []
        for i in range(count):
            kw = {"return_quads": True} if algos[i] == "dmc" else {}
            outs.append(mods[algos[i]](sdfs[i], defs[i], 0.0, **kw))
        return outs

    def batched():
        sdfs, defs = leaves()
        return bex(sdfs, algos, defs, return_quads=True)

    def step(fn):
        outs = fn()
        live = [(v, u) for (v, _), u in zip(outs, ups) if v.shape[0]]
        torch.autograd.backward([v for v, _ in live], grad_tensors=[u for _, u in live])

    with torch.no_grad():
        ups = [torch.randn(tuple(v.shape), generator=gen, device=dev) for v, _ in plain()]
    res = {}
    for name, fn in (("plain_loop", plain), ("batched", batched)):
        for _ in range(2):
            step(fn)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            step(fn)
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / reps
    return {
        "workload": f"{count} grids, N uniform over {{{','.join(str(x) for x in sizes)}}} (total {voxels} voxels), "
                    "sphere / rounded box / blobs / noise, deform U(-0.3,0.3), MC and DMC alternating, fwd+bwd",
        "voxels": int(voxels),
        "batched_gvoxel_s": round(voxels / res["batched"] / 1e9, 3),
        "plain_loop_gvoxel_s": round(voxels / res["plain_loop"] / 1e9, 3),
        "batched_ms": round(res["batched"] * 1e3, 3),
        "plain_loop_ms": round(res["plain_loop"] * 1e3, 3),
        "speedup": round(res["plain_loop"] / res["batched"], 3),
    }


def batched_step(sets, ups, steps, warmup):
    """The headline step (DiffMC + DiffDMC(quads) fwd+bwd on one grid) as one BatchExtractor call
    over the two extractions, against the plain per-module step, alternating the two input sets
    as the headline loop does; wall time over `steps` after `warmup`."""
    import torch

    import diso_amd

    n = sets[0][0].shape[0]
    mods = {"mc": diso_amd.DiffMC(torch.float32), "dmc": diso_amd.DiffDMC(torch.float32)}
    bex = diso_amd.BatchExtractor(torch.float32)

    def plain(i):
        sdf0, def0 = sets[i % 2][0], sets[i % 2][1]
        for name, mod in mods.items():
            sdf, deform = sdf0.detach().requires_grad_(True), def0.detach().requires_grad_(True)
            kw = {"return_quads": True} if name == "dmc" else {}
            v, _ = mod(sdf, deform, 0.0, **kw)
            torch.autograd.backward(v, grad_tensors=ups[i % 2][name])

    def batched(i):
        sdf0, def0 = sets[i % 2][0], sets[i % 2][1]
        sdfs = [sdf0.detach().requires_grad_(True) for _ in range(2)]
        defs = [def0.detach().requires_grad_(True) for _ in range(2)]
        outs = bex(sdfs, ["mc", "dmc"], defs, return_quads=True)
        torch.autograd.backward([outs[0][0], outs[1][0]], grad_tensors=[ups[i % 2]["mc"], ups[i % 2]["dmc"]])

    res = {}
    for name, fn in (("plain", plain), ("batched", batched)):
        for i in range(warmup):
            fn(i)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            fn(i)
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / steps
    vox = 2 * n ** 3
    return {"workload": f"the headline step (DiffMC + DiffDMC(quads) fwd+bwd, N={n}) as one BatchExtractor call "
                        "over its two extractions vs the plain per-module calls",
            "batched_gvoxel_s": round(vox / res["batched"] / 1e9, 3),
            "plain_gvoxel_s": round(vox / res["plain"] / 1e9, 3),
            "batched_ms": round(res["batched"] * 1e3, 3), "plain_ms": round(res["plain"] * 1e3, 3),
            "speedup": round(res["plain"] / res["batched"], 3)}


def learnable_fit(dev, n, n_stored, steps):
    """The differentiation-at-scale workload (diso_amd.fit): a learnable fp32 sdf (sphere r=0.4 +
    N(0, 0.01)) and deform (U(-0.1, 0.1)) fitted by Adam (lr 1e-3) to a sphere of radius 0.3
    through DiffDMC vertices. At n^3 with the recompute backward and at n_stored^3 with the stored
    classification: a few eager steps, then `steps` replays of the graph-captured step, timed with
    HIP events on the replay stream."""
    import torch

    from diso_amd.fit import SdfFit

    out = {}
    for label, size, recompute in (("recompute", n, True), ("stored", n_stored, False)):
        torch.cuda.reset_peak_memory_stats(dev)
        fit = SdfFit(size, device=dev, seed=0, recompute=recompute)
        fit.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            fit.step()
        torch.cuda.synchronize()
        eager_ms = (time.perf_counter() - t0) / 3 * 1e3
        fit.capture(warmup=3)
        torch.cuda.synchronize()
        loss0 = float(fit.loss)
        st = torch.cuda.current_stream(dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(steps):
            fit.replay()
        e1.record(st)
        torch.cuda.synchronize()
        fit.check()
        ms = e0.elapsed_time(e1) / steps
        out[label] = {"n": size, "recompute_backward": recompute, "graph_ms_per_step": round(ms, 3),
                      "eager_ms_per_step": round(eager_ms, 3), "gvoxel_s": round(size ** 3 / (ms / 1e3) / 1e9, 3),
                      "loss_before": round(loss0, 6), "loss_after": round(float(fit.loss), 6),
                      "dmc_verts": int(fit.counts[0]), "vert_capacity": fit.ex.vert_capacity,
                      "peak_alloc_gib": round(torch.cuda.max_memory_allocated(dev) / 2 ** 30, 2)}
        del fit
        torch.cuda.empty_cache()
    out["workload"] = (f"learnable sdf sphere r=0.4 + N(0,0.01), deform U(-0.1,0.1), Adam lr 1e-3, "
                       f"loss mean((|v-0.5|-0.3)^2) over DiffDMC verts (normalize=True), {steps} graph-captured "
                       f"steps per grid after 7 eager ones")
    return out


def serving(dev, count, load, seed, workers, budget_gib=200):
    """The serving workload (diso_amd.workloads.request_stream): `count` requests, N log-uniform
    over 32..384, mixed families, MC/DMC 1:1, 30 % with a backward. The plain path serves them one
    call at a time in arrival order; the server (diso_amd.serve, continuous batching under a
    `budget_gib` memory budget) takes them as they arrive. Capacity: every request at once (the
    plain path back to back). Latency (result ready - arrival) under Poisson arrivals at `load`
    times the plain path's capacity."""
    import torch

    import diso_amd
    from diso_amd import workloads
    from diso_amd.serve import Request, Server

    grids, reqs = workloads.request_stream(count, 1.0, seed=seed, device=dev)  # arrivals rescaled below
    mods = {"mc": diso_amd.DiffMC(torch.float32), "dmc": diso_amd.DiffDMC(torch.float32)}
    st = torch.cuda.current_stream(dev)
    voxels = sum(grids[i][3] ** 3 for _, i, _, _ in reqs)

    def plain(i, algo, bwd):
        sdf, deform, _, _ = grids[i]
        s, d = sdf.detach().requires_grad_(bwd), deform.detach().requires_grad_(bwd)
        v, f = mods[algo](s, d, return_quads=True) if algo == "dmc" else mods[algo](s, d)
        if bwd and v.shape[0]:
            torch.autograd.grad([v], [s, d], grad_outputs=[torch.randn_like(v)])
        st.synchronize()

    def wait_until(t0, t):
        dt = t - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)

    def run_plain(arrivals):
        lat = []
        t0 = time.perf_counter()
        for t, (_, i, a, b) in zip(arrivals, reqs):
            wait_until(t0, t)
            plain(i, a, b)
            lat.append(time.perf_counter() - t0 - t)
        return time.perf_counter() - t0, np.array(lat) * 1e3

    def run_served(arrivals):
        # results are dropped as they complete (a client consumes them): only in-flight work holds memory
        finished = threading.Event()
        state = {"left": count, "error": None}
        lock = threading.Lock()

        def on_done(fut):
            with lock:
                if fut.exception() is not None and state["error"] is None:
                    state["error"] = fut.exception()
                state["left"] -= 1
                if state["left"] == 0:
                    finished.set()

        with Server(budget_bytes=budget_gib << 30, workers=workers, device=dev) as srv:
            t0 = time.perf_counter()
            for t, (_, i, a, b) in zip(arrivals, reqs):
                wait_until(t0, t)
                sdf, deform, _, _ = grids[i]
                up = (lambda nv: torch.randn((nv, 3), device=dev)) if b else None
                srv.submit(Request(sdf, a, deform, upstream=up)).add_done_callback(on_done)
            finished.wait()
            span = time.perf_counter() - t0
            if state["error"] is not None:
                raise state["error"]
            return span, np.array(srv.latencies) * 1e3, srv.batches

    for _, i, a, b in reqs[:20]:
        plain(i, a, b)
    now = [0.0] * count
    plain_cap, _ = run_plain(now)
    served_cap, _, _ = run_served(now)
    rate = load * count / plain_cap
    arrivals = [t / rate for t, *_ in reqs]  # unit-rate Poisson times, rescaled
    plain_span, plain_lat = run_plain(arrivals)
    served_span, served_lat, nb = run_served(arrivals)
    pct = lambda a, q: round(float(np.percentile(a, q)), 3)
    return {
        "workload": f"{count} requests, N log-uniform 32..384 (pool of 48 grids), sphere / rounded box / blobs / noise, "
                    "deform U(-0.3,0.3), MC/DMC 1:1, 30% with backward, 200 GiB budget; capacity with every request "
                    f"queued at once, latency under Poisson arrivals at {rate:.0f}/s ({load} x the plain capacity)",
        "voxels": int(voxels),
        "rate_per_s": round(rate, 1),
        "capacity_gvoxel_s": {"served": round(voxels / served_cap / 1e9, 3), "plain": round(voxels / plain_cap / 1e9, 3),
                              "speedup": round(plain_cap / served_cap, 3)},
        "served": {"p50_ms": pct(served_lat, 50), "p99_ms": pct(served_lat, 99), "batches": nb,
                   "gvoxel_s": round(voxels / served_span / 1e9, 3), "workers": workers},
        "plain": {"p50_ms": pct(plain_lat, 50), "p99_ms": pct(plain_lat, 99),
                  "gvoxel_s": round(voxels / plain_span / 1e9, 3)},
    }


def cpu_baseline(sdf_np, deform_np, n, seed, reps):
    from diso_amd import workloads
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    times = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        for algo in (O.MC, O.DMC):
            ex = O.Extractor(algo, np.float32, threads=threads)
            v, _ = ex.forward(sdf_np, deform_np, 0.0, True, return_quads=True)
            g = np.random.default_rng(seed + 1).standard_normal(v.shape).astype(np.float32)
            ex.backward(g)
        dt = time.perf_counter() - t0
        if r > 0:  # first run is a warm-up
            times.append(dt)
    med = float(np.median(times))
    return {
        "value": round(2 * n ** 3 / med / 1e9, 4),
        "unit": "Gvoxel/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full workload: MC+DMC fwd+bwd on the {n}^3 grid, median of {reps} runs after 1 warm-up "
                  f"({med:.2f} s per step; OpenMP oracle built -O3, {threads} threads)",
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model(),
        "runs_s": [round(t, 3) for t in times],
    }


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _per_call(mod, sdf0, mode, reps):
    """End-to-end host time per call (synchronised) of `mod` on sdf0, forward alone or
    forward+backward; p50 and p99 over `reps` calls after 3 warm-up calls, and the output sizes."""
    import torch

    ts = []
    for r in range(reps + 3):
        sdf = sdf0.detach().requires_grad_(mode == "fwd_bwd")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        v, f = mod(sdf, None, 0.0)
        if mode == "fwd_bwd":
            v.sum().backward()
        torch.cuda.synchronize()
        if r >= 3:
            ts.append(time.perf_counter() - t0)
    return (round(float(np.percentile(ts, 50)) * 1e3, 3), round(float(np.percentile(ts, 99)) * 1e3, 3),
            int(v.shape[0]), int(f.shape[0]))


def round_cube(dev, reps, n=64):
    """For context beside the reference's README table (README.md:79-94, RTX 4090, grid size not
    stated): a rounded box (half-extent 0.35, radius 0.1) at n^3, no deform, one DiffMC and one
    DiffDMC call at a time through the public API, forward alone and forward+backward, end to end
    (host time per call, synchronised); p50 over `reps` calls."""
    import torch

    import diso_amd
    from diso_amd import workloads

    sdf0 = torch.tensor(workloads.round_box(n).astype(np.float32), device=dev)
    out = {"grid": n, "workload": f"rounded box half-extent 0.35 radius 0.1 at {n}^3, no deform, end-to-end per call"}
    for name, mod in (("mc", diso_amd.DiffMC(torch.float32)), ("dmc", diso_amd.DiffDMC(torch.float32))):
        res = {}
        for mode in ("fwd", "fwd_bwd"):
            res[mode + "_p50_ms"], res[mode + "_p99_ms"], res["verts"], res["faces"] = _per_call(mod, sdf0, mode, reps)
        out[name] = res
    return out


def sphere_64(dev, reps):
    """The reference example's input (diso/__init__.py:48-61 usage, test/example.py): |p| - 0.4 on
    a 64^3 grid over [-0.5, 0.5]^3, fp32, DiffMC forward (normalize=True), end to end per call."""
    import torch

    import diso_amd
    from diso_amd import workloads

    sdf0 = torch.tensor(workloads.sphere(64).astype(np.float32), device=dev)
    p50, p99, nv, nf = _per_call(diso_amd.DiffMC(torch.float32), sdf0, "fwd", reps)
    return {"workload": "sphere |p| - 0.4 at 64^3, DiffMC forward, end to end per call", "fwd_p50_ms": p50,
            "fwd_p99_ms": p99, "verts": nv, "faces": nf}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "diso_amd benchmarks one GPU; runs on several GPUs are not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    import diso_amd
    from diso_amd import _lib, autotune, workloads

    dev = torch.device("cuda:0")
    n = args.n
    sets = [make_inputs(n, args.seed + i, dev) for i in range(2)]
    mods = {"mc": diso_amd.DiffMC(torch.float32), "dmc": diso_amd.DiffDMC(torch.float32)}

    # Upstream gradients, drawn after V is known (per input set and algorithm).
    ups = []
    for sdf, deform, rng, _, _ in sets:
        d = {}
        for name, mod in mods.items():
            kw = {"return_quads": True} if name == "dmc" else {}
            with torch.no_grad():
                v, _ = mod(sdf, deform, 0.0, **kw)
            d[name] = torch.tensor(workloads.upstream(rng, v.shape[0]), device=dev)
        ups.append(d)

    def step(i):
        sdf0, def0 = sets[i % 2][0], sets[i % 2][1]
        for name, mod in mods.items():
            sdf = sdf0.detach().requires_grad_(True)
            deform = def0.detach().requires_grad_(True)
            kw = {"return_quads": True} if name == "dmc" else {}
            v, f = mod(sdf, deform, 0.0, **kw)
            torch.autograd.backward(v, grad_tensors=ups[i % 2][name])

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    per = []
    st = torch.cuda.current_stream()
    t0 = time.perf_counter()
    for i in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        step(i)
        e1.record(st)
        per.append((e0, e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = np.array([a.elapsed_time(b) for a, b in per])
    ms_per_step = wall / args.steps * 1e3
    value = 2 * n ** 3 / (ms_per_step / 1e3) / 1e9

    # Per-kernel timing and roofline on input set 0.
    sdf, deform = sets[0][0], sets[0][1]
    tv = touched_vertices(sdf)
    flags = _lib.FLAG_VIRTUAL_PAD | _lib.FLAG_SHIFT | _lib.FLAG_NORMALIZE | _lib.FLAG_INDEX64
    kern = {}
    for name, mod in mods.items():
        t, counted = time_kernels(mod._ex, sdf, deform, flags, ups[0][name], args.kernel_reps)
        k = 3 if name == "mc" else 4
        for kind, ms in t.items():
            b = algorithmic_bytes(kind, n, 4, tv, counted.n_verts, counted.n_faces, k)
            kern[f"{name}_{kind}"] = {"ms": round(ms, 4), "bytes": int(b), "GBs": round(b / (ms / 1e3) / 1e9, 1)}
        kern[f"{name}_counts"] = {"verts": counted.n_verts, "faces": counted.n_faces}
    timed = {k: v for k, v in kern.items() if "ms" in v}
    dom = max(timed, key=lambda k: timed[k]["ms"])
    d = timed[dom]
    step_bytes = sum(v["bytes"] for v in timed.values())  # both algorithms' count + emit + backward
    copy = copy_calibration(dev, args.copy_mib, args.kernel_reps)

    out = {
        "metric": "Gvoxel/s fwd+bwd DiffMC/DMC 256^3 fp32, batched vs plain; p50/p99 ms",
        "value": round(value, 3),
        "unit": "Gvoxel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "p50_ms": round(float(np.percentile(step_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(step_ms, 99)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain DiffMC + DiffDMC(quads) fwd+bwd, rounded box + U(-0.05,0.05) noise, "
                               f"deform U(-0.3,0.3), N={n}", "grid": n, "grids_per_step": 2,
                   "mode": "plain per-grid API", "touched_vertices": tv,
                   "launch_params": dict(autotune.applied(), bucket=str(autotune.BUCKET_SIZES[
                       _lib.lib().diso_param_bucket(n, n, n, _lib.FLAG_VIRTUAL_PAD)]), buckets=autotune.current())},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": d["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(d["GBs"] / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(dom, n),
                     "traffic_source": f"profiles/hbm_traffic_bench_n{n}.txt",
                     "step_bytes": int(step_bytes),
                     "step_frac": round(step_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels": kern,
        "copy_calibration": copy,
    }
    out["batched_step"] = batched_step(sets, ups, args.steps, args.warmup)
    if args.cube_reps > 0:
        out["sphere_64"] = sphere_64(dev, args.cube_reps)
        out["round_cube_64"] = round_cube(dev, args.cube_reps)
    if args.batch_count > 0:
        del sets, mods
        torch.cuda.empty_cache()
        out["batched"] = batched_vs_plain(dev, args.batch_count, args.seed, args.batch_reps)
    if args.batch_small_count > 0:
        torch.cuda.empty_cache()
        out["batched_small"] = batched_vs_plain(dev, args.batch_small_count, args.seed + 1, args.batch_reps,
                                                sizes=(16, 24, 32, 48, 64))
    if args.serve_requests > 0:
        torch.cuda.empty_cache()
        out["serving"] = serving(dev, args.serve_requests, args.serve_load, args.seed, args.serve_workers)
    if args.learn_n > 0:
        torch.cuda.empty_cache()
        out["learn"] = learnable_fit(dev, args.learn_n, args.learn_stored_n, args.learn_steps)
    if not args.no_cpu_baseline:
        sdf_np, def_np, _ = workloads.noisy_round_box(n, seed=args.seed)
        out["cpu_baseline"] = cpu_baseline(sdf_np, def_np, n, args.seed, args.cpu_reps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
