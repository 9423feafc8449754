"""Serving loop: continuous batching of extraction requests on one GPU.

Requests (a grid already on the device, MC or DMC, optionally a backward with an upstream
gradient) are queued by any thread. Each worker owns a HIP stream and a BatchExtractor; when free,
it takes the queued requests that share the oldest one's settings (up to `max_batch`) and runs
them as one batch, so work that arrives while other batches are on the device is admitted as soon
as a worker and memory free up, and the batches of different workers overlap on the device.

Admission is by a device-memory budget, in two steps. Taking a request charges its counting
workspace (known from its dims). The batch's counting pass then gives every request's exact
vertex and face counts, hence its exact footprint (workspace, outputs, and for a backward the
upstream gradient, scratch and input gradients); the requests that fit the free budget, in queue
order, run, and the others go back to the head of the queue. A request larger than the whole
budget runs alone. Results (and latencies, submit -> result ready) come back through futures.

Each result equals what DiffMC / DiffDMC return for that request alone, bit for bit: a batch runs
the same kernels per grid (diso_amd/batch.py).
"""
import collections
import concurrent.futures
import threading
import time

import torch

from . import _lib
from .batch import BatchExtractor
from .engine import Extractor, split_quads

__all__ = ["Request", "Server"]


class Request:
    """One extraction: `sdf` [X,Y,Z] (and `deform` [X,Y,Z,3] or None) on the server's device,
    `algo` "mc" / "dmc". With `upstream` ([V,3], or a callable V -> [V,3] tensor) the result also
    holds the gradients with respect to sdf and deform."""

    def __init__(self, sdf, algo, deform=None, isovalue=0.0, normalize=True, return_quads=True, upstream=None):
        self.sdf, self.algo, self.deform = sdf, algo, deform
        self.isovalue, self.normalize, self.return_quads = float(isovalue), bool(normalize), bool(return_quads)
        self.upstream = upstream
        self.future = concurrent.futures.Future()
        self.t_submit = None
        self.t_done = None
        self.charge = 0

    def key(self):
        """Requests batched together share these (one BatchExtractor call)."""
        return (self.isovalue, self.normalize, self.return_quads)

    def workspace(self):
        """Device bytes of the counting pass's workspace: a 16-bit word per padded position and a
        16-byte base per 64 positions."""
        X, Y, Z = (int(d) for d in self.sdf.shape)
        pos = (X + 2) * (Y + 2) * (Z + 2)
        return 2 * pos + pos // 4 + 4096

    def footprint(self, n_verts, n_faces):
        """Device bytes the engine holds for this request given its counts: workspace, vertices,
        faces (DMC quads, plus triangles and split scratch without `return_quads`), DMC vertex
        info, and with a backward the upstream gradient, the DMC edge-gradient scratch and the
        input gradients."""
        s = self.sdf.element_size()
        dmc = self.algo == "dmc"
        out = 3 * s * n_verts + (32 if dmc else 24) * n_faces + (n_verts if dmc else 0)
        if dmc and not self.return_quads:
            out += 49 * n_faces + 16384
        grads = 0
        if self.upstream is not None:
            grads = self.sdf.numel() * s + (self.deform.numel() * s if self.deform is not None else 0)
            grads += 3 * s * n_verts + (3 * s * n_faces if dmc else 0)
        return self.workspace() + out + grads


class Server:
    def __init__(self, dtype=torch.float32, device="cuda:0", budget_bytes=200 << 30, max_batch=64, workers=4,
                 reference_compat=False, max_batch_voxels=1 << 20):
        self.dtype = dtype
        self.device = torch.device(device)
        self.budget = int(budget_bytes)
        self.max_batch = int(max_batch)
        # A batch runs pass by pass over all its grids, so its inputs should stay cache-resident
        # between the passes: grids are batched up to this many voxels; a request that reaches it
        # alone runs through the per-grid engine.
        self.max_batch_voxels = int(max_batch_voxels)
        self.reference_compat = reference_compat
        self._queue = collections.deque()
        self._cv = threading.Condition()
        self._reserved = 0
        self.peak_reserved = 0
        self._closed = False
        self.latencies = []
        self.batches = 0
        self._threads = [threading.Thread(target=self._work, daemon=True, name=f"diso-serve-{i}")
                         for i in range(workers)]
        _lib.lib()
        for t in self._threads:
            t.start()

    def submit(self, req):
        if req.sdf.device != self.device or req.sdf.dtype != self.dtype:
            raise ValueError("request grid must be on the server's device and of its dtype")
        req.t_submit = time.perf_counter()
        with self._cv:
            if self._closed:
                raise RuntimeError("server closed")
            self._queue.append(req)
            self._cv.notify()
        return req.future

    def close(self):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- worker side

    def _admit(self):
        """Pops the requests the next batch counts: the oldest request (if its workspace fits the
        free budget, or nothing else holds memory), then queued requests with the same settings
        whose workspace fits, up to `max_batch` requests and `max_batch_voxels` voxels. Charges
        their workspaces."""
        first = self._queue[0]
        free = self.budget - self._reserved
        if first.workspace() > free and self._reserved > 0:
            return []  # wait for memory
        take = [self._queue.popleft()]
        used = first.workspace()
        voxels = first.sdf.numel()
        skipped = []
        # look at a bounded window of the queue (a long queue costs nothing per admission)
        while (self._queue and len(take) < self.max_batch and voxels < self.max_batch_voxels
               and len(skipped) < 2 * self.max_batch):
            r = self._queue.popleft()
            if (r.key() == first.key() and used + r.workspace() <= free
                    and voxels + r.sdf.numel() <= self.max_batch_voxels):
                take.append(r)
                used += r.workspace()
                voxels += r.sdf.numel()
            else:
                skipped.append(r)
        self._queue.extendleft(reversed(skipped))
        for r in take:
            r.charge = r.workspace()
        self._reserved += used
        self.peak_reserved = max(self.peak_reserved, self._reserved)
        return take

    def _fit(self, batch, sizes):
        """With the exact counts: charges the footprints of the longest queue-order prefix of the
        batch that fits the free budget and returns the rest to the head of the queue, uncharged.
        The counting workspace is already released, so the batch holds nothing while it waits
        for other batches to free memory; with nothing else running, the first request runs
        whatever its size. Called with the lock held."""
        self._reserved -= sum(r.charge for r in batch)
        for r in batch:
            r.charge = 0
        while True:
            free = self.budget - self._reserved
            run, need = [], 0
            for r, (nv, nf, _) in zip(batch, sizes):
                f = r.footprint(nv, nf)
                if need + f > free:
                    break
                run.append(r)
                need += f
            if run or self._reserved == 0:
                break
            self._cv.wait()
        if not run:
            run, need = batch[:1], batch[0].footprint(*sizes[0][:2])
        back = batch[len(run):]
        self._reserved += need
        for r, (nv, nf, _) in zip(run, sizes):
            r.charge = r.footprint(nv, nf)
        self._queue.extendleft(reversed(back))
        self.peak_reserved = max(self.peak_reserved, self._reserved)
        self._cv.notify_all()
        return run

    def _release(self, reqs):
        with self._cv:
            self._reserved -= sum(r.charge for r in reqs)
            for r in reqs:
                r.charge = 0
            self._cv.notify_all()

    def _work(self):
        stream = torch.cuda.Stream(device=self.device)  # default priority
        ex = BatchExtractor(self.dtype, reference_compat=self.reference_compat)
        single = {"mc": Extractor(_lib.MC, self.dtype), "dmc": Extractor(_lib.DMC, self.dtype)}
        while True:
            with self._cv:
                while True:
                    if self._queue:
                        batch = self._admit()
                        if batch:
                            break
                    elif self._closed:
                        return
                    self._cv.wait()
            try:
                if len(batch) == 1:
                    with torch.cuda.device(self.device), torch.cuda.stream(stream):
                        results = [self._run_one(single[batch[0].algo], batch[0])]
                        stream.synchronize()
                    self._finish(batch, results)
                    continue
                first = batch[0]
                with torch.cuda.device(self.device), torch.cuda.stream(stream):
                    counted = ex.count([r.sdf for r in batch], [r.algo for r in batch], first.isovalue,
                                       first.normalize)
                n_all = len(batch)
                with self._cv:
                    batch = self._fit(batch, counted.sizes)
                with torch.cuda.device(self.device), torch.cuda.stream(stream):
                    # the counting pass is reused unless part of the batch went back to the queue
                    results = self._run(ex, batch, counted if len(batch) == n_all else None)
                    stream.synchronize()
                self._finish(batch, results)
            except Exception as e:  # reported to every request of the batch
                for r in batch:
                    if not r.future.done():
                        r.future.set_exception(e)
            finally:
                self._release(batch)

    def _finish(self, batch, results):
        now = time.perf_counter()
        with self._cv:
            self.batches += 1
            for r in batch:
                r.t_done = now
                self.latencies.append(now - r.t_submit)
        for r, res in zip(batch, results):
            r.future.set_result(res)

    def _run_one(self, ex, r):
        """A request alone, through the per-grid engine (the passes DiffMC / DiffDMC run, without
        autograd), admitted on its exact counts."""
        dmc = r.algo == "dmc"
        flags = _lib.FLAG_VIRTUAL_PAD | _lib.FLAG_SHIFT | _lib.FLAG_INDEX64
        if r.normalize:
            flags |= _lib.FLAG_NORMALIZE
        if self.reference_compat and dmc:
            flags |= _lib.FLAG_REFERENCE_COMPAT
        grid = r.sdf.detach().contiguous()
        deform = None if r.deform is None else r.deform.detach().contiguous()
        counted = ex.count(grid, r.isovalue, flags)
        if counted.has_nan:
            raise RuntimeError("diso_amd: sdf grid contains NaN values")
        empty = counted.empty_grid
        nv, nf = (0, 0) if empty else (counted.n_verts, counted.n_faces)
        with self._cv:
            self._fit([r], [(nv, nf, counted.flags)])
        if empty:  # diso/__init__.py:49-50,103-104
            return (torch.zeros((0, 3), dtype=self.dtype, device=grid.device),
                    torch.zeros((0, 4 if dmc else 3), dtype=torch.int32, device=grid.device))
        verts, faces = ex.emit(grid, deform, counted, flags)
        if dmc and not r.return_quads:
            faces = split_quads(verts, faces)
        if r.upstream is None or nv == 0:
            return verts, faces
        up = r.upstream(nv) if callable(r.upstream) else r.upstream
        gs = torch.empty_like(grid)
        gd = None if deform is None else torch.empty_like(deform)
        ex.backward(grid, deform, counted, flags, up.to(self.dtype), gs, gd)
        return verts, faces, gs, gd

    def _run(self, ex, batch, counted=None):
        first = batch[0]
        want = [r.upstream is not None for r in batch]
        sdfs = [r.sdf.detach().requires_grad_(w) for r, w in zip(batch, want)]
        defs = [None if r.deform is None else r.deform.detach().requires_grad_(w) for r, w in zip(batch, want)]
        outs = ex(sdfs, [r.algo for r in batch], defs, isovalue=first.isovalue, normalize=first.normalize,
                  return_quads=first.return_quads, counted=counted)
        grads = [None] * len(batch)
        live = [i for i, w in enumerate(want) if w and outs[i][0].shape[0]]
        if live:
            ups = []
            for i in live:
                u = batch[i].upstream
                ups.append(u(outs[i][0].shape[0]) if callable(u) else u)
            inputs = []
            for i in live:
                inputs.append(sdfs[i])
                if defs[i] is not None:
                    inputs.append(defs[i])
            g = torch.autograd.grad([outs[i][0] for i in live], inputs, grad_outputs=ups, allow_unused=True)
            k = 0
            for i in live:
                gs = g[k]
                k += 1
                gd = None
                if defs[i] is not None:
                    gd = g[k]
                    k += 1
                grads[i] = (gs, gd)
        return [(v.detach(), f) if gr is None else (v.detach(), f, gr[0], gr[1]) for (v, f), gr in zip(outs, grads)]

    def latency_ms(self, q):
        import numpy as np

        return float(np.percentile(np.array(self.latencies) * 1e3, q)) if self.latencies else None
