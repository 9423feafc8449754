"""Serving loop: continuous batching of extraction requests on one GPU.

Requests (a grid already on the device, MC or DMC, optionally a backward with an upstream
gradient) are queued by any thread. Each worker owns a HIP stream and a BatchExtractor; when free,
it takes every admissible queued request (up to `max_batch`) and runs them as one batch, so work
that arrives while other batches are on the device is admitted as soon as a worker and memory free
up, and the batches of different workers overlap on the device. Admission is by a device-memory
budget: a request is charged its workspace, its outputs and gradients, estimated before the count
pass and corrected from the count pass's exact totals; a request larger than the whole budget runs
alone. Results (and latencies, submit -> result ready) come back through futures.

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

    def sizes(self):
        """Device bytes the engine holds for this request: (workspace, outputs, gradients), the
        outputs sized for a dense surface (at most 12 B per padded position for vertices and
        faces) until the count pass gives the exact ones."""
        X, Y, Z = (int(d) for d in self.sdf.shape)
        s = self.sdf.element_size()
        pos = (X + 2) * (Y + 2) * (Z + 2)
        grads = (self.sdf.numel() + (self.deform.numel() if self.deform is not None else 0)) * s \
            if self.upstream is not None else 0
        return 3 * pos, 12 * pos * s // 4, grads

    def estimate(self):
        return sum(self.sizes())


class Server:
    def __init__(self, dtype=torch.float32, device="cuda:0", budget_bytes=200 << 30, max_batch=64, workers=2,
                 reference_compat=False):
        self.dtype = dtype
        self.device = torch.device(device)
        self.budget = int(budget_bytes)
        self.max_batch = int(max_batch)
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
        """Pops the requests the next batch runs: the oldest request always (alone if it exceeds
        the budget), then every queued request with the same key that fits the remaining budget."""
        first = self._queue[0]
        free = self.budget - self._reserved
        if first.estimate() > free and self._reserved > 0:
            return []  # wait for memory
        take = [self._queue.popleft()]
        used = first.estimate()
        keep = collections.deque()
        while self._queue and len(take) < self.max_batch:
            r = self._queue.popleft()
            if r.key() == first.key() and used + r.estimate() <= max(free, first.estimate()):
                take.append(r)
                used += r.estimate()
            else:
                keep.append(r)
        keep.extend(self._queue)
        self._queue = keep
        for r in take:
            r.charge = r.estimate()
        self._reserved += used
        self.peak_reserved = max(self.peak_reserved, self._reserved)
        return take

    def _release(self, reqs):
        with self._cv:
            self._reserved -= sum(r.charge for r in reqs)
            self._cv.notify_all()

    def _work(self):
        stream = torch.cuda.Stream(device=self.device)  # default priority
        ex = BatchExtractor(self.dtype, reference_compat=self.reference_compat)
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
                with torch.cuda.device(self.device), torch.cuda.stream(stream):
                    results = self._run(ex, batch)
                    stream.synchronize()
                now = time.perf_counter()
                with self._cv:
                    self.batches += 1
                    for r in batch:
                        r.t_done = now
                        self.latencies.append(now - r.t_submit)
                for r, res in zip(batch, results):
                    r.future.set_result(res)
            except Exception as e:  # reported to every request of the batch
                for r in batch:
                    if not r.future.done():
                        r.future.set_exception(e)
            finally:
                self._release(batch)

    def _run(self, ex, batch):
        first = batch[0]
        want = [r.upstream is not None for r in batch]
        sdfs = [r.sdf.detach().requires_grad_(w) for r, w in zip(batch, want)]
        defs = [None if r.deform is None else r.deform.detach().requires_grad_(w) for r, w in zip(batch, want)]
        outs = ex(sdfs, [r.algo for r in batch], defs, isovalue=first.isovalue, normalize=first.normalize,
                  return_quads=first.return_quads)
        # the count pass fixed the output sizes: charge those instead of the estimate
        with self._cv:
            for r, (v, f) in zip(batch, outs):
                ws, _, gr = r.sizes()
                exact = ws + v.numel() * v.element_size() + f.numel() * f.element_size() + gr
                self._reserved += exact - r.charge
                r.charge = exact
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
