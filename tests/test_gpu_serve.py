"""Serving loop (needs a GPU): requests submitted from several threads and batched by two
workers give the same results, bit for bit, as each request extracted alone, and equal the
oracle; a tight memory budget serialises the work without changing results."""
import threading

import numpy as np
import pytest
import torch

import diso_amd
from diso_amd.serve import Request, Server
from tests.helpers import check_against_oracle, inputs

pytestmark = pytest.mark.gpu
DEV = "cuda:0"

MIX = [((24, 24, 24), "noisy_box", "mc", True, True), ((33, 20, 70), "noise", "dmc", True, True),
       ((40, 9, 130), "blobs", "mc", False, False), ((64, 64, 64), "sphere", "dmc", False, True),
       ((5, 3, 66), "noise", "dmc", True, False), ((31, 47, 29), "sphere", "mc", True, True),
       ((48, 40, 72), "noise", "dmc", True, True), ((17, 33, 70), "noise", "mc", False, True),
       ((8, 8, 8), "sphere", "mc", False, False), ((50, 50, 50), "noisy_box", "dmc", True, True)]


def _requests(seed):
    reqs, alone = [], []
    for i, (dims, kind, algo, with_deform, backward) in enumerate(MIX):
        sdf_np, def_np = inputs(dims, kind, seed + i, np.float32, with_deform)
        sdf = torch.tensor(sdf_np, device=DEV)
        deform = None if def_np is None else torch.tensor(def_np, device=DEV)
        gen = torch.Generator(device=DEV).manual_seed(seed + i)
        up = (lambda n, g=gen: torch.randn((n, 3), generator=g, device=DEV)) if backward else None
        reqs.append(Request(sdf, algo, deform, upstream=up))
        alone.append((sdf, deform, algo, backward, seed + i))
    return reqs, alone


def _alone(sdf, deform, algo, backward, seed):
    s = sdf.detach().requires_grad_(backward)
    d = None if deform is None else deform.detach().requires_grad_(backward)
    mod = diso_amd.DiffMC() if algo == "mc" else diso_amd.DiffDMC()
    v, f = mod(s, d, return_quads=True) if algo == "dmc" else mod(s, d)
    if not backward or not v.shape[0]:
        return v.detach(), f, None, None
    g = torch.randn((v.shape[0], 3), generator=torch.Generator(device=DEV).manual_seed(seed), device=DEV)
    gs, *gd = torch.autograd.grad([v], [s] + ([d] if d is not None else []), grad_outputs=[g])
    return v.detach(), f, gs, gd[0] if gd else None


@pytest.mark.parametrize("budget,workers", [(200 << 30, 2), (4 << 20, 2), (200 << 30, 4)])
def test_served_results_equal_alone(budget, workers):
    """Four workers run counting passes concurrently on four streams (each pass's single-pass scan
    must not wait on blocks of its own launch that cannot run)."""
    print(f"case serve budget={budget} workers={workers}", flush=True)
    reqs, alone = _requests(100)
    with Server(budget_bytes=budget, workers=workers, max_batch=4) as srv:
        threads = [threading.Thread(target=lambda rs=reqs[k::3]: [srv.submit(r) for r in rs]) for k in range(3)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        results = [r.future.result(timeout=120) for r in reqs]
        assert srv.batches >= 1 and len(srv.latencies) == len(reqs)
        if budget < (64 << 20):
            fp = [r.footprint(res[0].shape[0], res[1].shape[0]) for r, res in zip(reqs, results)]
            assert srv.peak_reserved <= max(budget, max(fp))
        assert srv.latency_ms(50) is not None and srv.latency_ms(99) >= srv.latency_ms(50)
    for res, a in zip(results, alone):
        v, f, gs, gd = _alone(*a)
        assert torch.equal(res[0], v) and torch.equal(res[1], f)
        if gs is not None:
            assert torch.equal(res[2], gs)
            if gd is not None:
                assert torch.equal(res[3], gd)
        sdf, deform, algo, backward, seed = a
        up = None
        if gs is not None:
            up = torch.randn((v.shape[0], 3), generator=torch.Generator(device=DEV).manual_seed(seed), device=DEV)
        check_against_oracle(algo, sdf.cpu().numpy(), None if deform is None else deform.cpu().numpy(), res[0], res[1],
                             up, None if up is None else res[2], None if up is None else res[3])


def test_concurrent_streams_match_sequential():
    """Regression test of the single-pass scan under concurrency: four threads, each on its own
    stream, run DiffMC / DiffDMC forward+backward on 96^3 grids at the same time; every result
    equals the sequential one."""
    print("case concurrent streams", flush=True)
    jobs = []
    for i in range(8):
        sdf_np, def_np = inputs((96, 96, 96), "noise" if i % 2 else "blobs", 200 + i, np.float32, True)
        jobs.append((torch.tensor(sdf_np, device=DEV), torch.tensor(def_np, device=DEV), "dmc" if i % 3 else "mc", 300 + i))
    want = [_alone(s, d, a, True, seed) for s, d, a, seed in jobs]
    got = [None] * len(jobs)
    errors = []

    def worker(k):
        try:
            st = torch.cuda.Stream()
            with torch.cuda.stream(st):
                for rep in range(3):
                    for i in range(k, len(jobs), 4):
                        got[i] = _alone(*jobs[i][:3], True, jobs[i][3])
                st.synchronize()
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for (v, f, gs, gd), (v2, f2, gs2, gd2) in zip(want, got):
        assert torch.equal(v, v2) and torch.equal(f, f2) and torch.equal(gs, gs2) and torch.equal(gd, gd2)
