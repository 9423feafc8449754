"""Admission logic of the serving loop, on the host (no GPU, no worker threads): footprints from
the counts, the workspace charge when a batch is taken, the batch limits, the exact charge after
the counting pass with the rest returned to the head of the queue, and release."""
import torch

from diso_amd.serve import Request, Server


def _req(n, algo="mc", backward=False, deform=False):
    sdf = torch.zeros((n, n, n))
    d = torch.zeros((n, n, n, 3)) if deform else None
    return Request(sdf, algo, d, upstream=(lambda v: torch.zeros((v, 3))) if backward else None)


def _server(budget, **kw):
    return Server(device="cpu", budget_bytes=budget, workers=0, **kw)  # no workers: only the bookkeeping


def test_footprint_counts_every_buffer():
    r = _req(10, "dmc", backward=True, deform=True)
    pos = 12 ** 3
    ws = 2 * pos + pos // 4 + 4096
    assert r.workspace() == ws
    nv, nf = 100, 90
    # vertices, quads, vinfo; upstream, edge scratch, sdf and deform gradients
    assert r.footprint(nv, nf) == ws + 12 * nv + 32 * nf + nv + (1000 * 4 + 3000 * 4) + 12 * nv + 12 * nf
    r = _req(10, "mc")
    assert r.footprint(nv, nf) == ws + 12 * nv + 24 * nf
    r = _req(10, "dmc")
    r.return_quads = False  # triangles and split scratch on top of the quads
    assert r.footprint(nv, nf) == ws + 12 * nv + 32 * nf + nv + 49 * nf + 16384


def test_admit_takes_same_settings_within_limits():
    srv = _server(1 << 40, max_batch=3, max_batch_voxels=1 << 20)
    a, b, c, d, e = _req(32), _req(32), _req(64), _req(32), _req(32)
    c.isovalue = 0.5  # other settings: not batched with a
    for r in (a, b, c, d, e):
        srv._queue.append(r)
    take = srv._admit()
    assert take == [a, b, d]  # max_batch 3, c skipped and kept in place
    assert list(srv._queue) == [c, e]
    assert srv._reserved == sum(r.workspace() for r in take) and all(r.charge == r.workspace() for r in take)
    # a voxel cap: the second 96^3 grid does not fit next to the first
    srv = _server(1 << 40, max_batch_voxels=100 ** 3)
    f, g = _req(96), _req(96)
    srv._queue.extend([f, g])
    assert srv._admit() == [f] and list(srv._queue) == [g]


def test_admit_waits_for_memory_unless_idle():
    big = _req(64)
    srv = _server(big.workspace() - 1)
    srv._queue.append(big)
    assert srv._admit() == [big]  # nothing else holds memory: runs alone, over budget
    srv._queue.append(_req(64))
    assert srv._admit() == []  # now it must wait
    srv._release([big])
    assert srv._reserved == 0 and len(srv._admit()) == 1


def test_fit_charges_exact_sizes_and_returns_the_rest():
    reqs = [_req(16) for _ in range(4)]
    sizes = [(1000, 2000, 3), (1000, 2000, 3), (10 ** 7, 2 * 10 ** 7, 3), (10, 20, 3)]
    fp = [r.footprint(nv, nf) for r, (nv, nf, _) in zip(reqs, sizes)]
    srv = _server(fp[0] + fp[1] + fp[2] // 2)
    srv._queue.extend(reqs)
    batch = srv._admit()
    assert batch == reqs
    with srv._cv:  # called with the lock held
        run = srv._fit(batch, sizes)
    assert run == reqs[:2]  # the third does not fit: it and everything after go back, in order
    assert list(srv._queue) == reqs[2:]
    assert srv._reserved == fp[0] + fp[1] and [r.charge for r in reqs] == [fp[0], fp[1], 0, 0]
    srv._release(run)
    assert srv._reserved == 0
    # alone, a request larger than the budget still runs
    batch = srv._admit()
    assert batch[0] is reqs[2]
    with srv._cv:
        run = srv._fit(batch, sizes[2:])
    assert run == [reqs[2]] and srv._reserved == fp[2] and list(srv._queue) == [reqs[3]]
