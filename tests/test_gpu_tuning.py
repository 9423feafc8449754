"""The per-bucket autotuner on the GPU: a one-bucket tuning run (graph-timed launches of count,
emission and backward) applies and saves a valid choice, leaves the other buckets alone, and
results stay those of the default parameters."""
import json

import numpy as np
import pytest
import torch

import diso_amd
from diso_amd import autotune, workloads

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _extract(sdf_np):
    out = []
    for mod in (diso_amd.DiffMC(), diso_amd.DiffDMC()):
        sdf = torch.tensor(sdf_np, device=DEV, requires_grad=True)
        v, f = mod(sdf)
        (v * v).sum().backward()
        out += [v.detach().cpu().numpy(), f.cpu().numpy(), sdf.grad.cpu().numpy()]
    return out


def test_tune_one_bucket_saves_and_applies(tmp_path):
    before = autotune.current()
    sdf_np = workloads.sphere(32, radius=0.3).astype(np.float32)
    ref = _extract(sdf_np)
    path = tmp_path / "tuning.json"
    try:
        print("case tune sizes=(32,) reps=1", flush=True)
        table = autotune.tune(sizes=(32,), reps=1, save=True, path=str(path), verbose=False)
        cur = autotune.current()
        assert cur == table
        assert cur["32"]["run"] in autotune.RUNS
        assert cur["32"]["count_blocks_per_cu"] in autotune.COUNT_BLOCKS_PER_CU
        for key in ("64", "128", "256", "512"):
            assert cur[key] == before[key]
        saved = json.loads(path.read_text())
        (entry,) = saved.values()
        assert entry["buckets"] == table
        assert set(entry["measured"]) == {"32"}
        got = _extract(sdf_np)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)
        autotune.apply(before)
        assert autotune.load(str(path)) == table and autotune.current() == table
    finally:
        autotune.apply(before)
