"""The reference's import names (needs a GPU): code written against `diso` (diso/__init__.py:1-9,
src/pybind.cpp:419-447) runs unchanged and gets the HIP engine's results, which equal the oracle."""
import numpy as np
import pytest
import torch

from tests.helpers import check_against_oracle, inputs

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def test_reference_import_names_run_the_hip_engine():
    print("case drop-in import", flush=True)
    from diso import DiffDMC, DiffMC
    import diso._C

    import diso_amd

    assert DiffMC is diso_amd.DiffMC and DiffDMC is diso_amd.DiffDMC
    sdf_np, def_np = inputs((30, 28, 33), "noisy_box", 1, np.float32, True)
    for name, mod in (("mc", DiffMC(dtype=torch.float32)), ("dmc", DiffDMC(dtype=torch.float32))):
        sdf = torch.tensor(sdf_np, device=DEV, requires_grad=True)
        deform = torch.tensor(def_np, device=DEV, requires_grad=True)
        kw = {"return_quads": True} if name == "dmc" else {}
        v, f = mod(sdf, deform, isovalue=0.0, normalize=True, **kw)
        up = torch.randn(tuple(v.shape), generator=torch.Generator(device=DEV).manual_seed(5), device=DEV)
        torch.autograd.backward(v, grad_tensors=up)
        check_against_oracle(name, sdf_np, def_np, v, f, up, sdf.grad, deform.grad)
    b = diso._C.CUMCFloat()
    g = torch.nn.functional.pad(torch.tensor(sdf_np, device=DEV), (1, 1, 1, 1, 1, 1), value=1.0)
    v, f = b.forward(g.contiguous(), 0.0)
    assert f.dtype == torch.int32 and v.shape[1] == 3 and f.shape[0] > 0
