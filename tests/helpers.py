"""Shared inputs and comparisons for the tests."""
import numpy as np

from diso_amd import workloads


def field(dims, kind, seed):
    """A [X,Y,Z] float64 sdf of the given family on an anisotropic grid."""
    rng = np.random.default_rng(seed)
    X, Y, Z = dims
    axes = [np.linspace(-0.5, 0.5, n) for n in dims]
    p = np.stack(np.meshgrid(*axes, indexing="ij"), -1)
    if kind == "sphere":
        return np.linalg.norm(p, axis=-1) - 0.37
    if kind == "noisy_box":
        q = np.abs(p) - 0.3 + 0.1
        f = np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(-1), 0) - 0.1
        return f + rng.uniform(-0.05, 0.05, size=dims)
    if kind == "noise":
        return rng.uniform(-1.0, 1.0, size=dims)
    if kind == "blobs":
        f = np.zeros(dims)
        for _ in range(6):
            c = rng.uniform(-0.3, 0.3, 3)
            s = rng.uniform(0.05, 0.15)
            f += np.exp(-((p - c) ** 2).sum(-1) / (2 * s * s))
        return f - 0.5
    raise ValueError(kind)


def inputs(dims, kind, seed, dtype, with_deform, deform_amp=0.3):
    rng = np.random.default_rng(seed + 1000)
    sdf = field(dims, kind, seed).astype(dtype)
    deform = rng.uniform(-deform_amp, deform_amp, size=tuple(dims) + (3,)).astype(dtype) if with_deform else None
    return sdf, deform


# Elementwise bound of the gradients, per dtype of the HIP result: |got - want| <= rtol * |want| +
# atol * max|want|. The HIP gather sums the same terms as the oracle in another order, and in fp32
# takes 1/(d1 - d0) from the hardware reciprocal (within 1 ulp) instead of three correctly rounded
# divisions: each term is then within a few ulp, so an element is off by a few ulp of its terms'
# magnitudes, which bounds it relative to itself unless its terms cancel, and to the field's scale
# where they do.
ELEMENT_TOL = {np.dtype(np.float32): (1e-4, 1e-6), np.dtype(np.float64): (1e-10, 1e-13)}


def grad_close(got, want, rel):
    """|got - want| <= rel * max(|want|) (gradients accumulate the same terms in a different order
    than the oracle, so they are compared against the field's scale), and elementwise within
    ELEMENT_TOL of each value."""
    scale = max(float(np.abs(want).max()), 1e-30)
    diff = np.abs(got.astype(np.float64) - want.astype(np.float64))
    err = float(diff.max())
    assert err <= rel * scale, f"max abs err {err:.3e} > {rel:.1e} * scale {scale:.3e}"
    rtol, atol = ELEMENT_TOL[np.dtype(got.dtype)]
    bound = rtol * np.abs(want.astype(np.float64)) + atol * scale
    worst = int(np.argmax(diff / bound))
    assert diff.flat[worst] <= bound.flat[worst], (
        f"element {worst}: |{got.flat[worst]!r} - {want.flat[worst]!r}| > {rtol:.0e} * |want| + {atol:.0e} * scale "
        f"{scale:.3e} (ratio {float(diff.flat[worst] / bound.flat[worst]):.3f})")
    return err / scale


def check_against_oracle(algo, sdf_np, def_np, verts, faces, up=None, grad_sdf=None, grad_deform=None, iso=0.0,
                         normalize=True, return_quads=True, compat=False):
    """An extraction's outputs (torch tensors) against the oracle on the same inputs: faces and
    vertices bit-exact, gradients (for upstream `up`) within 2e-5 (fp32) / 1e-12 (fp64) of the
    gradient's scale."""
    from oracle import oracle as O

    dtype = sdf_np.dtype.type
    ref = O.Extractor(O.MC if algo in ("mc", O.MC) else O.DMC, dtype)
    rv, rf = ref.forward(sdf_np, def_np, iso, normalize, return_quads=return_quads)
    assert tuple(faces.shape) == rf.shape, (tuple(faces.shape), rf.shape)
    assert np.array_equal(faces.cpu().numpy(), rf), "faces differ from the oracle"
    assert np.array_equal(verts.detach().cpu().numpy(), rv), "vertices differ from the oracle"
    if up is None or ref.empty:
        return
    tol = 2e-5 if dtype == np.float32 else 1e-12
    gs, gd = ref.backward(up.detach().cpu().numpy().astype(dtype), reference_compat=compat)
    grad_close(grad_sdf.cpu().numpy(), gs, tol)
    if def_np is not None:
        grad_close(grad_deform.cpu().numpy(), gd, tol)
