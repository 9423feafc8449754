"""Fitting a learnable sdf through DiffDMC, eagerly or as a captured HIP graph.

The sdf [N,N,N] and the deform [N,N,N,3] are parameters: a sphere of radius 0.4 plus N(0, noise)
(noise is a standard deviation) and U(-deform_amp, deform_amp). Adam (lr 1e-3, capturable)
minimises mean((|v - c| - 0.3)^2) over the DMC vertices v, where the extraction normalises
(v in [0,1]^3) and c = (0.5, 0.5, 0.5) recentres them. A step is forward, backward and optimizer
step. The extraction runs with fixed capacities (diso_amd.fixed), so a step has no host
synchronisation and no data-dependent shape: `capture()` records one step in a graph after an
eager warm-up, and `replay()` re-runs it with a single launch from the host.

Every reduction of the loss is staged over rows of 256 elements, so that no single sum spans
more than one workgroup; vertices past the count are masked out (they are zero).
"""
import torch

from .fixed import FixedExtractor

__all__ = ["SdfFit"]


def _staged_sum(x):
    """Sum of a 1-D tensor whose length is a multiple of 256**k for every stage k used."""
    while x.numel() > 256:
        x = x.view(-1, 256).sum(1)
    return x.sum()


def _capacity(need):
    """Smallest multiple of 256**k (k >= 1, 256**(k+1) >= need) that holds `need` rows."""
    k = 1
    while 256 ** (k + 1) < need:
        k += 1
    unit = 256 ** k
    return max(unit, -(-need // unit) * unit)


class SdfFit:
    def __init__(self, n, device="cuda:0", seed=0, noise=0.01, deform_amp=0.1, lr=1e-3, recompute=True,
                 headroom=1.5):
        dev = torch.device(device)
        self.n = n
        g = torch.Generator(device=dev).manual_seed(seed)
        a = torch.linspace(-0.5, 0.5, n, dtype=torch.float64, device=dev)
        x, y, z = torch.meshgrid(a, a, a, indexing="ij")
        sdf = torch.sqrt(x * x + y * y + z * z) - 0.4
        del x, y, z
        sdf += noise * torch.randn((n, n, n), generator=g, dtype=torch.float64, device=dev)
        deform = (torch.rand((n, n, n, 3), generator=g, dtype=torch.float64, device=dev) * 2 - 1) * deform_amp
        self.sdf = torch.nn.Parameter(sdf.float())
        self.deform = torch.nn.Parameter(deform.float())
        del sdf, deform
        self.opt = torch.optim.Adam([self.sdf, self.deform], lr=lr, capturable=True)
        # capacities from the initial surface, with headroom for how it changes while fitting
        probe = FixedExtractor("dmc", torch.float32, 1, 1)
        with torch.no_grad():
            _, _, counts = probe(self.sdf, self.deform)
        nv, nf, _ = (int(c) for c in counts.tolist())
        self.ex = FixedExtractor("dmc", torch.float32, _capacity(int(nv * headroom) + 1),
                                 int(nf * headroom) + 1, recompute_backward=recompute)
        self.rows = torch.arange(self.ex.vert_capacity, device=dev)
        self.loss = torch.zeros((), dtype=torch.float32, device=dev)  # last step's loss
        self.counts = torch.zeros(3, dtype=torch.int64, device=dev)  # last step's (verts, quads, flags)
        self.graph = None

    def _step(self):
        verts, _, counts = self.ex(self.sdf, self.deform)
        n = counts[0]
        d = verts - 0.5
        r = torch.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
        e = torch.where(self.rows < n, (r - 0.3) * (r - 0.3), torch.zeros((), dtype=r.dtype, device=r.device))
        loss = _staged_sum(e) / n.clamp_min(1).to(e.dtype)
        loss.backward()
        self.opt.step()
        self.loss.copy_(loss.detach())
        self.counts.copy_(counts)

    def step(self):
        """One eager step."""
        self.opt.zero_grad(set_to_none=True)
        self._step()

    def capture(self, warmup=3):
        """Warm-up steps on a side stream, then one step recorded in a graph (that recording does
        not run it)."""
        s = torch.cuda.Stream(device=self.sdf.device)  # default priority
        s.wait_stream(torch.cuda.current_stream(self.sdf.device))
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self.step()
        torch.cuda.current_stream(self.sdf.device).wait_stream(s)
        self.opt.zero_grad(set_to_none=True)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._step()

    def replay(self):
        self.graph.replay()

    def check(self):
        """Raises if the last step overflowed a capacity or met a NaN (synchronises)."""
        return self.ex.check(self.counts)
