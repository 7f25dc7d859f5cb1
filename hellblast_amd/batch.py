"""Mixed-geometry batches in one HIP-graph launch (no reference counterpart).

Maps of different nside / lmax are grouped by geometry; each group is one
batched call of its plan (matrix-core Legendre kernels for FP32 groups of 2+
maps), with tiles from the autotuner.  ``capture()`` records map2alm and
alm2map of every group into one HIP graph, each group on its own stream
branch (forked from and joined back to the capturing stream), so one
``replay()`` launches the whole batch and the runtime overlaps the groups'
kernels.  Inputs and outputs live in static device buffers owned by the
batch (``maps``, ``alm``, ``out``), as graph replay requires.
"""
from collections import OrderedDict

import torch

from .hp_compat import _plan
from .plan import nalm


class MixedBatch:
    """map2alm + alm2map of many HEALPix maps of mixed nside / lmax.

    specs: one (nside, lmax) per map (lmax None -> 2*nside); order is kept
    by ``map(i)`` / ``alm_of(i)`` / ``output(i)``.
    """

    def __init__(self, specs, precision="f32", tune=True, weights="dz", phi0="reference", alias="reference",
                 device=None):
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self.specs = [(int(n), 2 * int(n) if l is None else int(l)) for n, l in
                      ((s, None) if isinstance(s, int) else s for s in specs)]
        self.groups = OrderedDict()
        for i, key in enumerate(self.specs):
            self.groups.setdefault(key, []).append(i)
        self.where = {}
        self.plans, self.maps, self.alm, self.out, self.tiles = {}, {}, {}, {}, {}
        with torch.cuda.device(self.device):
            for key, idx in self.groups.items():
                nside, lmax = key
                plan = _plan(nside, lmax, lmax, precision, weights, phi0, alias)
                rdt = torch.float64 if plan.alm_dtype == torch.complex128 else torch.float32
                n = len(idx)
                self.maps[key] = torch.zeros((n, plan.npix), dtype=rdt, device=self.device)
                self.alm[key] = torch.zeros((n, nalm(lmax)), dtype=plan.alm_dtype, device=self.device)
                self.out[key] = torch.zeros((n, plan.npix), dtype=rdt, device=self.device)
                for j, i in enumerate(idx):
                    self.where[i] = (key, j)
                if tune:
                    from .autotune import tune as _tune
                    self.tiles[key] = _tune(plan, n)["tiles"]
                    plan = plan.with_tiles(**self.tiles[key])
                self.plans[key] = plan
        self.graph = None
        self.streams = [torch.cuda.Stream(device=self.device) for _ in self.groups]

    # -- static buffers, by map index --------------------------------------
    def map(self, i):
        key, j = self.where[i]
        return self.maps[key][j]

    def alm_of(self, i):
        key, j = self.where[i]
        return self.alm[key][j]

    def output(self, i):
        key, j = self.where[i]
        return self.out[key][j]

    def set_maps(self, maps):
        """Copy one input map per spec into the static buffers."""
        if len(maps) != len(self.specs):
            raise ValueError("one map per spec expected")
        for i, m in enumerate(maps):
            self.map(i).copy_(torch.as_tensor(m, device=self.device))

    @property
    def nmaps(self):
        return len(self.specs)

    # -- execution ----------------------------------------------------------
    def _group(self, key):
        plan = self.plans[key]
        plan.map2alm(self.maps[key], out=self.alm[key])
        plan.alm2map(self.alm[key], dtype=self.out[key].dtype, out=self.out[key])

    def _forked(self):
        main = torch.cuda.current_stream(self.device)
        for key, s in zip(self.groups, self.streams):
            s.wait_stream(main)
            with torch.cuda.stream(s):
                self._group(key)
        for s in self.streams:
            main.wait_stream(s)

    def run_eager(self):
        """The same work without a graph (forked streams, joined)."""
        with torch.cuda.device(self.device):
            self._forked()

    def capture(self):
        """Record every group's map2alm + alm2map into one HIP graph."""
        with torch.cuda.device(self.device):
            self.run_eager()                      # first use outside capture
            torch.cuda.synchronize(self.device)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._forked()
            self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()
