"""Mixed-geometry batches (no reference counterpart; the reference
transforms one map per call, ``sht_raw.py:30-57`` / ``isht.py:5-26``).

Maps of different nside / lmax are grouped by geometry.  Two execution modes
(``mode="auto"``, the default, times both once and keeps the faster):

- ``"persistent"``: one libhbx batch object (``MultiPlan``,
  ``hbx_multi_*``) over all groups, a fixed number of launches per call: the
  ring stage in fused launches over every group's rings; each Legendre
  direction ONE persistent kernel whose workgroups pull (group, m-block,
  ring-pair block, map block) tiles from work queues, so the m-blocks of a
  large map and the tiles of the small maps share the whole GPU.  Tile
  shapes (ring-pair slots, maps per synthesis tile, tile granularity) come
  from the autotuner (``autotune.tune_multi``).  The faster form for many
  small groups (whose own launches cannot fill the GPU).
- ``"grouped"``: each group is one batched call of its plan (matrix-core
  Legendre kernels for FP32 groups of 3+ maps), each group on its own stream
  branch.

``capture()`` records map2alm and alm2map of the whole batch into one HIP
graph, so one ``replay()`` launches everything.  Inputs, outputs and the
workspace live in static device buffers owned by the batch (``maps``,
``alm``, ``out``), as graph replay requires.
"""
import ctypes
from collections import OrderedDict

import torch

from . import _abi
from .hp_compat import _plan
from .plan import _map_dtype_code, _stream_ptr, nalm

MULTI_TILE_FIELDS = ("valu_slots", "syn_maps", "tiles_per_wave")


class MultiPlan:
    """A libhbx mixed-geometry batch (``hbx_multi``): ``plans[g]`` with
    ``counts[g]`` maps each, transformed by persistent Legendre launches.

    All plans must be spin 0, of one precision and on one device.  Tiles
    (0 = library default): ``valu_slots`` (analysis ring-pair slots of 64
    pairs per lane: 1, 2, 4, 8), ``syn_maps`` (maps per synthesis tile
    sharing each recursion step: 1, 2, 4), ``tiles_per_wave`` (analysis tiles
    per resident wave that the m ranges are cut for).
    """

    def __init__(self, plans, counts, **tiles):
        if len(plans) != len(counts) or not plans:
            raise ValueError("one map count per plan expected")
        if len(plans) > _abi.HBX_MULTI_MAX_GROUPS:
            raise ValueError(f"at most {_abi.HBX_MULTI_MAX_GROUPS} groups")
        for k in tiles:
            if k not in MULTI_TILE_FIELDS:
                raise ValueError(f"unknown tile {k!r}")
        self.plans = list(plans)
        self.counts = [int(c) for c in counts]
        self.device = self.plans[0].device
        self.alm_dtype = self.plans[0].alm_dtype
        n = len(self.plans)
        hs = (ctypes.c_void_p * n)(*[p._h.value for p in self.plans])
        cs = (ctypes.c_int32 * n)(*self.counts)
        t = _abi.hbx_multi_tiles(**{k: int(tiles.get(k, 0)) for k in MULTI_TILE_FIELDS})
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _abi.check(_abi.lib().hbx_multi_create(hs, cs, n, ctypes.byref(t), ctypes.byref(h)))
        self._h = h
        info = _abi.hbx_multi_info()
        _abi.check(_abi.lib().hbx_multi_get_info(self._h, ctypes.byref(info)))
        self.info = {f: int(getattr(info, f)) for f, _ in info._fields_ if f != "reserved"}
        self.tiles = {k: self.info[k] for k in MULTI_TILE_FIELDS}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                with torch.cuda.device(self.device):
                    _abi.lib().hbx_multi_destroy(h)
            except Exception:
                pass
            self._h = None

    def workspace(self):
        b = ctypes.c_uint64()
        _abi.check(_abi.lib().hbx_multi_workspace_bytes(self._h, ctypes.byref(b)))
        return torch.empty(max(int(b.value), 1), dtype=torch.uint8, device=self.device)

    def _check(self, tensors, ncols, dtypes, what):
        if len(tensors) != len(self.plans):
            raise ValueError(f"one {what} tensor per group expected")
        for g, (t, n) in enumerate(zip(tensors, self.counts)):
            if not isinstance(t, torch.Tensor) or t.device != self.device or not t.is_contiguous():
                raise ValueError(f"{what} of group {g} must be a contiguous tensor on {self.device}")
            if t.dim() != 2 or t.shape[0] != n or t.shape[1] != ncols[g] or t.dtype not in dtypes:
                raise ValueError(f"{what} of group {g}: [{n}, {ncols[g]}] of {dtypes} expected, "
                                 f"got {tuple(t.shape)} {t.dtype}")

    def _arrays(self, tensors):
        n = len(tensors)
        ptrs = (ctypes.c_void_p * n)(*[t.data_ptr() for t in tensors])
        strides = (ctypes.c_int64 * n)(*[t.shape[1] for t in tensors])
        return ptrs, strides

    def map2alm(self, maps, alm, ws):
        """maps[g]: [counts[g], npix_g] (FP32 or FP64, one dtype) -> alm[g]."""
        self._check(maps, [p.npix for p in self.plans], (maps[0].dtype,) if maps else (), "maps")
        self._check(alm, [p.nalm for p in self.plans], (self.alm_dtype,), "alm")
        mp, ms = self._arrays(maps)
        ap, as_ = self._arrays(alm)
        with torch.cuda.device(self.device):
            _abi.check(_abi.lib().hbx_multi_map2alm(self._h, mp, ms, _map_dtype_code(maps[0].dtype), ap, as_,
                                                    ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                                                    _stream_ptr(self.device)))
        return alm

    def alm2map(self, alm, maps, ws):
        """alm[g] -> maps[g] (dtype of maps[0], the same for every group)."""
        self._check(alm, [p.nalm for p in self.plans], (self.alm_dtype,), "alm")
        self._check(maps, [p.npix for p in self.plans], (maps[0].dtype,) if maps else (), "maps")
        ap, as_ = self._arrays(alm)
        mp, ms = self._arrays(maps)
        with torch.cuda.device(self.device):
            _abi.check(_abi.lib().hbx_multi_alm2map(self._h, ap, as_, mp, ms, _map_dtype_code(maps[0].dtype),
                                                    ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                                                    _stream_ptr(self.device)))
        return maps


class MixedBatch:
    """map2alm + alm2map of many HEALPix maps of mixed nside / lmax.

    specs: one (nside, lmax) per map (lmax None -> 2*nside); order is kept
    by ``map(i)`` / ``alm_of(i)`` / ``output(i)``.  mode: "persistent" (one
    persistent Legendre launch per direction over every group), "grouped"
    (one batched plan call per group) or "auto" (the faster of the two for
    these specs, timed once and cached by the autotuner).  multi_tiles: tiles of the persistent
    launch (default: the autotuner's choice when ``tune``, else the library
    defaults).
    """

    def __init__(self, specs, precision="f32", tune=True, weights="dz", phi0="reference", alias="reference",
                 device=None, mode="auto", multi_tiles=None):
        if mode not in ("persistent", "grouped", "auto"):
            raise ValueError("mode must be 'persistent', 'grouped' or 'auto'")
        if mode == "auto":
            from .autotune import tune_mixed_mode
            mode = tune_mixed_mode(specs, precision, weights, phi0, alias, device)
        self.mode = mode
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
                if tune and mode == "grouped":
                    from .autotune import tune as _tune
                    self.tiles[key] = _tune(plan, n)["tiles"]
                    plan = plan.with_tiles(**self.tiles[key])
                self.plans[key] = plan
            self.multi = None
            if mode == "persistent":
                keys = list(self.groups)
                plans = [self.plans[k] for k in keys]
                counts = [len(self.groups[k]) for k in keys]
                if multi_tiles is None and tune:
                    from .autotune import tune_multi
                    multi_tiles = tune_multi(plans, counts)["tiles"]
                self.multi = MultiPlan(plans, counts, **(multi_tiles or {}))
                self.multi_tiles = dict(self.multi.tiles)
                self.ws = self.multi.workspace()
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
        if self.multi is not None:
            keys = list(self.groups)
            self.multi.map2alm([self.maps[k] for k in keys], [self.alm[k] for k in keys], self.ws)
            self.multi.alm2map([self.alm[k] for k in keys], [self.out[k] for k in keys], self.ws)
            return
        main = torch.cuda.current_stream(self.device)
        for key, s in zip(self.groups, self.streams):
            s.wait_stream(main)
            with torch.cuda.stream(s):
                self._group(key)
        for s in self.streams:
            main.wait_stream(s)

    def run_eager(self):
        """The same work without a graph (persistent launches, or forked
        streams joined)."""
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
