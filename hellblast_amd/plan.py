"""Transform plans: ring geometry + lmax/mmax/precision bound to one device.

A plan owns the device tables of one geometry (ring offsets, ring-pair
cos theta, FFT chirp spectra, lambda recursion coefficients and switch-on
seeds).  Plans are cached per (device, geometry, lmax, mmax, precision).
Execute methods take and return torch tensors on the plan's device and
enqueue on torch's current stream; they allocate only the output and a
workspace through torch's caching allocator.
"""
import ctypes
import hashlib
import math
from collections import OrderedDict

import numpy as np
import torch

from . import _abi

PRECISIONS = {"f64": _abi.HBX_F64, "float64": _abi.HBX_F64, "double": _abi.HBX_F64,
              "f32": _abi.HBX_F32, "float32": _abi.HBX_F32, "single": _abi.HBX_F32}


def precision_code(precision):
    if isinstance(precision, int):
        if precision in (_abi.HBX_F64, _abi.HBX_F32):
            return precision
    key = str(precision).lower().replace("torch.", "")
    if key not in PRECISIONS:
        raise ValueError(f"precision must be 'f64' or 'f32', got {precision!r}")
    return PRECISIONS[key]


def nalm(lmax, mmax=None):
    mmax = lmax if mmax is None else mmax
    return (mmax + 1) * (lmax + 1) - mmax * (mmax + 1) // 2


class Geometry:
    """Host description of a ring set (hbx_geom); arrays are NumPy."""

    def __init__(self, nphi, theta=None, cos_theta=None, sin_theta=None, weight=None, phi0=None,
                 mcut=None, ring_start=None):
        self.nphi = np.ascontiguousarray(np.asarray(nphi, dtype=np.int32).ravel())
        nr = self.nphi.size
        if theta is not None:
            th = np.asarray(theta, dtype=np.float64).ravel()
            if th.size != nr:
                raise ValueError("theta and nphi sizes differ")
            cos_theta = np.cos(th) if cos_theta is None else cos_theta
            sin_theta = np.sin(th) if sin_theta is None else sin_theta
        if cos_theta is None or sin_theta is None:
            raise ValueError("theta (or cos_theta and sin_theta) required")
        self.cos_theta = np.ascontiguousarray(np.asarray(cos_theta, dtype=np.float64).ravel())
        self.sin_theta = np.ascontiguousarray(np.asarray(sin_theta, dtype=np.float64).ravel())
        if self.cos_theta.size != nr or self.sin_theta.size != nr:
            raise ValueError("cos/sin theta and nphi sizes differ")
        if ring_start is None:
            ring_start = np.zeros(nr + 1, dtype=np.int64)
            ring_start[1:] = np.cumsum(self.nphi.astype(np.int64))
        self.ring_start = np.ascontiguousarray(np.asarray(ring_start, dtype=np.int64).ravel())
        if self.ring_start.size != nr + 1:
            raise ValueError("ring_start must have nrings+1 entries")
        self.weight = None if weight is None else np.ascontiguousarray(np.asarray(weight, dtype=np.float64).ravel())
        self.phi0 = None if phi0 is None else np.ascontiguousarray(np.asarray(phi0, dtype=np.float64).ravel())
        self.mcut = None if mcut is None else np.ascontiguousarray(np.asarray(mcut, dtype=np.int32).ravel())
        for name in ("weight", "phi0", "mcut"):
            a = getattr(self, name)
            if a is not None and a.size != nr:
                raise ValueError(f"{name} must have one entry per ring")

    @property
    def nrings(self):
        return int(self.nphi.size)

    @property
    def npix(self):
        return int(self.ring_start[-1])

    def key(self):
        h = hashlib.sha1()
        for a in (self.ring_start, self.nphi, self.cos_theta, self.sin_theta, self.weight, self.phi0, self.mcut):
            h.update(b"|" if a is None else a.tobytes())
        return h.hexdigest()

    def as_struct(self):
        """hbx_geom pointing into this object's arrays (keep self alive)."""
        def p(a, ct):
            return None if a is None else a.ctypes.data_as(ctypes.POINTER(ct))
        g = _abi.hbx_geom()
        g.nrings = self.nrings
        g.ring_start = p(self.ring_start, ctypes.c_int64)
        g.nphi = p(self.nphi, ctypes.c_int32)
        g.cos_theta = p(self.cos_theta, ctypes.c_double)
        g.sin_theta = p(self.sin_theta, ctypes.c_double)
        g.phi0 = p(self.phi0, ctypes.c_double)
        g.weight = p(self.weight, ctypes.c_double)
        g.mcut = p(self.mcut, ctypes.c_int32)
        return g

    def validate(self, lmax, mmax):
        _abi.check(_abi.lib().hbx_validate_geom(ctypes.byref(self.as_struct()), int(lmax), int(mmax)))


def _needs_grad(t):
    return torch.is_grad_enabled() and isinstance(t, torch.Tensor) and t.requires_grad


def _stream_ptr(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _map_dtype_code(dtype):
    if dtype == torch.float64:
        return _abi.HBX_DT_F64
    if dtype == torch.float32:
        return _abi.HBX_DT_F32
    raise TypeError(f"maps must be float64 or float32, got {dtype}")


def phase_map_elems(nrings, mmax):
    """Complex elements between consecutive maps of a phase buffer: m-blocks
    of 8 m, each 2 * pairs8 * 8 elements (pairs8 = ceil(nrings/2) rounded up
    to 8), the larger of the two layouts of phase_index()."""
    pairs8 = ((nrings + 1) // 2 + 7) // 8 * 8
    return (mmax + 8) // 8 * 2 * pairs8 * 8


def phase_index(nrings, mmax, direction):
    """Offsets [ring][m] (m <= mmax), in complex elements, of one map's phase
    coefficients.  "analysis" (ring analysis -> Legendre analysis):
    [m/8][ring][m%8].  "synthesis" (Legendre synthesis -> ring synthesis):
    [m/8][half][pair/8][m%8][pair%8], ring r < npairs being half 0 of ring
    pair r and ring r >= npairs half 1 of pair nrings-1-r (npairs =
    ceil(nrings/2); halves pairs8 * 8 elements apart).  Maps lie
    phase_map_elems() apart in both."""
    nr = int(nrings)
    m = np.arange(int(mmax) + 1, dtype=np.int64)[None, :]
    r = np.arange(nr, dtype=np.int64)[:, None]
    if direction == "analysis":
        return (m >> 3) * (nr * 8) + r * 8 + (m & 7)
    if direction != "synthesis":
        raise ValueError(f"direction must be 'analysis' or 'synthesis', not {direction!r}")
    npairs = (nr + 1) // 2
    p8 = (npairs + 7) // 8 * 8
    half = (r >= npairs).astype(np.int64)
    p = np.where(half == 1, nr - 1 - r, r)
    return (m >> 3) * (2 * p8 * 8) + half * (p8 * 8) + (p >> 3) * 64 + (m & 7) * 8 + (p & 7)


class Plan:
    """A libhbx plan for one geometry, lmax, mmax, spin and precision on one GPU.

    spin = 0 plans serve map2alm/alm2map (and adjoints); spin >= 1 plans
    serve the *_spin methods, which take (Q, U) map pairs and (E, B) alm
    pairs.
    """

    def __init__(self, geom, lmax, mmax=None, precision="f64", device=None, spin=0):
        mmax = lmax if mmax is None else int(mmax)
        self.geom = geom
        self.lmax, self.mmax = int(lmax), int(mmax)
        self.spin = int(spin)
        self.precision = precision_code(precision)
        if not torch.cuda.is_available():
            geom.validate(self.lmax, self.mmax)
            raise RuntimeError("hellblast_amd transforms need a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None
                                   else torch.device(device).index or 0)
        L = _abi.lib()
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _abi.check(L.hbx_plan_create(ctypes.byref(geom.as_struct()), self.lmax, self.mmax, self.spin,
                                         self.precision, 0, ctypes.byref(handle)))
        self._h = handle
        info = _abi.hbx_plan_info()
        _abi.check(L.hbx_plan_get_info(self._h, ctypes.byref(info)))
        self.npix = int(info.npix)
        self.nalm = int(info.nalm)
        self.device_bytes = int(info.device_bytes)
        self.fp64_band_m = int(info.fp64_band_m)
        self.alm_dtype = torch.complex128 if self.precision == _abi.HBX_F64 else torch.complex64
        self._req_tiles = {k: 0 for k in self.TILE_FIELDS}
        self._derived = {}
        self._root = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                with torch.cuda.device(self.device):
                    _abi.lib().hbx_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    # ------------------------------------------------------------------
    TILE_FIELDS = ("valu_slots", "mfma_waves", "mfma_min_maps", "spin_slots", "mfma_kind", "mfma_groups")

    def with_tiles(self, **tiles):
        """This plan with other tiles (0 = default; unknown names raise).

        Plans are immutable: the result is another plan (``hbx_plan_derive``)
        that shares this plan's device tables, cached per tile set on this
        plan, so callers that tune tiles never change the kernels or the
        summation order of anyone else's plan.
        """
        req = dict(self._req_tiles)
        for k, v in tiles.items():
            if k not in self.TILE_FIELDS:
                raise ValueError(f"unknown tile {k!r}")
            req[k] = int(v)
        key = tuple(req[k] for k in self.TILE_FIELDS)
        if key == tuple(self._req_tiles[k] for k in self.TILE_FIELDS):
            return self
        root = self._root if self._root is not None else self
        hit = root._derived.get(key)
        if hit is not None:
            return hit
        t = _abi.hbx_tiles(**req)
        handle = ctypes.c_void_p()
        _abi.check(_abi.lib().hbx_plan_derive(self._h, ctypes.byref(t), ctypes.byref(handle)))
        q = object.__new__(Plan)
        q.__dict__.update({k: v for k, v in self.__dict__.items() if k not in ("_h", "_derived", "_root")})
        q._h = handle
        q._req_tiles = req
        q._derived = {}
        q._root = root
        root._derived[key] = q
        return q

    def tiles(self, effective=True):
        """Effective tiles (library defaults resolved), or the requested ones."""
        if not effective:
            return dict(self._req_tiles)
        t = _abi.hbx_tiles()
        _abi.check(_abi.lib().hbx_plan_get_tiles(self._h, ctypes.byref(t)))
        return {k: int(getattr(t, k)) for k in self.TILE_FIELDS}

    def workspace_bytes(self, nmaps):
        b = ctypes.c_uint64()
        _abi.check(_abi.lib().hbx_workspace_bytes(self._h, int(nmaps), ctypes.byref(b)))
        return int(b.value)

    def workspace(self, nmaps):
        return torch.empty(max(self.workspace_bytes(nmaps), 1), dtype=torch.uint8, device=self.device)

    def _as_maps(self, maps):
        t = torch.as_tensor(maps)
        if t.device != self.device:
            t = t.to(self.device)
        if t.dtype not in (torch.float64, torch.float32):
            t = t.to(torch.float64)
        squeeze = t.dim() == 1
        t = t.reshape(-1, t.shape[-1]) if not squeeze else t.reshape(1, -1)
        if t.shape[-1] != self.npix:
            raise ValueError(f"map has {t.shape[-1]} pixels, plan expects {self.npix}")
        return t.contiguous(), squeeze

    def _as_alm(self, alm):
        t = torch.as_tensor(alm)
        if t.device != self.device:
            t = t.to(self.device)
        if t.dtype != self.alm_dtype:
            t = t.to(self.alm_dtype)
        squeeze = t.dim() == 1
        t = t.reshape(1, -1) if squeeze else t.reshape(-1, t.shape[-1])
        if t.shape[-1] != self.nalm:
            raise ValueError(f"alm has {t.shape[-1]} coefficients, plan expects {self.nalm}")
        return t.contiguous(), squeeze

    def _out(self, out, shape, dtype):
        """A new output tensor, or the caller's ``out`` after checking that the
        kernels can write ``shape`` of ``dtype`` into it."""
        if out is None:
            return torch.empty(shape, dtype=dtype, device=self.device)
        if not isinstance(out, torch.Tensor):
            raise TypeError("out must be a torch tensor")
        if out.device != self.device:
            raise ValueError(f"out is on {out.device}, the plan on {self.device}")
        if out.dtype != dtype:
            raise ValueError(f"out has dtype {out.dtype}, expected {dtype}")
        if not out.is_contiguous():
            raise ValueError("out must be contiguous")
        if out.numel() != math.prod(shape) or out.dim() < 1 or out.shape[-1] != shape[-1]:
            raise ValueError(f"out has shape {tuple(out.shape)}; {tuple(shape)} (any leading shape) expected")
        return out

    def _buffer(self, buf, what):
        """Byte size of a caller's phase buffer / workspace (device, contiguous)."""
        if not isinstance(buf, torch.Tensor) or buf.device != self.device or not buf.is_contiguous():
            raise ValueError(f"{what} must be a contiguous tensor on {self.device}")
        return buf.numel() * buf.element_size()

    def _analysis(self, fn, maps, out=None):
        m, squeeze = self._as_maps(maps)
        nb = m.shape[0]
        alm = self._out(out, (nb, self.nalm), self.alm_dtype)
        ws = self.workspace(nb)
        with torch.cuda.device(self.device):
            _abi.check(fn(self._h, nb, ctypes.c_void_p(m.data_ptr()), _map_dtype_code(m.dtype), m.shape[1],
                          ctypes.c_void_p(alm.data_ptr()), alm.shape[-1], ctypes.c_void_p(ws.data_ptr()),
                          ws.numel(), _stream_ptr(self.device)))
        if out is not None:
            return out
        return alm[0] if squeeze else alm

    def _synthesis(self, fn, alm, dtype=torch.float64, out=None):
        a, squeeze = self._as_alm(alm)
        nb = a.shape[0]
        _map_dtype_code(dtype)
        maps = self._out(out, (nb, self.npix), dtype)
        ws = self.workspace(nb)
        with torch.cuda.device(self.device):
            _abi.check(fn(self._h, nb, ctypes.c_void_p(a.data_ptr()), a.shape[1], ctypes.c_void_p(maps.data_ptr()),
                          _map_dtype_code(maps.dtype), maps.shape[-1], ctypes.c_void_p(ws.data_ptr()),
                          ws.numel(), _stream_ptr(self.device)))
        if out is not None:
            return out
        return maps[0] if squeeze else maps

    # Public transforms are differentiable: an input that requires grad (with
    # grad mode on and no ``out=``) goes through hellblast_amd.autograd, whose
    # backward runs the transpose kernels.
    def map2alm(self, maps, out=None):
        """Analysis of [npix] or [nmaps, npix] maps -> m-major alm."""
        if out is None and _needs_grad(maps):
            from .autograd import _Map2Alm
            return _Map2Alm.apply(maps, self)
        return self._map2alm_raw(maps, out)

    def alm2map(self, alm, dtype=torch.float64, out=None):
        """Synthesis of [nalm] or [nmaps, nalm] alm -> maps."""
        if out is None and _needs_grad(alm):
            from .autograd import _Alm2Map
            return _Alm2Map.apply(alm, self, dtype)
        return self._alm2map_raw(alm, dtype, out)

    def alm2map_adjoint(self, maps, out=None):
        """Transpose of alm2map (complex-gradient convention)."""
        if out is None and _needs_grad(maps):
            from .autograd import _Alm2MapAdjoint
            return _Alm2MapAdjoint.apply(maps, self)
        return self._alm2map_adjoint_raw(maps, out)

    def map2alm_adjoint(self, alm, dtype=torch.float64, out=None):
        """Transpose of map2alm (complex-gradient convention)."""
        if out is None and _needs_grad(alm):
            from .autograd import _Map2AlmAdjoint
            return _Map2AlmAdjoint.apply(alm, self, dtype)
        return self._map2alm_adjoint_raw(alm, dtype, out)

    def _map2alm_raw(self, maps, out=None):
        return self._analysis(_abi.lib().hbx_map2alm, maps, out)

    def _alm2map_raw(self, alm, dtype=torch.float64, out=None):
        return self._synthesis(_abi.lib().hbx_alm2map, alm, dtype, out)

    def _alm2map_adjoint_raw(self, maps, out=None):
        return self._analysis(_abi.lib().hbx_alm2map_adjoint, maps, out)

    def _map2alm_adjoint_raw(self, alm, dtype=torch.float64, out=None):
        return self._synthesis(_abi.lib().hbx_map2alm_adjoint, alm, dtype, out)

    # ------------------------------------------------------------------
    # spin transforms: [..., 2, npix] (Q, U) <-> [..., 2, nalm] (E, B)
    # ------------------------------------------------------------------
    def _pairs(self, t, n):
        squeeze = t.dim() == 2
        if t.dim() < 2 or t.shape[-2] != 2 or t.shape[-1] != n:
            raise ValueError(f"expected [..., 2, {n}] (a pair per polarisation field), got {tuple(t.shape)}")
        return t.reshape(-1, 2, n).contiguous(), squeeze

    def _spin_analysis(self, fn, maps, out=None):
        t = torch.as_tensor(maps)
        t = t.to(self.device) if t.device != self.device else t
        if t.dtype not in (torch.float64, torch.float32):
            t = t.to(torch.float64)
        m, squeeze = self._pairs(t, self.npix)
        npol = m.shape[0]
        alm = self._out(out, (npol, 2, self.nalm), self.alm_dtype)
        ws = self.workspace(npol)
        with torch.cuda.device(self.device):
            _abi.check(fn(self._h, npol, ctypes.c_void_p(m.data_ptr()), _map_dtype_code(m.dtype), self.npix,
                          ctypes.c_void_p(alm.data_ptr()), self.nalm, ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                          _stream_ptr(self.device)))
        if out is not None:
            return out
        return alm[0] if squeeze else alm

    def _spin_synthesis(self, fn, alm, dtype=torch.float64, out=None):
        t = torch.as_tensor(alm)
        t = t.to(self.device) if t.device != self.device else t
        t = t.to(self.alm_dtype) if t.dtype != self.alm_dtype else t
        a, squeeze = self._pairs(t, self.nalm)
        npol = a.shape[0]
        _map_dtype_code(dtype)
        maps = self._out(out, (npol, 2, self.npix), dtype)
        ws = self.workspace(npol)
        with torch.cuda.device(self.device):
            _abi.check(fn(self._h, npol, ctypes.c_void_p(a.data_ptr()), self.nalm, ctypes.c_void_p(maps.data_ptr()),
                          _map_dtype_code(maps.dtype), self.npix, ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                          _stream_ptr(self.device)))
        if out is not None:
            return out
        return maps[0] if squeeze else maps

    def map2alm_spin(self, maps, out=None):
        """(Q, U) [..., 2, npix] -> (E, B) [..., 2, nalm], healpy convention."""
        if out is None and _needs_grad(maps):
            from .autograd import _Map2AlmSpin
            return _Map2AlmSpin.apply(maps, self)
        return self._map2alm_spin_raw(maps, out)

    def alm2map_spin(self, alm, dtype=torch.float64, out=None):
        """(E, B) [..., 2, nalm] -> (Q, U) [..., 2, npix]."""
        if out is None and _needs_grad(alm):
            from .autograd import _Alm2MapSpin
            return _Alm2MapSpin.apply(alm, self, dtype)
        return self._alm2map_spin_raw(alm, dtype, out)

    def alm2map_spin_adjoint(self, maps, out=None):
        """Transpose of alm2map_spin (complex-gradient convention)."""
        if out is None and _needs_grad(maps):
            from .autograd import _Alm2MapSpinAdjoint
            return _Alm2MapSpinAdjoint.apply(maps, self)
        return self._alm2map_spin_adjoint_raw(maps, out)

    def map2alm_spin_adjoint(self, alm, dtype=torch.float64, out=None):
        """Transpose of map2alm_spin (complex-gradient convention)."""
        if out is None and _needs_grad(alm):
            from .autograd import _Map2AlmSpinAdjoint
            return _Map2AlmSpinAdjoint.apply(alm, self, dtype)
        return self._map2alm_spin_adjoint_raw(alm, dtype, out)

    def _map2alm_spin_raw(self, maps, out=None):
        return self._spin_analysis(_abi.lib().hbx_map2alm_spin, maps, out)

    def _alm2map_spin_raw(self, alm, dtype=torch.float64, out=None):
        return self._spin_synthesis(_abi.lib().hbx_alm2map_spin, alm, dtype, out)

    def _alm2map_spin_adjoint_raw(self, maps, out=None):
        return self._spin_analysis(_abi.lib().hbx_alm2map_spin_adjoint, maps, out)

    def _map2alm_spin_adjoint_raw(self, alm, dtype=torch.float64, out=None):
        return self._spin_synthesis(_abi.lib().hbx_map2alm_spin_adjoint, alm, dtype, out)

    # ------------------------------------------------------------------
    # stage entry points: maps <-> phase buffer <-> alm
    # ------------------------------------------------------------------
    @property
    def phase_map_elems(self):
        """Complex elements between consecutive maps of a phase buffer."""
        return phase_map_elems(self.geom.nrings, self.mmax)

    def phase_index(self, direction):
        """Offsets [ring][m] of one map's phase coefficients (phase_index())."""
        return phase_index(self.geom.nrings, self.mmax, direction)

    def phase_buffer(self, nmaps):
        b = ctypes.c_uint64()
        _abi.check(_abi.lib().hbx_phase_bytes(self._h, int(nmaps), ctypes.byref(b)))
        return torch.empty(max(int(b.value), 1), dtype=torch.uint8, device=self.device)

    def ring_analysis(self, maps, phase):
        m, _ = self._as_maps(maps)
        pb = self._buffer(phase, "phase")
        with torch.cuda.device(self.device):
            _abi.check(_abi.lib().hbx_ring_analysis(
                self._h, m.shape[0], ctypes.c_void_p(m.data_ptr()), _map_dtype_code(m.dtype), m.shape[1],
                ctypes.c_void_p(phase.data_ptr()), pb, _stream_ptr(self.device)))
        return phase

    def legendre_analysis(self, phase, nmaps, out=None, ws=None):
        alm = self._out(out, (int(nmaps), self.nalm), self.alm_dtype)
        ws = self.workspace(nmaps) if ws is None else ws
        pb, wb = self._buffer(phase, "phase"), self._buffer(ws, "ws")
        with torch.cuda.device(self.device):
            _abi.check(_abi.lib().hbx_legendre_analysis(
                self._h, int(nmaps), ctypes.c_void_p(phase.data_ptr()), pb,
                ctypes.c_void_p(alm.data_ptr()), self.nalm, ctypes.c_void_p(ws.data_ptr()), wb,
                _stream_ptr(self.device)))
        return alm

    def legendre_synthesis(self, alm, phase, ws=None):
        a, _ = self._as_alm(alm)
        ws = self.workspace(a.shape[0]) if ws is None else ws
        pb, wb = self._buffer(phase, "phase"), self._buffer(ws, "ws")
        with torch.cuda.device(self.device):
            _abi.check(_abi.lib().hbx_legendre_synthesis(
                self._h, a.shape[0], ctypes.c_void_p(a.data_ptr()), a.shape[1], ctypes.c_void_p(phase.data_ptr()),
                pb, ctypes.c_void_p(ws.data_ptr()), wb, _stream_ptr(self.device)))
        return phase

    def ring_synthesis(self, phase, nmaps, dtype=torch.float64, out=None):
        _map_dtype_code(dtype)
        maps = self._out(out, (int(nmaps), self.npix), dtype)
        pb = self._buffer(phase, "phase")
        with torch.cuda.device(self.device):
            _abi.check(_abi.lib().hbx_ring_synthesis(
                self._h, int(nmaps), ctypes.c_void_p(phase.data_ptr()), pb,
                ctypes.c_void_p(maps.data_ptr()), _map_dtype_code(maps.dtype), self.npix,
                _stream_ptr(self.device)))
        return maps


class KernelTimer:
    """HIP-event timing of library kernels (hbx_timer_arm / hbx_timer_read).

    ``with KernelTimer("legendre_analysis", "legendre_synthesis") as t: ...``
    then ``t.result["legendre_analysis"] == (launches, total_ms)``.  The
    events are recorded on the stream each kernel is launched on.
    """

    IDS = {"legendre_analysis": _abi.HBX_TIMER_LEGENDRE_ANALYSIS,
           "legendre_synthesis": _abi.HBX_TIMER_LEGENDRE_SYNTHESIS,
           "ring_analysis": _abi.HBX_TIMER_RING_ANALYSIS,
           "ring_synthesis": _abi.HBX_TIMER_RING_SYNTHESIS,
           "spin_analysis": _abi.HBX_TIMER_SPIN_ANALYSIS,
           "spin_synthesis": _abi.HBX_TIMER_SPIN_SYNTHESIS}

    def __init__(self, *kernels):
        self.kernels = kernels
        self.result = {}

    def __enter__(self):
        mask = 0
        for k in self.kernels:
            mask |= 1 << self.IDS[k]
        _abi.check(_abi.lib().hbx_timer_arm(mask))
        return self

    def __exit__(self, *exc):
        L = _abi.lib()
        for k in self.kernels:
            n = ctypes.c_int32()
            ms = ctypes.c_double()
            _abi.check(L.hbx_timer_read(self.IDS[k], ctypes.byref(n), ctypes.byref(ms)))
            self.result[k] = (int(n.value), float(ms.value))
        _abi.check(L.hbx_timer_arm(0))
        return False

    def mean_ms(self, kernel):
        n, ms = self.result[kernel]
        return ms / n if n else float("nan")


_CACHE = OrderedDict()
_CACHE_MAX = 8


def get_plan(geom, lmax, mmax=None, precision="f64", device=None, spin=0):
    """Cached Plan for (device, geometry, lmax, mmax, precision, spin)."""
    mmax = lmax if mmax is None else int(mmax)
    dev = torch.cuda.current_device() if (device is None and torch.cuda.is_available()) else device
    key = (str(dev), geom.key(), int(lmax), int(mmax), precision_code(precision), int(spin))
    p = _CACHE.get(key)
    if p is not None:
        _CACHE.move_to_end(key)
        return p
    p = Plan(geom, lmax, mmax, precision, device, spin)
    _CACHE[key] = p
    while len(_CACHE) > _CACHE_MAX:
        _CACHE.popitem(last=False)
    return p


def clear_cache():
    _CACHE.clear()


def alm2cl(alm, lmax, mmax=None):
    """C_l of [nalm] or [nmaps, nalm] m-major alm on the GPU (FP64 result)."""
    mmax = lmax if mmax is None else int(mmax)
    a = torch.as_tensor(alm)
    if not a.is_cuda:
        raise ValueError("alm2cl expects a device tensor")
    if a.dtype not in (torch.complex128, torch.complex64):
        a = a.to(torch.complex128)
    squeeze = a.dim() == 1
    a = (a.reshape(1, -1) if squeeze else a.reshape(-1, a.shape[-1])).contiguous()
    if a.shape[-1] != nalm(lmax, mmax):
        raise ValueError("alm size does not match lmax/mmax")
    nb = a.shape[0]
    cl = torch.empty((nb, lmax + 1), dtype=torch.float64, device=a.device)
    dt = _abi.HBX_DT_F64 if a.dtype == torch.complex128 else _abi.HBX_DT_F32
    with torch.cuda.device(a.device):
        _abi.check(_abi.lib().hbx_alm2cl(int(lmax), int(mmax), nb, ctypes.c_void_p(a.data_ptr()), dt,
                                         a.shape[-1], ctypes.c_void_p(cl.data_ptr()), lmax + 1,
                                         _stream_ptr(a.device)))
    return cl[0] if squeeze else cl


def lmax_from_size(size, mmax=None):
    """lmax of an m-major alm array with mmax = lmax (healpy Alm.getlmax)."""
    lmax = (math.isqrt(1 + 8 * int(size)) - 3) // 2
    if nalm(lmax) != size:
        raise ValueError(f"{size} is not the size of a full-triangle m-major alm array")
    return lmax
