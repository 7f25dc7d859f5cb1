"""RING pixel <-> angle on the GPU (drop-in for ``coords_ring.RingMapperGPU``).

Reference: ``coords_ring.py:18-68`` over the kernels of ``coords_ring.cu:9-54``.
The reference's rule is kept (ring = z-edge cell of cos(theta), pixel = the
nearest centre in phi with the reference's phi_center0); both directions
run as HIP kernels of libhbx (``hbx_ang2pix_ring``, ``hbx_pix2ang_ring``).
``phi0="healpix"`` uses HEALPix's ring offsets instead.
"""
import ctypes
import math

import torch

from . import _abi
from .plan import _stream_ptr
from .ring_tables import HealpixRingTables


def _dev_f64(x, device):
    return torch.as_tensor(x, dtype=torch.float64, device=device).contiguous()


class RingMapperGPU:
    """``RingMapperGPU(nside)``: ``ang2pix(theta, phi)``, ``pix2ang(ipix)``,
    ``weights_dz()`` (``coords_ring.py:26-56``)."""

    def __init__(self, nside: int, device=None, phi0="reference"):
        self.tables = HealpixRingTables(nside, device=device)
        if phi0 not in ("reference", "healpix"):
            raise ValueError("phi0 must be 'reference' or 'healpix'")
        t = self.tables
        self._phi0 = t.phi_center0 if phi0 == "reference" else t.phi0_healpix
        self._nphi = t.nphi.to(torch.int32).contiguous()
        self._ring_start = t.ring_start.to(torch.int64).contiguous()
        self.device = t.device
        if self.device.type != "cuda":
            raise RuntimeError("RingMapperGPU needs a GPU device (libhbx kernels)")

    def ang2pix(self, theta, phi):
        th = _dev_f64(theta, self.device)
        ph = _dev_f64(phi, self.device)
        if th.shape != ph.shape:
            raise ValueError("theta and phi shapes differ")
        flat_t, flat_p = th.reshape(-1), ph.reshape(-1)
        out = torch.empty(flat_t.shape, dtype=torch.int64, device=self.device)
        t = self.tables
        with torch.cuda.device(self.device):
            _abi.check(_abi.lib().hbx_ang2pix_ring(
                flat_t.numel(), ctypes.c_void_p(flat_t.data_ptr()), ctypes.c_void_p(flat_p.data_ptr()), t.Nr,
                ctypes.c_void_p(t.z_edges.data_ptr()), ctypes.c_void_p(self._ring_start.data_ptr()),
                ctypes.c_void_p(self._nphi.data_ptr()), ctypes.c_void_p(self._phi0.data_ptr()),
                ctypes.c_void_p(out.data_ptr()), _stream_ptr(self.device)))
        return out.reshape(th.shape)

    def pix2ang(self, ipix):
        p = torch.as_tensor(ipix, dtype=torch.int64, device=self.device).contiguous()
        flat = p.reshape(-1)
        theta = torch.empty(flat.shape, dtype=torch.float64, device=self.device)
        phi = torch.empty(flat.shape, dtype=torch.float64, device=self.device)
        t = self.tables
        with torch.cuda.device(self.device):
            _abi.check(_abi.lib().hbx_pix2ang_ring(
                flat.numel(), ctypes.c_void_p(flat.data_ptr()), t.Nr, ctypes.c_void_p(t.theta_centers.data_ptr()),
                ctypes.c_void_p(self._ring_start.data_ptr()), ctypes.c_void_p(self._nphi.data_ptr()),
                ctypes.c_void_p(self._phi0.data_ptr()), ctypes.c_void_p(theta.data_ptr()),
                ctypes.c_void_p(phi.data_ptr()), _stream_ptr(self.device)))
        return theta.reshape(p.shape), phi.reshape(p.shape)

    def weights_dz(self):
        return self.tables.weights_dz()


def vec2pix_ring_gpu(x, y, z, mapper: RingMapperGPU):
    """Unit vector -> RING pixel (``coords_ring.py:59-68``)."""
    dev = mapper.device
    x, y, z = (_dev_f64(v, dev) for v in (x, y, z))
    theta = torch.arccos(torch.clamp(z, -1.0, 1.0))
    phi = torch.atan2(y, x)
    phi = torch.where(phi < 0, phi + 2.0 * math.pi, phi)
    return mapper.ang2pix(theta, phi)
