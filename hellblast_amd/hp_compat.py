"""healpy-style flat-map wrappers for the SHT path (reference: ``hp_compat.py:24-102``).

Maps are flat RING-ordered arrays ([npix] or batched [nmaps, npix]); alm are
m-major.  Defaults follow the reference: lmax = 3*nside (``hp_compat.py:45``),
midpoint-Delta-z weights (``ring_tables.py:39-40``), phi0 = 0 and the
reference aliasing rule.  ``nest=True`` is rejected instead of being ignored
silently as the reference does (``hp_compat.py:40-54``).
"""
import math
from functools import lru_cache

import numpy as np
import torch

from . import _util
from .plan import Geometry, get_plan, lmax_from_size, nalm
from .ring_tables import healpix_geometry


def _nside_from_npix(npix: int) -> int:
    """``hp_compat._nside_from_npix`` (hp_compat.py:24-29)."""
    nside = int(math.isqrt(int(npix) // 12))
    if 12 * nside * nside != npix:
        raise ValueError("npix is not 12*nside^2")
    return nside


@lru_cache(maxsize=64)
def healpix_geom(nside, mmax, weights="dz", phi0="reference", alias="reference"):
    """Geometry of a HEALPix RING map for the plan cache."""
    g = healpix_geometry(nside)
    nphi = g["nphi"]
    if weights == "dz":
        w = g["z_edges"][:-1] - g["z_edges"][1:]
    elif weights == "area":
        w = 2.0 * nphi.astype(np.float64) / g["npix"]
    else:
        raise ValueError("weights must be 'dz' or 'area'")
    if phi0 in (None, "reference", "zero"):
        ph = None
    elif phi0 == "healpix":
        ph = g["phi0_healpix"]
    else:
        raise ValueError("phi0 must be 'reference' or 'healpix'")
    return Geometry(nphi, theta=g["theta_centers"], weight=2.0 * math.pi * w, phi0=ph,
                    mcut=_util.alias_mcut(alias, nphi, mmax), ring_start=g["ring_start"])


def _check_nest(nest):
    if nest:
        raise NotImplementedError("NESTED maps are not supported by the SHT path; reorder to RING first")


def _plan(nside, lmax, mmax, precision, weights, phi0, alias, spin=0):
    return get_plan(healpix_geom(int(nside), int(mmax), weights, phi0, alias), int(lmax), int(mmax), precision,
                    spin=spin)


def map2alm(m, lmax=None, mmax=None, nest=False, iter=0, *, precision="f64", weights="dz",
            phi0="reference", alias="reference"):
    """Analysis of a flat RING map (``hp_compat.py:40-48``), m-major output.

    ``iter > 0`` applies Jacobi refinement a <- a + A(f - S a).
    """
    _check_nest(nest)
    t = _util.flat_maps(m, m.shape[-1] if hasattr(m, "shape") else len(m))
    nside = _nside_from_npix(t.shape[-1])
    lmax = 3 * nside if lmax is None else int(lmax)
    mmax = lmax if mmax is None else int(mmax)
    plan = _plan(nside, lmax, mmax, precision, weights, phi0, alias)
    alm = plan.map2alm(t)
    for _ in range(int(iter)):
        resid = t - plan.alm2map(alm, dtype=t.dtype)
        alm = alm + plan.map2alm(resid)
    return alm


def alm2map(alm, nside, lmax=None, nest=False, mmax=None, *, precision="f64", dtype=torch.float64,
            phi0="reference", alias="reference"):
    """Synthesis of m-major alm to a flat RING map (``hp_compat.py:50-54``)."""
    _check_nest(nest)
    a = torch.as_tensor(alm)
    if lmax is None:
        lmax = lmax_from_size(a.shape[-1])
    mmax = lmax if mmax is None else int(mmax)
    if a.shape[-1] != nalm(int(lmax), mmax):
        raise ValueError("alm size does not match lmax/mmax (m-major layout expected)")
    plan = _plan(nside, lmax, mmax, precision, "dz", phi0, alias)
    return plan.alm2map(a, dtype=dtype)


def anafast(m, lmax=None, mmax=None, nest=False, iter=0, **kw):
    """C_l of map2alm(m) (``anafast.py:5-7`` on a flat map)."""
    from .alm_ops import alm2cl
    alm = map2alm(m, lmax=lmax, mmax=mmax, nest=nest, iter=iter, **kw)
    nside = _nside_from_npix(m.shape[-1])
    lmax = 3 * nside if lmax is None else int(lmax)
    return alm2cl(alm, lmax=lmax, mmax=lmax if mmax is None else mmax)


def alm2cl(alm, lmax=None, mmax=None):
    from .alm_ops import alm2cl as _a
    return _a(alm, lmax=lmax, mmax=mmax)


# ---------- spin transforms ----------
def _check_spin(spin):
    spin = int(spin)
    if spin < 1 or spin > 8:
        raise ValueError("spin must be in 1..8 (the reference accepts 1 or 2)")
    return spin


def _wants_eb(spin, eb):
    """The reference's pair convention: spin 1 -> (1_a, -1_a) (hp_compat.py:66-67,
    81-82), spin 2 -> (E, B) (hp_compat.py:64-65, 79-80); ``eb`` overrides it
    (True: always (E, B); False: always the spin pair (s_a, -s_a))."""
    return spin != 1 if eb is None else bool(eb)


def map2alm_spin(Q, U, spin=2, lmax=None, mmax=None, nest=False, *, eb=None, precision="f64", weights="dz",
                 phi0="reference", alias="reference"):
    """Spin-weighted analysis of flat RING Q, U maps (``hp_compat.py:57-69``), m-major.

    Returns (E, B) for spin 2 and, as the reference's spin=1 branch
    (``map2alm_spin1_gpu``, ``hp_compat.py:66-67``), the pair (1_a, -1_a)
    for spin 1, with s_a = -(E + iB) and -s_a = -(-1)^s (E - iB).  Spins
    3..8 (no reference counterpart) return (E, B).  ``eb=True`` / ``False``
    selects (E, B) / the spin pair for any spin.  Batched [nmaps, npix]
    inputs give [nmaps, nalm] outputs.
    """
    from .spin import eb_to_spin_pair
    _check_nest(nest)
    spin = _check_spin(spin)
    q = _util.flat_maps(Q, Q.shape[-1] if hasattr(Q, "shape") else len(Q))
    u = _util.flat_maps(U, q.shape[-1], dtype=q.dtype)
    nside = _nside_from_npix(q.shape[-1])
    lmax = 3 * nside if lmax is None else int(lmax)
    mmax = lmax if mmax is None else int(mmax)
    plan = _plan(nside, lmax, mmax, precision, weights, phi0, alias, spin)
    out = plan.map2alm_spin(torch.stack([q, u], dim=-2))
    E, B = out[..., 0, :], out[..., 1, :]
    return (E, B) if _wants_eb(spin, eb) else eb_to_spin_pair(E, B, spin)


def alm2map_spin(Aplus, Aminus, spin=2, nside=None, lmax=None, nest=False, mmax=None, *, eb=None,
                 precision="f64", dtype=torch.float64, phi0="reference", alias="reference"):
    """Flat RING (Q, U) from m-major spin alm (``hp_compat.py:71-84``).

    (Aplus, Aminus) are (E, B) for spin 2 (``qu_from_eb_gpu``) and the pair
    (1_a, -1_a) for spin 1 (``alm2map_spin1_gpu``), as in the reference;
    ``eb`` overrides this as in ``map2alm_spin``.  nside defaults to
    ceil(lmax/3) as in the reference (``hp_compat.py:72-75``).
    """
    from .spin import spin_pair_to_eb
    _check_nest(nest)
    spin = _check_spin(spin)
    a = torch.as_tensor(Aplus)
    b = torch.as_tensor(Aminus)
    if lmax is None:
        lmax = lmax_from_size(a.shape[-1])
    mmax = lmax if mmax is None else int(mmax)
    if a.shape != b.shape or a.shape[-1] != nalm(int(lmax), mmax):
        raise ValueError("Aplus and Aminus must both be m-major alm of lmax/mmax")
    nside = max(1, -(-int(lmax) // 3)) if nside is None else int(nside)
    plan = _plan(nside, lmax, mmax, precision, "dz", phi0, alias, spin)
    dev = plan.device
    a, b = a.to(dev), b.to(dev)
    if not _wants_eb(spin, eb):
        a, b = spin_pair_to_eb(a, b, spin)
    qu = plan.alm2map_spin(torch.stack([a, b], dim=-2), dtype=dtype)
    return qu[..., 0, :], qu[..., 1, :]


def iqu2eb(Q, U, lmax=None, nest=False, **kw):
    """(E, B) of a spin-2 field (``hp_compat.py:87-94``)."""
    return map2alm_spin(Q, U, 2, lmax=lmax, nest=nest, eb=True, **kw)


def eb2iqu(E, B, nside=None, lmax=None, nest=False, **kw):
    """(Q, U) of spin-2 (E, B) (``hp_compat.py:96-102``)."""
    return alm2map_spin(E, B, 2, nside=nside, lmax=lmax, nest=nest, eb=True, **kw)


# ---------- alm utilities (``hp_compat.py:113-120``) ----------
def almxfl(alm, fl, lmax=None, mmax=None):
    """a_lm f_l on m-major alm (``alm_ops.py:7-17``)."""
    from .alm_ops import almxfl as _almxfl
    return _almxfl(alm, fl, lmax=lmax, mmax=mmax)


def synalm(cl, seed=None, lmax=None, mmax=None):
    """Gaussian m-major alm with spectrum cl (``alm_ops.py:34-54``; drawn with
    NumPy ``default_rng(seed)``)."""
    from .alm_ops import synalm as _synalm
    return _synalm(cl, lmax=lmax, mmax=mmax, seed=seed)


# ---------- RING geometry (``hp_compat.py:131-149``) ----------
def _check_nest_pix(nest):
    if nest:
        raise NotImplementedError("NESTED pixel indexing is not part of this package; use nest=False (RING)")


def ang2pix(nside, theta, phi, nest=False):
    """RING pixel of (theta, phi) (``coords_ring.cu:9-28``); NESTED is not built."""
    _check_nest_pix(nest)
    from .coords_ring import RingMapperGPU
    return RingMapperGPU(int(nside)).ang2pix(theta, phi)


def pix2ang(nside, ipix, nest=False):
    """(theta, phi) of RING pixels (``coords_ring.cu:30-54``); NESTED is not built."""
    _check_nest_pix(nest)
    from .coords_ring import RingMapperGPU
    return RingMapperGPU(int(nside)).pix2ang(ipix)


def pix2vec(nside, ipix, nest=False):
    """Unit vectors of RING pixel centres (``hp_compat.py:143-149``)."""
    th, ph = pix2ang(nside, ipix, nest=nest)
    return torch.sin(th) * torch.cos(ph), torch.sin(th) * torch.sin(ph), torch.cos(th)


# ---------- smoothing and synfast (``hp_compat.py:228-236``) ----------
def smooth_map(m, nside, lmax=None, fwhm=None, bl=None, use_pixel_window=True, nest=False, **kw):
    from .beams import smooth_map as _smooth_map
    return _smooth_map(m, nside, lmax=lmax, fwhm=fwhm, bl=bl, use_pixel_window=use_pixel_window, nest=nest, **kw)


def synfast(cls, nside, lmax=None, fwhm=None, use_pixel_window=True, seed=None, sigmaT=None, sigmaQ=None,
            sigmaU=None, return_alm=False, **kw):
    from .synfast_tqu import synfast_tqu
    return synfast_tqu(cls, nside, lmax=lmax, fwhm=fwhm, use_pixel_window=use_pixel_window, seed=seed,
                       sigmaT=sigmaT, sigmaQ=sigmaQ, sigmaU=sigmaU, return_alm=return_alm, **kw)


# ---------- FITS helpers (``hp_compat.py:204-225``) ----------
def read_map(path, hdu=1, to_gpu=True, dtype=np.float64):
    from .io_fits import read_map_fits
    return read_map_fits(path, hdu=hdu, to_gpu=to_gpu, dtype=dtype)


def write_map(path, m, header=None, overwrite=True):
    from .io_fits import write_map_fits
    return write_map_fits(path, m, header=header, overwrite=overwrite)


def read_alm(path, hdu=1, to_gpu=True, dtype=np.complex128):
    from .io_fits import read_alm_fits
    return read_alm_fits(path, hdu=hdu, to_gpu=to_gpu, dtype=dtype)


def write_alm(path, alm, header=None, overwrite=True):
    from .io_fits import write_alm_fits
    return write_alm_fits(path, alm, header=header, overwrite=overwrite)


def read_tqu(path, to_gpu=True):
    from .io_fits_multi import read_tqu_fits
    return read_tqu_fits(path, to_gpu=to_gpu)


def write_tqu(path, T=None, Q=None, U=None, nside=None, ordering="RING", coordsys="G", overwrite=True,
              extra_header=None):
    from .io_fits_multi import write_tqu_fits
    return write_tqu_fits(path, T=T, Q=Q, U=U, nside=nside, ordering=ordering, coordsys=coordsys,
                          overwrite=overwrite, extra_header=extra_header)
