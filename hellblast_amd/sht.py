"""Scalar analysis (drop-in for the reference's ``sht.map2alm_gpu``).

Reference: ``sht.py:1-4`` binds ``map2alm_gpu`` to ``sht_raw`` or
``sht_fallback`` at import time.  Here it binds to the HIP path only; a
missing ``libhbx.so`` raises at the first call.
"""
import math

import numpy as np

from . import _util
from .plan import Geometry, get_plan


def make_geometry(thetas, nphi_per_ring, weights=None, phi0=None, alias="reference", mmax=0,
                  two_pi=True):
    nphi = _util.host_array(nphi_per_ring, np.int32).ravel()
    th = _util.host_array(thetas, np.float64).ravel()
    w = None
    if weights is not None:
        w = _util.host_array(weights, np.float64).ravel() * (2.0 * math.pi if two_pi else 1.0)
    ph = None if phi0 is None else _util.host_array(phi0, np.float64).ravel()
    return Geometry(nphi, theta=th, weight=w, phi0=ph, mcut=_util.alias_mcut(alias, nphi, mmax))


def map2alm_gpu(rings, thetas, nphi_per_ring, weights_dz, lmax, mmax=None, *, phi0=None,
                alias="reference", precision="f64"):
    """a_lm = 2 pi sum_r w_r lambda_lm(cos theta_r) F_r(m)  (sht_fallback.py:36-65).

    rings: list of per-ring arrays (reference form) or a flat [npix] /
    [nmaps, npix] tensor in ring order.  Returns m-major complex alm
    (complex128, or complex64 with ``precision='f32'``) on the GPU.
    ``phi0`` gives ring longitudes (default 0, as in the reference);
    ``alias='fold'`` uses the periodic healpy aliasing instead of the
    reference's m < nphi rule (sht_raw.py:51).
    """
    mmax = lmax if mmax is None else int(mmax)
    geom = make_geometry(thetas, nphi_per_ring, weights_dz, phi0, alias, mmax)
    plan = get_plan(geom, int(lmax), mmax, precision)
    return plan.map2alm(_util.flat_maps(rings, geom.npix))
