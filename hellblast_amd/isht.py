"""Scalar synthesis (drop-in for the reference's ``isht.alm2map_gpu``).

Reference: ``isht.py:5-26``.  The reference overwrites F[m] and F[nphi-m]
whenever nphi <= 2 mmax (``isht.py:18-23``); here every m < mcut_r is
accumulated into bin m mod nphi, which is the exact adjoint-consistent sum
and equals the reference on rings with nphi > 2 mmax.
"""
import numpy as np
import torch

from . import _util
from .plan import get_plan, nalm
from .sht import make_geometry


def alm2map_gpu(alm, thetas, nphi_per_ring, lmax, mmax=None, dtype=torch.float64, *, phi0=None,
                alias="reference", precision="f64", flat=False):
    """f_r(phi) = Re sum_m G_r(m) e^{i m phi}, G_r(m) = sum_l lambda_lm a_lm.

    alm: m-major [nalm] or [nmaps, nalm].  Returns a list of per-ring views
    (reference form) or, with ``flat=True``, the flat [npix] / [nmaps, npix]
    tensor.
    """
    mmax = lmax if mmax is None else int(mmax)
    geom = make_geometry(thetas, nphi_per_ring, None, phi0, alias, mmax)
    plan = get_plan(geom, int(lmax), mmax, precision)
    a = torch.as_tensor(alm)
    if a.shape[-1] != nalm(int(lmax), mmax):
        raise ValueError("alm size does not match lmax/mmax (m-major layout expected)")
    dt = dtype if isinstance(dtype, torch.dtype) else torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
    out = plan.alm2map(a, dtype=dt)
    if flat:
        return out
    return list(torch.split(out, geom.nphi.tolist(), dim=-1))
