"""Gaussian beam transfer functions (``smoothing.py:5-14``), m-major alm."""
import math

import torch

from . import _util
from .alm_ops import almxfl
from .plan import lmax_from_size


def gaussian_beam_fl(lmax, fwhm_rad, device=None):
    """b_l = exp(-l(l+1) sigma^2 / 2), sigma = fwhm / sqrt(8 ln 2), l = 0..lmax (FP64)."""
    sigma = float(fwhm_rad) / math.sqrt(8.0 * math.log(2.0))
    dev = device if device is not None else _util.device()
    ell = torch.arange(int(lmax) + 1, dtype=torch.float64, device=dev)
    return torch.exp(-0.5 * ell * (ell + 1.0) * sigma * sigma)


def smooth_alm_gaussian(alm, fwhm_rad):
    """alm * b_l, lmax from the m-major size (``smoothing.py:11-14``)."""
    a = torch.as_tensor(alm)
    lmax = lmax_from_size(a.shape[-1])
    return almxfl(a, gaussian_beam_fl(lmax, fwhm_rad, device=a.device))
