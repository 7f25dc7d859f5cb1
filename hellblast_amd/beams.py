"""Beam and pixel-window filtering of maps and alm (``beams.py:6-47``), on the
SHT of this package (m-major alm, flat RING maps)."""
import torch

from . import _util
from .alm_ops import almxfl
from .smoothing import gaussian_beam_fl


def pixel_window_fl(nside, lmax, pol=False, device=None):
    """Pixel window p_l (``beams.py:6-22``): healpy's table when healpy is
    importable, else ones (the reference's own fallback).  pol=True returns
    (p_T, p_P); otherwise (p_T, None)."""
    dev = device if device is not None else _util.device()
    try:
        import healpy as hp  # optional, CPU
        w = hp.sphtfunc.pixwin(int(nside), lmax=int(lmax), pol=pol)
        if pol:
            return (torch.as_tensor(w[0], dtype=torch.float64, device=dev),
                    torch.as_tensor(w[1], dtype=torch.float64, device=dev))
        return torch.as_tensor(w, dtype=torch.float64, device=dev), None
    except Exception:
        ones = torch.ones(int(lmax) + 1, dtype=torch.float64, device=dev)
        return (ones, ones.clone()) if pol else (ones, None)


def smooth_alm_with_beam(alm, lmax, fwhm=None, bl=None, pixwin=None):
    """alm * F_l, F_l = b_l p_l (``beams.py:24-36``): b_l given, else a Gaussian
    of fwhm (radians) if given, else 1; p_l applied when given."""
    a = torch.as_tensor(alm)
    fl = torch.ones(int(lmax) + 1, dtype=torch.float64, device=a.device)
    if bl is not None:
        fl = fl * torch.as_tensor(bl, dtype=torch.float64, device=a.device)
    elif fwhm is not None:
        fl = fl * gaussian_beam_fl(lmax, fwhm, device=a.device)
    if pixwin is not None:
        fl = fl * torch.as_tensor(pixwin, dtype=torch.float64, device=a.device)
    return almxfl(a, fl, lmax=lmax)


def smooth_map(m, nside, lmax=None, fwhm=None, bl=None, use_pixel_window=True, nest=False, **kw):
    """map2alm -> alm * b_l p_l -> alm2map (``beams.py:38-47``); lmax defaults
    to 3 nside.  Extra keywords (precision, ...) go to both transforms."""
    from .hp_compat import alm2map, map2alm
    nside = int(nside)
    lmax = 3 * nside if lmax is None else int(lmax)
    pw = pixel_window_fl(nside, lmax)[0] if use_pixel_window else None
    alm = map2alm(m, lmax=lmax, nest=nest, **kw)
    alm_sm = smooth_alm_with_beam(alm, lmax, fwhm=fwhm, bl=bl, pixwin=pw)
    return alm2map(alm_sm, nside, lmax=lmax, nest=nest, **kw)
