"""T, Q, U maps from spectra (``synfast_tqu.py:69-128``), on the SHT of this
package.

Draws (host NumPy ``default_rng(seed)``, m-major alm): per l, T and E from
the Cholesky factor of [[C_TT, C_TE], [C_TE, C_EE]] (``synfast_tqu.py:27-58``),
B independent; m = 0 real with variance C, m > 0 real and imaginary parts
with variance C/2 each.  (The reference's T/E block gives its m > 0 parts
variance C each, twice its own scalar ``synalm`` and healpy's; the C/2
convention is kept here for T, E and B alike.)  Then beam and pixel window,
T = alm2map(a_T), (Q, U) = alm2map_spin(a_E, a_B) and optional white noise.
"""
import numpy as np
import torch

from . import _util
from .alm_ops import almxfl, synalm
from .beams import pixel_window_fl
from .plan import nalm
from .smoothing import gaussian_beam_fl


def _cholesky_te(ctt, cee, cte):
    """L11, L21, L22 per l (``synfast_tqu.py:36-43``; degenerate l -> diagonal)."""
    ok = (ctt > 0) & (cee > 0) & (ctt * cee - cte * cte > 0)
    l11 = np.sqrt(np.maximum(ctt, 0.0))
    l21 = np.where(ok, cte / np.where(ok, l11, 1.0), 0.0)
    l22 = np.where(ok, np.sqrt(np.maximum(cee - l21 * l21, 0.0)), np.sqrt(np.maximum(cee, 0.0)))
    return l11, l21, l22


def _synalm_te(ctt, cee, cbb, cte, lmax, rng):
    """(a_T, a_E, a_B) m-major complex128 on the host."""
    l11, l21, l22 = _cholesky_te(ctt, cee, cte)
    n = nalm(lmax)
    aT = np.zeros(n, dtype=np.complex128)
    aE = np.zeros(n, dtype=np.complex128)
    aB = np.zeros(n, dtype=np.complex128)
    i0 = 0
    for m in range(lmax + 1):
        ls = np.arange(m, lmax + 1)
        k = ls.size
        if m == 0:
            z1, z2, zb = rng.standard_normal(k), rng.standard_normal(k), rng.standard_normal(k)
            s = 1.0
        else:
            z1 = rng.standard_normal(k) + 1j * rng.standard_normal(k)
            z2 = rng.standard_normal(k) + 1j * rng.standard_normal(k)
            zb = rng.standard_normal(k) + 1j * rng.standard_normal(k)
            s = np.sqrt(0.5)
        aT[i0:i0 + k] = s * l11[ls] * z1
        aE[i0:i0 + k] = s * (l21[ls] * z1 + l22[ls] * z2)
        aB[i0:i0 + k] = s * np.sqrt(np.maximum(cbb[ls], 0.0)) * zb
        i0 += k
    return aT, aE, aB


def synfast_tqu(cls, nside, lmax=None, fwhm=None, use_pixel_window=True, seed=None, sigmaT=None, sigmaQ=None,
                sigmaU=None, return_alm=False, *, precision="f64", dtype=torch.float64, device=None):
    """cls: {'tt'} for T only, or {'tt', 'ee', 'bb', 'te'}; returns flat RING
    (T, Q, U) [and the filtered m-major (a_T, a_E, a_B) with return_alm]."""
    from .hp_compat import alm2map, alm2map_spin
    dev = device if device is not None else _util.device()
    get = lambda k: None if cls.get(k) is None else np.asarray(  # noqa: E731
        cls[k].cpu().numpy() if isinstance(cls[k], torch.Tensor) else cls[k], dtype=np.float64)
    ctt, cee, cbb, cte = get("tt"), get("ee"), get("bb"), get("te")
    if lmax is None:
        lmax = next(v for v in (ctt, cee, cbb, cte) if v is not None).size - 1
    lmax = int(lmax)
    rng = np.random.default_rng(seed)
    if cee is None and cbb is None and cte is None:
        aT = synalm(ctt[:lmax + 1], lmax, seed=rng, device=dev)
        aE = torch.zeros_like(aT)
        aB = torch.zeros_like(aT)
    else:
        if ctt is None or cee is None or cbb is None or cte is None:
            raise ValueError("For polarized synfast provide tt, ee, bb, te spectra.")
        aT, aE, aB = (torch.as_tensor(a, device=dev) for a in
                      _synalm_te(ctt[:lmax + 1], cee[:lmax + 1], cbb[:lmax + 1], cte[:lmax + 1], lmax, rng))
    bl = gaussian_beam_fl(lmax, fwhm, device=dev) if fwhm is not None else None
    pT, pP = pixel_window_fl(int(nside), lmax, pol=True, device=dev) if use_pixel_window else (None, None)

    def filt(a, p):
        if bl is None and p is None:
            return a
        fl = torch.ones(lmax + 1, dtype=torch.float64, device=dev)
        if bl is not None:
            fl = fl * bl
        if p is not None:
            fl = fl * p
        return almxfl(a, fl, lmax=lmax)

    aT, aE, aB = filt(aT, pT), filt(aE, pP), filt(aB, pP)
    T = alm2map(aT, int(nside), lmax=lmax, precision=precision, dtype=dtype)
    Q, U = alm2map_spin(aE, aB, 2, nside=int(nside), lmax=lmax, precision=precision, dtype=dtype)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(rng.integers(0, 2 ** 62)))
    for name, sig in (("T", sigmaT), ("Q", sigmaQ), ("U", sigmaU)):
        if sig is not None and sig > 0:
            noise = float(sig) * torch.randn(T.shape, generator=gen, device=dev, dtype=T.dtype)
            if name == "T":
                T = T + noise
            elif name == "Q":
                Q = Q + noise
            else:
                U = U + noise
    if return_alm:
        return T, Q, U, aT, aE, aB
    return T, Q, U
