"""alm utilities in the healpy m-major layout (reference: ``alm_ops.py``).

The reference stores alm at ``l*(l+1)+m`` in an array of length (lmax+1)^2
(``alm_ops.py:4-5``); the transforms here use healpy's m-major packing
idx = m(2 lmax + 1 - m)/2 + l.  ``to_reference_layout`` /
``from_reference_layout`` convert between the two by pure gathers.
"""
import numpy as np
import torch

from . import _util
from .plan import alm2cl as _alm2cl_gpu
from .plan import lmax_from_size, nalm


def alm_index(l, m):
    """Reference layout index l(l+1)+m (``alm_ops.py:4-5``)."""
    return l * (l + 1) + m


def getidx(lmax, l, m):
    """m-major index m(2 lmax + 1 - m)/2 + l."""
    return m * (2 * lmax + 1 - m) // 2 + l


def getsize(lmax, mmax=None):
    return nalm(lmax, mmax)


def getlmax(size, mmax=None):
    return lmax_from_size(size, mmax)


def _lm_arrays(lmax, mmax):
    ls, ms = [], []
    for m in range(mmax + 1):
        ls.append(np.arange(m, lmax + 1))
        ms.append(np.full(lmax + 1 - m, m))
    return np.concatenate(ls), np.concatenate(ms)


def alm2cl(alm, lmax=None, mmax=None):
    """C_l = (|a_l0|^2 + 2 sum_{m>=1} |a_lm|^2)/(2l+1) (``alm_ops.py:19-32``)."""
    a = torch.as_tensor(alm)
    if lmax is None:
        lmax = lmax_from_size(a.shape[-1])
    mmax = lmax if mmax is None else mmax
    if not a.is_cuda:
        a = a.to(_util.device())
    return _alm2cl_gpu(a, int(lmax), int(mmax))


def to_reference_layout(alm, lmax, mmax=None):
    """m-major [..., nalm] -> reference l(l+1)+m layout [..., (lmax+1)^2]."""
    mmax = lmax if mmax is None else mmax
    a = torch.as_tensor(alm)
    l, m = _lm_arrays(lmax, mmax)
    dst = torch.as_tensor(alm_index(l, m), device=a.device)
    src = torch.as_tensor(getidx(lmax, l, m), device=a.device)
    out = torch.zeros(a.shape[:-1] + ((lmax + 1) ** 2,), dtype=a.dtype, device=a.device)
    out[..., dst] = a[..., src]
    return out


def from_reference_layout(alm_ref, lmax, mmax=None):
    """Reference l(l+1)+m layout -> m-major (m <= mmax)."""
    mmax = lmax if mmax is None else mmax
    a = torch.as_tensor(alm_ref)
    l, m = _lm_arrays(lmax, mmax)
    src = torch.as_tensor(alm_index(l, m), device=a.device)
    dst = torch.as_tensor(getidx(lmax, l, m), device=a.device)
    out = torch.zeros(a.shape[:-1] + (nalm(lmax, mmax),), dtype=a.dtype, device=a.device)
    out[..., dst] = a[..., src]
    return out


def almxfl(alm, fl, lmax=None, mmax=None):
    """a_lm * f_l (``alm_ops.py:7-17``) on m-major alm."""
    a = torch.as_tensor(alm)
    if lmax is None:
        lmax = lmax_from_size(a.shape[-1])
    mmax = lmax if mmax is None else mmax
    l, _ = _lm_arrays(lmax, mmax)
    f = torch.as_tensor(np.asarray(fl, dtype=np.float64) if not isinstance(fl, torch.Tensor) else fl,
                        device=a.device).to(torch.float64)
    return a * f[torch.as_tensor(l, device=a.device)].to(a.real.dtype)


def synalm(cl, lmax=None, mmax=None, seed=None, device=None):
    """Gaussian alm with spectrum C_l (``alm_ops.py:34-54``), m-major.

    a_l0 ~ N(0, C_l) real; Re, Im a_lm ~ N(0, C_l/2).  Drawn on the host
    with ``numpy.random.default_rng(seed)`` (the reference's CuPy stream is
    not reproducible outside CuPy).
    """
    cl = np.asarray(cl, dtype=np.float64)
    lmax = cl.size - 1 if lmax is None else int(lmax)
    mmax = lmax if mmax is None else int(mmax)
    rng = np.random.default_rng(seed)
    out = np.zeros(nalm(lmax, mmax), dtype=np.complex128)
    for m in range(mmax + 1):
        i0 = getidx(lmax, m, m)
        c = np.maximum(cl[m:lmax + 1], 0.0)
        n = lmax - m + 1
        if m == 0:
            out[i0:i0 + n] = rng.standard_normal(n) * np.sqrt(c)
        else:
            s = np.sqrt(c / 2.0)
            out[i0:i0 + n] = rng.standard_normal(n) * s + 1j * rng.standard_normal(n) * s
    dev = device if device is not None else (_util.device() if torch.cuda.is_available() else "cpu")
    return torch.as_tensor(out, device=dev)
