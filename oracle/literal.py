"""Literal NumPy restatements of the reference transforms (cp -> np).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

These follow the reference source text statement by statement and are valid
only where the reference is finite: l + m <= 170 for the host norm
(``sht_raw.py:26-28``) and m <= 150 for the FP64 P_mm (``legendre.cu:13-21``),
i.e. lmax <= 85.  Output is the reference layout l(l+1)+m
(``alm_ops.py:4-5``).  The only deviation: ``np.math.factorial`` (removed in
NumPy 2, ``sht_fallback.py:8``) is ``math.factorial``.

Timing (the CPU baseline): ``norm=alm_norm_safe`` replaces the reference's
``float(factorial(l+m))``, which raises for l+m >= 171, by an exact integer
division (same statements and cost, no exception); ``m_stride`` visits every
k-th m only.  Above lmax 85 the results are not meaningful (P_mm overflows at
m >= 151, the norm underflows), and only the elapsed time is used.
"""
import math

import numpy as np

from .legendre import plm_literal


def _alm_norm(l, m):
    # sht_raw.py:26-28 / sht_fallback.py:4-8
    return np.sqrt((2 * l + 1) / (4 * np.pi) * float(math.factorial(l - m)) / float(math.factorial(l + m)))


def alm_norm_safe(l, m):
    """``_alm_norm`` with the factorial ratio as one exact integer division
    (finite for every l, m; underflows to 0 where the ratio does)."""
    return np.sqrt((2 * l + 1) / (4 * np.pi) * (math.factorial(l - m) / math.factorial(l + m)))


def sht_fallback_map2alm(rings, thetas, nphi_per_ring, weights_dz, lmax, mmax=None, norm=_alm_norm, m_stride=1):
    """``sht_fallback.map2alm_gpu`` (sht_fallback.py:36-65)."""
    thetas = np.asarray(thetas, dtype=np.float64)
    nphi_per_ring = np.asarray(nphi_per_ring, dtype=np.int32)
    w = np.asarray(weights_dz, dtype=np.float64)
    if mmax is None:
        mmax = lmax
    Nr = int(thetas.size)
    Fm = []
    for r in range(Nr):                                       # :51-56
        f = np.asarray(rings[r], dtype=np.float64)
        nphi = int(nphi_per_ring[r])
        F = np.fft.fft(f) / nphi
        Fm.append([F[m] for m in range(min(mmax, nphi - 1) + 1)])
    alm = np.zeros(((lmax + 1) * (lmax + 1),), dtype=np.complex128)
    x = np.cos(thetas)
    for m in range(0, mmax + 1, m_stride):                    # :59-64
        Pl = plm_literal(x, lmax, m)
        Fcol = np.asarray([Fm[r][m] if m < len(Fm[r]) else 0.0 + 0.0j for r in range(Nr)])
        for l in range(m, lmax + 1):
            val = np.sum((2 * np.pi) * Fcol * Pl[:, l] * w)
            alm[l * (l + 1) + m] = norm(l, m) * val
    return alm


def sht_raw_map2alm(rings, thetas, nphi_per_ring, weights_dz, lmax, mmax=None):
    """``sht_raw.map2alm_gpu`` (sht_raw.py:30-57): no 2 pi, m <= max nphi//2."""
    thetas = np.asarray(thetas, dtype=np.float64)
    nphi_per_ring = np.asarray(nphi_per_ring, dtype=np.int32)
    w = np.asarray(weights_dz, dtype=np.float64)
    if mmax is None:
        mmax = lmax
    Nr = int(thetas.size)
    Fm_list = []
    mmax_eff = 0
    for r in range(Nr):                                       # :38-43
        f_r = np.asarray(rings[r], dtype=np.float64)
        nphi = int(nphi_per_ring[r])
        Fm_list.append(np.fft.fft(f_r) / nphi)
        mmax_eff = max(mmax_eff, min(mmax, nphi // 2))
    x = np.cos(thetas)
    alm = np.zeros(((lmax + 1) * (lmax + 1),), dtype=np.complex128)
    for m in range(0, mmax_eff + 1):                          # :46-56
        Plm = plm_literal(x, lmax, m)
        norms = np.asarray([_alm_norm(l, m) for l in range(m, lmax + 1)])
        Plm_block = Plm[:, m:(lmax + 1)]
        Fm_vec = np.asarray([Fm_list[r][m] if m < int(nphi_per_ring[r]) else 0.0 + 0.0j
                             for r in range(Nr)], dtype=np.complex128)[:, None]
        accum = np.sum(w[:, None] * Plm_block * Fm_vec, axis=0) * norms
        for idx_l, l in enumerate(range(m, lmax + 1)):
            alm[l * (l + 1) + m] = accum[idx_l]
    return alm


def isht_alm2map(alm, thetas, nphi_per_ring, lmax, mmax=None, dtype=np.float64, norm=_alm_norm):
    """``isht.alm2map_gpu`` (isht.py:5-26), overwrite semantics and real part.

    ``cp.asrealarray`` (isht.py:25) does not exist; the real part is assumed.
    """
    thetas = np.asarray(thetas, dtype=np.float64)
    nphi_per_ring = np.asarray(nphi_per_ring, dtype=np.int32)
    Nr = int(thetas.size)
    if mmax is None:
        mmax = lmax
    x = np.cos(thetas)
    blocks = {m: plm_literal(x, lmax, m)[:, m:lmax + 1] for m in range(mmax + 1)}
    rings = []
    for r in range(Nr):                                       # :15-25
        nphi = int(nphi_per_ring[r])
        F = np.zeros((nphi,), dtype=np.complex128)
        for m in range(0, min(mmax, nphi - 1) + 1):
            norms = np.asarray([norm(l, m) for l in range(m, lmax + 1)])
            coeff = np.sum(norms * blocks[m][r] * np.asarray([alm[l * (l + 1) + m] for l in range(m, lmax + 1)]))
            F[m] = coeff
            if m > 0:
                F[nphi - m] = np.conj(coeff)
        f = np.fft.ifft(F) * nphi
        rings.append(np.real(f).astype(dtype))
    return rings
