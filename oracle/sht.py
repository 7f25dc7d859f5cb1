"""FP64 scalar SHT on the CPU in the sht_fallback + isht convention.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Analysis  (``sht_fallback.py:36-65``):
    F_r(m) = [m < mcut_r] e^{-i m phi0_r} DFT_r[m mod nphi_r] / nphi_r
    a_lm   = 2 pi  sum_r w_r lambda_lm(cos theta_r) F_r(m)
Synthesis (``isht.py:5-26`` with bins accumulated instead of overwritten):
    G_r(m) = sum_l lambda_lm(cos theta_r) a_lm
    f_rj   = Re sum_{|m| < mcut_r} G_r(m) e^{i m (phi0_r + 2 pi j / nphi_r)}
alm2cl    (``alm_ops.py:19-32``):
    C_l = (|a_l0|^2 + 2 sum_{m>=1} |a_lm|^2) / (2l+1)

With phi0 = 0 and mcut_r = nphi_r this is exactly the reference rule of
``sht_raw.py:51`` / ``sht_fallback.py:55,63``.  Every ring is handled
individually (no north/south pairing), so the oracle also pins the GPU's
pairing logic.  alm are healpy m-major: idx(l,m) = m(2 lmax + 1 - m)/2 + l.
"""
import math

import numpy as np

from .legendre import lambda_all


# ---------------------------------------------------------------------------
# layouts
# ---------------------------------------------------------------------------

def nalm(lmax, mmax=None):
    mmax = lmax if mmax is None else mmax
    return (mmax + 1) * (lmax + 1) - mmax * (mmax + 1) // 2


def mmajor_index(l, m, lmax):
    """healpy ``Alm.getidx``: m (2 lmax + 1 - m) / 2 + l."""
    return m * (2 * lmax + 1 - m) // 2 + l


def ref_index(l, m):
    """``alm_ops.alm_index`` (alm_ops.py:4-5): l (l+1) + m."""
    return l * (l + 1) + m


def mmajor_to_ref(alm, lmax, mmax=None):
    """Gather m-major alm into the reference layout of length (lmax+1)^2."""
    mmax = lmax if mmax is None else mmax
    out = np.zeros((lmax + 1) ** 2, dtype=np.complex128)
    for m in range(mmax + 1):
        l = np.arange(m, lmax + 1)
        out[ref_index(l, m)] = alm[mmajor_index(l, m, lmax)]
    return out


def ref_to_mmajor(alm_ref, lmax, mmax=None):
    mmax = lmax if mmax is None else mmax
    out = np.zeros(nalm(lmax, mmax), dtype=np.complex128)
    for m in range(mmax + 1):
        l = np.arange(m, lmax + 1)
        out[mmajor_index(l, m, lmax)] = alm_ref[ref_index(l, m)]
    return out


# ---------------------------------------------------------------------------
# ring phase stage
# ---------------------------------------------------------------------------

def ring_phase(flat, ring_start, nphi, mmax, phi0=None, mcut=None):
    """F[r, m] for m = 0..mmax (``sht_raw.py:38-43,51``)."""
    nr = len(nphi)
    F = np.zeros((nr, mmax + 1), dtype=np.complex128)
    m = np.arange(mmax + 1)
    for r in range(nr):
        n = int(nphi[r])
        f = np.asarray(flat[ring_start[r]:ring_start[r] + n], dtype=np.float64)
        X = np.fft.fft(f) / n
        cut = n if mcut is None else int(mcut[r])
        v = X[m % n] * (m < cut)
        if phi0 is not None and phi0[r] != 0.0:
            v = v * np.exp(-1j * m * phi0[r])
        F[r] = v
    return F


def ring_synthesis(G, ring_start, nphi, mmax, phi0=None, mcut=None, dtype=np.float64):
    """Real ring values from G[r, m] (``isht.py:16-25``, bins accumulated)."""
    nr = len(nphi)
    out = np.zeros(int(ring_start[-1]), dtype=dtype)
    for r in range(nr):
        n = int(nphi[r])
        cut = n if mcut is None else int(mcut[r])
        top = min(mmax, cut - 1)
        g = G[r, :top + 1].copy()
        if phi0 is not None and phi0[r] != 0.0:
            g = g * np.exp(1j * np.arange(top + 1) * phi0[r])
        X = np.zeros(n, dtype=np.complex128)
        X[0] += g[0].real
        if top >= 1:
            ms = np.arange(1, top + 1)
            np.add.at(X, ms % n, g[1:])
            np.add.at(X, (-ms) % n, np.conj(g[1:]))
        f = np.fft.ifft(X) * n
        out[ring_start[r]:ring_start[r] + n] = f.real
    return out


# ---------------------------------------------------------------------------
# Legendre stage
# ---------------------------------------------------------------------------

def _trig(theta):
    theta = np.asarray(theta, dtype=np.float64)
    return np.cos(theta), np.sin(theta)


def legendre_analysis(F, theta, weight, lmax, mmax):
    """a_lm = sum_r weight_r lambda_lm(x_r) F[r, m]; weight includes 2 pi."""
    x, st = _trig(theta)
    alm = np.zeros(nalm(lmax, mmax), dtype=np.complex128)
    wF = F * np.asarray(weight, dtype=np.float64)[:, None]
    for m, lam in lambda_all(lmax, mmax, x, st):
        i0 = mmajor_index(m, m, lmax)
        alm[i0:i0 + lmax - m + 1] = lam @ wF[:, m]
    return alm


def legendre_synthesis(alm, theta, lmax, mmax):
    """G[r, m] = sum_l lambda_lm(x_r) a_lm."""
    x, st = _trig(theta)
    G = np.zeros((x.size, mmax + 1), dtype=np.complex128)
    for m, lam in lambda_all(lmax, mmax, x, st):
        i0 = mmajor_index(m, m, lmax)
        G[:, m] = lam.T @ alm[i0:i0 + lmax - m + 1]
    return G


def map2alm(flat, tabs, lmax, mmax=None, weights=None, phi0=None, mcut=None):
    """FP64 analysis of one flat RING map (sht_fallback convention, 2 pi)."""
    mmax = lmax if mmax is None else mmax
    w = tabs["weights_dz"] if weights is None else weights
    F = ring_phase(flat, tabs["ring_start"], tabs["nphi"], mmax, phi0, mcut)
    return legendre_analysis(F, tabs["theta_centers"], 2.0 * np.pi * np.asarray(w), lmax, mmax)


def alm2map(alm, tabs, lmax, mmax=None, phi0=None, mcut=None):
    """FP64 synthesis of one flat RING map."""
    mmax = lmax if mmax is None else mmax
    G = legendre_synthesis(np.asarray(alm, dtype=np.complex128), tabs["theta_centers"], lmax, mmax)
    return ring_synthesis(G, tabs["ring_start"], tabs["nphi"], mmax, phi0, mcut)


def alm2cl(alm, lmax, mmax=None):
    """``alm_ops.alm2cl`` (alm_ops.py:19-32) on m-major alm."""
    mmax = lmax if mmax is None else mmax
    cl = np.zeros(lmax + 1)
    for m in range(mmax + 1):
        i0 = mmajor_index(m, m, lmax)
        p = np.abs(alm[i0:i0 + lmax - m + 1]) ** 2
        cl[m:] += p if m == 0 else 2.0 * p
    return cl / (2.0 * np.arange(lmax + 1) + 1.0)


# ---------------------------------------------------------------------------
# adjoints (PyTorch complex-gradient convention: grad = dL/dRe + i dL/dIm)
# ---------------------------------------------------------------------------

def alm2map_adjoint(gmap, tabs, lmax, mmax=None, phi0=None, mcut=None):
    """Adjoint of ``alm2map`` w.r.t. (Re a, Im a): c_m * unweighted analysis.

    For f = S(a), dL/da_lm = c_m sum_p g_p lambda_lm(x_p) e^{-i m phi_p},
    c_0 = 1 (and Im part 0 for m = 0), c_{m>0} = 2.
    """
    mmax = lmax if mmax is None else mmax
    nphi = tabs["nphi"].astype(np.float64)
    # ring_phase divides by nphi; undo so the sum is unweighted
    F = ring_phase(gmap, tabs["ring_start"], tabs["nphi"], mmax, phi0, mcut) * nphi[:, None]
    a = legendre_analysis(F, tabs["theta_centers"], np.ones(len(nphi)), lmax, mmax)
    for m in range(mmax + 1):
        i0 = mmajor_index(m, m, lmax)
        if m == 0:
            a[i0:i0 + lmax + 1] = a[i0:i0 + lmax + 1].real
        else:
            a[i0:i0 + lmax - m + 1] *= 2.0
    return a


def map2alm_adjoint(galm, tabs, lmax, mmax=None, weights=None, phi0=None, mcut=None):
    """Adjoint of ``map2alm`` w.r.t. the real map, for a complex cotangent.

    L depends on a = A f through Re/Im; dL/df_p = Re sum_lm conj(g_lm) dA_lm/df_p
    = omega_p * S(g')_p with omega_p = 2 pi w_r / nphi_r and
    g'_l0 = g_l0, g'_lm = g_lm / 2 (m > 0).
    """
    mmax = lmax if mmax is None else mmax
    w = tabs["weights_dz"] if weights is None else weights
    g = np.asarray(galm, dtype=np.complex128).copy()
    for m in range(1, mmax + 1):
        i0 = mmajor_index(m, m, lmax)
        g[i0:i0 + lmax - m + 1] *= 0.5
    f = alm2map(g, tabs, lmax, mmax, phi0, mcut)
    omega = 2.0 * np.pi * np.asarray(w) / tabs["nphi"].astype(np.float64)
    rs = tabs["ring_start"]
    for r in range(tabs["Nr"]):
        f[rs[r]:rs[r + 1]] *= omega[r]
    return f


# ---------------------------------------------------------------------------
# error metrics
# ---------------------------------------------------------------------------

def alm_errors(alm, ref, lmax, mmax=None):
    """(eps_max, eps_rms): max_lm |da|/sqrt(C_l^ref) and ||da||/||a_ref||."""
    mmax = lmax if mmax is None else mmax
    cl = alm2cl(ref, lmax, mmax)
    d = np.asarray(alm) - np.asarray(ref)
    emax = 0.0
    for m in range(mmax + 1):
        i0 = mmajor_index(m, m, lmax)
        ls = np.arange(m, lmax + 1)
        c = cl[ls]
        ok = c > 0
        if ok.any():
            emax = max(emax, float(np.max(np.abs(d[i0:i0 + lmax - m + 1][ok]) / np.sqrt(c[ok]))))
    erms = float(np.linalg.norm(d) / max(np.linalg.norm(ref), 1e-300))
    return emax, erms


def synalm(cl, lmax, mmax=None, seed=0, rng=None):
    """Gaussian m-major alm: a_l0 ~ N(0, C_l) real, Re/Im a_lm ~ N(0, C_l/2).

    Same distribution as ``alm_ops.synalm`` (alm_ops.py:34-54), drawn with
    ``numpy.random.default_rng`` (the CuPy stream is not reproducible here).
    """
    mmax = lmax if mmax is None else mmax
    rng = np.random.default_rng(seed) if rng is None else rng
    cl = np.asarray(cl, dtype=np.float64)
    alm = np.zeros(nalm(lmax, mmax), dtype=np.complex128)
    for m in range(mmax + 1):
        i0 = mmajor_index(m, m, lmax)
        c = cl[m:lmax + 1]
        n = lmax - m + 1
        if m == 0:
            alm[i0:i0 + n] = rng.standard_normal(n) * np.sqrt(c)
        else:
            s = np.sqrt(c / 2.0)
            alm[i0:i0 + n] = rng.standard_normal(n) * s + 1j * rng.standard_normal(n) * s
    return alm
