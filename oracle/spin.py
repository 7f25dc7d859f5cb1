"""Spin-weighted SHT (spin s = 1, 2) on the CPU, FP64, healpy E/B convention.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Reference: spin.py:32-115 (map2alm_spin_gpu / eb_from_qu_gpu /
qu_from_eb_gpu) and spin1.py:5-69.  The reference's Wigner kernel has wrong
exponents (wigner_d_ms.cu:31-47) and its synthesis flips the sign of B
(spin.py:76-77 vs 84-85), so only its conventions are followed:
    (Q + iU) = sum s_a_lm  sY_lm,   (Q - iU) = sum -s_a_lm -sY_lm,
    s_a = -(E + iB),  -s_a = -(-1)^s (E - iB),
with sY_lm(theta, phi) = s_lambda_lm(theta) e^{i m phi} defined by Goldberg et
al. (1967):
    s_lambda_lm = (-1)^m sqrt((l+m)!(l-m)!(2l+1) / (4 pi (l+s)!(l-s)!))
                  sin^{2l}(theta/2) sum_r C(l-s, r) C(l+s, r+s-m)
                  (-1)^{l-r-s} cot^{2r+s-m}(theta/2)
(2Y_20 = 1/4 sqrt(15/(2 pi)) sin^2 theta).  With
lambda_pm = (s_lambda +- (-1)^s (-s)_lambda)/2 the transforms are
    E_lm = -2 pi sum_r w_r [lambda_+ F_Q + i lambda_- F_U]
    B_lm = -2 pi sum_r w_r [lambda_+ F_U - i lambda_- F_Q]
    G_Q(m) = -sum_l (lambda_+ E + i lambda_- B)
    G_U(m) = -sum_l (lambda_+ B - i lambda_- E)
followed by the scalar ring stage (oracle/sht.py) for Q and U.

Here s_lambda and (-s)_lambda are generated by two separate three-term
recursions in l (the Wigner-d recursion), each seeded by closed forms at
l0 = max(m, s); the GPU uses a coupled recursion for lambda_+- instead.
"""
import math

import numpy as np

from .sht import mmajor_index, nalm, ring_phase, ring_synthesis

_BIG = 2.0 ** 600


def goldberg_lambda(s, l, m, theta, mp=None):
    """s_lambda_lm(theta) by the explicit Goldberg sum (mpmath, any precision)."""
    import mpmath
    mp = mpmath if mp is None else mp
    th = mp.mpf(theta) / 2
    if l < max(abs(m), abs(s)):
        return mp.mpf(0)
    pref = (-1) ** m * mp.sqrt(mp.factorial(l + m) * mp.factorial(l - m) * (2 * l + 1)
                               / (4 * mp.pi * mp.factorial(l + s) * mp.factorial(l - s)))
    tot = mp.mpf(0)
    for r in range(0, l - s + 1):
        k = r + s - m
        if k < 0 or k > l + s:
            continue
        tot += mp.binomial(l - s, r) * mp.binomial(l + s, k) * (-1) ** (l - r - s) * mp.cot(th) ** (2 * r + s - m)
    return pref * mp.sin(th) ** (2 * l) * tot


def _seed_small_m(s, m, theta):
    """s_lambda_{l0,m} at l0 = |s| > m (closed form of the Goldberg sum)."""
    a = abs(s)
    sh, ch = np.sin(theta / 2), np.cos(theta / 2)
    base = math.sqrt(math.factorial(a + m) * math.factorial(a - m) * (2 * a + 1)
                     / (4 * math.pi * math.factorial(2 * a)))
    if s > 0:
        return (-1) ** m * base * math.comb(2 * a, a - m) * sh ** (a + m) * ch ** (a - m)
    return (-1) ** a * base * math.comb(2 * a, a + m) * sh ** (a - m) * ch ** (a + m)


def _seed_chain(s, mmax, theta):
    """Yield (m, mant, exp) with s_lambda_mm = mant * 2**exp for m = |s|..mmax.

    s_lambda_{|s||s|} = (-1)^s sqrt((2|s|+1)/(4 pi)) sin^{2|s|}(theta/2) (s > 0) or
    cos^{2|s|}(theta/2) (s < 0); then
    s_lambda_{m+1,m+1} = -sqrt((2m+3)(2m+2)/((m+1+s)(m+1-s))) sin(theta)/2 * s_lambda_mm.
    """
    a = abs(s)
    th = np.asarray(theta, dtype=np.float64)
    half = np.sin(th / 2) if s > 0 else np.cos(th / 2)
    v = np.full(th.shape, (-1) ** a * math.sqrt((2 * a + 1) / (4 * math.pi)))
    e = np.zeros(th.shape, dtype=np.int64)
    for _ in range(2 * a):
        v = v * half
        fm, fe = np.frexp(v)
        v, e = fm, e + fe
    yield a, v.copy(), e.copy()
    st = np.sin(th)
    for m in range(a, mmax):
        ratio = -math.sqrt((2 * m + 3) * (2 * m + 2) / float((m + 1 + s) * (m + 1 - s))) * 0.5
        v = v * ratio * st
        fm, fe = np.frexp(v)
        v, e = fm, e + fe
        yield m + 1, v.copy(), e.copy()


def _recur(s, m, l0, lmax, x, mant, ex):
    """s_lambda_lm for l = l0..lmax via the Wigner-d three-term recursion."""
    n = lmax - l0 + 1
    out = np.zeros((max(n, 0), x.size))
    if n <= 0:
        return out
    p1 = mant.astype(np.float64).copy()
    p2 = np.zeros_like(p1)
    e = ex.astype(np.int64).copy()
    out[0] = np.ldexp(p1, e)
    for l in range(l0, lmax):
        den = math.sqrt(float(((l + 1) ** 2 - m * m) * ((l + 1) ** 2 - s * s)))
        A = (2 * l + 1) * (l + 1) * math.sqrt((2 * l + 3) / (2 * l + 1)) / den
        B = m * s / float(l * (l + 1)) if l > 0 else 0.0
        C = ((l + 1) * math.sqrt(float((l * l - m * m) * (l * l - s * s))) * math.sqrt((2 * l + 3) / (2 * l - 1))
             / (l * den)) if l > 0 else 0.0
        new = A * (x + B) * p1 - C * p2
        big = np.abs(new) > _BIG
        if big.any():
            new = np.where(big, new / _BIG, new)
            p1 = np.where(big, p1 / _BIG, p1)
            e = np.where(big, e + 600, e)
        p2, p1 = p1, new
        out[l + 1 - l0] = np.ldexp(p1, e)
    return out


def spin_lambda_all(s, lmax, mmax, theta):
    """Yield (m, l0, lam_p, lam_m): s_lambda and (-s)_lambda for l = l0..lmax."""
    th = np.asarray(theta, dtype=np.float64)
    x = np.cos(th)
    a = abs(s)
    for m in range(0, min(a, mmax + 1)):
        lp = _recur(s, m, a, lmax, x, *np.frexp(_seed_small_m(s, m, th)))
        lm = _recur(-s, m, a, lmax, x, *np.frexp(_seed_small_m(-s, m, th)))
        yield m, a, lp, lm
    if mmax >= a:
        for (m, mp_, ep), (_, mm_, em) in zip(_seed_chain(s, mmax, th), _seed_chain(-s, mmax, th)):
            yield m, m, _recur(s, m, m, lmax, x, mp_, ep), _recur(-s, m, m, lmax, x, mm_, em)


def lambda_pm_all(s, lmax, mmax, theta):
    """Yield (m, l0, lambda_plus, lambda_minus) with the (-1)^s convention."""
    sg = (-1) ** s
    for m, l0, lp, lm in spin_lambda_all(s, lmax, mmax, theta):
        yield m, l0, 0.5 * (lp + sg * lm), 0.5 * (lp - sg * lm)


def map2alm_spin(q, u, tabs, lmax, mmax=None, spin=2, weights=None, phi0=None, mcut=None):
    """(E, B) m-major from flat Q, U maps (2 pi convention of sht_fallback)."""
    mmax = lmax if mmax is None else mmax
    w = np.asarray(tabs["weights_dz"] if weights is None else weights) * 2.0 * math.pi
    FQ = ring_phase(q, tabs["ring_start"], tabs["nphi"], mmax, phi0, mcut) * w[:, None]
    FU = ring_phase(u, tabs["ring_start"], tabs["nphi"], mmax, phi0, mcut) * w[:, None]
    E = np.zeros(nalm(lmax, mmax), dtype=np.complex128)
    B = np.zeros_like(E)
    for m, l0, lp, lmn in lambda_pm_all(spin, lmax, mmax, tabs["theta_centers"]):
        if lp.shape[0] == 0:
            continue
        i0 = mmajor_index(l0, m, lmax)
        n = lmax - l0 + 1
        E[i0:i0 + n] = -(lp @ FQ[:, m] + 1j * (lmn @ FU[:, m]))
        B[i0:i0 + n] = -(lp @ FU[:, m] - 1j * (lmn @ FQ[:, m]))
    return E, B


def alm2map_spin(E, B, tabs, lmax, mmax=None, spin=2, phi0=None, mcut=None):
    """Flat Q, U maps from m-major (E, B)."""
    mmax = lmax if mmax is None else mmax
    nr = tabs["Nr"]
    GQ = np.zeros((nr, mmax + 1), dtype=np.complex128)
    GU = np.zeros_like(GQ)
    E = np.asarray(E, dtype=np.complex128)
    B = np.asarray(B, dtype=np.complex128)
    for m, l0, lp, lmn in lambda_pm_all(spin, lmax, mmax, tabs["theta_centers"]):
        if lp.shape[0] == 0:
            continue
        i0 = mmajor_index(l0, m, lmax)
        n = lmax - l0 + 1
        e = E[i0:i0 + n]
        b = B[i0:i0 + n]
        GQ[:, m] = -(lp.T @ e + 1j * (lmn.T @ b))
        GU[:, m] = -(lp.T @ b - 1j * (lmn.T @ e))
    q = ring_synthesis(GQ, tabs["ring_start"], tabs["nphi"], mmax, phi0, mcut)
    u = ring_synthesis(GU, tabs["ring_start"], tabs["nphi"], mmax, phi0, mcut)
    return q, u


def _cm_fix(a, lmax, mmax):
    """m = 0 rows real, m > 0 rows doubled (adjoint of the m >= 0 storage)."""
    a = a.copy()
    for m in range(mmax + 1):
        i0 = mmajor_index(m, m, lmax)
        n = lmax - m + 1
        if m == 0:
            a[i0:i0 + n] = a[i0:i0 + n].real
        else:
            a[i0:i0 + n] *= 2.0
    return a


def alm2map_spin_adjoint(gq, gu, tabs, lmax, mmax=None, spin=2, phi0=None, mcut=None):
    """Adjoint of ``alm2map_spin`` (complex-gradient convention): c_m times
    the unweighted analysis -sum_p [lambda_+ g_Q + i lambda_- g_U] e^{-i m phi_p}
    (and likewise for B)."""
    mmax = lmax if mmax is None else mmax
    w = tabs["nphi"].astype(np.float64) / (2.0 * math.pi)
    E, B = map2alm_spin(gq, gu, tabs, lmax, mmax, spin, w, phi0, mcut)
    return _cm_fix(E, lmax, mmax), _cm_fix(B, lmax, mmax)


def map2alm_spin_adjoint(gE, gB, tabs, lmax, mmax=None, spin=2, weights=None, phi0=None, mcut=None):
    """Adjoint of ``map2alm_spin`` w.r.t. the real maps: omega_r times the
    synthesis of (g_E, g_B) with m > 0 rows halved, omega_r = 2 pi w_r / nphi_r."""
    mmax = lmax if mmax is None else mmax
    w = np.asarray(tabs["weights_dz"] if weights is None else weights)
    h = []
    for g in (gE, gB):
        g = np.asarray(g, dtype=np.complex128).copy()
        for m in range(1, mmax + 1):
            i0 = mmajor_index(m, m, lmax)
            g[i0:i0 + lmax - m + 1] *= 0.5
        h.append(g)
    q, u = alm2map_spin(h[0], h[1], tabs, lmax, mmax, spin, phi0, mcut)
    omega = 2.0 * math.pi * w / tabs["nphi"].astype(np.float64)
    rs = tabs["ring_start"]
    for r in range(tabs["Nr"]):
        q[rs[r]:rs[r + 1]] *= omega[r]
        u[rs[r]:rs[r + 1]] *= omega[r]
    return q, u


def synalm_eb(lmax, mmax=None, spin=2, seed=0):
    """Random (E, B) m-major with real m = 0 rows and zeros below l = max(m, spin)."""
    mmax = lmax if mmax is None else mmax
    rng = np.random.default_rng(seed)
    n = nalm(lmax, mmax)
    out = []
    for _ in range(2):
        a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        for m in range(mmax + 1):
            i0 = mmajor_index(m, m, lmax)
            lo = max(m, abs(spin))
            a[i0:i0 + lo - m] = 0.0
            if m == 0:
                a[i0:i0 + lmax + 1] = a[i0:i0 + lmax + 1].real
        out.append(a)
    return out[0], out[1]
