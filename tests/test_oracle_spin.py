"""Pins of the spin-weighted oracle (oracle/spin.py) on the CPU.

The reference's spin path cannot be used as a checker (its Wigner kernel,
wigner_d_ms.cu:31-47, has wrong exponents and its synthesis flips B,
spin.py:84), so the oracle is pinned by mathematics:
  - the Goldberg et al. (1967) sum evaluated in mpmath,
  - the closed form 2Y_20 = 1/4 sqrt(15 / 2 pi) sin^2 theta,
  - the spin-0 limit against the independent scalar oracle,
  - exact round trips on Gauss-Legendre grids (s = 1, 2),
  - the mirror-ring parity rule the GPU kernels rely on,
  - adjoint dot products and central differences.
"""
import math

import mpmath
import numpy as np
import pytest

from oracle import geometry as OG
from oracle import sht as OS
from oracle import spin as OSp


def _lam_table(s, lmax, mmax, theta):
    out = {}
    for m, l0, lp, lm in OSp.spin_lambda_all(s, lmax, mmax, np.atleast_1d(theta)):
        for j in range(lp.shape[0]):
            out[(l0 + j, m)] = (lp[j, 0], lm[j, 0])
    return out


@pytest.mark.parametrize("s", [1, 2, 3])
@pytest.mark.parametrize("theta", [0.05, 0.7, 1.9, 3.0])
def test_spin_lambda_vs_goldberg_mpmath(s, theta):
    mpmath.mp.dps = 40
    lmax = 40
    tab = _lam_table(s, lmax, lmax, theta)
    worst = 0.0
    for (l, m), (vp, vm) in tab.items():
        if (l + m) % 3:        # a third of the table keeps the test fast
            continue
        gp = float(OSp.goldberg_lambda(s, l, m, theta))
        gm = float(OSp.goldberg_lambda(-s, l, m, theta))
        scale = math.sqrt((2 * l + 1) / (4 * math.pi))
        worst = max(worst, abs(vp - gp) / scale, abs(vm - gm) / scale)
    assert worst < 1e-12


def test_spin2_y20_closed_form():
    th = np.linspace(0.01, 3.13, 17)
    for t in th:
        tab = _lam_table(2, 4, 2, t)
        want = 0.25 * math.sqrt(15.0 / (2 * math.pi)) * math.sin(t) ** 2
        assert abs(tab[(2, 0)][0] - want) < 1e-14
        assert abs(tab[(2, 0)][1] - want) < 1e-14      # (-2)Y_20 = 2Y_20


def test_spin0_limit_is_scalar():
    nside, lmax = 8, 20
    t = OG.ring_tables(nside)
    rng = np.random.default_rng(1)
    q, u = rng.standard_normal(t["npix"]), rng.standard_normal(t["npix"])
    E, B = OSp.map2alm_spin(q, u, t, lmax, spin=0)
    assert np.max(np.abs(E + OS.map2alm(q, t, lmax))) < 1e-13
    assert np.max(np.abs(B + OS.map2alm(u, t, lmax))) < 1e-13


def _gl_tabs(lmax):
    x, w = np.polynomial.legendre.leggauss(lmax + 1)
    x, w = x[::-1].copy(), w[::-1].copy()
    nphi = np.full(lmax + 1, 2 * lmax + 2, dtype=np.int32)
    rs = np.zeros(lmax + 2, dtype=np.int64)
    rs[1:] = np.cumsum(nphi)
    return {"Nr": lmax + 1, "npix": int(rs[-1]), "nphi": nphi, "ring_start": rs,
            "theta_centers": np.arccos(x), "weights_dz": w}


@pytest.mark.parametrize("s", [1, 2])
@pytest.mark.parametrize("lmax", [6, 23, 40])
def test_spin_gauss_legendre_round_trip(s, lmax):
    t = _gl_tabs(lmax)
    E, B = OSp.synalm_eb(lmax, spin=s, seed=lmax + s)
    q, u = OSp.alm2map_spin(E, B, t, lmax, spin=s)
    E2, B2 = OSp.map2alm_spin(q, u, t, lmax, spin=s)
    assert np.max(np.abs(E2 - E)) < 1e-12
    assert np.max(np.abs(B2 - B)) < 1e-12


def test_pure_e_has_no_b():
    lmax = 30
    t = _gl_tabs(lmax)
    E, _ = OSp.synalm_eb(lmax, spin=2, seed=5)
    q, u = OSp.alm2map_spin(E, np.zeros_like(E), t, lmax, spin=2)
    _, B2 = OSp.map2alm_spin(q, u, t, lmax, spin=2)
    assert np.max(np.abs(B2)) < 1e-12 * np.max(np.abs(E))


@pytest.mark.parametrize("s", [1, 2])
def test_mirror_parity_rule(s):
    """lambda_+(pi - th) = (-1)^q lambda_+(th), lambda_-(pi - th) = -(-1)^q lambda_-(th),
    q = (l - m + s) mod 2: the north/south pairing of the GPU kernels."""
    lmax = 30
    th = np.array([0.3, 1.1])
    a = {m: (l0, lp, lm) for m, l0, lp, lm in OSp.lambda_pm_all(s, lmax, lmax, th)}
    b = {m: (l0, lp, lm) for m, l0, lp, lm in OSp.lambda_pm_all(s, lmax, lmax, math.pi - th)}
    for m in a:
        l0, lp, lm = a[m]
        _, lp2, lm2 = b[m]
        sg = np.array([(-1.0) ** ((l - m + s) % 2) for l in range(l0, lmax + 1)])[:, None]
        assert np.max(np.abs(lp2 - sg * lp)) < 1e-12
        assert np.max(np.abs(lm2 + sg * lm)) < 1e-12


def test_spin_adjoint_dot_products():
    nside, lmax = 8, 20
    t = OG.ring_tables(nside)
    rng = np.random.default_rng(2)
    E, B = OSp.synalm_eb(lmax, spin=2, seed=1)
    gq, gu = rng.standard_normal(t["npix"]), rng.standard_normal(t["npix"])
    q, u = OSp.alm2map_spin(E, B, t, lmax, spin=2)
    aE, aB = OSp.alm2map_spin_adjoint(gq, gu, t, lmax, spin=2)
    lhs = q @ gq + u @ gu
    rhs = np.sum((np.conj(aE) * E).real + (np.conj(aB) * B).real)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    gE, gB = OSp.synalm_eb(lmax, spin=2, seed=3)
    E2, B2 = OSp.map2alm_spin(gq, gu, t, lmax, spin=2)
    lhs = np.sum((np.conj(gE) * E2).real + (np.conj(gB) * B2).real)
    q3, u3 = OSp.map2alm_spin_adjoint(gE, gB, t, lmax, spin=2)
    assert abs(lhs - (q3 @ gq + u3 @ gu)) < 1e-12 * abs(lhs)


def test_spin_adjoint_central_differences():
    nside, lmax = 4, 10
    t = OG.ring_tables(nside)
    rng = np.random.default_rng(3)
    E, B = OSp.synalm_eb(lmax, spin=2, seed=4)
    dq, du = rng.standard_normal(t["npix"]), rng.standard_normal(t["npix"])

    def loss(e, b):
        q, u = OSp.alm2map_spin(e, b, t, lmax, spin=2)
        return float((q - dq) @ (q - dq) + (u - du) @ (u - du))
    q, u = OSp.alm2map_spin(E, B, t, lmax, spin=2)
    gE, gB = OSp.alm2map_spin_adjoint(2 * (q - dq), 2 * (u - du), t, lmax, spin=2)
    h = 1e-6
    for k in rng.choice(OS.nalm(lmax), 6, replace=False):
        for unit in (1.0, 1j):
            for which in (0, 1):
                e = np.zeros_like(E)
                e[k] = unit * h
                args_p = (E + e, B) if which == 0 else (E, B + e)
                args_m = (E - e, B) if which == 0 else (E, B - e)
                fd = (loss(*args_p) - loss(*args_m)) / (2 * h)
                g = (gE if which == 0 else gB)[k]
                an = g.real if unit == 1.0 else g.imag
                assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))
