"""Pins of the CPU oracle (CPU only).

The reference ships no tests, fixtures or golden vectors, so the oracle is
pinned by: literal restatements of the reference's own formulas
(sht_fallback.py:36-65, sht_raw.py:30-57, isht.py:5-26, legendre.cu:2-35)
over their valid domain; closed forms; scipy / mpmath spherical harmonics;
Gauss-Legendre quadrature exactness; brute-force sums over Y_lm; adjoint
dot-product identities; and the committed golden fixtures.
"""
import math
import os

import mpmath
import numpy as np
import pytest
import scipy.special as sp

from oracle import geometry as OG
from oracle import legendre as OLg
from oracle import literal as OL
from oracle import sht as OS

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------------------
# geometry (ring_tables_healpix.cu:12-51)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("nside", [1, 2, 4, 8, 64, 2048])
def test_geometry_closed_form(nside):
    t = OG.ring_tables(nside)
    nr = 4 * nside - 1
    assert t["Nr"] == nr and t["npix"] == 12 * nside * nside
    assert abs(t["weights_dz"].sum() - 2.0) < 1e-12
    assert abs(t["weights_area"].sum() - 2.0) < 1e-12
    # mirror pairs
    assert np.array_equal(t["nphi"], t["nphi"][::-1])
    assert np.allclose(t["z_centers"], -t["z_centers"][::-1], atol=0, rtol=0) or \
        np.max(np.abs(t["z_centers"] + t["z_centers"][::-1])) < 1e-15
    assert np.allclose(t["weights_dz"], t["weights_dz"][::-1], rtol=0, atol=1e-15)
    # explicit rings
    for i in (1, nside - 1, nside, 2 * nside, 3 * nside, 4 * nside - 1):
        if i < 1 or i > nr:
            continue
        r = i - 1
        if i < nside:
            z, n = 1 - i * i / (3 * nside * nside), 4 * i
        elif i <= 3 * nside:
            z, n = (2 * nside - i) * 2 / (3 * nside), 4 * nside
        else:
            ip = 4 * nside - i
            z, n = -(1 - ip * ip / (3 * nside * nside)), 4 * ip
        assert abs(t["z_centers"][r] - z) < 1e-15 and t["nphi"][r] == n
    assert t["ring_start"][0] == 0 and np.all(np.diff(t["ring_start"]) == t["nphi"])
    # the reference phi_center0 has the opposite equatorial parity to HEALPix
    if nside >= 2:
        r = nside - 1          # ring i = nside
        dphi = 2 * math.pi / t["nphi"][r]
        assert abs(t["phi_center0"][r] - dphi) < 1e-15
        assert abs(t["phi0_healpix"][r] - 0.5 * dphi) < 1e-15


@pytest.mark.parametrize("nside", [1, 4, 64, 2048])
def test_product_geometry_matches_oracle(nside):
    from hellblast_amd.ring_tables import healpix_geometry
    g = healpix_geometry(nside)
    t = OG.ring_tables(nside)
    for k in ("z_centers", "theta_centers", "phi_center0", "phi0_healpix", "z_edges"):
        assert np.array_equal(g[k], t[k]), k
    for k in ("nphi", "ring_start"):
        assert np.array_equal(g[k], t[k]), k


# ---------------------------------------------------------------------------
# lambda_lm
# ---------------------------------------------------------------------------
def _lam(lmax, mmax, theta):
    th = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    return dict(OLg.lambda_all(lmax, mmax, np.cos(th), np.sin(th)))


def test_lambda_closed_forms():
    th = np.linspace(0.01, math.pi - 0.01, 37)
    lam = _lam(2, 2, th)
    c, s = np.cos(th), np.sin(th)
    assert np.allclose(lam[0][0], 1 / math.sqrt(4 * math.pi), rtol=1e-15, atol=0)
    assert np.allclose(lam[0][1], math.sqrt(3 / (4 * math.pi)) * c, rtol=1e-14, atol=1e-16)
    assert np.allclose(lam[1][0], -math.sqrt(3 / (8 * math.pi)) * s, rtol=1e-14, atol=0)
    assert np.allclose(lam[2][0], 0.25 * math.sqrt(15 / (2 * math.pi)) * s * s, rtol=1e-14, atol=0)
    assert np.allclose(lam[0][2], math.sqrt(5 / (16 * math.pi)) * (3 * c * c - 1), rtol=1e-13, atol=1e-15)


def test_lambda_vs_scipy():
    lmax = 200
    th = np.array([0.013, 0.3, 1.0, 1.5707963267948966, 2.2, 3.1])
    lam = _lam(lmax, lmax, th)
    worst = 0.0
    for m in range(0, lmax + 1, 7):
        for l in range(m, lmax + 1, 5):
            ref = np.real(sp.sph_harm_y(l, m, th, 0.0))
            worst = max(worst, float(np.max(np.abs(lam[m][l - m] - ref))))
    assert worst < 1e-12


@pytest.mark.parametrize("l,m,theta", [(4095, 0, 4e-4), (4095, 1, 0.01), (4095, 100, 0.05),
                                        (4095, 2000, 0.6), (4095, 4095, math.pi / 2), (3000, 2900, 1.3),
                                        (4095, 3000, 0.9)])
def test_lambda_vs_mpmath_high_l(l, m, theta):
    """lambda_lm at lmax 4095 near the poles, where sin^m underflows FP64."""
    mpmath.mp.dps = 40
    # evaluate at the exact value of the double cos(theta): the recursion's own
    # error is pinned, not the O(l * ulp / sin theta) rounding of the input
    ct = math.cos(theta)
    x = mpmath.mpf(ct)
    # normalised associated Legendre with Condon-Shortley phase
    norm = mpmath.sqrt((2 * l + 1) / (4 * mpmath.pi) * mpmath.factorial(l - m) / mpmath.factorial(l + m))
    ref = float(norm * mpmath.legenp(l, m, x, type=2))
    st = float(mpmath.sqrt(1 - x * x))
    lam = None
    for mm, mant, ex in OLg.lambda_mm_scaled(m, np.array([st])):
        if mm == m:
            lam = OLg.lambda_block(m, l, np.array([ct]), mant, ex)[l - m, 0]
    scale = math.sqrt((2 * l + 1) / (4 * math.pi))
    assert abs(lam - ref) <= 1e-10 * scale + 1e-300


def test_lambda_matches_literal_legendre_cu():
    """legendre.cu recursion x exact N_lm (sht_raw.py:26-28) over its valid domain."""
    lmax = 85
    th = np.linspace(0.02, math.pi - 0.02, 23)
    lam = _lam(lmax, lmax, th)
    x = np.cos(th)
    worst = 0.0
    for m in range(lmax + 1):
        P = OLg.plm_literal(x, lmax, m)
        for l in range(m, lmax + 1):
            ref = OLg.alm_norm_exact(l, m) * P[:, l]
            worst = max(worst, float(np.max(np.abs(lam[m][l - m] - ref))))
    assert worst < 1e-12


# ---------------------------------------------------------------------------
# transforms vs the literal reference formulas
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("nside,lmax", [(8, 20), (16, 30)])
def test_oracle_matches_literal_reference(nside, lmax):
    t = OG.ring_tables(nside)
    f = np.random.default_rng(1).standard_normal(t["npix"])
    rings = OG.flat_to_rings(f, t)
    a = OS.map2alm(f, t, lmax)
    fb = OL.sht_fallback_map2alm(rings, t["theta_centers"], t["nphi"], t["weights_dz"], lmax)
    assert np.max(np.abs(OS.mmajor_to_ref(a, lmax) - fb)) < 1e-13 * np.max(np.abs(fb))
    # sht_raw: no 2 pi and m <= min(mmax, max nphi // 2) (sht_raw.py:43)
    raw = OL.sht_raw_map2alm(rings, t["theta_centers"], t["nphi"], t["weights_dz"], lmax)
    mtop = min(lmax, int(t["nphi"].max()) // 2)
    keep = np.zeros((lmax + 1) ** 2, dtype=bool)
    for l in range(lmax + 1):
        for m in range(min(l, mtop) + 1):
            keep[l * (l + 1) + m] = True
    d = OS.mmajor_to_ref(a, lmax) - 2 * math.pi * raw
    assert np.max(np.abs(d[keep])) < 1e-13 * np.max(np.abs(fb))
    assert np.all(raw[~keep] == 0)
    alm = OS.synalm(np.ones(lmax + 1), lmax, seed=2)
    fo = OS.alm2map(alm, t, lmax)
    fl = OL.isht_alm2map(OS.mmajor_to_ref(alm, lmax), t["theta_centers"], t["nphi"], lmax)
    for r in range(t["Nr"]):
        if t["nphi"][r] > 2 * lmax:
            rs = t["ring_start"][r]
            assert np.max(np.abs(fo[rs:rs + t["nphi"][r]] - fl[r])) < 1e-13


def _gl_tabs(lmax):
    x, w = np.polynomial.legendre.leggauss(lmax + 1)
    x = x[::-1].copy()
    w = w[::-1].copy()
    nphi = np.full(lmax + 1, 2 * lmax + 2, dtype=np.int32)
    rs = np.zeros(lmax + 2, dtype=np.int64)
    rs[1:] = np.cumsum(nphi)
    return {"Nr": lmax + 1, "npix": int(rs[-1]), "nphi": nphi, "ring_start": rs,
            "theta_centers": np.arccos(x), "weights_dz": w}


@pytest.mark.parametrize("lmax", [8, 31, 64])
def test_gauss_legendre_exactness(lmax):
    t = _gl_tabs(lmax)
    alm = OS.synalm(np.ones(lmax + 1), lmax, seed=3)
    for m in range(lmax + 1):                      # a_l0 of a real map is real
        pass
    alm[:lmax + 1] = alm[:lmax + 1].real
    back = OS.map2alm(OS.alm2map(alm, t, lmax), t, lmax)
    assert np.max(np.abs(back - alm)) < 1e-12


@pytest.mark.parametrize("alias", ["reference", "fold"])
def test_synthesis_brute_force(alias):
    """alm2map == sum_{|m| < mcut_r} a_lm Y_lm(theta, phi) at every pixel."""
    nside, lmax = 4, 20
    t = OG.ring_tables(nside)
    alm = OS.synalm(np.ones(lmax + 1), lmax, seed=4)
    phi0 = t["phi0_healpix"]
    mcut = np.full(t["Nr"], lmax + 1) if alias == "fold" else None
    f = OS.alm2map(alm, t, lmax, phi0=phi0, mcut=mcut)
    for r in range(t["Nr"]):
        n = t["nphi"][r]
        cut = lmax + 1 if alias == "fold" else n
        phis = phi0[r] + 2 * math.pi * np.arange(n) / n
        val = np.zeros(n)
        for m in range(0, min(lmax, cut - 1) + 1):
            for l in range(m, lmax + 1):
                y = sp.sph_harm_y(l, m, t["theta_centers"][r], phis)
                a = alm[OS.mmajor_index(l, m, lmax)]
                val += (a * y).real if m == 0 else 2 * (a * y).real
        rs = t["ring_start"][r]
        assert np.max(np.abs(f[rs:rs + n] - val)) < 1e-12


def test_analysis_alias_rules_differ_only_on_small_rings():
    """Reference rule (m < nphi) vs periodic fold: equal when every nphi > mmax."""
    nside, lmax = 8, 30
    t = OG.ring_tables(nside)
    f = np.random.default_rng(5).standard_normal(t["npix"])
    a_ref = OS.map2alm(f, t, lmax)
    a_fold = OS.map2alm(f, t, lmax, mcut=np.full(t["Nr"], lmax + 1))
    assert np.max(np.abs(a_ref - a_fold)) > 1e-6           # polar rings nphi=4i <= 30 fold
    small = OS.map2alm(f, t, 3)
    small_fold = OS.map2alm(f, t, 3, mcut=np.full(t["Nr"], 4))
    assert np.max(np.abs(small - small_fold)) < 1e-15


def test_phi0_rotation_closed_form():
    """Shifting every ring by delta multiplies a_lm by e^{-i m delta}."""
    nside, lmax = 8, 16
    t = OG.ring_tables(nside)
    f = np.random.default_rng(6).standard_normal(t["npix"])
    d = 0.123
    a0 = OS.map2alm(f, t, lmax)
    a1 = OS.map2alm(f, t, lmax, phi0=np.full(t["Nr"], d))
    for m in range(lmax + 1):
        i0 = OS.mmajor_index(m, m, lmax)
        sl = slice(i0, i0 + lmax - m + 1)
        assert np.max(np.abs(a1[sl] - a0[sl] * np.exp(-1j * m * d))) < 1e-13


# ---------------------------------------------------------------------------
# adjoints, layouts, alm2cl
# ---------------------------------------------------------------------------
def test_adjoint_dot_products():
    nside, lmax = 8, 20
    t = OG.ring_tables(nside)
    rng = np.random.default_rng(7)
    a = OS.synalm(np.ones(lmax + 1), lmax, seed=8)
    g = rng.standard_normal(t["npix"])
    lhs = np.dot(OS.alm2map(a, t, lmax), g)
    rhs = np.sum((np.conj(OS.alm2map_adjoint(g, t, lmax)) * a).real)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    ga = OS.synalm(np.ones(lmax + 1), lmax, seed=9)
    lhs = np.sum((np.conj(ga) * OS.map2alm(g, t, lmax)).real)
    rhs = np.dot(OS.map2alm_adjoint(ga, t, lmax), g)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_adjoint_central_differences():
    nside, lmax = 4, 10
    t = OG.ring_tables(nside)
    rng = np.random.default_rng(10)
    a = OS.synalm(np.ones(lmax + 1), lmax, seed=11)
    d = rng.standard_normal(t["npix"])

    def loss(aa):
        r = OS.alm2map(aa, t, lmax) - d
        return float(np.dot(r, r))
    grad = OS.alm2map_adjoint(2 * (OS.alm2map(a, t, lmax) - d), t, lmax)
    h = 1e-6
    for k in rng.choice(OS.nalm(lmax), 10, replace=False):
        for unit in (1.0, 1j):
            e = np.zeros_like(a)
            e[k] = unit * h
            fd = (loss(a + e) - loss(a - e)) / (2 * h)
            an = grad[k].real if unit == 1.0 else grad[k].imag
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))


def test_layout_converters():
    lmax, mmax = 12, 9
    a = OS.synalm(np.ones(lmax + 1), lmax, mmax, seed=12)
    ref = OS.mmajor_to_ref(a, lmax, mmax)
    assert np.array_equal(OS.ref_to_mmajor(ref, lmax, mmax), a)
    for l in range(lmax + 1):
        for m in range(min(l, mmax) + 1):
            assert ref[l * (l + 1) + m] == a[m * (2 * lmax + 1 - m) // 2 + l]


def test_alm2cl_closed_form():
    lmax = 6
    a = np.zeros(OS.nalm(lmax), dtype=complex)
    a[OS.mmajor_index(4, 0, lmax)] = 3.0
    a[OS.mmajor_index(4, 2, lmax)] = 1 + 2j
    cl = OS.alm2cl(a, lmax)
    assert abs(cl[4] - (9 + 2 * 5) / 9) < 1e-15 and np.count_nonzero(cl) == 1


# ---------------------------------------------------------------------------
# golden fixtures (written by tests/golden/make_golden.py from the oracle)
# ---------------------------------------------------------------------------
def test_golden_fixtures_regenerate():
    path = os.path.join(GOLDEN, "scalar_small.npz")
    if not os.path.exists(path):
        pytest.skip("fixtures not generated")
    from tests.golden.make_golden import scalar_cases
    with np.load(path, allow_pickle=False) as z:
        for name, arr in scalar_cases().items():
            assert np.allclose(z[name], arr, rtol=0, atol=1e-13 * max(1.0, np.max(np.abs(arr)))), name
