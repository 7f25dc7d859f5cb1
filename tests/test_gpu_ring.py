"""GPU parity of the ring phase stage (hbx_ring_analysis / hbx_ring_synthesis)
at nside 2048, mmax 4095: every FFT size class of the stage runs here (the
4097 equatorial rings of 8192 pixels as 4096-point FFTs, Bluestein polar
rings up to a 8192-point inner FFT, the small polar classes, aliasing for
rings with nphi <= mmax), checked ring by ring against the oracle's NumPy
FFTs (oracle.sht.ring_phase / ring_synthesis).

Bars: FP64 per-ring relative error <= 1e-12; FP32 <= 2e-6 (relative to the
ring's largest coefficient or pixel).
"""
import numpy as np
import pytest
import torch

from oracle import sht as OS

pytestmark = pytest.mark.gpu

NSIDE, LMAX = 2048, 4095
TOL = {"f64": 1e-12, "f32": 2e-6}


def _setup(precision, phi0="healpix"):
    from hellblast_amd.hp_compat import healpix_geom
    from hellblast_amd.plan import get_plan
    geom = healpix_geom(NSIDE, LMAX, "dz", phi0, "reference")
    return geom, get_plan(geom, LMAX, LMAX, precision)


def _phase_flat(plan, phase, nmaps):
    """[map][element] complex view of a phase buffer (maps phase_map_elems apart)."""
    cdt = torch.complex128 if plan.alm_dtype == torch.complex128 else torch.complex64
    csize = 16 if cdt == torch.complex128 else 8
    n = nmaps * plan.phase_map_elems
    return phase[:n * csize].view(cdt).reshape(nmaps, plan.phase_map_elems)


def _phase_get(plan, phase, nmaps, direction):
    """[map][ring][m] coefficients read through the plan's layout index."""
    ix = torch.as_tensor(plan.phase_index(direction).ravel(), device=phase.device)
    v = _phase_flat(plan, phase, nmaps)[:, ix]
    return v.reshape(nmaps, plan.geom.nrings, plan.mmax + 1).cpu().numpy()


def _phase_put(plan, phase, G, direction):
    """Write [map][ring][m] coefficients G through the plan's layout index."""
    nmaps = G.shape[0]
    flat = _phase_flat(plan, phase, nmaps)
    ix = torch.as_tensor(plan.phase_index(direction).ravel(), device=phase.device)
    flat[:, ix] = torch.as_tensor(G.reshape(nmaps, -1)).to(device=phase.device, dtype=flat.dtype)


def _per_ring_rel(got, ref):
    """max over rings of max |got - ref| / max |ref| along the last axis."""
    d = np.max(np.abs(got - ref), axis=-1)
    s = np.maximum(np.max(np.abs(ref), axis=-1), 1e-300)
    return float(np.max(d / s))


@pytest.mark.parametrize("precision,phi0,other", [("f64", "healpix", False), ("f64", "reference", False),
                                                  ("f32", "healpix", False), ("f32", "reference", False),
                                                  ("f64", "healpix", True), ("f32", "healpix", True)])
def test_ring_analysis_all_classes(gpu, precision, phi0, other):
    """other: maps of the other float type (FP32 maps into an FP64 plan, FP64
    maps into an FP32 plan)."""
    print(f"[case] ring analysis nside={NSIDE} mmax={LMAX} {precision} phi0={phi0} other-dtype={other}", flush=True)
    geom, plan = _setup(precision, phi0)
    nmaps = 2
    rng = np.random.default_rng(7)
    rdt = torch.float64 if (precision == "f64") != other else torch.float32
    maps = torch.as_tensor(rng.standard_normal((nmaps, plan.npix)), dtype=rdt, device="cuda")
    phase = plan.phase_buffer(nmaps)
    plan.ring_analysis(maps, phase)
    got_all = _phase_get(plan, phase, nmaps, "analysis")
    f = maps.double().cpu().numpy()
    for b in range(nmaps):
        ref = OS.ring_phase(f[b], geom.ring_start, geom.nphi, LMAX, geom.phi0, geom.mcut) * geom.weight[:, None]
        err = _per_ring_rel(got_all[b], ref)
        print(f"map {b}: per-ring rel err {err:.2e}", flush=True)
        assert err <= TOL[precision]


@pytest.mark.parametrize("precision,other", [("f64", False), ("f32", False), ("f64", True), ("f32", True)])
def test_ring_synthesis_all_classes(gpu, precision, other):
    """other: pixels written in the other float type (an FP64 plan's FP32
    output is held to FP32 rounding, 1e-7)."""
    print(f"[case] ring synthesis nside={NSIDE} mmax={LMAX} {precision} other-dtype={other}", flush=True)
    geom, plan = _setup(precision)
    nmaps = 2
    rng = np.random.default_rng(8)
    nr = plan.geom.nrings
    G = rng.standard_normal((nmaps, nr, LMAX + 1)) + 1j * rng.standard_normal((nmaps, nr, LMAX + 1))
    phase = plan.phase_buffer(nmaps)
    _phase_put(plan, phase, G, "synthesis")
    rdt = torch.float64 if (precision == "f64") != other else torch.float32
    tol = 1e-7 if (precision == "f64" and other) else TOL[precision]
    got = plan.ring_synthesis(phase, nmaps, dtype=rdt).double().cpu().numpy()
    Gr = G
    if precision == "f32":
        Gr = Gr.astype(np.complex64).astype(np.complex128)     # the values the kernel reads
    for b in range(nmaps):
        ref = OS.ring_synthesis(Gr[b], geom.ring_start, geom.nphi, LMAX, geom.phi0, geom.mcut)
        worst = 0.0
        for r in range(nr):
            s, e = geom.ring_start[r], geom.ring_start[r + 1]
            worst = max(worst, float(np.max(np.abs(got[b, s:e] - ref[s:e])) / np.max(np.abs(ref[s:e]))))
        print(f"map {b}: per-ring rel err {worst:.2e}", flush=True)
        assert worst <= tol


def test_f32_gate_full_size(gpu):
    """FP32 map2alm at nside 2048, lmax 4095 within eps_max <= 1e-5 of FP64 on
    a red-spectrum map (C_l ~ 1/(l+1)^2): the top coefficients (l = m = 4095)
    are the most sensitive to FP32 ring-FFT errors that are coherent across
    the equatorial rings."""
    print("[case] f32 gate nside=2048 lmax=4095", flush=True)
    import math
    from hellblast_amd.hp_compat import _plan
    from hellblast_amd.plan import nalm
    dev = torch.device("cuda", 0)
    p64 = _plan(NSIDE, LMAX, LMAX, "f64", "dz", "reference", "reference")
    p32 = _plan(NSIDE, LMAX, LMAX, "f32", "dz", "reference", "reference")
    n = nalm(LMAX)
    ms = torch.repeat_interleave(torch.arange(LMAX + 1, device=dev), torch.arange(LMAX + 1, 0, -1, device=dev))
    ls = torch.arange(n, device=dev) - torch.repeat_interleave(
        torch.cumsum(torch.arange(LMAX + 1, 0, -1, device=dev), 0) - torch.arange(LMAX + 1, 0, -1, device=dev),
        torch.arange(LMAX + 1, 0, -1, device=dev)) + ms
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    sd = 1.0 / (ls.double() + 1.0)
    s = torch.where(ms == 0, sd, sd / math.sqrt(2.0))
    re = torch.randn(n, generator=g, device=dev, dtype=torch.float64)
    im = torch.randn(n, generator=g, device=dev, dtype=torch.float64)
    alm = torch.complex(re * s, torch.where(ms == 0, torch.zeros_like(im), im * s))
    m_in = p64.alm2map(alm[None])[0].to(torch.float32)
    ref = p64.map2alm(m_in.double()[None])[0]
    p = ref.real ** 2 + ref.imag ** 2
    cw = torch.where(ms == 0, p, 2.0 * p)
    cl = torch.zeros(LMAX + 1, dtype=torch.float64, device=dev).index_add_(0, ls, cw)
    cl /= 2.0 * torch.arange(LMAX + 1, device=dev, dtype=torch.float64) + 1.0
    for name, a in (("single", p32.map2alm(m_in[None])[0]), ("batch", p32.map2alm(m_in[None].repeat(16, 1))[0])):
        e = float(((a.to(torch.complex128) - ref).abs() / cl[ls].sqrt()).max())
        print(f"{name}: eps_max {e:.2e}", flush=True)
        assert e <= 1e-5


def test_ring_stage_nside4096_f32(gpu):
    """FP32 ring analysis and synthesis at nside 4096 (mmax 2047): the
    16384-point class (equatorial rings of 16384 pixels, 1024 threads) and
    Bluestein rings up to a 16384-point inner FFT, against NumPy."""
    print("[case] ring stage nside=4096 f32", flush=True)
    from hellblast_amd.hp_compat import healpix_geom
    from hellblast_amd.plan import get_plan
    nside, mmax = 4096, 2047
    geom = healpix_geom(nside, mmax, "dz", "healpix", "reference")
    plan = get_plan(geom, mmax, mmax, "f32")
    rng = np.random.default_rng(9)
    f = rng.standard_normal(plan.npix)
    phase = plan.phase_buffer(1)
    plan.ring_analysis(torch.as_tensor(f, dtype=torch.float32, device="cuda")[None], phase)
    nr = plan.geom.nrings
    got = _phase_get(plan, phase, 1, "analysis")[0]
    f32 = f.astype(np.float32).astype(np.float64)
    ref = OS.ring_phase(f32, geom.ring_start, geom.nphi, mmax, geom.phi0, geom.mcut) * geom.weight[:, None]
    err = _per_ring_rel(got, ref)
    print(f"analysis per-ring rel err {err:.2e}", flush=True)
    assert err <= TOL["f32"]
    G = (rng.standard_normal((nr, mmax + 1)) + 1j * rng.standard_normal((nr, mmax + 1))).astype(np.complex64)
    _phase_put(plan, phase, G[None], "synthesis")
    maps = plan.ring_synthesis(phase, 1, dtype=torch.float32)[0].double().cpu().numpy()
    refm = OS.ring_synthesis(G.astype(np.complex128), geom.ring_start, geom.nphi, mmax, geom.phi0, geom.mcut)
    worst = 0.0
    for r in range(0, nr, 7):
        s, e = geom.ring_start[r], geom.ring_start[r + 1]
        worst = max(worst, float(np.max(np.abs(maps[s:e] - refm[s:e])) / np.max(np.abs(refm[s:e]))))
    print(f"synthesis per-ring rel err (every 7th ring) {worst:.2e}", flush=True)
    assert worst <= TOL["f32"]


@pytest.mark.parametrize("lmax", [40, 63, 100])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_gauss_legendre_round_trip_any_ring_length(gpu, lmax, precision):
    """Synthesis then analysis on a Gauss-Legendre grid (lmax + 1 rings of
    nphi = 2 lmax + 2 pixels: half-lengths 41 (odd, Bluestein), 64 (power of
    two) and 101) returns the band-limited input: quadrature exact, so the
    round trip is limited by rounding only."""
    import math
    from hellblast_amd.plan import Geometry, get_plan
    print(f"[case] GL round trip lmax={lmax} {precision}", flush=True)
    x, w = np.polynomial.legendre.leggauss(lmax + 1)
    x, w = x[::-1].copy(), w[::-1].copy()
    nphi = np.full(lmax + 1, 2 * lmax + 2, dtype=np.int32)
    geom = Geometry(nphi, cos_theta=x, sin_theta=np.sqrt((1 - x) * (1 + x)), weight=2 * math.pi * w)
    plan = get_plan(geom, lmax, lmax, precision)
    alm = OS.synalm(1.0 / (np.arange(lmax + 1) + 1.0) ** 2, lmax, seed=lmax)
    a = torch.as_tensor(alm, device="cuda")
    if precision == "f32":
        a = a.to(torch.complex64)
    back = plan.map2alm(plan.alm2map(a[None].repeat(3, 1), dtype=torch.float64 if precision == "f64" else torch.float32))
    err = float((back - a[None]).abs().max() / a.abs().max())
    print(f"round trip rel err {err:.2e}", flush=True)
    assert err <= (1e-12 if precision == "f64" else 2e-6)
