"""GPU parity of the spin-weighted path (hbx_*_spin via the C ABI) against
the FP64 oracle (oracle/spin.py, pinned in test_oracle_spin.py).

FP64 bar: relative max error <= 1e-11.  FP32 plans (FP32 data and
contraction -- on the matrix cores for batches --, FP64 recursion, FP64
band m < 8 in the analysis): relative max
error <= 2e-5 of the largest coefficient at nside 32, and eps_max =
max |da_lm| / sqrt(C_l) <= 1e-5 for E and for B at nside 2048, lmax 4095
(4 Q/U pairs, C_l^BB = 0.1 C_l^EE).
"""
import math

import numpy as np
import pytest
import torch

from oracle import geometry as OG
from oracle import sht as OS
from oracle import spin as OSp

pytestmark = pytest.mark.gpu

F64_TOL = 1e-11
F32_TOL = 2e-5


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _plan(nside, lmax, mmax=None, spin=2, precision="f64", phi0="reference", alias="reference"):
    from hellblast_amd.hp_compat import _plan
    mmax = lmax if mmax is None else mmax
    return _plan(nside, lmax, mmax, precision, "dz", phi0, alias, spin)


def _qu(tabs, seed, n=None):
    rng = np.random.default_rng(seed)
    shape = (2, tabs["npix"]) if n is None else (n, 2, tabs["npix"])
    return rng.standard_normal(shape)


CASES = [(1, 2, 2), (2, 5, 2), (4, 8, 1), (8, 16, 2), (8, 24, 1), (16, 40, 2), (32, 64, 2), (16, 33, 3)]


@pytest.mark.parametrize("nside,lmax,s", CASES)
def test_map2alm_spin_f64_matches_oracle(gpu, nside, lmax, s):
    print(f"[case] map2alm_spin f64 nside={nside} lmax={lmax} s={s}", flush=True)
    tabs = OG.ring_tables(nside)
    qu = _qu(tabs, 10 + nside + s)
    E, B = OSp.map2alm_spin(qu[0], qu[1], tabs, lmax, spin=s)
    got = _plan(nside, lmax, spin=s).map2alm_spin(torch.as_tensor(qu, device="cuda")).cpu().numpy()
    assert _rel(got[0], E) <= F64_TOL
    assert _rel(got[1], B) <= F64_TOL


@pytest.mark.parametrize("nside,lmax,s", CASES)
def test_alm2map_spin_f64_matches_oracle(gpu, nside, lmax, s):
    print(f"[case] alm2map_spin f64 nside={nside} lmax={lmax} s={s}", flush=True)
    tabs = OG.ring_tables(nside)
    E, B = OSp.synalm_eb(lmax, spin=s, seed=20 + nside)
    q, u = OSp.alm2map_spin(E, B, tabs, lmax, spin=s)
    eb = torch.as_tensor(np.stack([E, B]), device="cuda")
    got = _plan(nside, lmax, spin=s).alm2map_spin(eb).cpu().numpy()
    assert _rel(got[0], q) <= F64_TOL
    assert _rel(got[1], u) <= F64_TOL


@pytest.mark.parametrize("phi0,alias", [("healpix", "reference"), ("reference", "fold")])
def test_spin_phase_options_and_mmax(gpu, phi0, alias):
    nside, lmax, mmax, s = 16, 47, 30, 2
    print(f"[case] spin options phi0={phi0} alias={alias}", flush=True)
    tabs = OG.ring_tables(nside)
    ph = tabs["phi0_healpix"] if phi0 == "healpix" else None
    mc = np.full(tabs["Nr"], mmax + 1) if alias == "fold" else None
    plan = _plan(nside, lmax, mmax, spin=s, phi0=phi0, alias=alias)
    qu = _qu(tabs, 5)
    E, B = OSp.map2alm_spin(qu[0], qu[1], tabs, lmax, mmax, spin=s, phi0=ph, mcut=mc)
    got = plan.map2alm_spin(torch.as_tensor(qu, device="cuda")).cpu().numpy()
    assert _rel(got[0], E) <= F64_TOL and _rel(got[1], B) <= F64_TOL
    E, B = OSp.synalm_eb(lmax, mmax, spin=s, seed=6)
    q, u = OSp.alm2map_spin(E, B, tabs, lmax, mmax, spin=s, phi0=ph, mcut=mc)
    got = plan.alm2map_spin(torch.as_tensor(np.stack([E, B]), device="cuda")).cpu().numpy()
    assert _rel(got[0], q) <= F64_TOL and _rel(got[1], u) <= F64_TOL


def test_spin_multi_group_matches_oracle(gpu):
    """nside 256: 512 ring pairs, so the analysis splits the rings over two
    groups and reduces their partial sums."""
    nside, lmax, s = 256, 200, 2
    print(f"[case] spin multi-group nside={nside} lmax={lmax}", flush=True)
    tabs = OG.ring_tables(nside)
    qu = _qu(tabs, 7)
    E, B = OSp.map2alm_spin(qu[0], qu[1], tabs, lmax, spin=s)
    plan = _plan(nside, lmax, spin=s)
    got = plan.map2alm_spin(torch.as_tensor(qu, device="cuda")).cpu().numpy()
    assert _rel(got[0], E) <= F64_TOL and _rel(got[1], B) <= F64_TOL
    Ea, Ba = OSp.synalm_eb(lmax, spin=s, seed=8)
    q, u = OSp.alm2map_spin(Ea, Ba, tabs, lmax, spin=s)
    gm = plan.alm2map_spin(torch.as_tensor(np.stack([Ea, Ba]), device="cuda")).cpu().numpy()
    assert _rel(gm[0], q) <= F64_TOL and _rel(gm[1], u) <= F64_TOL


def test_spin_adjoints_match_oracle(gpu):
    nside, lmax, s = 16, 40, 2
    print("[case] spin adjoints", flush=True)
    tabs = OG.ring_tables(nside)
    plan = _plan(nside, lmax, spin=s)
    g = _qu(tabs, 9)
    aE, aB = OSp.alm2map_spin_adjoint(g[0], g[1], tabs, lmax, spin=s)
    got = plan.alm2map_spin_adjoint(torch.as_tensor(g, device="cuda")).cpu().numpy()
    assert _rel(got[0], aE) <= F64_TOL and _rel(got[1], aB) <= F64_TOL
    gE, gB = OSp.synalm_eb(lmax, spin=s, seed=10)
    q, u = OSp.map2alm_spin_adjoint(gE, gB, tabs, lmax, spin=s)
    got = plan.map2alm_spin_adjoint(torch.as_tensor(np.stack([gE, gB]), device="cuda")).cpu().numpy()
    assert _rel(got[0], q) <= F64_TOL and _rel(got[1], u) <= F64_TOL


def test_spin_batched_equals_single_and_deterministic(gpu):
    nside, lmax, s = 32, 64, 2
    print("[case] spin batched", flush=True)
    tabs = OG.ring_tables(nside)
    plan = _plan(nside, lmax, spin=s)
    qu = torch.as_tensor(_qu(tabs, 11, n=3), device="cuda")
    batch = plan.map2alm_spin(qu)
    for b in range(3):
        assert torch.equal(batch[b], plan.map2alm_spin(qu[b]))
    assert torch.equal(batch, plan.map2alm_spin(qu))
    maps = plan.alm2map_spin(batch)
    for b in range(3):
        assert torch.equal(maps[b], plan.alm2map_spin(batch[b]))


@pytest.mark.parametrize("s", [1, 2])
def test_spin_f32_against_f64_oracle(gpu, s):
    nside, lmax = 32, 64
    print(f"[case] spin f32 s={s}", flush=True)
    tabs = OG.ring_tables(nside)
    qu = _qu(tabs, 12)
    E, B = OSp.map2alm_spin(qu[0], qu[1], tabs, lmax, spin=s)
    plan = _plan(nside, lmax, spin=s, precision="f32")
    got = plan.map2alm_spin(torch.as_tensor(qu, device="cuda", dtype=torch.float32)).cpu().numpy()
    assert _rel(got[0], E) <= F32_TOL and _rel(got[1], B) <= F32_TOL
    Ea, Ba = OSp.synalm_eb(lmax, spin=s, seed=13)
    q, u = OSp.alm2map_spin(Ea, Ba, tabs, lmax, spin=s)
    gm = plan.alm2map_spin(torch.as_tensor(np.stack([Ea, Ba]), device="cuda"), dtype=torch.float32).cpu().numpy()
    assert _rel(gm[0], q) <= F32_TOL and _rel(gm[1], u) <= F32_TOL


def test_spin_f32_matrix_core_analysis(gpu):
    """FP32 spin analysis on the matrix cores (128 ring pairs per workgroup,
    4 Q/U pairs per product) at nside 128 (256 ring pairs: two groups whose
    partial sums are reduced) on 5 Q/U pairs (a second, partly filled pair
    group): against the FP64 oracle; equal to the VALU kernel (selected by
    mfma_min_maps above the batch) to FP32 rounding, for the weighted and
    the unweighted (adjoint) analysis; one pair alone gives the bits it gets
    inside the batch."""
    nside, lmax, s = 128, 200, 2
    print("[case] spin f32 matrix-core analysis nside=128 lmax=200 5 pairs", flush=True)
    tabs = OG.ring_tables(nside)
    qu = _qu(tabs, 31, n=5)
    # every batch, one pair included, on the matrix cores
    plan = _plan(nside, lmax, spin=s, precision="f32").with_tiles(mfma_min_maps=2)
    _matrix_core_analysis_checks(plan, qu, tabs, lmax, s)


def _matrix_core_analysis_checks(plan, qu, tabs, lmax, s):
    x = torch.as_tensor(qu, device="cuda", dtype=torch.float32)
    got = plan.map2alm_spin(x).cpu().numpy()
    q32 = qu.astype(np.float32).astype(np.float64)
    for b in (0, 4):
        E, B = OSp.map2alm_spin(q32[b, 0], q32[b, 1], tabs, lmax, spin=s)
        print(f"pair {b}: E {_rel(got[b, 0], E):.2e} B {_rel(got[b, 1], B):.2e}", flush=True)
        assert _rel(got[b, 0], E) <= F32_TOL and _rel(got[b, 1], B) <= F32_TOL
    one = plan.map2alm_spin(x[4:5]).cpu().numpy()
    assert np.array_equal(one[0], got[4])
    adj = plan.alm2map_spin_adjoint(x).cpu().numpy()
    vplan = plan.with_tiles(mfma_min_maps=1000)
    valu = vplan.map2alm_spin(x).cpu().numpy()
    vadj = vplan.alm2map_spin_adjoint(x).cpu().numpy()
    print(f"matrix vs VALU: {_rel(got, valu):.2e}, adjoint {_rel(adj, vadj):.2e}", flush=True)
    assert _rel(got, valu) <= 4e-6 and _rel(adj, vadj) <= 4e-6


def test_spin_f32_matrix_core_synthesis(gpu):
    """FP32 spin synthesis on the matrix cores (32 ring pairs per wave, 8
    polarisation pairs per product) at nside 128, lmax 200 for s = 1, 2, 3
    (s = 1, 3: the rows m < s start at an l of the other parity) on 5 Q/U
    pairs (a partly filled pair group): against the FP64 oracle; equal to
    the VALU kernel to FP32 rounding, for the synthesis and its weighted
    transpose (map2alm_spin_adjoint); one pair alone gives the bits it gets
    inside the batch."""
    nside, lmax = 128, 200
    tabs = OG.ring_tables(nside)
    for s in (1, 2, 3):
        print(f"[case] spin f32 matrix-core synthesis nside=128 lmax=200 s={s} 5 pairs", flush=True)
        # every batch, one pair included, on the matrix cores
        plan = _plan(nside, lmax, spin=s, precision="f32").with_tiles(mfma_min_maps=2)
        eb = np.stack([np.stack(OSp.synalm_eb(lmax, spin=s, seed=40 + b)) for b in range(5)])
        a = torch.as_tensor(eb, device="cuda").to(torch.complex64)
        got = plan.alm2map_spin(a, dtype=torch.float32).double().cpu().numpy()
        e32 = eb.astype(np.complex64).astype(np.complex128)
        for b in (0, 4):
            q, u = OSp.alm2map_spin(e32[b, 0], e32[b, 1], tabs, lmax, spin=s)
            print(f"pair {b}: Q {_rel(got[b, 0], q):.2e} U {_rel(got[b, 1], u):.2e}", flush=True)
            assert _rel(got[b, 0], q) <= F32_TOL and _rel(got[b, 1], u) <= F32_TOL
        one = plan.alm2map_spin(a[4:5], dtype=torch.float32).double().cpu().numpy()
        assert np.array_equal(one[0], got[4])
        adj = plan.map2alm_spin_adjoint(a, dtype=torch.float32).double().cpu().numpy()
        vplan = plan.with_tiles(mfma_min_maps=1000)
        valu = vplan.alm2map_spin(a, dtype=torch.float32).double().cpu().numpy()
        vadj = vplan.map2alm_spin_adjoint(a, dtype=torch.float32).double().cpu().numpy()
        print(f"matrix vs VALU: {_rel(got, valu):.2e}, adjoint {_rel(adj, vadj):.2e}", flush=True)
        assert _rel(got, valu) <= 4e-6 and _rel(adj, vadj) <= 4e-6


@pytest.mark.parametrize("nside,lmax,s", [(1, 2, 2), (2, 5, 2), (4, 8, 1), (8, 24, 3), (16, 40, 2)])
def test_spin_f32_small_sizes(gpu, nside, lmax, s):
    """FP32 spin batches at small sizes: mmax < 8 (the analysis is all FP64
    band, the synthesis runs the matrix kernel on every m), one partly used
    ring-pair tile, the self-paired equator ring; 2 Q/U pairs."""
    print(f"[case] spin f32 small nside={nside} lmax={lmax} s={s}", flush=True)
    tabs = OG.ring_tables(nside)
    # the matrix kernels, which 2 pairs would not get by default
    plan = _plan(nside, lmax, spin=s, precision="f32").with_tiles(mfma_min_maps=2)
    _small_size_checks(plan, tabs, nside, lmax, s)


def _small_size_checks(plan, tabs, nside, lmax, s):
    qu = _qu(tabs, 70 + nside, n=2)
    got = plan.map2alm_spin(torch.as_tensor(qu, device="cuda", dtype=torch.float32)).cpu().numpy()
    q32 = qu.astype(np.float32).astype(np.float64)
    eb = np.stack([np.stack(OSp.synalm_eb(lmax, spin=s, seed=80 + b)) for b in range(2)])
    gm = plan.alm2map_spin(torch.as_tensor(eb, device="cuda").to(torch.complex64),
                           dtype=torch.float32).double().cpu().numpy()
    e32 = eb.astype(np.complex64).astype(np.complex128)
    for b in range(2):
        E, B = OSp.map2alm_spin(q32[b, 0], q32[b, 1], tabs, lmax, spin=s)
        assert _rel(got[b, 0], E) <= F32_TOL and _rel(got[b, 1], B) <= F32_TOL
        q, u = OSp.alm2map_spin(e32[b, 0], e32[b, 1], tabs, lmax, spin=s)
        assert _rel(gm[b, 0], q) <= F32_TOL and _rel(gm[b, 1], u) <= F32_TOL


@pytest.mark.parametrize("s", [1, 3])
def test_spin_f32_matrix_cores_options(gpu, s):
    """The FP32 matrix-core spin kernels on the other spins (l0 = max(m, s)),
    mmax < lmax with mmax + 1 not a multiple of 8 (a partial last m-block),
    HEALPix phi0 and the periodic alias fold, 3 Q/U pairs (one partly filled
    pair group): analysis and synthesis against the FP64 oracle."""
    nside, lmax, mmax = 64, 150, 100
    print(f"[case] spin f32 matrix cores options s={s} nside=64 lmax=150 mmax=100", flush=True)
    tabs = OG.ring_tables(nside)
    ph = tabs["phi0_healpix"]
    mc = np.full(tabs["Nr"], mmax + 1)
    plan = _plan(nside, lmax, mmax, spin=s, precision="f32", phi0="healpix", alias="fold")
    qu = _qu(tabs, 50 + s, n=3)
    got = plan.map2alm_spin(torch.as_tensor(qu, device="cuda", dtype=torch.float32)).cpu().numpy()
    q32 = qu.astype(np.float32).astype(np.float64)
    for b in (0, 2):
        E, B = OSp.map2alm_spin(q32[b, 0], q32[b, 1], tabs, lmax, mmax, spin=s, phi0=ph, mcut=mc)
        print(f"analysis pair {b}: E {_rel(got[b, 0], E):.2e} B {_rel(got[b, 1], B):.2e}", flush=True)
        assert _rel(got[b, 0], E) <= F32_TOL and _rel(got[b, 1], B) <= F32_TOL
    eb = np.stack([np.stack(OSp.synalm_eb(lmax, mmax, spin=s, seed=60 + b)) for b in range(3)])
    a = torch.as_tensor(eb, device="cuda").to(torch.complex64)
    gm = plan.alm2map_spin(a, dtype=torch.float32).double().cpu().numpy()
    e32 = eb.astype(np.complex64).astype(np.complex128)
    for b in (0, 2):
        q, u = OSp.alm2map_spin(e32[b, 0], e32[b, 1], tabs, lmax, mmax, spin=s, phi0=ph, mcut=mc)
        print(f"synthesis pair {b}: Q {_rel(gm[b, 0], q):.2e} U {_rel(gm[b, 1], u):.2e}", flush=True)
        assert _rel(gm[b, 0], q) <= F32_TOL and _rel(gm[b, 1], u) <= F32_TOL


def _gl_geometry(lmax):
    from hellblast_amd.plan import Geometry
    x, w = np.polynomial.legendre.leggauss(lmax + 1)
    x, w = x[::-1].copy(), w[::-1].copy()
    nphi = np.full(lmax + 1, 2 * lmax + 2, dtype=np.int32)
    return Geometry(nphi, cos_theta=x, sin_theta=np.sqrt((1 - x) * (1 + x)), weight=2 * math.pi * w)


@pytest.mark.parametrize("s", [1, 2])
def test_spin_gauss_legendre_round_trip_large(gpu, s):
    """Size-independent property at lmax 1023: synthesis then analysis on a
    Gauss-Legendre grid returns the input (exact quadrature)."""
    from hellblast_amd.plan import get_plan
    lmax = 1023
    print(f"[case] spin GL round trip lmax={lmax} s={s}", flush=True)
    plan = get_plan(_gl_geometry(lmax), lmax, lmax, "f64", spin=s)
    E, B = OSp.synalm_eb(lmax, spin=s, seed=14)
    eb = torch.as_tensor(np.stack([E, B]), device="cuda")
    back = plan.map2alm_spin(plan.alm2map_spin(eb))
    assert float((back - eb).abs().max()) <= 1e-10 * float(eb.abs().max())


def test_spin_hp_compat_wrappers(gpu):
    from hellblast_amd import hp_compat
    from hellblast_amd.spin import eb_from_qu_gpu, map2alm_spin_gpu, qu_from_eb_gpu, spin_pair_to_eb
    nside, lmax = 8, 20
    print("[case] spin wrappers", flush=True)
    tabs = OG.ring_tables(nside)
    qu = _qu(tabs, 15)
    E, B = OSp.map2alm_spin(qu[0], qu[1], tabs, lmax, spin=2)
    q_t = torch.as_tensor(qu[0], device="cuda")
    u_t = torch.as_tensor(qu[1], device="cuda")
    e1, b1 = hp_compat.map2alm_spin(q_t, u_t, spin=2, lmax=lmax)
    assert _rel(e1.cpu().numpy(), E) <= F64_TOL and _rel(b1.cpu().numpy(), B) <= F64_TOL
    e2, b2 = hp_compat.iqu2eb(q_t, u_t, lmax=lmax)
    assert torch.equal(e1, e2) and torch.equal(b1, b2)
    rings_q = [q_t[tabs["ring_start"][r]:tabs["ring_start"][r + 1]] for r in range(tabs["Nr"])]
    rings_u = [u_t[tabs["ring_start"][r]:tabs["ring_start"][r + 1]] for r in range(tabs["Nr"])]
    e3, b3 = eb_from_qu_gpu(rings_q, rings_u, tabs["theta_centers"], tabs["nphi"], tabs["weights_dz"], lmax)
    assert torch.equal(e1, e3) and torch.equal(b1, b3)
    a2, am2 = map2alm_spin_gpu(rings_q, rings_u, tabs["theta_centers"], tabs["nphi"], tabs["weights_dz"], lmax)
    e4, b4 = spin_pair_to_eb(a2, am2, 2)
    assert torch.allclose(e4, e1, rtol=0, atol=1e-14) and torch.allclose(b4, b1, rtol=0, atol=1e-14)
    q, u = OSp.alm2map_spin(E, B, tabs, lmax, spin=2)
    Q1, U1 = hp_compat.eb2iqu(torch.as_tensor(E), torch.as_tensor(B), nside=nside, lmax=lmax)
    assert _rel(Q1.cpu().numpy(), q) <= F64_TOL and _rel(U1.cpu().numpy(), u) <= F64_TOL
    Qr, Ur = qu_from_eb_gpu(torch.as_tensor(E), torch.as_tensor(B), tabs["theta_centers"], tabs["nphi"], lmax)
    assert torch.equal(torch.cat(Qr), Q1) and torch.equal(torch.cat(Ur), U1)


def test_spin_plan_kind_errors(gpu):
    from hellblast_amd._abi import HbxValueError
    scalar = _plan(8, 16, spin=0)
    spin = _plan(8, 16, spin=2)
    with pytest.raises(HbxValueError):
        scalar.map2alm_spin(torch.zeros(2, 768, device="cuda", dtype=torch.float64))
    with pytest.raises(HbxValueError):
        spin.map2alm(torch.zeros(768, device="cuda", dtype=torch.float64))


def test_spin_f32_gate_full_size(gpu):
    """Spin-2 FP32 analysis at nside 2048, lmax 4095 on 4 Q/U pairs:
    eps_max <= 1e-5 against the FP64 plan for E and for B, with
    C_l^EE = 1/(l+1)^2 (l >= 2) and C_l^BB = 0.1 C_l^EE (B, ten times weaker,
    is the sensitive one)."""
    print("[case] spin f32 gate nside=2048 lmax=4095", flush=True)
    nside, lmax = 2048, 4095
    dev = torch.device("cuda", 0)
    p64 = _plan(nside, lmax, spin=2, precision="f64")
    p32 = _plan(nside, lmax, spin=2, precision="f32")
    n = (lmax + 1) * (lmax + 2) // 2
    rows = torch.arange(lmax + 1, 0, -1, device=dev)
    ms = torch.repeat_interleave(torch.arange(lmax + 1, device=dev), rows)
    ls = torch.arange(n, device=dev) - torch.repeat_interleave(torch.cumsum(rows, 0) - rows, rows) + ms
    g = torch.Generator(device=dev)
    g.manual_seed(21)

    def draw(amp):
        sd = torch.where(ls < 2, torch.zeros(n, device=dev, dtype=torch.float64), amp / (ls.double() + 1.0))
        sc = torch.where(ms == 0, sd, sd / math.sqrt(2.0))
        re = torch.randn(n, generator=g, device=dev, dtype=torch.float64)
        im = torch.randn(n, generator=g, device=dev, dtype=torch.float64)
        return torch.complex(re * sc, torch.where(ms == 0, torch.zeros_like(im), im * sc))

    eb = torch.stack([draw(1.0), draw(math.sqrt(0.1))])[None]
    qu32 = p64.alm2map_spin(eb).to(torch.float32)
    ref = p64.map2alm_spin(qu32.double())[0]
    got = p32.map2alm_spin(qu32.repeat(4, 1, 1))
    for k, name in enumerate("EB"):
        r = ref[k]
        pw = r.real ** 2 + r.imag ** 2
        cl = torch.zeros(lmax + 1, dtype=torch.float64, device=dev).index_add_(0, ls, torch.where(ms == 0, pw, 2 * pw))
        cl /= 2.0 * torch.arange(lmax + 1, device=dev, dtype=torch.float64) + 1.0
        ok = cl[ls] > 0
        for b in (0, 3):
            e = float(((got[b, k].to(torch.complex128) - r).abs()[ok] / cl[ls][ok].sqrt()).max())
            print(f"{name} pair {b}: eps_max {e:.2e}", flush=True)
            assert e <= 1e-5


def test_default_matrix_thresholds(gpu):
    """Effective defaults: scalar FP32 batches of 3 or more maps and spin FP32
    batches of 3 or more Q/U pairs (6 maps) use the matrix-core kernels."""
    print("[case] default matrix thresholds", flush=True)
    assert _plan(8, 16, precision="f32", spin=0).tiles()["mfma_min_maps"] == 3
    assert _plan(8, 16, precision="f32", spin=2).tiles()["mfma_min_maps"] == 6
