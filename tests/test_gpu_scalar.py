"""GPU parity of the scalar SHT path (libhbx.so via the C ABI) against the oracle.

FP64 bar: relative max error <= 1e-11 against the FP64 CPU oracle (which is
itself pinned to the reference formulas in test_oracle.py).  FP32 bar
(the FP32 gate): eps_max <= 1e-5 against the FP64 oracle.
"""
import math

import numpy as np
import pytest
import torch

from oracle import geometry as OG
from oracle import literal as OL
from oracle import sht as OS

pytestmark = pytest.mark.gpu

F64_TOL = 1e-11


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _plan(nside, lmax, mmax=None, precision="f64", phi0="reference", alias="reference", weights="dz"):
    from hellblast_amd.hp_compat import _plan
    mmax = lmax if mmax is None else mmax
    return _plan(nside, lmax, mmax, precision, weights, phi0, alias)


def _oracle_kw(tabs, phi0, alias, mmax):
    ph = tabs["phi0_healpix"] if phi0 == "healpix" else None
    mc = np.full(tabs["Nr"], mmax + 1) if alias == "fold" else None
    return ph, mc


CASES = [(1, 2), (2, 5), (4, 8), (8, 16), (8, 24), (16, 40), (32, 64), (64, 128)]


@pytest.mark.parametrize("nside,lmax", CASES)
def test_map2alm_f64_matches_oracle(gpu, nside, lmax):
    print(f"[case] map2alm f64 nside={nside} lmax={lmax}", flush=True)
    tabs = OG.ring_tables(nside)
    rng = np.random.default_rng(100 + nside)
    f = rng.standard_normal(tabs["npix"])
    ref = OS.map2alm(f, tabs, lmax)
    got = _plan(nside, lmax).map2alm(torch.as_tensor(f, device="cuda")).cpu().numpy()
    assert _rel(got, ref) <= F64_TOL


@pytest.mark.parametrize("nside,lmax", CASES)
def test_alm2map_f64_matches_oracle(gpu, nside, lmax):
    print(f"[case] alm2map f64 nside={nside} lmax={lmax}", flush=True)
    tabs = OG.ring_tables(nside)
    alm = OS.synalm(1.0 / (np.arange(lmax + 1) + 1.0) ** 2, lmax, seed=200 + nside)
    ref = OS.alm2map(alm, tabs, lmax)
    got = _plan(nside, lmax).alm2map(torch.as_tensor(alm, device="cuda")).cpu().numpy()
    assert _rel(got, ref) <= F64_TOL


@pytest.mark.parametrize("phi0,alias", [("healpix", "reference"), ("reference", "fold"), ("healpix", "fold")])
@pytest.mark.parametrize("nside,lmax,mmax", [(8, 24, 24), (16, 47, 30), (32, 95, 95)])
def test_phase_options_match_oracle(gpu, nside, lmax, mmax, phi0, alias):
    print(f"[case] options nside={nside} lmax={lmax} mmax={mmax} phi0={phi0} alias={alias}", flush=True)
    tabs = OG.ring_tables(nside)
    ph, mc = _oracle_kw(tabs, phi0, alias, mmax)
    rng = np.random.default_rng(7)
    f = rng.standard_normal(tabs["npix"])
    plan = _plan(nside, lmax, mmax, phi0=phi0, alias=alias)
    ref = OS.map2alm(f, tabs, lmax, mmax, phi0=ph, mcut=mc)
    got = plan.map2alm(torch.as_tensor(f, device="cuda")).cpu().numpy()
    assert _rel(got, ref) <= F64_TOL
    alm = OS.synalm(np.ones(lmax + 1), lmax, mmax, seed=8)
    refm = OS.alm2map(alm, tabs, lmax, mmax, phi0=ph, mcut=mc)
    gotm = plan.alm2map(torch.as_tensor(alm, device="cuda")).cpu().numpy()
    assert _rel(gotm, refm) <= F64_TOL


def test_batched_equals_single(gpu):
    nside, lmax = 16, 40
    tabs = OG.ring_tables(nside)
    rng = np.random.default_rng(3)
    f = torch.as_tensor(rng.standard_normal((5, tabs["npix"])), device="cuda")
    plan = _plan(nside, lmax)
    batch = plan.map2alm(f)
    for i in range(5):
        assert torch.equal(batch[i], plan.map2alm(f[i]))
    maps = plan.alm2map(batch)
    for i in range(5):
        assert torch.equal(maps[i], plan.alm2map(batch[i]))


def test_f64_deterministic(gpu):
    nside, lmax = 32, 80
    tabs = OG.ring_tables(nside)
    f = torch.as_tensor(np.random.default_rng(4).standard_normal(tabs["npix"]), device="cuda")
    plan = _plan(nside, lmax)
    a1 = plan.map2alm(f)
    a2 = plan.map2alm(f)
    assert torch.equal(a1, a2)
    assert torch.equal(plan.alm2map(a1), plan.alm2map(a2))


def test_adjoints_dot_product(gpu):
    nside, lmax = 16, 40
    tabs = OG.ring_tables(nside)
    rng = np.random.default_rng(5)
    plan = _plan(nside, lmax, phi0="healpix")
    a = torch.as_tensor(OS.synalm(np.ones(lmax + 1), lmax, seed=9), device="cuda")
    g = torch.as_tensor(rng.standard_normal(tabs["npix"]), device="cuda")
    lhs = torch.dot(plan.alm2map(a), g).item()
    adj = plan.alm2map_adjoint(g)
    rhs = torch.sum((adj.conj() * a).real).item()
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)
    ga = torch.as_tensor(OS.synalm(np.ones(lmax + 1), lmax, seed=10), device="cuda")
    lhs2 = torch.sum((ga.conj() * plan.map2alm(g)).real).item()
    rhs2 = torch.dot(plan.map2alm_adjoint(ga), g).item()
    assert abs(lhs2 - rhs2) <= 1e-11 * abs(lhs2)
    # and against the oracle's transposes
    ref = OS.alm2map_adjoint(g.cpu().numpy(), tabs, lmax, phi0=tabs["phi0_healpix"])
    assert _rel(adj.cpu().numpy(), ref) <= F64_TOL
    ref2 = OS.map2alm_adjoint(ga.cpu().numpy(), tabs, lmax, phi0=tabs["phi0_healpix"])
    assert _rel(plan.map2alm_adjoint(ga).cpu().numpy(), ref2) <= F64_TOL


def test_alm2cl_matches_oracle(gpu):
    from hellblast_amd import alm2cl
    lmax, mmax = 60, 45
    alm = OS.synalm(np.ones(lmax + 1), lmax, mmax, seed=11)
    got = alm2cl(torch.as_tensor(alm, device="cuda"), lmax=lmax, mmax=mmax).cpu().numpy()
    assert _rel(got, OS.alm2cl(alm, lmax, mmax)) <= 1e-14


def test_reference_signature_matches_literal_fallback(gpu):
    """map2alm_gpu(rings, thetas, nphi, weights_dz, lmax) == sht_fallback restated literally."""
    from hellblast_amd import alm_ops, map2alm_gpu
    nside, lmax = 16, 30
    tabs = OG.ring_tables(nside)
    f = np.random.default_rng(12).standard_normal(tabs["npix"])
    rings = OG.flat_to_rings(f, tabs)
    ref = OL.sht_fallback_map2alm(rings, tabs["theta_centers"], tabs["nphi"], tabs["weights_dz"], lmax)
    got = map2alm_gpu([torch.as_tensor(r) for r in rings], tabs["theta_centers"], tabs["nphi"],
                      tabs["weights_dz"], lmax)
    got_ref = alm_ops.to_reference_layout(got, lmax).cpu().numpy()
    assert _rel(got_ref, ref) <= F64_TOL
    # sht_raw (the reference's dispatched path) differs by 2 pi and truncates m at max nphi//2
    raw = OL.sht_raw_map2alm(rings, tabs["theta_centers"], tabs["nphi"], tabs["weights_dz"], lmax)
    assert _rel(got_ref, 2.0 * math.pi * raw) <= F64_TOL


def test_reference_synthesis_matches_literal_isht(gpu):
    """alm2map_gpu equals isht restated literally on rings with nphi > 2 mmax."""
    from hellblast_amd import alm2map_gpu, alm_ops
    nside, lmax = 16, 30
    tabs = OG.ring_tables(nside)
    alm = OS.synalm(1.0 / (np.arange(lmax + 1) + 1.0) ** 2, lmax, seed=13)
    ref = OL.isht_alm2map(OS.mmajor_to_ref(alm, lmax), tabs["theta_centers"], tabs["nphi"], lmax)
    got = alm2map_gpu(torch.as_tensor(alm, device="cuda"), tabs["theta_centers"], tabs["nphi"], lmax)
    for r in range(tabs["Nr"]):
        if tabs["nphi"][r] > 2 * lmax:
            assert _rel(got[r].cpu().numpy(), ref[r]) <= F64_TOL


def test_config1_roundtrip(gpu):
    """C1: nside 64, lmax 128, alm2map then map2alm (iter=0) vs the oracle."""
    from hellblast_amd import hp_compat
    nside, lmax = 64, 128
    cl = 1.0 / (np.arange(lmax + 1) + 1.0) ** 2
    alm = OS.synalm(cl, lmax, seed=1)
    alm[0] = 0.0
    tabs = OG.ring_tables(nside)
    m = hp_compat.alm2map(torch.as_tensor(alm, device="cuda"), nside, lmax=lmax)
    emap = _rel(m.cpu().numpy(), OS.alm2map(alm, tabs, lmax))
    a2 = hp_compat.map2alm(m, lmax=lmax)
    ref2 = OS.map2alm(m.cpu().numpy(), tabs, lmax)
    ealm = _rel(a2.cpu().numpy(), ref2)
    rt_max, rt_rms = OS.alm_errors(a2.cpu().numpy(), alm, lmax)
    print(f"C1: GPU vs oracle: alm2map rel {emap:.2e}, map2alm rel {ealm:.2e}; "
          f"round trip alm -> map -> alm (iter 0): eps_max {rt_max:.3e} eps_rms {rt_rms:.3e}", flush=True)
    assert emap <= F64_TOL and ealm <= F64_TOL


@pytest.mark.parametrize("nside,lmax", [(32, 64), (64, 128), (128, 256)])
def test_f32_within_gate(gpu, nside, lmax):
    tabs = OG.ring_tables(nside)
    alm = OS.synalm((np.arange(lmax + 1) + 10.0) ** -2, lmax, seed=14)
    ref_map = OS.alm2map(alm, tabs, lmax)
    ref_alm = OS.map2alm(ref_map, tabs, lmax)
    p32 = _plan(nside, lmax, precision="f32")
    m32 = p32.alm2map(torch.as_tensor(alm, device="cuda"), dtype=torch.float32)
    emap = _rel(m32.double().cpu().numpy(), ref_map)
    a32 = p32.map2alm(torch.as_tensor(ref_map, device="cuda"))
    emax, erms = OS.alm_errors(a32.cpu().numpy().astype(np.complex128), ref_alm, lmax)
    print(f"f32 nside={nside}: map rel {emap:.2e} alm eps_max {emax:.2e} eps_rms {erms:.2e}")
    assert emap <= 1e-5 and emax <= 1e-5


@pytest.mark.parametrize("nmaps", [2, 5, 16, 20])
@pytest.mark.parametrize("nside,lmax", [(16, 40), (64, 128)])
def test_f32_batched_synthesis_matches_oracle(gpu, nmaps, nside, lmax):
    """FP32 batches run the matrix-core synthesis; every map must match the oracle."""
    print(f"[case] f32 batched synthesis nmaps={nmaps} nside={nside} lmax={lmax}", flush=True)
    tabs = OG.ring_tables(nside)
    rng = np.random.default_rng(30 + nmaps)
    alms = np.stack([OS.synalm((np.arange(lmax + 1) + 10.0) ** -2, lmax, rng=rng) for _ in range(nmaps)])
    p32 = _plan(nside, lmax, precision="f32", phi0="healpix")
    got = p32.alm2map(torch.as_tensor(alms, device="cuda"), dtype=torch.float64).cpu().numpy()
    for b in range(nmaps):
        ref = OS.alm2map(alms[b], tabs, lmax, phi0=tabs["phi0_healpix"])
        assert _rel(got[b], ref) <= 2e-6, b
    # the transpose of map2alm takes the same route
    ga = torch.as_tensor(alms, device="cuda")
    adj = p32.map2alm_adjoint(ga, dtype=torch.float64).cpu().numpy()
    for b in range(nmaps):
        ref = OS.map2alm_adjoint(alms[b], tabs, lmax, phi0=tabs["phi0_healpix"])
        assert _rel(adj[b], ref) <= 2e-6, b


@pytest.mark.parametrize("nmaps", [2, 5, 16, 20])
@pytest.mark.parametrize("nside,lmax", [(16, 40), (64, 128)])
def test_f32_batched_analysis_matches_oracle(gpu, nmaps, nside, lmax):
    """FP32 batches run the matrix-core analysis; every map must match the oracle."""
    print(f"[case] f32 batched analysis nmaps={nmaps} nside={nside} lmax={lmax}", flush=True)
    tabs = OG.ring_tables(nside)
    rng = np.random.default_rng(40 + nmaps)
    maps = rng.standard_normal((nmaps, tabs["npix"]))
    p32 = _plan(nside, lmax, precision="f32", phi0="healpix")
    got = p32.map2alm(torch.as_tensor(maps, device="cuda")).cpu().numpy()
    for b in range(nmaps):
        ref = OS.map2alm(maps[b], tabs, lmax, phi0=tabs["phi0_healpix"])
        assert _rel(got[b], ref) <= 2e-6, b
    adj = p32.alm2map_adjoint(torch.as_tensor(maps, device="cuda")).cpu().numpy()
    for b in range(nmaps):
        ref = OS.alm2map_adjoint(maps[b], tabs, lmax, phi0=tabs["phi0_healpix"])
        assert _rel(adj[b], ref) <= 2e-6, b


@pytest.mark.parametrize("nside,lmax", [(512, 300), (1024, 256)])
def test_f32_batched_multi_group_matches_f64(gpu, nside, lmax):
    """Ring-pair groups > 1 (partial sums + reduce): FP32 batch vs the FP64 path."""
    print(f"[case] f32 multi-group nside={nside} lmax={lmax}", flush=True)
    nmaps = 3
    rng = np.random.default_rng(50)
    npix = 12 * nside * nside
    maps = torch.as_tensor(rng.standard_normal((nmaps, npix)), device="cuda")
    p32 = _plan(nside, lmax, precision="f32")
    p64 = _plan(nside, lmax)
    a32 = p32.map2alm(maps)
    a64 = p64.map2alm(maps)
    for b in range(nmaps):
        assert _rel(a32[b].cpu().numpy(), a64[b].cpu().numpy()) <= 2e-6
    m32 = p32.alm2map(a64.to(torch.complex64), dtype=torch.float64)
    m64 = p64.alm2map(a64)
    for b in range(nmaps):
        assert _rel(m32[b].cpu().numpy(), m64[b].cpu().numpy()) <= 2e-6


def test_config2_single_map_f32_vs_f64(gpu):
    """C2: one float64 map at nside 1024, lmax 2047: seeded Gaussian alm with
    C_l = 1/(l(l+1)) for l >= 2 (C_0 = C_1 = 0) plus white pixel noise
    sigma = 0.01.  Forward, inverse and anafast in FP32 (the single-map
    path) against the FP64 plan (itself pinned to the oracle at small sizes):
    alm eps_max <= 1e-5, C_l relative error <= 1e-5, map relative error
    <= 1e-5; FP64 round trip alm -> map -> alm reported."""
    from hellblast_amd.plan import alm2cl
    nside, lmax = 1024, 2047
    print(f"[case] config 2 nside={nside} lmax={lmax}", flush=True)
    ell = np.arange(lmax + 1, dtype=np.float64)
    cl = np.zeros(lmax + 1)
    cl[2:] = 1.0 / (ell[2:] * (ell[2:] + 1.0))
    alm = OS.synalm(cl, lmax, seed=22)
    p64 = _plan(nside, lmax, precision="f64")
    p32 = _plan(nside, lmax, precision="f32")
    a_in = torch.as_tensor(alm, device="cuda")
    rng = np.random.default_rng(23)
    m = p64.alm2map(a_in) + torch.as_tensor(0.01 * rng.standard_normal(p64.npix), device="cuda")
    ref = p64.map2alm(m)
    got = p32.map2alm(m)
    emax, erms = OS.alm_errors(got.cpu().numpy().astype(np.complex128), ref.cpu().numpy(), lmax)
    cl64 = alm2cl(ref, lmax).cpu().numpy()
    cl32 = alm2cl(got.to(torch.complex128), lmax).cpu().numpy()
    ecl = float(np.max(np.abs(cl32[2:] - cl64[2:]) / cl64[2:]))
    map64 = p64.alm2map(ref)
    map32 = p32.alm2map(got, dtype=torch.float64)
    emap = float((map32 - map64).abs().max() / map64.abs().max())
    rt = p64.map2alm(p64.alm2map(a_in))
    ert = float((rt - a_in).abs().max() / a_in.abs().max())
    print(f"f32 vs f64: alm eps_max {emax:.2e} eps_rms {erms:.2e}, C_l max rel {ecl:.2e}, "
          f"map rel {emap:.2e}; f64 round trip (iter 0) {ert:.2e}", flush=True)
    assert emax <= 1e-5 and ecl <= 1e-5 and emap <= 1e-5


def test_map2alm_iter_matches_oracle_jacobi(gpu):
    """hp_compat.map2alm(..., iter=k): a <- a + A(f - S a) from a = A f, equal
    to the oracle's iteration, and the alm -> map -> alm round trip error
    falls ~14x per step (band-limited input at lmax = 2 nside - 1, below the
    Nyquist bin of the 4 nside-pixel rings)."""
    from hellblast_amd import hp_compat
    nside, lmax = 16, 31
    print(f"[case] map2alm iter nside={nside} lmax={lmax}", flush=True)
    tabs = OG.ring_tables(nside)
    alm = OS.synalm(1.0 / (np.arange(lmax + 1) + 1.0) ** 2, lmax, seed=31)
    f = OS.alm2map(alm, tabs, lmax)
    errs = []
    for k in (0, 1, 3):
        got = hp_compat.map2alm(torch.as_tensor(f, device="cuda"), lmax=lmax, iter=k).cpu().numpy()
        ref = OS.map2alm(f, tabs, lmax)
        for _ in range(k):
            ref = ref + OS.map2alm(f - OS.alm2map(ref, tabs, lmax), tabs, lmax)
        assert _rel(got, ref) <= F64_TOL
        errs.append(_rel(got, alm))
    print("round trip error by iter:", errs, flush=True)
    assert errs[1] < 0.2 * errs[0] and errs[2] < 1e-3 * errs[0]


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_empty_batches(gpu, precision):
    """Batches of zero maps / pairs / pixels run nothing and return empty
    outputs of the right shape on every entry point."""
    print(f"[case] empty batches {precision}", flush=True)
    from hellblast_amd.coords_ring import RingMapperGPU
    from hellblast_amd.hp_compat import _plan
    nside, lmax = 8, 16
    rdt = torch.float64 if precision == "f64" else torch.float32
    p = _plan(nside, lmax, lmax, precision, "dz", "reference", "reference")
    ps = _plan(nside, lmax, lmax, precision, "dz", "reference", "reference", 2)
    npix = 12 * nside * nside
    a = p.map2alm(torch.zeros(0, npix, device="cuda", dtype=rdt))
    assert a.shape == (0, p.nalm) and a.dtype == p.alm_dtype
    m = p.alm2map(torch.zeros(0, p.nalm, device="cuda", dtype=p.alm_dtype), dtype=rdt)
    assert m.shape == (0, npix)
    assert p.alm2map_adjoint(torch.zeros(0, npix, device="cuda", dtype=rdt)).shape == (0, p.nalm)
    assert ps.map2alm_spin(torch.zeros(0, 2, npix, device="cuda", dtype=rdt)).shape == (0, 2, ps.nalm)
    assert ps.alm2map_spin(torch.zeros(0, 2, ps.nalm, device="cuda", dtype=ps.alm_dtype),
                           dtype=rdt).shape == (0, 2, npix)
    mp = RingMapperGPU(nside)
    assert mp.ang2pix(torch.zeros(0, dtype=torch.float64), torch.zeros(0, dtype=torch.float64)).numel() == 0
    th, ph = mp.pix2ang(torch.zeros(0, dtype=torch.int64))
    assert th.numel() == 0 and ph.numel() == 0
