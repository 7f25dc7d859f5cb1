"""FP32 batches through the split-fp16 matrix-core kernels, checked row by
row against the FP64 plan (the FP64 plan is pinned to the oracle at <= 1e-11
in test_gpu_scalar.py).  The contraction they replace is the per-(m, l) ring
sum of ``sht_raw.py:53`` / ``sht_fallback.py:63`` (analysis) and the per-ring
l-sum of ``isht.py:15-21`` (synthesis).

Bars, for every row of every batch:
  * alm: eps_max = max |da_lm| / sqrt(C_l^ref) <= 1e-5 (C_l of the FP64 alm);
  * maps: max |dmap| <= 1e-5 max |map|;
  * C_l (anafast): max |dC_l| / C_l <= 1e-5;
  * power-of-two scaling (2^s inputs) gives 2^s times the unscaled result to
    rel 1e-6 of the largest coefficient, over the whole FP32 exponent range.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GATE = 1e-5


def _plans(nside, lmax):
    from hellblast_amd.hp_compat import _plan
    p64 = _plan(nside, lmax, lmax, "f64", "dz", "reference", "reference")
    p32 = _plan(nside, lmax, lmax, "f32", "dz", "reference", "reference")
    return p64, p32


def _lm(lmax, dev):
    rows = torch.arange(lmax + 1, 0, -1, device=dev)
    ms = torch.repeat_interleave(torch.arange(lmax + 1, device=dev), rows)
    ls = torch.arange(int(rows.sum()), device=dev) - torch.repeat_interleave(torch.cumsum(rows, 0) - rows, rows) + ms
    return ls, ms


def _draw(cl_of_l, lmax, nmaps, seed, dev):
    """Gaussian m-major alm [nmaps, nalm] with spectrum cl_of_l(l) (a_l0 real)."""
    ls, ms = _lm(lmax, dev)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    sd = torch.sqrt(cl_of_l(ls.double()))
    s = torch.where(ms == 0, sd, sd / math.sqrt(2.0))
    re = torch.randn((nmaps, ls.numel()), generator=g, device=dev, dtype=torch.float64)
    im = torch.randn((nmaps, ls.numel()), generator=g, device=dev, dtype=torch.float64)
    return torch.complex(re * s, torch.where(ms == 0, torch.zeros_like(im), im * s))


def _cl(a, ls, ms, lmax):
    p = a.real ** 2 + a.imag ** 2
    cw = torch.where(ms == 0, p, 2.0 * p)
    cl = torch.zeros(lmax + 1, dtype=torch.float64, device=a.device).index_add_(0, ls, cw)
    return cl / (2.0 * torch.arange(lmax + 1, device=a.device, dtype=torch.float64) + 1.0)


def _eps_max(a, ref, ls, ms, lmax):
    cl = _cl(ref, ls, ms, lmax)
    c = cl[ls]
    ok = c > 0
    return float(((a.to(torch.complex128) - ref).abs()[ok] / c[ok].sqrt()).max())


def test_h16_full_size_distinct_maps(gpu):
    """16 distinct maps at nside 2048, lmax 4095 (the bench workload) through
    the split-fp16 analysis (8 ring-pair groups of 512 pairs) and synthesis:
    every row against the FP64 plan.  A map-index slip in the batched
    kernels would hit some row."""
    nside, lmax, nb = 2048, 4095, 16
    print(f"[case] h16 distinct maps nside={nside} lmax={lmax} batch={nb}", flush=True)
    dev = torch.device("cuda", 0)
    p64, p32 = _plans(nside, lmax)
    ls, ms = _lm(lmax, dev)
    alm = _draw(lambda l: 1.0 / (l + 1.0) ** 2, lmax, nb, 41, dev)
    maps32 = torch.empty((nb, p32.npix), dtype=torch.float32, device=dev)
    for b in range(nb):
        maps32[b] = p64.alm2map(alm[b])
    a32 = p32.map2alm(maps32)
    worst = 0.0
    for b in range(nb):
        ref = p64.map2alm(maps32[b].double())
        e = _eps_max(a32[b], ref, ls, ms, lmax)
        worst = max(worst, e)
        assert e <= GATE, (b, e)
    print(f"analysis eps_max worst row {worst:.2e}", flush=True)
    del a32
    alm32 = alm.to(torch.complex64)
    m32 = p32.alm2map(alm32, dtype=torch.float32)
    worst = 0.0
    for b in range(nb):
        ref = p64.alm2map(alm32[b].to(torch.complex128))
        e = float((m32[b].double() - ref).abs().max() / ref.abs().max())
        worst = max(worst, e)
        assert e <= GATE, (b, e)
    print(f"synthesis map err worst row {worst:.2e}", flush=True)


def test_h16_beam_smoothed_synthesis_full_size(gpu):
    """A 1 degree FWHM Gaussian-beam-smoothed alm set (C_l = b_l^2/(l+1)^2:
    each m row spans hundreds of decades, and rows with m > ~2000 underflow
    to zero in complex64) through the split-fp16 synthesis, 4 maps at nside
    2048, lmax 4095, against the FP64 plan of the same complex64 alm."""
    nside, lmax, nb = 2048, 4095, 4
    print(f"[case] h16 beam-smoothed synthesis nside={nside} batch={nb}", flush=True)
    dev = torch.device("cuda", 0)
    p64, p32 = _plans(nside, lmax)
    sigma = math.radians(1.0) / math.sqrt(8.0 * math.log(2.0))
    alm = _draw(lambda l: torch.exp(-l * (l + 1.0) * sigma ** 2) / (l + 1.0) ** 2, lmax, nb, 43, dev)
    alm32 = alm.to(torch.complex64)
    m32 = p32.alm2map(alm32, dtype=torch.float32)
    for b in range(nb):
        ref = p64.alm2map(alm32[b].to(torch.complex128))
        e = float((m32[b].double() - ref).abs().max() / ref.abs().max())
        print(f"row {b}: map err {e:.2e}", flush=True)
        assert e <= GATE


def test_h16_c2_spectrum_with_noise_batch(gpu):
    """C2's input (C_l = 1/(l(l+1)) for l >= 2, plus white pixel noise of
    sigma 0.01) at nside 1024, lmax 2047, as a batch of 4 maps (the matrix
    cores, unlike the single-map C2 test): alm eps_max and the anafast C_l of
    every row against the FP64 plan."""
    nside, lmax, nb = 1024, 2047, 4
    print(f"[case] h16 C2 spectrum + noise nside={nside} batch={nb}", flush=True)
    from hellblast_amd.plan import alm2cl
    dev = torch.device("cuda", 0)
    p64, p32 = _plans(nside, lmax)
    ls, ms = _lm(lmax, dev)
    alm = _draw(lambda l: torch.where(l >= 2, 1.0 / (l * (l + 1.0)).clamp(min=1.0), torch.zeros_like(l)),
                lmax, nb, 2, dev)
    g = torch.Generator(device=dev)
    g.manual_seed(22)
    maps = p64.alm2map(alm) + 0.01 * torch.randn((nb, p64.npix), generator=g, device=dev, dtype=torch.float64)
    maps32 = maps.to(torch.float32)
    assert nb >= p32.tiles()["mfma_min_maps"]
    a32 = p32.map2alm(maps32)
    cl32 = alm2cl(a32, lmax)
    for b in range(nb):
        ref = p64.map2alm(maps32[b].double())
        e = _eps_max(a32[b], ref, ls, ms, lmax)
        cref = alm2cl(ref, lmax)
        ok = cref > 0
        ecl = float(((cl32[b] - cref).abs()[ok] / cref[ok]).max())
        print(f"row {b}: eps_max {e:.2e}  C_l err {ecl:.2e}", flush=True)
        assert e <= GATE and ecl <= GATE


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize("nside,lmax,shift", [(64, 128, -110), (64, 128, 110), (256, 511, -100), (256, 511, 100)])
def test_h16_power_of_two_scaled_transforms(gpu, nside, lmax, shift):
    """Maps and alm scaled by 2^shift through the split-fp16 kernels (4 maps)
    give 2^shift times the unscaled transforms, and stay within the gate of
    the FP64 plan: the per-map operand scaling covers the FP32 range.  The
    shifts stay inside what the FP32 ring stage itself represents: its
    coefficients of 2^-110 maps are normal numbers only at small nside, and
    a Bluestein ring of an inner FFT of M points holds sums up to
    M^1.5 max|map| (2^+110 maps overflow FP32 at nside 256, M = 1024).  The
    Legendre stages alone are checked to 2^+-120 below."""
    print(f"[case] h16 scaled 2^{shift} nside={nside} lmax={lmax}", flush=True)
    dev = torch.device("cuda", 0)
    p64, p32 = _plans(nside, lmax)
    ls, ms = _lm(lmax, dev)
    nb = 4
    alm = _draw(lambda l: 1.0 / (l + 1.0) ** 2, lmax, nb, 50 + nside, dev)
    maps32 = p64.alm2map(alm).to(torch.float32)
    sc = 2.0 ** shift
    a1 = p32.map2alm(maps32)
    a2 = p32.map2alm(maps32 * sc)
    assert torch.isfinite(torch.view_as_real(a2)).all()
    assert _rel(a2.to(torch.complex128) / sc, a1.to(torch.complex128)) <= 1e-6
    for b in range(nb):
        ref = p64.map2alm(maps32[b].double() * sc)
        assert _eps_max(a2[b], ref, ls, ms, lmax) <= GATE
    alm32 = alm.to(torch.complex64)
    m1 = p32.alm2map(alm32, dtype=torch.float32)
    m2 = p32.alm2map(alm32 * sc, dtype=torch.float32)
    assert torch.isfinite(m2).all()
    assert _rel(m2.double() / sc, m1.double()) <= 1e-6
    for b in range(nb):
        ref = p64.alm2map((alm32[b] * sc).to(torch.complex128))
        assert _rel(m2[b].double(), ref) <= GATE


@pytest.mark.parametrize("shift", [-110, -100, 100, 120])
def test_h16_legendre_stages_scale_exactly(gpu, shift):
    """The Legendre stages alone (stage entry points, nside 256, lmax 511, 4
    maps) on a phase buffer / alm scaled by 2^shift: the result is 2^shift
    times the unscaled one to FP32 rounding (rows below 2^-86 lost their low
    halves before the scale exponent was unclamped)."""
    nside, lmax, nb = 256, 511, 4
    print(f"[case] h16 Legendre stages 2^{shift}", flush=True)
    dev = torch.device("cuda", 0)
    p64, p32 = _plans(nside, lmax)
    alm = _draw(lambda l: 1.0 / (l + 1.0) ** 2, lmax, nb, 60, dev)
    maps32 = p64.alm2map(alm).to(torch.float32)
    phase = p32.phase_buffer(nb)
    p32.ring_analysis(maps32, phase)
    # FP32 phase coefficients first, then the FP64 band rows m < 8
    main = nb * p32.phase_map_elems * 8
    band = (main + 255) // 256 * 256
    f32v = phase[:main].view(torch.float32)
    f64v = phase[band:band + nb * p32.geom.nrings * 8 * 16].view(torch.float64)
    a1 = p32.legendre_analysis(phase, nb)
    sc = 2.0 ** shift
    f32v.mul_(sc)
    f64v.mul_(sc)
    a2 = p32.legendre_analysis(phase, nb)
    assert torch.isfinite(torch.view_as_real(a2)).all()
    assert _rel(a2.to(torch.complex128) / sc, a1.to(torch.complex128)) <= 1e-6
    alm32 = (alm * 2.0 ** -8).to(torch.complex64)         # keep 2^120 alm finite in FP32
    ph1 = p32.phase_buffer(nb).zero_()
    p32.legendre_synthesis(alm32, ph1)
    g1 = ph1[:main].view(torch.float32).double()
    ph2 = p32.phase_buffer(nb).zero_()
    p32.legendre_synthesis(alm32 * sc, ph2)
    g2 = ph2[:main].view(torch.float32).double()
    assert torch.isfinite(g2).all()
    assert _rel(g2 / sc, g1) <= 1e-6
