"""Differentiation: torch.autograd through the adjoint kernels.

Bars: gradients within 1e-5 relative of central finite differences of the
FP64 forward (h = 1e-6); transposes exact to 1e-12 (FP64) by the
dot-product identity; torch.autograd.gradcheck / gradgradcheck at small
sizes (complex inputs included); FP32 gradients within 1e-5 of FP64.
"""
import math

import numpy as np
import pytest
import torch

from oracle import geometry as OG
from oracle import sht as OS
from oracle import spin as OSp

pytestmark = pytest.mark.gpu


def _plan(nside, lmax, precision="f64", spin=0):
    from hellblast_amd.hp_compat import _plan
    return _plan(nside, lmax, lmax, precision, "dz", "reference", "reference", spin)


def _mask(tabs, deg=20.0):
    """w_p = 0 within |latitude| < deg, else 1 (per pixel)."""
    z = np.cos(tabs["theta_centers"])
    wr = (np.abs(z) >= math.sin(math.radians(deg))).astype(np.float64)
    return np.repeat(wr, tabs["nphi"])


def _rdot(a, b):
    return float((torch.conj(a) * b).real.sum()) if a.is_complex() else float((a * b).sum())


def test_masked_chi2_gradient_matches_finite_differences(gpu):
    """16 alm at nside 512, lmax 1023, C_l = 1/(l+1)^2; L = sum_p w_p (alm2map(a)_p - d_p)^2."""
    nside, lmax, B = 512, 1023, 16
    print(f"[case] masked chi2 FD nside={nside} lmax={lmax} B={B}", flush=True)
    tabs = OG.ring_tables(nside)
    plan = _plan(nside, lmax)
    cl = 1.0 / (np.arange(lmax + 1) + 1.0) ** 2
    a_np = np.stack([OS.synalm(cl, lmax, seed=400 + b) for b in range(B)])
    d_alm = np.stack([OS.synalm(cl, lmax, seed=500 + b) for b in range(B)])
    rng = np.random.default_rng(4)
    dev = torch.device("cuda")
    w = torch.as_tensor(_mask(tabs), device=dev)
    d = plan.alm2map(torch.as_tensor(d_alm, device=dev)) + 0.1 * torch.as_tensor(
        rng.standard_normal((B, tabs["npix"])), device=dev)

    a = torch.as_tensor(a_np, device=dev).requires_grad_(True)
    loss = ((plan.alm2map(a) - d) ** 2 * w).sum()
    loss.backward()
    g = a.grad.detach()

    def loss_b(b, ab):
        r = plan.alm2map(ab) - d[b]
        return float(((r * r) * w).sum())

    h = 1e-6
    nlm = plan.nalm
    picks = [(int(rng.integers(B)), int(rng.integers(nlm)), int(rng.integers(2))) for _ in range(200)]
    worst = 0.0
    gscale = float(g.abs().max())
    with torch.no_grad():
        for b, i, part in picks:
            e = torch.zeros(nlm, dtype=torch.complex128, device=dev)
            e[i] = h if part == 0 else 1j * h
            fd = (loss_b(b, a[b] + e) - loss_b(b, a[b] - e)) / (2 * h)
            an = float(g[b, i].real if part == 0 else g[b, i].imag)
            worst = max(worst, abs(fd - an) / max(abs(an), 1e-3 * gscale))
    print(f"worst FD relative error {worst:.2e}")
    assert worst <= 1e-5


def test_transposes_at_c4_size(gpu):
    nside, lmax, B = 512, 1023, 4
    print("[case] transposes nside 512", flush=True)
    tabs = OG.ring_tables(nside)
    plan = _plan(nside, lmax)
    dev = torch.device("cuda")
    rng = np.random.default_rng(5)
    a = torch.as_tensor(np.stack([OS.synalm(np.ones(lmax + 1), lmax, seed=10 + b) for b in range(B)]), device=dev)
    f = torch.as_tensor(rng.standard_normal((B, tabs["npix"])), device=dev)
    lhs = _rdot(plan.alm2map(a), f)
    rhs = _rdot(a, plan.alm2map_adjoint(f))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    lhs = _rdot(a, plan.map2alm(f))
    rhs = _rdot(plan.map2alm_adjoint(a), f)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_f32_gradient_matches_f64(gpu):
    nside, lmax, B = 256, 511, 4
    print("[case] f32 gradient", flush=True)
    tabs = OG.ring_tables(nside)
    dev = torch.device("cuda")
    rng = np.random.default_rng(6)
    f = torch.as_tensor(rng.standard_normal((B, tabs["npix"])), device=dev)
    g64 = _plan(nside, lmax).alm2map_adjoint(f)
    g32 = _plan(nside, lmax, "f32").alm2map_adjoint(f.float()).to(torch.complex128)
    assert float((g32 - g64).abs().max()) <= 1e-5 * float(g64.abs().max())
    alm = torch.as_tensor(np.stack([OS.synalm(np.ones(lmax + 1), lmax, seed=20 + b) for b in range(B)]),
                          device=dev)
    m64 = _plan(nside, lmax).map2alm_adjoint(alm)
    m32 = _plan(nside, lmax, "f32").map2alm_adjoint(alm.to(torch.complex64)).double()
    assert float((m32 - m64).abs().max()) <= 1e-5 * float(m64.abs().max())


def test_gradcheck_scalar_small(gpu):
    print("[case] gradcheck scalar", flush=True)
    nside, lmax = 2, 5
    plan = _plan(nside, lmax)
    dev = torch.device("cuda")
    a = torch.as_tensor(OS.synalm(np.ones(lmax + 1), lmax, seed=1), device=dev).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda x: plan.alm2map(x), (a,), eps=1e-6, atol=1e-8, rtol=1e-6)
    assert torch.autograd.gradgradcheck(lambda x: plan.alm2map(x), (a,), eps=1e-6, atol=1e-8, rtol=1e-6)
    f = torch.randn(12 * nside * nside, dtype=torch.float64, device=dev, requires_grad=True)
    assert torch.autograd.gradcheck(lambda x: plan.map2alm(x), (f,), eps=1e-6, atol=1e-8, rtol=1e-6)
    assert torch.autograd.gradgradcheck(lambda x: plan.map2alm(x), (f,), eps=1e-6, atol=1e-8, rtol=1e-6)


def test_gradcheck_spin_small(gpu):
    print("[case] gradcheck spin", flush=True)
    nside, lmax = 2, 5
    plan = _plan(nside, lmax, spin=2)
    dev = torch.device("cuda")
    E, B = OSp.synalm_eb(lmax, spin=2, seed=2)
    eb = torch.as_tensor(np.stack([E, B]), device=dev).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda x: plan.alm2map_spin(x), (eb,), eps=1e-6, atol=1e-8, rtol=1e-6)
    qu = torch.randn(2, 12 * nside * nside, dtype=torch.float64, device=dev, requires_grad=True)
    assert torch.autograd.gradcheck(lambda x: plan.map2alm_spin(x), (qu,), eps=1e-6, atol=1e-8, rtol=1e-6)


def test_eb_separation_gradient(gpu):
    """E/B separation: L = sum w (alm2map_spin(E, B) - d)^2; gradient vs
    central differences on 40 coefficients (FP64, nside 64, lmax 128)."""
    nside, lmax = 64, 128
    print("[case] E/B separation FD", flush=True)
    tabs = OG.ring_tables(nside)
    plan = _plan(nside, lmax, spin=2)
    dev = torch.device("cuda")
    rng = np.random.default_rng(7)
    # C^EE = 1/(l+1)^2, C^BB = 0.1 C^EE (unit-variance draws scaled per l)
    ls = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)])
    amp = 1.0 / (ls + 1.0)
    E, B = OSp.synalm_eb(lmax, spin=2, seed=8)
    E, B = E * amp, B * amp * math.sqrt(0.1)
    E2, B2 = OSp.synalm_eb(lmax, spin=2, seed=9)
    w = torch.as_tensor(_mask(tabs), device=dev)
    d = plan.alm2map_spin(torch.as_tensor(np.stack([E2 * amp, B2 * amp * math.sqrt(0.1)]), device=dev))
    eb = torch.as_tensor(np.stack([E, B]), device=dev).requires_grad_(True)
    loss = ((plan.alm2map_spin(eb) - d) ** 2 * w).sum()
    loss.backward()
    g = eb.grad.detach()

    def lossf(x):
        r = plan.alm2map_spin(x) - d
        return float((r * r * w).sum())

    h = 1e-6
    worst = 0.0
    gscale = float(g.abs().max())
    with torch.no_grad():
        for _ in range(40):
            c, i, part = int(rng.integers(2)), int(rng.integers(plan.nalm)), int(rng.integers(2))
            e = torch.zeros_like(eb)
            e[c, i] = h if part == 0 else 1j * h
            fd = (lossf(eb + e) - lossf(eb - e)) / (2 * h)
            an = float(g[c, i].real if part == 0 else g[c, i].imag)
            worst = max(worst, abs(fd - an) / max(abs(an), 1e-3 * gscale))
    print(f"worst FD relative error {worst:.2e}")
    assert worst <= 1e-5


def test_eb_separation_gradient_c5_size(gpu):
    """E/B separation at C5's size (nside 2048, lmax 4095, FP64 forward):
    L = sum_p w_p |alm2map_spin(E, B)_p - d_p|^2 with the |b| < 20 deg mask,
    C^EE = 1/(l+1)^2, C^BB = 0.1 C^EE; the autograd gradient (through the
    spin adjoint kernels) against central differences on 40 coefficients of
    E and B.  The difference L(a + e) - L(a - e) is summed as
    sum_p w_p (r_+ - r_-)(r_+ + r_-), the same number as the difference of
    the two sums in exact arithmetic, so that the 5e7-pixel sums do not
    cancel at h = 1e-6."""
    nside, lmax = 2048, 4095
    print(f"[case] E/B separation FD nside={nside} lmax={lmax}", flush=True)
    plan = _plan(nside, lmax, spin=2)
    dev = torch.device("cuda")
    n = plan.nalm
    rows = torch.arange(lmax + 1, 0, -1, device=dev)
    ms = torch.repeat_interleave(torch.arange(lmax + 1, device=dev), rows)
    ls = torch.arange(n, device=dev) - torch.repeat_interleave(torch.cumsum(rows, 0) - rows, rows) + ms
    gen = torch.Generator(device=dev)
    gen.manual_seed(55)

    def draw(amp):
        sd = amp / (ls.double() + 1.0)
        sc = torch.where(ms == 0, sd, sd / math.sqrt(2.0))
        re = torch.randn(n, generator=gen, device=dev, dtype=torch.float64)
        im = torch.randn(n, generator=gen, device=dev, dtype=torch.float64)
        return torch.complex(re * sc, torch.where(ms == 0, torch.zeros_like(im), im * sc))

    eb = torch.stack([draw(1.0), draw(math.sqrt(0.1))]).requires_grad_(True)
    with torch.no_grad():
        d = plan.alm2map_spin(torch.stack([draw(1.0), draw(math.sqrt(0.1))]))
    w = torch.as_tensor(_mask(OG.ring_tables(nside)), device=dev)
    assert w.numel() == plan.npix
    loss = ((plan.alm2map_spin(eb) - d) ** 2 * w).sum()
    loss.backward()
    g = eb.grad.detach()
    rng = np.random.default_rng(12)
    h = 1e-6
    worst = 0.0
    gscale = float(g.abs().max())
    with torch.no_grad():
        for _ in range(40):
            c, i, part = int(rng.integers(2)), int(rng.integers(n)), int(rng.integers(2))
            e = torch.zeros_like(eb)
            e[c, i] = h if part == 0 else 1j * h
            rp = plan.alm2map_spin(eb + e) - d
            rm = plan.alm2map_spin(eb - e) - d
            fd = float((w * (rp - rm) * (rp + rm)).sum()) / (2 * h)
            an = float(g[c, i].real if part == 0 else g[c, i].imag)
            worst = max(worst, abs(fd - an) / max(abs(an), 1e-3 * gscale))
    print(f"worst FD relative error {worst:.2e}", flush=True)
    assert worst <= 1e-5


def test_spin2_transpose_c5_size_f32(gpu):
    """4 Q/U pairs at nside 2048, lmax 4095 (FP32 plan): the spin adjoint is
    the transpose of the spin synthesis to FP32 accuracy."""
    nside, lmax, P = 2048, 4095, 4
    print("[case] spin-2 transpose nside 2048 f32", flush=True)
    plan = _plan(nside, lmax, "f32", spin=2)
    dev = torch.device("cuda")
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    eb = torch.complex(torch.randn((P, 2, plan.nalm), generator=gen, device=dev),
                       torch.randn((P, 2, plan.nalm), generator=gen, device=dev))
    f = torch.randn((P, 2, plan.npix), generator=gen, device=dev)
    lhs = float((plan.alm2map_spin(eb, dtype=torch.float32).double() * f.double()).sum())
    adj = plan.alm2map_spin_adjoint(f)
    rhs = float((torch.conj(eb.to(torch.complex128)) * adj.to(torch.complex128)).real.sum())
    print(f"lhs {lhs:.6e} rhs {rhs:.6e}")
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
