#!/usr/bin/env python
"""SHT throughput on one MI355X: nside 2048, lmax = mmax = 4095.

One step = map2alm + alm2map of a batch of B synthetic maps (one "pair"
per map) through the HIP path, inputs resident in HBM.  The headline value
is FP32-batched pairs/s; the plain FP64 drop-in (B = 1) is timed in the same
run and reported beside it, together with the FP32-vs-FP64 alm error (the
1e-5 gate), the Legendre-kernel roofline and the CPU oracle timed on a
bounded sample.  Prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SHT pairs/s nside2048 lmax4095: FP32-batched vs FP64 at ≤1e-5 rel err"
PEAK_F32 = 157.3   # TFLOP/s, MI355X FP32 (matrix / packed vector)
PEAK_F64 = 78.6    # TFLOP/s, MI355X FP64 vector


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--nside", type=int, default=2048)
    ap.add_argument("--lmax", type=int, default=None)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def synth_alm(torch, lmax, nmaps, seed, dtype, device):
    """Gaussian alm with C_l = 1/(l+1)^2 (a_l0 real), drawn on the device."""
    from hellblast_amd.plan import nalm
    n = nalm(lmax)
    ls = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)])
    ms = np.concatenate([np.full(lmax + 1 - m, m) for m in range(lmax + 1)])
    l_t = torch.as_tensor(ls, device=device)
    sd = (1.0 / (l_t.double() + 1.0))
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    re = torch.randn((nmaps, n), generator=g, device=device, dtype=torch.float64)
    im = torch.randn((nmaps, n), generator=g, device=device, dtype=torch.float64)
    m0 = torch.as_tensor(ms == 0, device=device)
    s = torch.where(m0, sd, sd / math.sqrt(2.0))
    a = torch.complex(re * s, torch.where(m0, torch.zeros_like(im), im * s))
    return a.to(dtype), l_t


def eps_metrics(torch, a, ref, l_t, lmax):
    """eps_max = max |da|/sqrt(C_l^ref), eps_rms = ||da||/||ref|| (FP64 on device)."""
    a = a.to(torch.complex128)
    ref = ref.to(torch.complex128)
    p = ref.real ** 2 + ref.imag ** 2
    cw = 2.0 * p
    cw[:lmax + 1] = p[:lmax + 1]          # m-major: the first lmax+1 entries are m = 0
    cl = torch.zeros(lmax + 1, dtype=torch.float64, device=a.device)
    cl.index_add_(0, l_t, cw)
    cl /= (2.0 * torch.arange(lmax + 1, device=a.device, dtype=torch.float64) + 1.0)
    d = (a - ref).abs()
    c = cl[l_t]
    ok = c > 0
    emax = float((d[ok] / c[ok].sqrt()).max())
    erms = float(torch.linalg.vector_norm(a - ref) / torch.linalg.vector_norm(ref))
    return emax, erms


def time_steps(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(steps):
        fn()
    t1.record()
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps


def time_stage(torch, fn, reps=3):
    """Average duration of one stage launch, HIP events on the current stream."""
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def cpu_baseline(nside, lmax, seconds):
    """FP64 NumPy oracle (vectorised over rings, loop over m) on a bounded sample.

    Ring stage: numpy FFT on a sample of rings, scaled by pixel count.
    Legendre stage: lambda recursion + contraction for sampled m, scaled by
    the triangle (lmax-m+1) x rings.  One thread (BLAS pinned to 1).
    """
    from threadpoolctl import threadpool_limits
    from oracle import geometry as OG
    from oracle import legendre as OLg
    from oracle import sht as OS

    with threadpool_limits(1):
        t = OG.ring_tables(nside)
        nr, npix = t["Nr"], t["npix"]
        rng = np.random.default_rng(0)
        f = rng.standard_normal(npix)
        # ring stage sample: every 64th ring
        rings = np.arange(0, nr, 64)
        sel_pix = int(t["nphi"][rings].sum())
        sub = {"ring_start": np.concatenate([[0], np.cumsum(t["nphi"][rings])]), "nphi": t["nphi"][rings]}
        fs = np.concatenate([f[t["ring_start"][r]:t["ring_start"][r + 1]] for r in rings])
        t0 = time.perf_counter()
        F = OS.ring_phase(fs, sub["ring_start"], sub["nphi"], lmax)
        OS.ring_synthesis(F, sub["ring_start"], sub["nphi"], lmax)
        t_ring = (time.perf_counter() - t0) * npix / sel_pix
        # Legendre sample: a few m over all rings, both directions
        x = np.cos(t["theta_centers"])
        st = np.sin(t["theta_centers"])
        wF = rng.standard_normal(nr) + 1j * rng.standard_normal(nr)
        budget = max(seconds - t_ring * sel_pix / npix, 2.0)
        done_units, spent = 0.0, 0.0
        samples = 0
        targets = list(range(0, lmax + 1, max(1, (lmax + 1) // 16)))
        gen = OLg.lambda_mm_scaled(lmax, st)
        for m, mant, ex in gen:
            if m not in targets:
                continue
            a0 = np.zeros(lmax - m + 1, dtype=complex)
            t0 = time.perf_counter()
            lam = OLg.lambda_block(m, lmax, x, mant, ex)
            a0 = lam @ wF                       # analysis contraction
            g = lam.T @ a0                      # synthesis contraction
            dt = time.perf_counter() - t0
            spent += dt
            done_units += (lmax - m + 1) * nr
            samples += 1
            if spent > budget:
                break
        per_unit = spent / done_units
        nlm = (lmax + 1) * (lmax + 2) // 2
        t_leg = per_unit * nlm * nr
        total = t_ring + t_leg
    return {
        "value": 1.0 / total, "unit": "pairs/s", "cores": 1, "kind": "port",
        "sample": (f"FP64 NumPy oracle (oracle/sht.py, oracle/legendre.py), 1 thread: ring FFTs on "
                   f"{len(rings)} of {nr} rings and lambda recursion + both contractions for {samples} of "
                   f"{lmax + 1} m over all rings, extrapolated by pixel count and by (lmax-m+1) x rings; "
                   f"{spent + t_ring * sel_pix / npix:.1f} s measured"),
        "seconds_per_pair": total,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported in this build "
              "(multi-GPU sharding is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    from hellblast_amd.hp_compat import _plan
    from hellblast_amd.plan import nalm

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    nside = args.nside
    lmax = 2 * nside - 1 if args.lmax is None else args.lmax
    B = args.batch
    npairs = 2 * nside
    nlm = nalm(lmax)
    w_dir = 4.0 * npairs * nlm          # algorithmic flops per map per direction

    log(f"plans nside={nside} lmax={lmax}")
    t0 = time.time()
    p64 = _plan(nside, lmax, lmax, "f64", "dz", "reference", "reference")
    p32 = _plan(nside, lmax, lmax, "f32", "dz", "reference", "reference")
    torch.cuda.synchronize()
    log(f"plans built in {time.time() - t0:.1f}s")

    # ---------------- inputs
    alm64, l_t = synth_alm(torch, lmax, 1, 5, torch.complex128, dev)
    map64 = p64.alm2map(alm64)                                   # [1, npix] f64
    alm32b, _ = synth_alm(torch, lmax, B, 6, torch.complex64, dev)
    maps32 = p32.alm2map(alm32b, dtype=torch.float32)            # [B, npix] f32
    torch.cuda.synchronize()

    # ---------------- accuracy gate: FP32 vs FP64 on the same input map
    # (row 0 of the timed FP32 batch, so the batched matrix-core path is checked)
    m_in = map64[0].to(torch.float32)
    maps32[0].copy_(m_in)
    a_ref = p64.map2alm(m_in.double())
    a_32 = p32.map2alm(maps32)[0]
    emax, erms = eps_metrics(torch, a_32, a_ref, l_t, lmax)
    a_32s = p32.map2alm(m_in)                                     # single map: VALU path
    emax_single, _ = eps_metrics(torch, a_32s, a_ref, l_t, lmax)
    m32 = p32.alm2map(alm64.to(torch.complex64), dtype=torch.float64)
    map_err = float((m32 - map64).abs().max() / map64.abs().max())
    log(f"FP32 vs FP64: eps_max {emax:.2e} (single-map path {emax_single:.2e}) eps_rms {erms:.2e} map {map_err:.2e}")

    # ---------------- FP32 batched (headline)
    out_alm = torch.empty((B, nlm), dtype=torch.complex64, device=dev)
    out_map = torch.empty((B, p32.npix), dtype=torch.float32, device=dev)

    def step32():
        p32.map2alm(maps32, out=out_alm)
        p32.alm2map(alm32b, dtype=torch.float32, out=out_map)

    log(f"timing FP32 batch={B} steps={args.steps}")
    ms32 = time_steps(torch, step32, args.steps, args.warmup)
    pairs32 = B * 1000.0 / ms32

    # ---------------- FP64 plain drop-in
    out64a = torch.empty((1, nlm), dtype=torch.complex128, device=dev)
    out64m = torch.empty((1, p64.npix), dtype=torch.float64, device=dev)

    def step64():
        p64.map2alm(map64, out=out64a)
        p64.alm2map(alm64, out=out64m)

    log(f"timing FP64 steps={args.fp64_steps}")
    ms64 = time_steps(torch, step64, args.fp64_steps, 1)
    pairs64 = 1000.0 / ms64

    # ---------------- per-stage kernel timing (FP32 batched)
    ph = p32.phase_buffer(B)
    ws = p32.workspace(B)
    t_ra = time_stage(torch, lambda: p32.ring_analysis(maps32, ph))
    t_la = time_stage(torch, lambda: p32.legendre_analysis(ph, B, out=out_alm, ws=ws))
    t_ls = time_stage(torch, lambda: p32.legendre_synthesis(alm32b, ph, ws))
    t_rs = time_stage(torch, lambda: p32.ring_synthesis(ph, B, dtype=torch.float32, out=out_map))
    stages = {"ring_analysis_ms": t_ra, "legendre_analysis_ms": t_la, "legendre_synthesis_ms": t_ls,
              "ring_synthesis_ms": t_rs}
    log("stages (ms):", json.dumps({k: round(v, 3) for k, v in stages.items()}))
    dom_name, dom_ms = max((("legendre_analysis", t_la), ("legendre_synthesis", t_ls)), key=lambda kv: kv[1])
    achieved = w_dir * B / (dom_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_F32, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_F32, 4), "traffic": None, "kernel": dom_name,
                "flops_per_launch": w_dir * B, "launch_ms": round(dom_ms, 4)}

    ph64 = p64.phase_buffer(1)
    ws64 = p64.workspace(1)
    t64_la = time_stage(torch, lambda: p64.legendre_analysis(p64.ring_analysis(map64, ph64), 1, out=out64a, ws=ws64), 2)
    f64_ach = w_dir / (t64_la * 1e-3) / 1e12

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(nside, lmax, args.cpu_seconds)

    res = {
        "metric": METRIC, "value": round(pairs32, 3), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms32, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"nside{nside} lmax{lmax} FP32-batched map2alm+alm2map (1 pair per map)",
                   "nside": nside, "lmax": lmax, "mmax": lmax, "batch": B,
                   "fp64_pairs_per_s": round(pairs64, 3), "fp64_ms_per_pair": round(ms64, 3),
                   "fp32_over_fp64": round(pairs32 / pairs64, 3),
                   "eps_max_f32_vs_f64": emax, "eps_rms_f32_vs_f64": erms, "eps_max_f32_single_map": emax_single,
                   "map_rel_err_f32": map_err, "gate_1e-5": bool(emax <= 1e-5),
                   "stages_ms_f32_batch": {k: round(v, 3) for k, v in stages.items()},
                   "fp64_ring+legendre_analysis_tflops": round(f64_ach, 3),
                   "fp64_frac_of_78.6": round(f64_ach / PEAK_F64, 4),
                   "algorithmic_flops_per_map_per_direction": w_dir},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
