#!/usr/bin/env python
"""SHT throughput on one MI355X: nside 2048, lmax = mmax = 4095.

One step = map2alm + alm2map of a batch of B synthetic maps (one "pair"
per map) through the HIP path, inputs resident in HBM.  The headline value
is FP32-batched pairs/s; the plain FP64 drop-in (B = 1) is timed in the same
run and reported beside it, together with the FP32-vs-FP64 alm error (the
1e-5 gate), the roofline of the dominant kernel (its mean launch time from
HIP events recorded around it on its own stream over the timed steps), the
ring stage's HBM rate and the CPU oracle timed on a bounded sample.  Prints
ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B (32)]
    python bench.py --spin      (spin-2: 4 Q/U pairs FP32 vs 1 pair FP64)
    python bench.py --mixed     (64 maps of mixed nside in one HIP graph)
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
PEAK_F16 = 2500.0  # TFLOP/s, MI355X FP16 matrix, dense
PEAK_F64 = 78.6    # TFLOP/s, MI355X FP64 vector


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--nside", type=int, default=2048)
    ap.add_argument("--lmax", type=int, default=None)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-full", action="store_true",
                    help="time the CPU oracle at C2 (whole pair) and the target (every 16th m); no GPU")
    ap.add_argument("--configs", action="store_true",
                    help="C1 (nside 64, lmax 128) and C2 (nside 1024, lmax 2047) single-map figures")
    ap.add_argument("--spin", action="store_true",
                    help="spin-2 workload: 4 Q/U pairs FP32 vs 1 pair FP64 at nside 2048, lmax 4095")
    ap.add_argument("--mixed-mode", default="auto", choices=["auto", "persistent", "grouped"],
                    help="with --mixed: one persistent Legendre launch per direction over every group, one "
                         "batched call per group, or auto: time both and report the faster")
    ap.add_argument("--mixed", action="store_true",
                    help="instead of the headline: 64 FP32 maps of mixed nside (128..2048, lmax 2 nside) "
                         "in one HIP graph with autotuned tiles; prints its own JSON line")
    ap.add_argument("--mixed-workload", default="nside128-2048", choices=["nside128-2048", "groups16"],
                    help="with --mixed: the 64-map batch, or 16 FP32 maps at one nside in 16 (nside, lmax) "
                         "groups, lmax = nside (1 + k/16), k = 0..15 (one map each)")
    ap.add_argument("--mixed-nside", type=int, default=128,
                    help="with --mixed-workload groups16: the maps' nside (a power of two, at least 16)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB -> 1 GiB elementwise copy kernel (HBM counter calibration)")
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


def time_steps(torch, fn, steps, warmup, timer=None):
    """ms per step over exactly ``steps`` steps (after ``warmup``), bracketed by
    device synchronisation; ``timer`` (a KernelTimer) is armed over the same
    timed region."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    if timer is not None:
        timer.__enter__()
    t0.record()
    for _ in range(steps):
        fn()
    t1.record()
    torch.cuda.synchronize()
    if timer is not None:
        timer.__exit__(None, None, None)
    return t0.elapsed_time(t1) / steps


PROFILE_DIR = os.path.join(ROOT, "profiles")


def measured_traffic(kernel, workload):
    """HBM bytes per launch of ``kernel`` on ``workload`` from the committed
    rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration), or None when no summary covers this workload."""
    path = os.path.join(PROFILE_DIR, f"{kernel}_{workload}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def _cores():
    """Host cores visible to this process (BLAS uses all of them)."""
    return {"os_cpu_count": os.cpu_count(), "sched_getaffinity": len(os.sched_getaffinity(0))}


def _nlm(lmax):
    return (lmax + 1) * (lmax + 2) // 2


# literal CPU fallback sizes: (nside, m stride of the analysis, ring stride of
# the synthesis); lmax = 2 nside, the first is C1 (nside 64, lmax 128)
LITERAL_SIZES = ((64, 1, 2), (128, 1, 16), (256, 2, 128))


def cpu_literal(nside_t, lmax_t):
    """The reference's CPU fallback as written: sht_fallback.map2alm_gpu
    (sht_fallback.py:36-65) and isht.alm2map_gpu (isht.py:5-26), restated
    statement by statement with NumPy for CuPy (oracle/literal.py), timed at
    nside 64 (C1), 128 and 256 with lmax = 2 nside.  The analysis visits every
    m (every 2nd at nside 256, scaled by the m count); the synthesis, whose
    per-ring loop dominates, runs on every k-th ring (scaled by the ring
    count).  Per direction t = c N_lm N_r is fitted by least squares through
    the origin and extrapolated to the target: an extrapolation, not a
    measurement at the target."""
    from oracle import geometry as OG
    from oracle import literal as OL
    rows, u, ta, ts = [], [], [], []
    old = np.seterr(all="ignore")          # the literal P_mm overflows above m = 150
    try:
        for nside, mst, rst in LITERAL_SIZES:
            lmax = 2 * nside
            t = OG.ring_tables(nside)
            rs, nphi, th = t["ring_start"], np.asarray(t["nphi"]), np.asarray(t["theta_centers"])
            nr = t["Nr"]
            f = np.random.default_rng(1).standard_normal(t["npix"])
            rings = [f[rs[r]:rs[r + 1]] for r in range(nr)]
            t0 = time.perf_counter()
            alm = OL.sht_fallback_map2alm(rings, th, nphi, t["weights_dz"], lmax, norm=OL.alm_norm_safe,
                                          m_stride=mst)
            t_a = (time.perf_counter() - t0) * (lmax + 1) / len(range(0, lmax + 1, mst))
            sub = np.arange(rst // 2, nr, rst)
            alm = np.random.default_rng(2).standard_normal(alm.size) + 0j
            t0 = time.perf_counter()
            OL.isht_alm2map(alm, th[sub], nphi[sub], lmax, norm=OL.alm_norm_safe)
            t_s = (time.perf_counter() - t0) * nr / len(sub)
            rows.append({"nside": nside, "lmax": lmax, "map2alm_s": round(t_a, 3), "alm2map_s": round(t_s, 3),
                         "map2alm_m_stride": mst, "alm2map_ring_stride": rst})
            u.append(_nlm(lmax) * nr)
            ta.append(t_a)
            ts.append(t_s)
    finally:
        np.seterr(**old)
    u = np.asarray(u, dtype=np.float64)
    ca = float(np.dot(u, ta) / np.dot(u, u))
    cs = float(np.dot(u, ts) / np.dot(u, u))
    ut = _nlm(lmax_t) * (4 * nside_t - 1)
    t_pair = (ca + cs) * ut
    return {"pairs_per_s_extrapolated": 1.0 / t_pair, "seconds_per_pair_extrapolated": round(t_pair, 1),
            "fit": "t = c N_lm N_r per direction, least squares through the origin",
            "c_map2alm_s": ca, "c_alm2map_s": cs, "measured": rows, "cores": 1,
            "note": "literal sht_fallback + isht restatement (oracle/literal.py); numpy elementwise, one thread"}


def cpu_vectorised(nside, lmax, seconds, m_stride=8, ring_stride=64):
    """The FP64 NumPy oracle's pair (oracle/sht.py map2alm + alm2map: a loop
    over m, NumPy over rings, BLAS for the contractions) with BLAS on every
    core.  The ring stage runs on every ring_stride-th ring (scaled by pixel
    count); the Legendre stage (the lambda recursion once per direction, as
    the oracle's analysis and synthesis each run it, plus both contractions)
    on every m_stride-th m in increasing m until ``seconds`` are spent,
    scaled per (l, ring) unit.  Strides of 1 and no time limit are the whole
    pair.  Logs progress every 30 s."""
    from oracle import geometry as OG
    from oracle import legendre as OLg
    from oracle import sht as OS
    t = OG.ring_tables(nside)
    nr, npix = t["Nr"], t["npix"]
    rng = np.random.default_rng(0)
    f = rng.standard_normal(npix)
    rings = np.arange(0, nr, ring_stride)
    sel_pix = int(t["nphi"][rings].sum())
    sub_start = np.concatenate([[0], np.cumsum(t["nphi"][rings])])
    fs = np.concatenate([f[t["ring_start"][r]:t["ring_start"][r + 1]] for r in rings])
    t0 = time.perf_counter()
    F = OS.ring_phase(fs, sub_start, t["nphi"][rings], lmax)
    OS.ring_synthesis(F, sub_start, t["nphi"][rings], lmax)
    t_ring_s = time.perf_counter() - t0
    t_ring = t_ring_s * npix / sel_pix
    x = np.cos(t["theta_centers"])
    st = np.sin(t["theta_centers"])
    wF = rng.standard_normal(nr) + 1j * rng.standard_normal(nr)
    done, spent, samples, last = 0.0, 0.0, 0, time.perf_counter()
    for m, mant, ex in OLg.lambda_mm_scaled(lmax, st):
        if m % m_stride:
            continue
        t0 = time.perf_counter()
        lam = OLg.lambda_block(m, lmax, x, mant, ex)
        a0 = lam @ wF                           # analysis contraction
        lam = OLg.lambda_block(m, lmax, x, mant, ex)
        lam.T @ a0                              # synthesis contraction (its own recursion)
        t1 = time.perf_counter()
        spent += t1 - t0
        done += (lmax - m + 1) * nr
        samples += 1
        if t1 - last > 30.0:
            log(f"  cpu oracle nside {nside}: m = {m}, {spent:.0f} s")
            last = t1
        if spent > seconds:
            break
    t_leg = spent / done * _nlm(lmax) * nr
    whole = m_stride == 1 and ring_stride == 1 and samples == lmax + 1
    return {"seconds_per_pair": t_ring + t_leg, "measured_s": round(spent + t_ring_s, 1),
            "ring_stage_s": round(t_ring, 2), "legendre_stage_s": round(t_leg, 2),
            "sample": (f"whole pair at nside {nside}, lmax {lmax}" if whole else
                       f"ring FFTs on {len(rings)} of {nr} rings (scaled by pixel count); lambda recursion "
                       f"(once per direction) + both contractions for {samples} m (every {m_stride}th from 0) "
                       f"over all rings, scaled by (lmax-m+1) x rings")}


def cpu_baseline(nside, lmax, seconds):
    """cpu_baseline for the JSON line: the vectorised FP64 oracle on a bounded
    sample at the target (value), with the literal reference fallback's fitted
    extrapolation beside it."""
    # 32 m spread over 0..lmax: short rows (large m) cost more per (l, ring)
    # unit than long ones, so a sample of small m only would flatter the CPU
    v = cpu_vectorised(nside, lmax, seconds, m_stride=max(1, (lmax + 1) // 32))
    lit = cpu_literal(nside, lmax)
    cores = _cores()
    return {
        "value": 1.0 / v["seconds_per_pair"], "unit": "pairs/s", "cores": cores["sched_getaffinity"],
        "kind": "port",
        "sample": ("FP64 NumPy oracle (oracle/sht.py, oracle/legendre.py), BLAS on all cores, numpy elementwise "
                   "single-threaded: " + v["sample"] + f"; {v['measured_s']} s measured"),
        "seconds_per_pair": round(v["seconds_per_pair"], 1), **cores,
        "literal_reference_fallback": lit,
    }


def cpu_spin_baseline(nside, lmax, seconds, m_stride=8):
    """cpu_baseline of the spin-2 line: the FP64 NumPy oracle's Q/U pair
    (oracle/spin.py map2alm_spin + alm2map_spin), BLAS on every core, on a
    bounded sample: Q and U ring FFTs on every 64th ring (scaled by pixel
    count); for every m_stride-th m (in increasing m until ``seconds`` are
    spent) the +2 and -2 recursions once per direction, as the oracle's
    analysis and synthesis each run them, plus the four contractions of each
    direction, scaled per (l, ring) unit."""
    from oracle import geometry as OG
    from oracle import sht as OS
    from oracle import spin as OSp
    t = OG.ring_tables(nside)
    nr, npix = t["Nr"], t["npix"]
    rng = np.random.default_rng(0)
    rings = np.arange(0, nr, 64)
    sel_pix = int(t["nphi"][rings].sum())
    sub_start = np.concatenate([[0], np.cumsum(t["nphi"][rings])])
    t0 = time.perf_counter()
    for _ in range(2):                                  # Q and U
        fs = rng.standard_normal(sel_pix)
        F = OS.ring_phase(fs, sub_start, t["nphi"][rings], lmax)
        OS.ring_synthesis(F, sub_start, t["nphi"][rings], lmax)
    t_ring_s = time.perf_counter() - t0
    t_ring = t_ring_s * npix / sel_pix
    th = np.asarray(t["theta_centers"], dtype=np.float64)
    x = np.cos(th)
    fq = rng.standard_normal(nr) + 1j * rng.standard_normal(nr)
    fu = rng.standard_normal(nr) + 1j * rng.standard_normal(nr)
    done, spent, samples = 0.0, 0.0, 0
    for (m, mp_, ep), (_, mm_, em) in zip(OSp._seed_chain(2, lmax, th), OSp._seed_chain(-2, lmax, th)):
        if m % m_stride:
            continue
        t0 = time.perf_counter()
        for direction in range(2):
            lp = OSp._recur(2, m, m, lmax, x, mp_, ep)
            lm = OSp._recur(-2, m, m, lmax, x, mm_, em)
            lplus, lminus = 0.5 * (lp + lm), 0.5 * (lp - lm)
            if direction == 0:
                E = -(lplus @ fq + 1j * (lminus @ fu))
                B = -(lplus @ fu - 1j * (lminus @ fq))
            else:
                lplus.T @ E + 1j * (lminus.T @ B)
                lplus.T @ B - 1j * (lminus.T @ E)
        spent += time.perf_counter() - t0
        done += (lmax - m + 1) * nr
        samples += 1
        if spent > seconds:
            break
    t_leg = spent / done * _nlm(lmax) * nr
    total = t_ring + t_leg
    cores = _cores()
    return {"value": 1.0 / total, "unit": "pairs/s", "cores": cores["sched_getaffinity"], "kind": "port",
            "sample": (f"FP64 NumPy spin oracle (oracle/spin.py), BLAS on all cores, numpy elementwise "
                       f"single-threaded, one Q/U pair: Q and U ring FFTs on {len(rings)} of {nr} rings (scaled "
                       f"by pixel count); +-2 recursions (once per direction) and the contractions for {samples} m "
                       f"(every {m_stride}th from 0) over all rings, scaled by (lmax-m+1) x rings; "
                       f"{spent + t_ring_s:.1f} s measured"),
            "seconds_per_pair": round(total, 1), **cores}


def cpu_full_main(args):
    """--cpu-full: the vectorised oracle's whole pair at C2 (nside 1024, lmax
    2047) and, at the target, every 16th m over all rings (a full target pair
    takes about an hour of NumPy); no GPU."""
    nside = args.nside
    lmax = 2 * nside - 1 if args.lmax is None else args.lmax
    log("vectorised oracle: full pair at nside 1024, lmax 2047")
    c2 = cpu_vectorised(1024, 2047, float("inf"), m_stride=1, ring_stride=1)
    log(f"  {c2['seconds_per_pair']:.1f} s")
    log(f"vectorised oracle: target nside {nside}, lmax {lmax}, every 16th m")
    tgt = cpu_vectorised(nside, lmax, float("inf"), m_stride=16, ring_stride=1)
    log(f"  {tgt['seconds_per_pair']:.1f} s per pair (extrapolated from {tgt['measured_s']} s)")
    log("literal fallback: nside 64/128/256")
    lit = cpu_literal(nside, lmax)
    print(json.dumps({"cpu_full": True, **_cores(), "vectorised_c2_full": c2, "vectorised_target": tgt,
                      "literal_reference_fallback": lit}), flush=True)


def configs_main(args):
    """--configs: throughput at C1 (nside 64, lmax 128, one FP64 map: the
    reference's CPU-runnable case) and C2 (nside 1024, lmax 2047, one map,
    FP64 and FP32), with C2's FP32-vs-FP64 errors and the CPU baselines at C1
    (the literal reference fallback and the vectorised oracle, whole pairs).
    The accuracy against the oracle is checked by tests/ (C1:
    test_config1_roundtrip, C2: test_config2_single_map_f32_vs_f64)."""
    import torch
    from hellblast_amd.hp_compat import _plan
    from hellblast_amd.plan import alm2cl, nalm
    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    out = {"configs": True}

    def pair_ms(plan, maps, alm, dtype, reps):
        om = torch.empty_like(maps)
        oa = torch.empty_like(alm)
        def step():
            plan.map2alm(maps, out=oa)
            plan.alm2map(alm, dtype=dtype, out=om)
        return time_steps(torch, step, reps, 2)

    # C1
    nside, lmax = 64, 128
    p64 = _plan(nside, lmax, lmax, "f64", "dz", "reference", "reference")
    alm, l_t = synth_alm(torch, lmax, 1, 1, torch.complex128, dev)
    alm[:, 0] = 0
    maps = p64.alm2map(alm)
    ms = pair_ms(p64, maps, alm, torch.float64, 50)
    rt = p64.map2alm(maps)
    emax, erms = eps_metrics(torch, rt[0], alm[0], l_t, lmax)
    out["C1"] = {"nside": nside, "lmax": lmax, "fp64_pairs_per_s": round(1000.0 / ms, 1), "ms_per_pair": round(ms, 4),
                 "round_trip_eps_max": emax, "round_trip_eps_rms": erms}
    log(f"C1: {1000.0 / ms:.1f} FP64 pairs/s, round trip eps_max {emax:.2e}")
    if not args.no_cpu:
        from oracle import geometry as OG
        from oracle import literal as OL
        from oracle import sht as OS
        t = OG.ring_tables(nside)
        f = maps[0].cpu().numpy()
        rings = [f[t["ring_start"][r]:t["ring_start"][r + 1]] for r in range(t["Nr"])]
        a_np = alm[0].cpu().numpy()
        old = np.seterr(all="ignore")
        t0 = time.perf_counter()
        OL.sht_fallback_map2alm(rings, t["theta_centers"], t["nphi"], t["weights_dz"], lmax, norm=OL.alm_norm_safe)
        OL.isht_alm2map(OS.mmajor_to_ref(a_np, lmax), t["theta_centers"], t["nphi"], lmax, norm=OL.alm_norm_safe)
        t_lit = time.perf_counter() - t0
        np.seterr(**old)
        t0 = time.perf_counter()
        OS.alm2map(OS.map2alm(f, t, lmax), t, lmax)
        t_vec = time.perf_counter() - t0
        out["C1"]["cpu"] = {"literal_fallback_s_per_pair": round(t_lit, 3), "vectorised_oracle_s_per_pair": round(t_vec, 3),
                            **_cores(), "note": "whole pairs; literal: sht_fallback + isht restated, one thread; "
                                                "vectorised: BLAS on all cores"}
        log(f"C1 CPU: literal {t_lit:.2f} s, vectorised {t_vec:.2f} s per pair")
    # C2
    nside, lmax = 1024, 2047
    p64 = _plan(nside, lmax, lmax, "f64", "dz", "reference", "reference")
    p32 = _plan(nside, lmax, lmax, "f32", "dz", "reference", "reference")
    n = nalm(lmax)
    ls = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)])
    l_t = torch.as_tensor(ls, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(2)
    lf = l_t.double()
    sd = torch.where(l_t >= 2, 1.0 / torch.sqrt((lf * (lf + 1.0)).clamp(min=1.0)), torch.zeros_like(lf))
    m0 = torch.as_tensor(np.concatenate([np.full(lmax + 1 - m, m) for m in range(lmax + 1)]) == 0, device=dev)
    s_ = torch.where(m0, sd, sd / math.sqrt(2.0))
    re = torch.randn((1, n), generator=g, device=dev, dtype=torch.float64)
    im = torch.randn((1, n), generator=g, device=dev, dtype=torch.float64)
    alm = torch.complex(re * s_, torch.where(m0, torch.zeros_like(im), im * s_))
    maps = p64.alm2map(alm) + 0.01 * torch.randn((1, p64.npix), generator=g, device=dev, dtype=torch.float64)
    ms64 = pair_ms(p64, maps, alm, torch.float64, 10)
    ms32 = pair_ms(p32, maps.float(), alm.to(torch.complex64), torch.float32, 10)
    a64 = p64.map2alm(maps)[0]
    a32 = p32.map2alm(maps)[0]
    emax, erms = eps_metrics(torch, a32, a64, l_t, lmax)
    c64, c32 = alm2cl(a64, lmax), alm2cl(a32.to(torch.complex128), lmax)
    ecl = float(((c32[2:] - c64[2:]).abs() / c64[2:]).max())
    m64 = p64.alm2map(a64)
    m32 = p32.alm2map(a32, dtype=torch.float64)
    emap = float((m32 - m64).abs().max() / m64.abs().max())
    rt = p64.map2alm(p64.alm2map(alm))
    rt_max, rt_rms = eps_metrics(torch, rt[0], alm[0], l_t, lmax)
    out["C2"] = {"nside": nside, "lmax": lmax, "fp64_pairs_per_s": round(1000.0 / ms64, 2),
                 "fp32_single_map_pairs_per_s": round(1000.0 / ms32, 2), "fp32_over_fp64": round(ms64 / ms32, 3),
                 "eps_max_f32_vs_f64": emax, "eps_rms_f32_vs_f64": erms, "cl_max_rel_f32_vs_f64": ecl,
                 "map_rel_f32_vs_f64": emap, "fp64_round_trip_eps_max": rt_max, "fp64_round_trip_eps_rms": rt_rms}
    log(f"C2: FP64 {1000.0 / ms64:.2f} pairs/s, FP32 {1000.0 / ms32:.2f} pairs/s, eps_max {emax:.2e}")
    print(json.dumps(out), flush=True)


def mixed_main(args):
    """64 FP32 maps, nside {2048:1, 1024:2, 512:5, 256:14, 128:42} (seeded
    shuffle), lmax = 2 nside, alm ~ C_l = (l+10)^-2: map2alm + alm2map of the
    whole batch per step, replayed from one HIP graph (MixedBatch).  With
    --mixed-workload groups16: 16 maps at nside N (--mixed-nside, default
    128), lmax N + k N/16 (16 groups of one map), the same spectrum."""
    import numpy as np
    import torch
    from hellblast_amd.batch import MixedBatch
    from hellblast_amd.hp_compat import _plan
    from hellblast_amd.plan import KernelTimer, nalm
    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(3)
    many = args.mixed_workload == "groups16"
    alloc = {2048: 1, 1024: 2, 512: 5, 256: 14, 128: 42}
    ns = args.mixed_nside
    if many:
        if ns < 16 or ns & (ns - 1):
            raise SystemExit("--mixed-nside must be a power of two, at least 16")
        sides = [ns] * 16
        specs = [(ns, ns + ns // 16 * k) for k in range(16)]
    else:
        sides = [n for n, k in alloc.items() for _ in range(k)]
        rng.shuffle(sides)
        specs = [(n, 2 * n) for n in sides]
    t0 = time.time()
    modes = ["persistent", "grouped"] if args.mixed_mode == "auto" else [args.mixed_mode]
    maps = []
    for i, (nside, lmax) in enumerate(specs):
        n = nalm(lmax)
        ls = torch.cat([torch.arange(m, lmax + 1, device=dev) for m in range(lmax + 1)])
        m0 = torch.zeros(n, dtype=torch.bool, device=dev)
        m0[:lmax + 1] = True
        g = torch.Generator(device=dev)
        g.manual_seed(1000 + i)
        sd = 1.0 / (ls.double() + 10.0)
        sc = torch.where(m0, sd, sd / math.sqrt(2.0))
        re = torch.randn(n, generator=g, device=dev, dtype=torch.float64)
        im = torch.randn(n, generator=g, device=dev, dtype=torch.float64)
        alm = torch.complex(re * sc, torch.where(m0, torch.zeros_like(im), im * sc))
        maps.append(_plan(nside, lmax, lmax, "f64", "dz", "reference", "reference").alm2map(alm[None])[0].float())

    def timed(fn, k):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(k):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    runs = {}
    for mode in modes:
        batch = MixedBatch(specs, precision="f32", tune=True, mode=mode)
        log(f"mixed batch of {len(specs)} maps ({mode}) planned and tuned in {time.time() - t0:.1f}s: "
            f"{batch.tiles} {getattr(batch, 'multi_tiles', '')}")
        batch.set_maps(maps)
        batch.capture()
        torch.cuda.synchronize()
        r = {"graph_ms": timed(batch.replay, args.steps), "eager_ms": timed(batch.run_eager, args.steps)}
        if batch.multi is not None:
            info = batch.multi.info
            with KernelTimer("legendre_analysis", "legendre_synthesis") as kt:
                for _ in range(args.steps):
                    batch.run_eager()
                torch.cuda.synchronize()
            r["timers"] = {"persistent_analysis_ms": round(kt.mean_ms("legendre_analysis"), 3),
                           "persistent_synthesis_ms": round(kt.mean_ms("legendre_synthesis"), 3),
                           "analysis_tiles": info["analysis_tiles"], "synthesis_tiles": info["synthesis_tiles"],
                           "analysis_workgroups": info["analysis_workgroups"],
                           "synthesis_workgroups": info["synthesis_workgroups"]}
            r["tiles"] = batch.multi_tiles
        else:
            r["tiles"] = {f"nside{k[0]}_lmax{k[1]}": v for k, v in batch.tiles.items()}
        worst = 0.0
        for i, (nside, lmax) in enumerate(specs):
            p64 = _plan(nside, lmax, lmax, "f64", "dz", "reference", "reference")
            ref = p64.map2alm(batch.map(i).double()[None])[0]
            ls = torch.cat([torch.arange(m, lmax + 1, device=dev) for m in range(lmax + 1)])
            emax, _ = eps_metrics(torch, batch.alm_of(i), ref, ls, lmax)
            worst = max(worst, emax)
        r["eps"] = worst
        log(f"{mode}: graph {r['graph_ms']:.3f} ms, eager {r['eager_ms']:.3f} ms, eps_max {worst:.2e} "
            f"{r.get('timers', '')}")
        runs[mode] = r
        del batch
        torch.cuda.empty_cache()
    # each group alone: Legendre kernels of its plan (HIP events), with the
    # VALU kernels (4 slots, the persistent tiles' arithmetic) and with the
    # plan's default kernels (matrix cores for 3+ maps)
    groups = {}
    for nside, count in ({} if many else alloc).items():
        lmax = 2 * nside
        p = _plan(nside, lmax, lmax, "f32", "dz", "reference", "reference")
        mp = torch.randn((count, p.npix), device=dev)
        row = {"maps": count}
        for name, q in (("valu", p.with_tiles(valu_slots=4, mfma_min_maps=100000)), ("default", p)):
            a = q.map2alm(mp)
            o = torch.empty_like(mp)
            with KernelTimer("legendre_analysis", "legendre_synthesis") as kt:
                for _ in range(args.steps):
                    q.map2alm(mp, out=a)
                    q.alm2map(a, dtype=torch.float32, out=o)
                torch.cuda.synchronize()
            row[f"{name}_analysis_ms"] = round(kt.mean_ms("legendre_analysis"), 3)
            row[f"{name}_synthesis_ms"] = round(kt.mean_ms("legendre_synthesis"), 3)
        groups[f"nside{nside}_lmax{lmax}"] = row
    log(f"groups alone: {groups}")
    best = min(runs, key=lambda k: runs[k]["graph_ms"])
    rb = runs[best]
    npix = sum(12 * n * n for n in sides)
    metric = (f"mixed-batch SHT pairs/s: 16 FP32 maps, nside {ns}, 16 groups lmax {specs[0][1]}..{specs[-1][1]}, "
              "one HIP graph" if many else
              "mixed-batch SHT pairs/s: 64 FP32 maps, nside 128..2048, lmax 2 nside, one HIP graph")
    workload = (f"16 maps nside {ns}, lmax {ns}+{ns // 16}k (k=0..15, one map per group), C_l=(l+10)^-2, "
                "map2alm+alm2map per map" if many else
                "64 maps nside {2048:1,1024:2,512:5,256:14,128:42}, lmax 2 nside, C_l=(l+10)^-2, map2alm+alm2map "
                "per map")
    res = {"metric": metric,
           "value": round(len(specs) / (rb["graph_ms"] / 1e3), 3), "unit": "pairs/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(rb["graph_ms"], 3),
           "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": workload,
                      "pixels_per_step": npix, "mode": best, "eager_ms_per_step": round(rb["eager_ms"], 3),
                      "graph_over_eager": round(rb["eager_ms"] / rb["graph_ms"], 3), "eps_max_worst_map": rb["eps"],
                      "tiles": rb["tiles"],
                      "modes": {k: {"graph_ms_per_step": round(v["graph_ms"], 3),
                                    "eager_ms_per_step": round(v["eager_ms"], 3), "eps_max_worst_map": v["eps"],
                                    **v.get("timers", {})} for k, v in runs.items()},
                      "groups_alone": groups}}
    print(json.dumps(res), flush=True)


def spin_main(args):
    """Spin-2 Q/U <-> E/B at nside 2048, lmax 4095: one step = map2alm_spin +
    alm2map_spin of 4 Q/U pairs in FP32 (FP32 maps, alm and contraction on
    the matrix cores; FP64 recursion), the FP64 drop-in (1 pair) timed
    beside it.  E/B drawn with C_l^EE = 1/(l+1)^2 (l >= 2), C_l^BB = 0.1
    C_l^EE; Q/U = FP64 synthesis rounded to FP32; eps_max of E and B against
    the FP64 analysis of the same FP32 maps."""
    import torch
    from hellblast_amd.hp_compat import _plan
    from hellblast_amd.plan import KernelTimer, nalm
    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    nside = args.nside
    lmax = 2 * nside - 1 if args.lmax is None else args.lmax
    npol = 4
    t0 = time.time()
    p64 = _plan(nside, lmax, lmax, "f64", "dz", "reference", "reference", 2)
    p32 = _plan(nside, lmax, lmax, "f32", "dz", "reference", "reference", 2)
    torch.cuda.synchronize()
    log(f"spin plans built in {time.time() - t0:.1f}s")
    n = nalm(lmax)
    rows = torch.arange(lmax + 1, 0, -1, device=dev)
    ms_ = torch.repeat_interleave(torch.arange(lmax + 1, device=dev), rows)
    ls = torch.arange(n, device=dev) - torch.repeat_interleave(torch.cumsum(rows, 0) - rows, rows) + ms_
    g = torch.Generator(device=dev)
    g.manual_seed(21)

    def draw(amp):
        sd = torch.where(ls < 2, torch.zeros(n, device=dev, dtype=torch.float64), amp / (ls.double() + 1.0))
        sc = torch.where(ms_ == 0, sd, sd / math.sqrt(2.0))
        re = torch.randn(n, generator=g, device=dev, dtype=torch.float64)
        im = torch.randn(n, generator=g, device=dev, dtype=torch.float64)
        return torch.complex(re * sc, torch.where(ms_ == 0, torch.zeros_like(im), im * sc))

    eb = torch.stack([torch.stack([draw(1.0), draw(math.sqrt(0.1))]) for _ in range(npol)])
    qu = p64.alm2map_spin(eb).to(torch.float32)                  # [4, 2, npix]
    eb32 = eb.to(torch.complex64)
    got = p32.map2alm_spin(qu)
    eps = {"E": 0.0, "B": 0.0}
    for b in range(npol):
        ref = p64.map2alm_spin(qu[b].double())
        for k, name in enumerate("EB"):
            eps[name] = max(eps[name], eps_metrics(torch, got[b, k], ref[k], ls, lmax)[0])
    del got, ref
    log(f"FP32 vs FP64: eps_max E {eps['E']:.2e} B {eps['B']:.2e}")

    out_eb = torch.empty((npol, 2, n), dtype=torch.complex64, device=dev)
    out_qu = torch.empty_like(qu)

    def step32():
        p32.map2alm_spin(qu, out=out_eb)
        p32.alm2map_spin(eb32, dtype=torch.float32, out=out_qu)

    kt = KernelTimer("spin_analysis", "spin_synthesis")
    ms32 = time_steps(torch, step32, args.steps, args.warmup, kt)
    kms = {k: kt.mean_ms(k) for k in kt.kernels}
    q64 = qu[:1].double()
    out64a = torch.empty((1, 2, n), dtype=torch.complex128, device=dev)
    out64m = torch.empty((1, 2, p64.npix), dtype=torch.float64, device=dev)

    def step64():
        p64.map2alm_spin(q64, out=out64a)
        p64.alm2map_spin(eb[:1], out=out64m)

    kt64 = KernelTimer("spin_analysis", "spin_synthesis")
    ms64 = time_steps(torch, step64, args.fp64_steps, 1, kt64)
    kms64 = {k: kt64.mean_ms(k) for k in kt64.kernels}
    log("kernel means (ms): FP32", json.dumps({k: round(v, 3) for k, v in kms.items()}),
        "FP64", json.dumps({k: round(v, 3) for k, v in kms64.items()}))
    # algorithmic flops: 16 per (ring pair, l, m) per Q/U pair and direction;
    # the FP32 analysis kernel covers m >= 8 (the band m < 8 runs in FP64)
    npairs = 2 * nside
    mb = p32.fp64_band_m
    band = mb * (2 * lmax + 1 - mb) // 2 + mb
    flops = {"spin_analysis": 16.0 * npairs * (n - band) * npol, "spin_synthesis": 16.0 * npairs * n * npol}
    tf = {k: flops[k] / (kms[k] * 1e-3) / 1e12 for k in flops}
    dom = max(flops, key=lambda k: kms[k])
    names = {"spin_analysis": "spin_analysis_mfma_kernel", "spin_synthesis": "spin_synthesis_mfma_kernel"}
    pairs32, pairs64 = npol * 1000.0 / ms32, 1000.0 / ms64
    traffic = measured_traffic(names[dom], f"spin_nside{nside}_lmax{lmax}_b{npol}")
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (spin oracle)")
        cpu = cpu_spin_baseline(nside, lmax, args.cpu_seconds)
    res = {"metric": "spin-2 SHT pairs/s nside2048 lmax4095 (Q/U pairs): FP32-batched vs FP64",
           "value": round(pairs32, 3), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms32, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"nside{nside} lmax{lmax} spin-2: 4 Q/U pairs map2alm_spin+alm2map_spin, "
                                  "C_l^EE=1/(l+1)^2, C_l^BB=0.1 C_l^EE",
                      "fp64_pairs_per_s": round(pairs64, 3), "fp32_over_fp64": round(pairs32 / pairs64, 3),
                      "eps_max_E": eps["E"], "eps_max_B": eps["B"], "gate_1e-5": bool(max(eps.values()) <= 1e-5),
                      "kernel_ms_f32": {k: round(v, 4) for k, v in kms.items()},
                      "kernel_ms_f64": {k: round(v, 4) for k, v in kms64.items()},
                      "tflops_f32": {k: round(v, 3) for k, v in tf.items()}},
           "roofline": {"bound": "mfma", "achieved": round(tf[dom], 3), "peak": PEAK_F32, "unit": "TFLOP/s",
                        "frac": round(tf[dom] / PEAK_F32, 4), "traffic": traffic, "kernel": names[dom],
                        "flops_per_launch": flops[dom], "launch_ms": round(kms[dom], 4)},
           "cpu_baseline": cpu}
    print(json.dumps(res), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported in this build "
              "(this build runs one GPU per process; batches shard by map across processes)", file=sys.stderr)
        sys.exit(2)
    if args.cpu_full:
        return cpu_full_main(args)
    if args.configs:
        return configs_main(args)
    if args.mixed:
        return mixed_main(args)
    if args.spin:
        return spin_main(args)
    import torch
    from hellblast_amd import _abi
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

    # ---------------- accuracy gate on the timed batched calls, every row:
    # FP32 analysis of each row against the FP64 analysis of the same FP32 map
    # (row 0 is the FP64 map rounded to FP32), FP32 synthesis of each alm row
    # against the FP64 synthesis of the same complex64 alm
    m_in = map64[0].to(torch.float32)
    maps32[0].copy_(m_in)
    a32b = p32.map2alm(maps32)
    emax, erms = 0.0, 0.0
    for b in range(B):
        a_ref = p64.map2alm(maps32[b].double())
        e_b, r_b = eps_metrics(torch, a32b[b], a_ref, l_t, lmax)
        emax, erms = max(emax, e_b), max(erms, r_b)
    del a32b
    a_ref = p64.map2alm(m_in.double())
    a_32s = p32.map2alm(m_in)                                     # single map: VALU path
    emax_single, _ = eps_metrics(torch, a_32s, a_ref, l_t, lmax)
    m32b = p32.alm2map(alm32b, dtype=torch.float32)
    map_err = 0.0
    for b in range(B):
        ref = p64.alm2map(alm32b[b].to(torch.complex128))
        map_err = max(map_err, float((m32b[b].double() - ref).abs().max() / ref.abs().max()))
    del m32b, ref
    log(f"FP32 vs FP64 (worst of {B} rows): eps_max {emax:.2e} (single-map path {emax_single:.2e}) "
        f"eps_rms {erms:.2e} map {map_err:.2e}")

    # ---------------- FP32 batched (headline)
    from hellblast_amd.plan import KernelTimer
    out_alm = torch.empty((B, nlm), dtype=torch.complex64, device=dev)
    out_map = torch.empty((B, p32.npix), dtype=torch.float32, device=dev)

    def step32():
        p32.map2alm(maps32, out=out_alm)
        p32.alm2map(alm32b, dtype=torch.float32, out=out_map)

    log(f"timing FP32 batch={B} steps={args.steps}")
    kt = KernelTimer("legendre_analysis", "legendre_synthesis", "ring_analysis", "ring_synthesis")
    ms32 = time_steps(torch, step32, args.steps, args.warmup, kt)
    pairs32 = B * 1000.0 / ms32
    kms = {k: kt.mean_ms(k) for k in kt.kernels}
    log("kernel means over the timed region (ms):", json.dumps({k: round(v, 3) for k, v in kms.items()}),
        "launches:", json.dumps({k: v[0] for k, v in kt.result.items()}))

    # ---------------- FP64 plain drop-in
    out64a = torch.empty((1, nlm), dtype=torch.complex128, device=dev)
    out64m = torch.empty((1, p64.npix), dtype=torch.float64, device=dev)

    def step64():
        p64.map2alm(map64, out=out64a)
        p64.alm2map(alm64, out=out64m)

    log(f"timing FP64 steps={args.fp64_steps}")
    kt64 = KernelTimer("legendre_analysis", "legendre_synthesis", "ring_analysis", "ring_synthesis")
    ms64 = time_steps(torch, step64, args.fp64_steps, 1, kt64)
    pairs64 = 1000.0 / ms64
    kms64 = {k: kt64.mean_ms(k) for k in kt64.kernels}
    log("FP64 kernel means (ms):", json.dumps({k: round(v, 3) for k, v in kms64.items()}))

    # ---------------- roofline of the dominant kernel (FP32 batch, matrix cores)
    # analysis covers m >= mb (rows m < mb run in the FP64 band); synthesis covers every m.
    # Algorithmic flops: 2 per (ring, l, m, map, re/im) multiply-add.  The split-fp16
    # kernels run them on the FP16 matrix path as three fp16 products each (hi*hi +
    # hi*lo + lo*hi), so their peak is the FP16 matrix peak and the executed matrix
    # flops are three times the algorithmic ones.
    mb = p32.fp64_band_m
    band = mb * (2 * lmax + 1 - mb) // 2 + mb
    flops = {"legendre_analysis": 4.0 * npairs * (nlm - band) * B,
             "legendre_synthesis": 4.0 * npairs * nlm * B}
    split16 = p32.tiles()["mfma_kind"] == _abi.HBX_MFMA_F16X2
    if split16:
        names = {"legendre_analysis": "leg_analysis_h16_kernel", "legendre_synthesis": "leg_synthesis_h16_kernel"}
        peak, per_product = PEAK_F16, 3.0
    else:
        names = {"legendre_analysis": "leg_analysis_mfma_kernel", "legendre_synthesis": "leg_synthesis_mfma_kernel"}
        peak, per_product = PEAK_F32, 1.0
    dom = max(flops, key=lambda k: kms[k])
    achieved = flops[dom] / (kms[dom] * 1e-3) / 1e12
    workload = f"nside{nside}_lmax{lmax}_b{B}"
    traffic = measured_traffic(names[dom], workload)
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": traffic, "kernel": names[dom],
                "flops_per_launch": flops[dom], "launch_ms": round(kms[dom], 4),
                "executed_matrix_tflops": round(per_product * achieved, 3),
                "executed_matrix_frac": round(per_product * achieved / peak, 4)}
    other = [k for k in flops if k != dom][0]
    other_tf = flops[other] / (kms[other] * 1e-3) / 1e12
    # ring stage: HBM bound; bytes per map per direction = npix*s + Nr*(mmax+1)*2*s (s = 4 for FP32)
    nr = 4 * nside - 1
    ring_bytes = (p32.npix * 4 + nr * (lmax + 1) * 8) * B
    ring_gbs = {k: ring_bytes / (kms[k] * 1e-3) / 1e9 for k in ("ring_analysis", "ring_synthesis")}
    f64_ach = 4.0 * npairs * nlm / (kms64["legendre_analysis"] * 1e-3) / 1e12
    f64_syn = 4.0 * npairs * nlm / (kms64["legendre_synthesis"] * 1e-3) / 1e12

    if args.copy_calibration:
        log("copy calibration: 1 GiB -> 1 GiB elementwise kernel, 2 launches")
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(2):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

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
                   "map_rel_err_f32": map_err, "errors_over": f"every row of the timed batch of {B}",
                   "gate_1e-5": bool(emax <= 1e-5 and map_err <= 1e-5),
                   "legendre_matrix_operands": ("fp16 hi/lo pairs, 3 fp16 products per FP32 product, FP32 "
                                                "accumulation" if split16 else "FP32"),
                   "kernel_ms_f32_batch": {k: round(v, 4) for k, v in kms.items()},
                   "kernel_ms_f64": {k: round(v, 4) for k, v in kms64.items()},
                   f"{other}_tflops_f32": round(other_tf, 3),
                   "ring_stage_hbm_gbs_f32": {k: round(v, 1) for k, v in ring_gbs.items()},
                   "ring_stage_hbm_frac_of_8000": {k: round(v / 8000.0, 4) for k, v in ring_gbs.items()},
                   "fp64_legendre_analysis_tflops": round(f64_ach, 3),
                   "fp64_legendre_synthesis_tflops": round(f64_syn, 3),
                   "fp64_analysis_frac_of_78.6": round(f64_ach / PEAK_F64, 4),
                   "algorithmic_flops_per_map_per_direction": w_dir},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
