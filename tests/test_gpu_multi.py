"""Persistent mixed-geometry batches (hbx_multi_*, ``batch.MultiPlan``).

Bars: FP64 batches equal the oracle's map2alm / alm2map of every map to
1e-11 relative (the FP64 parity bar); FP32 and FP64 batches equal each
group's own per-plan VALU transform (same ring-pair slots) bit for bit, for
every synthesis map-block size and tile granularity; a captured HIP graph
of a batch replays the eager results bit for bit.  Edge cases: groups whose
lmax is below the FP32 plans' FP64 band (no persistent tiles), mmax < lmax,
odd map counts (partial synthesis map blocks), ragged group sizes.
"""
import numpy as np
import pytest
import torch

from oracle import geometry as OG
from oracle import sht as OS

pytestmark = pytest.mark.gpu

F64_TOL = 1e-11


def _plan(nside, lmax, mmax=None, precision="f64"):
    from hellblast_amd.hp_compat import _plan
    return _plan(nside, lmax, lmax if mmax is None else mmax, precision, "dz", "reference", "reference")


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300))


# (nside, lmax, mmax, maps): ragged groups, mmax < lmax, one tiny geometry
GROUPS = [(16, 32, 32, 3), (8, 20, 12, 1), (32, 64, 64, 2), (2, 4, 4, 2), (64, 128, 128, 5)]


def _inputs(groups, precision, seed):
    rdt = torch.float64 if precision == "f64" else torch.float32
    cdt = torch.complex128 if precision == "f64" else torch.complex64
    rng = np.random.default_rng(seed)
    plans, maps, alm = [], [], []
    for nside, lmax, mmax, n in groups:
        p = _plan(nside, lmax, mmax, precision)
        plans.append(p)
        maps.append(torch.as_tensor(rng.standard_normal((n, p.npix)), device="cuda").to(rdt))
        a = np.stack([OS.synalm(1.0 / (np.arange(lmax + 1) + 1.0) ** 2, lmax, mmax, seed=seed + 17 * k + nside)
                      for k in range(n)])
        alm.append(torch.as_tensor(a, device="cuda").to(cdt))
    return plans, maps, alm


def _run(mp, maps, alm_in):
    ws = mp.workspace()
    alm = [torch.full((m.shape[0], p.nalm), float("nan"), dtype=mp.alm_dtype, device="cuda")
           for m, p in zip(maps, mp.plans)]
    out = [torch.full_like(m, float("nan")) for m in maps]
    mp.map2alm(maps, alm, ws)
    mp.alm2map(alm_in, out, ws)
    torch.cuda.synchronize()
    return alm, out


def test_multi_f64_matches_oracle(gpu):
    from hellblast_amd.batch import MultiPlan
    print("[case] multi f64 vs oracle", flush=True)
    plans, maps, alm_in = _inputs(GROUPS, "f64", 11)
    mp = MultiPlan(plans, [m.shape[0] for m in maps])
    alm, out = _run(mp, maps, alm_in)
    for (nside, lmax, mmax, n), m, a, ain, o in zip(GROUPS, maps, alm, alm_in, out):
        tabs = OG.ring_tables(nside)
        for b in range(n):
            ref = OS.map2alm(m[b].cpu().numpy(), tabs, lmax, mmax)
            assert _rel(a[b].cpu().numpy(), ref) <= F64_TOL, (nside, b)
            refm = OS.alm2map(ain[b].cpu().numpy(), tabs, lmax, mmax)
            assert _rel(o[b].cpu().numpy(), refm) <= F64_TOL, (nside, b)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("slots", [1, 2, 4, 8])
def test_multi_equals_per_plan_valu_bit_for_bit(gpu, precision, slots):
    """Every group's alm and maps equal its plan's VALU path (valu_slots =
    slots, matrix cores off) bit for bit, for every synthesis map block and
    tile granularity."""
    from hellblast_amd.batch import MultiPlan
    print(f"[case] multi == per-plan VALU {precision} slots={slots}", flush=True)
    plans, maps, alm_in = _inputs(GROUPS, precision, 12)
    rdt = maps[0].dtype
    # the batch runs each FP32 group's FP64 band with its plan's kernels, so
    # it gets the same derived plans as the reference
    plans = [p.with_tiles(valu_slots=slots, mfma_min_maps=1000) for p in plans]
    refs_a, refs_m = [], []
    for q, m, a in zip(plans, maps, alm_in):
        refs_a.append(q.map2alm(m))
        refs_m.append(q.alm2map(a, dtype=rdt))
    counts = [m.shape[0] for m in maps]
    for syn_maps, tpw in ((1, 8), (2, 1), (4, 64)):
        mp = MultiPlan(plans, counts, valu_slots=slots, syn_maps=syn_maps, tiles_per_wave=tpw)
        assert mp.tiles == {"valu_slots": slots, "syn_maps": syn_maps, "tiles_per_wave": tpw}
        alm, out = _run(mp, maps, alm_in)
        for g in range(len(plans)):
            assert torch.equal(alm[g], refs_a[g]), (g, syn_maps, tpw)
            assert torch.equal(out[g], refs_m[g]), (g, syn_maps, tpw)


# 16 groups (the batch maximum): geometries whose FFT size classes repeat
# across groups, so the fused ring launches hold many segments each
GROUPS16 = [(1, 2, 2, 1), (2, 4, 4, 2), (2, 6, 3, 1), (4, 8, 8, 1), (4, 12, 6, 2), (8, 16, 16, 1),
            (8, 20, 12, 3), (8, 24, 24, 1), (16, 32, 32, 2), (16, 40, 20, 1), (16, 48, 48, 1), (32, 64, 64, 1),
            (32, 80, 40, 2), (32, 96, 96, 1), (64, 128, 128, 1), (64, 160, 100, 1)]


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_multi_16_groups_other_map_dtype_bit_for_bit(gpu, precision):
    """A batch of 16 groups whose maps have the other float type (FP64 maps
    into an FP32 batch, FP32 maps into an FP64 batch) equals each group's
    per-plan VALU transform with the same map type bit for bit."""
    from hellblast_amd.batch import MultiPlan
    print(f"[case] multi 16 groups {precision}, other map dtype", flush=True)
    plans, maps, alm_in = _inputs(GROUPS16, precision, 14)
    mdt = torch.float32 if precision == "f64" else torch.float64
    maps = [m.to(mdt) for m in maps]
    plans = [p.with_tiles(valu_slots=4, mfma_min_maps=1000) for p in plans]
    refs_a = [q.map2alm(m) for q, m in zip(plans, maps)]
    refs_m = [q.alm2map(a, dtype=mdt) for q, a in zip(plans, alm_in)]
    mp = MultiPlan(plans, [m.shape[0] for m in maps], valu_slots=4, syn_maps=2)
    alm, out = _run(mp, maps, alm_in)
    for g in range(len(plans)):
        assert out[g].dtype == mdt
        assert torch.equal(alm[g], refs_a[g]), g
        assert torch.equal(out[g], refs_m[g]), g


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_multi_large_geometries_bit_for_bit(gpu, precision):
    """Groups at nside 128..512, whose size classes run in fused launches of
    64 to 512 threads (Bluestein and power-of-two rings, several groups per
    launch), equal their per-plan VALU transforms bit for bit."""
    from hellblast_amd.batch import MultiPlan
    print(f"[case] multi nside 128..512 {precision}", flush=True)
    groups = [(256, 512, 512, 2), (512, 1023, 1023, 1), (128, 256, 200, 3), (512, 700, 700, 1)]
    plans, maps, alm_in = _inputs(groups, precision, 16)
    plans = [p.with_tiles(valu_slots=4, mfma_min_maps=1000) for p in plans]
    rdt = maps[0].dtype
    refs_a = [q.map2alm(m) for q, m in zip(plans, maps)]
    refs_m = [q.alm2map(a, dtype=rdt) for q, a in zip(plans, alm_in)]
    mp = MultiPlan(plans, [m.shape[0] for m in maps], valu_slots=4)
    alm, out = _run(mp, maps, alm_in)
    for g in range(len(plans)):
        assert torch.equal(alm[g], refs_a[g]), g
        assert torch.equal(out[g], refs_m[g]), g


def test_multi_shared_by_threads(gpu):
    """One FP32 batch used from two host threads on two streams at once, each
    with its own workspace (the batch's side stream and events are shared
    under its lock): each thread's alm and maps equal a lone call's."""
    import threading
    from hellblast_amd.batch import MultiPlan
    print("[case] multi batch from two threads", flush=True)
    plans, maps, alm_in = _inputs(GROUPS, "f32", 15)
    mp = MultiPlan(plans, [m.shape[0] for m in maps])
    ref_a, ref_o = _run(mp, maps, alm_in)
    outs = [None, None]
    errs = []

    def work(i):
        try:
            st = torch.cuda.Stream()
            with torch.cuda.stream(st):
                for _ in range(4):
                    outs[i] = _run(mp, maps, alm_in)
            st.synchronize()
        except Exception as e:          # reported by the main thread
            errs.append(e)

    ts = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs, errs
    for i in range(2):
        a, o = outs[i]
        for g in range(len(plans)):
            assert torch.equal(a[g], ref_a[g]), (i, g)
            assert torch.equal(o[g], ref_o[g]), (i, g)


def test_multi_f32_c3_shape_within_gate(gpu):
    """FP32 batch over nside 128..1024 (lmax 2 nside, C_l = (l+10)^-2, the
    mixed-batch spectrum): eps_max <= 1e-5 of every map against the FP64
    plan's analysis of the same FP32 map."""
    from hellblast_amd.batch import MultiPlan
    print("[case] multi f32 nside 128..1024 gate", flush=True)
    groups = [(1024, 2048, 1), (512, 1024, 2), (256, 512, 3), (128, 256, 7)]
    plans, maps = [], []
    for nside, lmax, n in groups:
        p64 = _plan(nside, lmax)
        cl = (np.arange(lmax + 1) + 10.0) ** -2
        a = torch.as_tensor(np.stack([OS.synalm(cl, lmax, seed=500 + 7 * k + nside) for k in range(n)]),
                            device="cuda")
        maps.append(p64.alm2map(a).float())
        plans.append(_plan(nside, lmax, precision="f32"))
    mp = MultiPlan(plans, [n for _, _, n in groups])
    ws = mp.workspace()
    alm = [torch.empty((m.shape[0], p.nalm), dtype=torch.complex64, device="cuda") for m, p in zip(maps, plans)]
    mp.map2alm(maps, alm, ws)
    torch.cuda.synchronize()
    worst = 0.0
    for (nside, lmax, n), m, a in zip(groups, maps, alm):
        ref = _plan(nside, lmax).map2alm(m.double()).cpu().numpy()
        cl = OS.alm2cl(ref[0], lmax)
        ls = np.concatenate([np.arange(mm, lmax + 1) for mm in range(lmax + 1)])
        for b in range(n):
            worst = max(worst, float(np.max(np.abs(a[b].cpu().numpy() - ref[b]) / np.sqrt(cl[ls]))))
    print(f"worst eps_max {worst:.2e}")
    assert worst <= 1e-5


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_multi_graph_replay_equals_eager(gpu, precision):
    from hellblast_amd.batch import MultiPlan
    print(f"[case] multi graph replay {precision}", flush=True)
    plans, maps, alm_in = _inputs(GROUPS + [(128, 256, 256, 3)], precision, 13)
    mp = MultiPlan(plans, [m.shape[0] for m in maps])
    ws = mp.workspace()
    alm = [torch.zeros((m.shape[0], p.nalm), dtype=mp.alm_dtype, device="cuda") for m, p in zip(maps, plans)]
    out = [torch.zeros_like(m) for m in maps]

    def step():
        mp.map2alm(maps, alm, ws)
        mp.alm2map(alm_in, out, ws)

    step()
    torch.cuda.synchronize()
    ref_a = [a.clone() for a in alm]
    ref_o = [o.clone() for o in out]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    for rep in range(3):
        for t in alm + out:
            t.zero_()
        g.replay()
        torch.cuda.synchronize()
        for k in range(len(plans)):
            assert torch.equal(alm[k], ref_a[k]), (rep, k)
            assert torch.equal(out[k], ref_o[k]), (rep, k)
        step()                                   # eager work between replays
        torch.cuda.synchronize()


def test_multi_validation(gpu):
    from hellblast_amd._abi import HbxValueError
    from hellblast_amd.batch import MultiPlan
    f32 = _plan(8, 16, precision="f32")
    f64 = _plan(8, 16)
    with pytest.raises(HbxValueError):
        MultiPlan([f32, f64], [1, 1])                  # mixed precision
    spin = __import__("hellblast_amd.hp_compat", fromlist=["_plan"])._plan(8, 16, 16, "f64", "dz", "reference",
                                                                           "reference", 2)
    with pytest.raises(HbxValueError):
        MultiPlan([spin], [1])                         # spin plan
    with pytest.raises(HbxValueError):
        MultiPlan([f64], [0])                          # empty group
    for bad in ({"valu_slots": 3}, {"syn_maps": 3}):
        with pytest.raises(HbxValueError):
            MultiPlan([f64], [1], **bad)
    with pytest.raises(ValueError):
        MultiPlan([f64], [1], no_such_tile=1)
    mp = MultiPlan([f64], [2])
    ws = mp.workspace()
    maps = [torch.zeros((2, f64.npix), dtype=torch.float64, device="cuda")]
    alm = [torch.zeros((2, f64.nalm), dtype=torch.complex128, device="cuda")]
    with pytest.raises(HbxValueError):
        mp.map2alm(maps, alm, ws[:16])                 # workspace too small
    with pytest.raises(ValueError):
        mp.map2alm([maps[0][:1]], alm, ws)             # wrong map count
    with pytest.raises(ValueError):
        mp.map2alm(maps, [alm[0].to(torch.complex64)], ws)


def test_mixed_batch_auto_mode_is_timed_and_cached(gpu, tmp_path, monkeypatch):
    """mode="auto" times the persistent and the grouped form once, keeps the
    faster, and reuses the cached choice; both forms give alm within FP32
    rounding of each other."""
    import json
    from hellblast_amd import autotune
    from hellblast_amd.batch import MixedBatch
    monkeypatch.setenv(autotune.CACHE_ENV, str(tmp_path / "tiles.json"))
    specs = [(32, 64), (64, 128), (32, 64), (16, 32)]
    print("[case] mixed batch auto mode", flush=True)
    b = MixedBatch(specs, precision="f32", mode="auto")
    assert b.mode in ("persistent", "grouped")
    db = json.load(open(tmp_path / "tiles.json"))
    entries = [v for k, v in db.items() if "|mixed-mode|" in k]
    assert len(entries) == 1 and set(entries[0]["timed"]) == {"persistent", "grouped"}
    assert entries[0]["mode"] == b.mode
    assert MixedBatch(specs, precision="f32", mode="auto").mode == b.mode
    gen = torch.Generator(device="cuda")
    gen.manual_seed(21)
    maps = [torch.randn(12 * n * n, generator=gen, device="cuda") for n, _ in specs]
    outs = {}
    for mode in ("persistent", "grouped"):
        mb = MixedBatch(specs, precision="f32", mode=mode, tune=False)
        mb.set_maps(maps)
        mb.run_eager()
        torch.cuda.synchronize()
        outs[mode] = [mb.alm_of(i).clone() for i in range(len(specs))]
    for a, c in zip(outs["persistent"], outs["grouped"]):
        assert float((a - c).abs().max()) <= 2e-6 * float(c.abs().max())
