"""Tiles, the autotuner and mixed-geometry batches in one HIP graph.

Bars: every tile set computes the same transform (FP64: 1e-13 relative,
FP32: 2e-6 relative to the default tiles); graph replay equals the eager run
bit for bit on every replay, with eager work between replays; every map of
the mixed batch is within eps_max 1e-5 of its FP64 transform.
"""
import numpy as np
import pytest
import torch

from oracle import sht as OS

pytestmark = pytest.mark.gpu


def _plan(nside, lmax, precision="f64", spin=0):
    from hellblast_amd.hp_compat import _plan
    return _plan(nside, lmax, lmax, precision, "dz", "reference", "reference", spin)


def _eps_max(a, ref, lmax):
    """max |a - ref| / sqrt(C_l(ref)) over (l, m) (m-major alm)."""
    a = np.asarray(a, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    cl = OS.alm2cl(ref, lmax)
    ls = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)])
    c = cl[ls]
    ok = c > 0
    return float(np.max(np.abs(a - ref)[ok] / np.sqrt(c[ok])))


def test_tiles_derive_get_and_validation(gpu):
    """Tiles make a derived plan (hbx_plan_derive); the base plan keeps its
    own tiles, so tuning never changes anyone else's kernels."""
    from hellblast_amd import _abi
    from hellblast_amd._abi import HbxValueError
    plan = _plan(32, 64, "f32")
    base = plan.tiles()
    assert base["mfma_waves"] == 4 and base["mfma_min_maps"] == 3
    assert base["mfma_kind"] == _abi.HBX_MFMA_F16X2
    d = plan.with_tiles(mfma_waves=4, valu_slots=2)
    t = d.tiles()
    assert t["mfma_waves"] == 4 and t["valu_slots"] == 2
    assert plan.tiles() == base
    assert plan.with_tiles(mfma_waves=4, valu_slots=2) is d          # cached per tile set
    assert d.with_tiles(mfma_waves=0, valu_slots=0).tiles() == base
    for bad in ({"mfma_waves": 5}, {"valu_slots": 3}, {"mfma_kind": 3},
                {"mfma_waves": 12, "mfma_kind": _abi.HBX_MFMA_F16X2}):
        with pytest.raises(HbxValueError):
            plan.with_tiles(**bad)
    with pytest.raises(ValueError):
        plan.with_tiles(no_such_tile=1)


def test_derived_plan_outlives_its_base(gpu):
    """A derived plan shares the base's device tables and keeps them alive."""
    import gc
    from hellblast_amd.hp_compat import healpix_geom
    from hellblast_amd.plan import Plan
    nside, lmax = 16, 32
    geom = healpix_geom(nside, lmax)
    base = Plan(geom, lmax, precision="f32")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(8)
    maps = torch.randn((4, base.npix), generator=gen, device="cuda", dtype=torch.float32)
    ref = base.map2alm(maps)
    d = base.with_tiles(mfma_min_maps=100)            # the VALU kernels
    del base
    gc.collect()
    got = d.map2alm(maps)
    assert float((got - ref).abs().max()) <= 2e-6 * float(ref.abs().max())


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_every_candidate_tile_gives_the_same_transform(gpu, precision):
    from hellblast_amd.autotune import candidates
    nside, lmax, n = 64, 128, 4
    print(f"[case] tile candidates {precision}", flush=True)
    plan = _plan(nside, lmax, precision)
    rdt = torch.float64 if precision == "f64" else torch.float32
    gen = torch.Generator(device="cuda")
    gen.manual_seed(3)
    maps = torch.randn((n, plan.npix), generator=gen, device="cuda", dtype=rdt)
    ref_a = plan.map2alm(maps)
    ref_m = plan.alm2map(ref_a, dtype=rdt)
    tol = 1e-13 if precision == "f64" else 2e-6
    for cand in candidates(plan, n):
        p = plan.with_tiles(**cand)
        a = p.map2alm(maps)
        m = p.alm2map(ref_a, dtype=rdt)
        assert float((a - ref_a).abs().max()) <= tol * float(ref_a.abs().max()), cand
        assert float((m - ref_m).abs().max()) <= tol * float(ref_m.abs().max()), cand


def test_autotune_caches_its_choice(gpu, tmp_path, monkeypatch):
    from hellblast_amd import autotune
    monkeypatch.setenv(autotune.CACHE_ENV, str(tmp_path / "tiles.json"))
    plan = _plan(64, 128, "f32")
    before = plan.tiles()
    first = autotune.tune(plan, 4, reps=1)
    assert (tmp_path / "tiles.json").exists()
    assert len(first["timed"]) == len(autotune.candidates(plan, 4))
    assert plan.tiles() == before                      # tuning leaves the plan alone
    second = autotune.tune(plan, 4, reps=1)
    assert second["tiles"] == first["tiles"]
    eff = plan.with_tiles(**first["tiles"]).tiles()
    for k, v in first["tiles"].items():
        if v:
            assert eff[k] == v


def test_synthesis_map_groups_agree(gpu):
    """Two 16-map groups per split-fp16 synthesis workgroup compute the same
    maps, bit for bit, as one group per workgroup (40 maps: a partial group)."""
    nside, lmax, n = 64, 128, 40
    print("[case] synthesis map groups", flush=True)
    plan = _plan(nside, lmax, "f32")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(4)
    alm = torch.complex(torch.randn((n, plan.nalm), generator=gen, device="cuda"),
                        torch.randn((n, plan.nalm), generator=gen, device="cuda"))
    m1 = plan.with_tiles(mfma_groups=1).alm2map(alm, dtype=torch.float32)
    m2 = plan.with_tiles(mfma_groups=2).alm2map(alm, dtype=torch.float32)
    assert torch.equal(m1, m2)


@pytest.mark.parametrize("nside,lmax,n,waves", [(64, 128, 40, 0), (128, 256, 64, 0), (512, 1023, 32, 0),
                                                (1024, 2047, 48, 0), (512, 1023, 48, 4), (1024, 2047, 32, 4)])
def test_analysis_map_groups_agree(gpu, nside, lmax, n, waves):
    """Two 16-map groups per split-fp16 analysis workgroup (one lambda tile
    shared by 32 maps: 8 waves over 256 ring pairs, or 16 waves over 512)
    compute the same alm, bit for bit, as one group per workgroup: nside 64
    with a partial map group (40 maps, one ring-pair group), nside 128 with
    64 maps (one ring-pair group: the automatic choice takes two map groups
    there), nside 512 (two ring-pair groups of 512 pairs: partial sums),
    nside 1024 (four ring-pair groups, 48 maps: a partial map group), and the
    256-pair geometry over several ring-pair groups (mfma_waves 4)."""
    print(f"[case] analysis map groups nside={nside} n={n} waves={waves}", flush=True)
    plan = _plan(nside, lmax, "f32")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(14 + nside)
    maps = torch.randn((n, plan.npix), generator=gen, device="cuda")
    a1 = plan.with_tiles(mfma_groups=1, mfma_waves=waves).map2alm(maps)
    a2 = plan.with_tiles(mfma_groups=2, mfma_waves=waves).map2alm(maps)
    assert torch.equal(a1, a2)
    if waves == 0:
        assert torch.equal(plan.map2alm(maps), a1)       # the automatic choice


def _c3_specs(rng):
    """64 maps, nside allocation {2048:1, 1024:2, 512:5, 256:14, 128:42}, shuffled."""
    alloc = {2048: 1, 1024: 2, 512: 5, 256: 14, 128: 42}
    specs = [n for n, k in alloc.items() for _ in range(k)]
    rng.shuffle(specs)
    return [(n, 2 * n) for n in specs]


@pytest.mark.parametrize("mode", ["persistent", "grouped"])
def test_mixed_batch_graph_replay_equals_eager(gpu, tmp_path, monkeypatch, mode):
    """The 64-map mixed batch (lmax = 2 nside, C_l = (l+10)^-2), FP32, tuned
    tiles, one graph launch (persistent Legendre launches over all groups, or
    one batched call per group): replays equal the eager run bit for bit and
    every map's alm is within eps_max 1e-5 of its FP64 transform."""
    from hellblast_amd import autotune
    from hellblast_amd.batch import MixedBatch
    monkeypatch.setenv(autotune.CACHE_ENV, str(tmp_path / "tiles.json"))
    rng = np.random.default_rng(3)
    specs = _c3_specs(rng)
    print(f"[case] mixed batch of {len(specs)} maps, {mode}", flush=True)
    batch = MixedBatch(specs, precision="f32", tune=True, mode=mode)
    # inputs: FP64 syntheses of C_l = (l+10)^-2 alm, stored as FP32 maps
    maps64 = []
    for i, (nside, lmax) in enumerate(specs):
        cl = (np.arange(lmax + 1) + 10.0) ** -2
        alm = torch.as_tensor(OS.synalm(cl, lmax, seed=1000 + i), device="cuda")
        maps64.append(_plan(nside, lmax).alm2map(alm))
    batch.set_maps([m.float() for m in maps64])
    batch.run_eager()
    torch.cuda.synchronize()
    eager_alm = [batch.alm_of(i).clone() for i in range(batch.nmaps)]
    eager_out = [batch.output(i).clone() for i in range(batch.nmaps)]
    batch.capture()
    for rep in range(3):
        for key in batch.alm:
            batch.alm[key].zero_()
            batch.out[key].zero_()
        batch.replay()
        torch.cuda.synchronize()
        for i in range(batch.nmaps):
            assert torch.equal(batch.alm_of(i), eager_alm[i]), (rep, i)
            assert torch.equal(batch.output(i), eager_out[i]), (rep, i)
        # eager work between replays
        _ = _plan(128, 256, "f32").map2alm(torch.randn((2, 12 * 128 * 128), device="cuda"))
        torch.cuda.synchronize()
    worst = 0.0
    for i, (nside, lmax) in enumerate(specs):
        ref = _plan(nside, lmax).map2alm(batch.map(i).double()).cpu().numpy()
        worst = max(worst, _eps_max(eager_alm[i].cpu().numpy(), ref, lmax))
    print(f"worst eps_max over the batch {worst:.2e}; tiles {batch.tiles} {getattr(batch, 'multi_tiles', '')}")
    assert worst <= 1e-5


def test_fp32_analysis_band_joins_caller_stream(gpu):
    """FP32 map2alm on a user stream: rows m < 8 come from the FP64 band,
    which runs on the plan's side stream; work queued on the user stream
    right after the call (a copy, no synchronisation between) sees them
    complete, and the spin analysis likewise."""
    print("[case] band join on a user stream", flush=True)
    p = _plan(256, 511, "f32")
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    maps = torch.randn((16, p.npix), generator=g, device="cuda")
    ref = p.map2alm(maps)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(3):
            out = p.map2alm(maps).clone()
    s.synchronize()
    assert torch.equal(out, ref)
    ps = _plan(256, 511, "f32", spin=2)
    qu = torch.randn((4, 2, ps.npix), generator=g, device="cuda")
    ref_eb = ps.map2alm_spin(qu)
    torch.cuda.synchronize()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        out_eb = ps.map2alm_spin(qu).clone()
    s.synchronize()
    assert torch.equal(out_eb, ref_eb)


def test_fp32_plan_shared_by_threads(gpu):
    """One FP32 plan used from two host threads on two streams at once (the
    side stream and its events are shared under the plan's lock): each
    thread's alm equal the same maps' alm computed alone."""
    import threading
    print("[case] FP32 plan from two threads", flush=True)
    p = _plan(256, 511, "f32")
    g = torch.Generator(device="cuda")
    g.manual_seed(6)
    maps = [torch.randn((16, p.npix), generator=g, device="cuda") for _ in range(2)]
    refs = [p.map2alm(m) for m in maps]
    torch.cuda.synchronize()
    outs = [None, None]
    errs = []

    def work(i):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for _ in range(4):
                    outs[i] = p.map2alm(maps[i])
            s.synchronize()
        except Exception as e:          # reported by the main thread
            errs.append(e)

    ts = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs, errs
    for i in range(2):
        assert torch.equal(outs[i], refs[i]), i
