"""Tile autotuner for libhbx plans (no reference counterpart).

For one plan and batch size it times map2alm + alm2map (spin plans: the
spin pair) under each candidate tile set (a plan derived with
``hbx_plan_derive``, sharing the plan's device tables) and keeps the
fastest.  Winners are cached in a JSON file keyed by device name,
geometry, lmax, mmax, precision, spin and batch size, so a tuned
configuration is reused across processes:
``$HBX_TUNE_CACHE`` or ``~/.cache/hellblast_amd/tiles.json``.

Tiles only change how work is split (ring pairs per workgroup or lane, and
the matrix-core / VALU switch), so every candidate computes the same
transform; results agree to rounding.
"""
import json
import os
import tempfile

import torch

CACHE_ENV = "HBX_TUNE_CACHE"


def cache_path():
    return os.environ.get(CACHE_ENV) or os.path.join(os.path.expanduser("~"), ".cache", "hellblast_amd",
                                                     "tiles.json")


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(db):
    path = cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=os.path.dirname(path), suffix=".tmp")
    with os.fdopen(fd, "w") as f:
        json.dump(db, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def key(plan, nmaps):
    dev = torch.cuda.get_device_name(plan.device)
    return (f"{dev}|{plan.geom.key()[:16]}|lmax{plan.lmax}|mmax{plan.mmax}|p{plan.precision}|"
            f"s{plan.spin}|n{int(nmaps)}")


def candidates(plan, nmaps):
    """Tile sets worth timing for this plan and batch size."""
    from . import _abi
    npairs = (plan.geom.nrings + 1) // 2
    # slots beyond what covers every ring pair with one wave only add idle lanes
    cover = 1
    while cover * 64 < npairs and cover < 16:
        cover *= 2
    slots = [s for s in (1, 2, 4, 8, 16) if s <= cover] or [1]
    if plan.spin:
        valu = [{"spin_slots": s} for s in (1, 2, 4) if s <= max(cover, 1)]
        if plan.precision == _abi.HBX_F64:
            return valu
        # FP32: the matrix-core analysis, or the VALU kernel (batches below
        # mfma_min_maps maps take it)
        return [{"mfma_min_maps": 2}] + [dict(t, mfma_min_maps=2 * nmaps + 1) for t in valu]
    if plan.precision == _abi.HBX_F64:
        return [{"valu_slots": s} for s in slots if s <= 8]
    out = []
    if nmaps >= 2:
        # matrix cores: split-fp16 or FP32 operands; 4 / 8 waves of 64 ring
        # pairs, or (FP32) 12 = 8 MFMA + 4 recursion waves over 512 pairs
        waves = [w for w in (4, 8, 12, 16) if w == 4 or min(w, 8) * 32 < npairs]
        out += [{"mfma_kind": _abi.HBX_MFMA_F16X2, "mfma_waves": w, "mfma_min_maps": 2} for w in waves if w != 12]
        out += [{"mfma_kind": _abi.HBX_MFMA_F32, "mfma_waves": w, "mfma_min_maps": 2} for w in waves if w != 16]
    # the VALU kernels (also for batches: mfma_min_maps above the batch size)
    out += [{"valu_slots": s, "mfma_min_maps": nmaps + 1} for s in slots]
    return out


def _time(plan, nmaps, reps):
    dev = plan.device
    gen = torch.Generator(device=dev)
    gen.manual_seed(1)
    rdt = torch.float64 if plan.alm_dtype == torch.complex128 else torch.float32
    if plan.spin:
        maps = torch.randn((nmaps, 2, plan.npix), generator=gen, device=dev, dtype=rdt)
        fwd = lambda: plan.alm2map_spin(plan.map2alm_spin(maps), dtype=rdt)   # noqa: E731
    else:
        maps = torch.randn((nmaps, plan.npix), generator=gen, device=dev, dtype=rdt)
        fwd = lambda: plan.alm2map(plan.map2alm(maps), dtype=rdt)   # noqa: E731
    fwd()
    torch.cuda.synchronize(dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fwd()
    e1.record()
    torch.cuda.synchronize(dev)
    return e0.elapsed_time(e1) / reps


def tune(plan, nmaps, reps=3, force=False, verbose=False):
    """The best tiles for (plan, nmaps), timing candidates unless cached.

    Returns the cache entry {"tiles", "ms", "timed"}.  ``plan`` itself is not
    changed (plans are immutable); run with ``plan.with_tiles(**entry["tiles"])``.
    """
    db = _load()
    k = key(plan, nmaps)
    if not force and k in db:
        return db[k]
    results = []
    for cand in candidates(plan, nmaps):
        ms = _time(plan.with_tiles(**{f: 0 for f in plan.TILE_FIELDS}).with_tiles(**cand), nmaps, reps)
        results.append((ms, cand))
        if verbose:
            print(f"[autotune] {k} {cand} {ms:.3f} ms", flush=True)
    ms, best = min(results, key=lambda r: r[0])
    entry = {"tiles": best, "ms": ms, "timed": [[m, c] for m, c in results]}
    db = _load()
    db[k] = entry
    _save(db)
    return entry


def multi_key(plans, counts):
    dev = torch.cuda.get_device_name(plans[0].device)
    groups = ",".join(f"{p.geom.key()[:12]}:l{p.lmax}:m{p.mmax}:n{int(c)}" for p, c in zip(plans, counts))
    return f"{dev}|multi|p{plans[0].precision}|{groups}"


def _time_multi(plans, counts, tiles, reps):
    from .batch import MultiPlan
    mp = MultiPlan(plans, counts, **tiles)
    dev = mp.device
    gen = torch.Generator(device=dev)
    gen.manual_seed(1)
    rdt = torch.float64 if mp.alm_dtype == torch.complex128 else torch.float32
    maps = [torch.randn((c, p.npix), generator=gen, device=dev, dtype=rdt) for p, c in zip(plans, counts)]
    alm = [torch.empty((c, p.nalm), device=dev, dtype=mp.alm_dtype) for p, c in zip(plans, counts)]
    out = [torch.empty_like(m) for m in maps]
    ws = mp.workspace()

    def step():
        mp.map2alm(maps, alm, ws)
        mp.alm2map(alm, out, ws)

    step()
    torch.cuda.synchronize(dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        step()
    e1.record()
    torch.cuda.synchronize(dev)
    return e0.elapsed_time(e1) / reps


def tune_multi(plans, counts, reps=3, force=False, verbose=False):
    """The best tiles of a persistent mixed-geometry batch (``MultiPlan``),
    timing candidates unless cached: analysis slots x tile granularity
    first, then maps per synthesis tile.  Returns {"tiles", "ms", "timed"}."""
    db = _load()
    k = multi_key(plans, counts)
    if not force and k in db:
        return db[k]
    from . import _abi
    slots = (2, 4, 8) if plans[0].precision == _abi.HBX_F32 else (4, 8)
    results = []
    for s in slots:
        for tpw in (4, 8, 16):
            cand = {"valu_slots": s, "syn_maps": 4, "tiles_per_wave": tpw}
            results.append((_time_multi(plans, counts, cand, reps), cand))
            if verbose:
                print(f"[autotune] {k} {cand} {results[-1][0]:.3f} ms", flush=True)
    best = dict(min(results, key=lambda r: r[0])[1])
    for sm in (1, 2):
        cand = dict(best, syn_maps=sm)
        results.append((_time_multi(plans, counts, cand, reps), cand))
        if verbose:
            print(f"[autotune] {k} {cand} {results[-1][0]:.3f} ms", flush=True)
    ms, best = min(results, key=lambda r: r[0])
    entry = {"tiles": best, "ms": ms, "timed": [[m, c] for m, c in results]}
    db = _load()
    db[k] = entry
    _save(db)
    return entry


def tune_mixed_mode(specs, precision="f32", weights="dz", phi0="reference", alias="reference", device=None,
                    reps=3, force=False):
    """"persistent" or "grouped", whichever runs a MixedBatch of these specs
    faster (graph replay of map2alm + alm2map), timed once and cached."""
    from .batch import MixedBatch
    dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
    norm = sorted((int(n), 2 * int(n) if lm is None else int(lm)) for n, lm in
                  ((s, None) if isinstance(s, int) else s for s in specs))
    k = (f"{torch.cuda.get_device_name(dev)}|mixed-mode|{precision}|{weights}|{phi0}|{alias}|"
         + ",".join(f"{n}:{lm}" for n, lm in norm))
    db = _load()
    if not force and k in db:
        return db[k]["mode"]
    timed = {}
    for mode in ("persistent", "grouped"):
        b = MixedBatch(specs, precision, True, weights, phi0, alias, device, mode=mode)
        b.capture()
        torch.cuda.synchronize(dev)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            b.replay()
        e1.record()
        torch.cuda.synchronize(dev)
        timed[mode] = e0.elapsed_time(e1) / reps
        del b
    best = min(timed, key=timed.get)
    db = _load()
    db[k] = {"mode": best, "timed": timed}
    _save(db)
    return best
