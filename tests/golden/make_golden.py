"""Write the small golden fixtures used by the CPU and GPU parity tests.

The reference ships no golden vectors, so these are produced by the CPU
oracle (itself pinned in tests/test_oracle.py) on seeded inputs:
    python -m tests.golden.make_golden
"""
import os

import numpy as np

from oracle import geometry as OG
from oracle import sht as OS

HERE = os.path.dirname(os.path.abspath(__file__))


def scalar_cases():
    out = {}
    for nside, lmax in ((8, 16), (16, 40)):
        t = OG.ring_tables(nside)
        rng = np.random.default_rng(1000 + nside)
        f = rng.standard_normal(t["npix"])
        out[f"map_n{nside}"] = f
        out[f"map2alm_n{nside}_l{lmax}"] = OS.map2alm(f, t, lmax)
        alm = OS.synalm(1.0 / (np.arange(lmax + 1) + 1.0) ** 2, lmax, seed=2000 + nside)
        out[f"alm_n{nside}_l{lmax}"] = alm
        out[f"alm2map_n{nside}_l{lmax}"] = OS.alm2map(alm, t, lmax)
        out[f"map2alm_fold_hpx_n{nside}_l{lmax}"] = OS.map2alm(
            f, t, lmax, phi0=t["phi0_healpix"], mcut=np.full(t["Nr"], lmax + 1))
    return out


if __name__ == "__main__":
    path = os.path.join(HERE, "scalar_small.npz")
    np.savez_compressed(path, **scalar_cases())
    print(path, os.path.getsize(path))
