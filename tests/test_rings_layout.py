"""Flat RING map <-> per-ring arrays (reference: udgrade.py:86-102,
hp_compat.py:31-37): ring r is the slice [ring_start[r], ring_start[r] +
nphi[r]) of the oracle's ring tables; unpacking makes views (no copies),
packing inverts it, wrong ring lengths are refused."""
import numpy as np
import pytest
import torch

from hellblast_amd.ring_tables import HealpixRingTables, pack_rings_to_flat, unpack_flat_to_rings
from oracle import geometry as OG


@pytest.mark.parametrize("nside", [1, 4, 64])
def test_unpack_pack_round_trip(nside):
    tabs = HealpixRingTables(nside, device="cpu")
    og = OG.ring_tables(nside)
    flat = torch.as_tensor(np.random.default_rng(nside).standard_normal((2, og["npix"])))
    rings = unpack_flat_to_rings(flat, tabs)
    assert len(rings) == og["Nr"]
    for r in (0, og["Nr"] // 2, og["Nr"] - 1):
        s, e = og["ring_start"][r], og["ring_start"][r + 1]
        assert torch.equal(rings[r], flat[:, s:e])
    assert torch.equal(pack_rings_to_flat(rings, tabs), flat)
    rings[0][0, 0] = 123.0                      # views of the flat map
    assert float(flat[0, 0]) == 123.0


def test_pack_rejects_wrong_rings():
    tabs = HealpixRingTables(2, device="cpu")
    rings = unpack_flat_to_rings(torch.zeros(tabs.npix), tabs)
    with pytest.raises(ValueError):
        pack_rings_to_flat(rings[:-1], tabs)
    bad = list(rings)
    bad[1] = torch.zeros(bad[1].numel() + 1)
    with pytest.raises(ValueError):
        pack_rings_to_flat(bad, tabs)
    with pytest.raises(ValueError):
        unpack_flat_to_rings(torch.zeros(tabs.npix + 1), tabs)
