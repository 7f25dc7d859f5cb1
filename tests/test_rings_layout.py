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


@pytest.mark.parametrize("nrings,mmax", [(7, 3), (8, 9), (255, 100), (8191, 255)])
@pytest.mark.parametrize("direction", ["analysis", "synthesis"])
def test_phase_index_is_a_layout(nrings, mmax, direction):
    """Every (ring, m) gets its own element inside one map's stride."""
    from hellblast_amd.plan import phase_index, phase_map_elems
    ix = phase_index(nrings, mmax, direction)
    assert ix.shape == (nrings, mmax + 1)
    assert ix.min() >= 0 and ix.max() < phase_map_elems(nrings, mmax)
    assert np.unique(ix).size == ix.size


@pytest.mark.parametrize("nrings", [7, 16, 8191])
def test_phase_index_line_structure(nrings):
    """Analysis: 8 m of one ring are consecutive.  Synthesis: 8 ring pairs
    (p = 8k .. 8k+7) of one m and one hemisphere are consecutive, ring
    nrings-1-p sitting pairs8 * 8 elements after ring p."""
    from hellblast_amd.plan import phase_index
    mmax = 23
    a = phase_index(nrings, mmax, "analysis")
    assert np.all(np.diff(a[:, 8:16], axis=1) == 1)
    s = phase_index(nrings, mmax, "synthesis")
    npairs = (nrings + 1) // 2
    p8 = (npairs + 7) // 8 * 8
    full = (npairs // 8) * 8
    if full:
        north = s[:full]                                  # rings 0 .. full-1 = pairs 0 .. full-1
        blocks = north.reshape(full // 8, 8, mmax + 1)
        assert np.all(np.diff(blocks, axis=1) == 1)
    for p in range(npairs):
        if nrings - 1 - p != p:
            assert np.all(s[nrings - 1 - p] - s[p] == p8 * 8)
