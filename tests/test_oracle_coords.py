"""Pins of the RING pixel <-> angle oracle (oracle/coords.py) on the CPU.

The reference's rule (coords_ring.py:18-24, coords_ring.cu:9-54) is pinned by
the HEALPix closed forms of the pixel centres, by the round trip
ang2pix(pix2ang(p)) = p over every pixel, and by a brute-force nearest-centre
search inside the z-edge ring cell.
"""
import math

import numpy as np
import pytest

from oracle import coords as OC
from oracle import geometry as OG


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16, 64])
def test_pix2ang_closed_form(nside):
    t = OG.ring_tables(nside)
    p = np.arange(t["npix"])
    th, ph = OC.pix2ang(p, t)
    for r in range(0, t["Nr"], max(1, t["Nr"] // 17)):
        i = r + 1                                        # HEALPix ring number
        if i < nside:
            z, n, s = 1 - i * i / (3 * nside * nside), 4 * i, 1
        elif i <= 3 * nside:
            z, n, s = (2 * nside - i) * 2 / (3 * nside), 4 * nside, (i - nside) & 1
        else:
            ip = 4 * nside - i
            z, n, s = -(1 - ip * ip / (3 * nside * nside)), 4 * ip, 1
        sl = slice(t["ring_start"][r], t["ring_start"][r + 1])
        j = np.arange(n)
        assert np.allclose(th[sl], math.acos(z), atol=1e-15)
        want = np.mod((1 - 0.5 * s) * 2 * math.pi / n + j * 2 * math.pi / n, 2 * math.pi)
        assert np.all(np.abs((ph[sl] - want + math.pi) % (2 * math.pi) - math.pi) < 1e-12)
        assert np.all((ph[sl] >= 0) & (ph[sl] < 2 * math.pi))


@pytest.mark.parametrize("nside", [1, 2, 8, 32, 64])
def test_round_trip_every_pixel(nside):
    t = OG.ring_tables(nside)
    p = np.arange(t["npix"])
    th, ph = OC.pix2ang(p, t)
    assert np.array_equal(OC.ang2pix(th, ph, t), p)
    th2, ph2 = OC.pix2ang(p, t, t["phi0_healpix"])
    assert np.array_equal(OC.ang2pix(th2, ph2, t, t["phi0_healpix"]), p)


def test_ang2pix_brute_force_nearest_centre():
    nside = 8
    t = OG.ring_tables(nside)
    rng = np.random.default_rng(5)
    n = 2000
    th = np.arccos(rng.uniform(-1, 1, n))
    ph = rng.uniform(-3 * math.pi, 5 * math.pi, n)
    got = OC.ang2pix(th, ph, t)
    z = np.cos(th)
    for k in range(n):
        r = [q for q in range(t["Nr"]) if t["z_edges"][q] >= z[k] > t["z_edges"][q + 1]][0]
        npr = t["nphi"][r]
        cen = t["phi_center0"][r] + np.arange(npr) * 2 * math.pi / npr
        dist = np.abs((ph[k] - cen + math.pi) % (2 * math.pi) - math.pi)
        assert got[k] == t["ring_start"][r] + int(np.argmin(dist))
