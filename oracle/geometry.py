"""HEALPix RING geometry, restated on the CPU with exact integer arithmetic.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows ``ring_tables_healpix.cu:12-51`` (z, theta, nphi, phi_center0,
ring_start, z_edges) and ``ring_tables.py:39-40`` (``weights_dz``).  The
reference kernel computes the prefix sum with an unsynchronised thread-0 loop
(``ring_tables_healpix.cu:40-51``); here it is a plain cumulative sum.
"""
import numpy as np


def ring_tables(nside):
    """Return the reference's per-ring tables for ``nside`` as NumPy arrays.

    Ring i = 1..4*nside-1 is stored at index r = i-1
    (``ring_tables_healpix.cu:15-39``).
    """
    nside = int(nside)
    if nside < 1 or (nside & (nside - 1)) != 0:
        raise ValueError("nside must be a power of 2")       # ring_tables.py:19
    nr = 4 * nside - 1
    i = np.arange(1, nr + 1, dtype=np.int64)
    north = i < nside
    south = i > 3 * nside
    ip = np.where(south, 4 * nside - i, i)                   # mirrored index
    z = np.empty(nr, dtype=np.float64)
    # polar caps: z = 1 - i^2/(3 nside^2)  (ring_tables_healpix.cu:17,31)
    cap = north | south
    zc = 1.0 - (ip[cap] * ip[cap]).astype(np.float64) / (3.0 * float(nside * nside))
    z[cap] = np.where(south[cap], -zc, zc)
    eq = ~cap
    # equator: z = (2 nside - i) * 2/(3 nside)  (ring_tables_healpix.cu:23)
    z[eq] = (2 * nside - i[eq]).astype(np.float64) * (2.0 / (3.0 * float(nside)))
    nphi = np.where(cap, 4 * ip, 4 * nside).astype(np.int32)
    dphi = 2.0 * np.pi / nphi.astype(np.float64)
    s_eq = ((i - nside) & 1).astype(np.float64)
    # reference phi_center0 (ring_tables_healpix.cu:20-34): polar 0.5*dphi,
    # equatorial (1 - 0.5*((i-nside)&1))*dphi.  Unused by the reference transforms.
    phi0_ref = np.where(cap, 0.5 * dphi, (1.0 - 0.5 * s_eq) * dphi)
    # HEALPix pixel centres: 0.5*dphi on polar rings, equatorial 0.5*dphi when
    # (i-nside) is even and 0 when odd.
    phi0_hpx = np.where(cap, 0.5 * dphi, np.where(s_eq == 0, 0.5 * dphi, 0.0))
    theta = np.arccos(np.clip(z, -1.0, 1.0))
    ring_start = np.zeros(nr + 1, dtype=np.int64)
    ring_start[1:] = np.cumsum(nphi.astype(np.int64))
    z_edges = np.empty(nr + 1, dtype=np.float64)
    z_edges[0] = 1.0
    z_edges[1:nr] = 0.5 * (z[:-1] + z[1:])
    z_edges[nr] = -1.0
    weights_dz = z_edges[:-1] - z_edges[1:]                  # ring_tables.py:39-40
    npix = int(ring_start[-1])
    assert npix == 12 * nside * nside                        # ring_tables.py:37
    weights_area = 2.0 * nphi.astype(np.float64) / float(npix)
    return {
        "nside": nside, "Nr": nr, "npix": npix,
        "z_centers": z, "theta_centers": theta, "nphi": nphi,
        "phi_center0": phi0_ref, "phi0_healpix": phi0_hpx,
        "ring_start": ring_start, "z_edges": z_edges,
        "weights_dz": weights_dz, "weights_area": weights_area,
    }


def flat_to_rings(flat, tabs):
    """``udgrade.unpack_flat_to_rings`` (udgrade.py:96-102)."""
    rs = tabs["ring_start"]
    return [np.asarray(flat[rs[r]:rs[r + 1]]) for r in range(tabs["Nr"])]


def rings_to_flat(rings):
    """``udgrade.pack_rings_to_flat`` (udgrade.py:86-94)."""
    return np.concatenate([np.asarray(r) for r in rings])
