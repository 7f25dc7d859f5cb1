"""NumPy restatement of the reference's RING pixel <-> angle rule.

Test infrastructure only (the product runs the HIP kernels of libhbx).
  ang2pix: ``coords_ring.py:18-24`` (ring = searchsorted over -z_edges) and
           ``coords_ring.cu:9-32`` (nearest phi centre, llround, wrap 2 pi)
  pix2ang: ``coords_ring.cu:33-54``
Pinned in tests/test_oracle.py by closed forms (pixel centres of the HEALPix
grid), the round trip ang2pix(pix2ang(p)) = p and a brute-force nearest-centre
search.
"""
import numpy as np

TWO_PI = 6.28318530717958647692528676655900577


def ring_of_z(z, z_edges):
    """coords_ring.py:18-24: searchsorted(-z_edges, -z, 'right') - 1, clipped."""
    nr = len(z_edges) - 1
    idx = np.searchsorted(-np.asarray(z_edges), -np.asarray(z, dtype=np.float64), side="right") - 1
    return np.clip(idx, 0, nr - 1)


def ang2pix(theta, phi, tabs, phi0=None):
    """coords_ring.cu:9-32 over the z-edge ring cells."""
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    p0 = tabs["phi_center0"] if phi0 is None else phi0
    r = ring_of_z(np.cos(theta), tabs["z_edges"])
    npr = tabs["nphi"][r].astype(np.int64)
    dphi = TWO_PI / npr
    d = np.fmod(phi - p0[r], TWO_PI)
    d = np.where(d < 0.0, d + TWO_PI, d)
    q = d / dphi                                      # >= 0
    j = np.floor(q)
    j = (j + (q - j >= 0.5)).astype(np.int64)         # llround, exactly (q - floor q is exact)
    j = np.where(j >= npr, j - npr, j)
    return tabs["ring_start"][r] + j


def pix2ang(ipix, tabs, phi0=None):
    """coords_ring.cu:33-54 (ring found by search instead of a linear scan)."""
    p = np.asarray(ipix, dtype=np.int64)
    p0 = tabs["phi_center0"] if phi0 is None else phi0
    rs = tabs["ring_start"]
    r = np.searchsorted(rs, p, side="right") - 1
    npr = tabs["nphi"][r].astype(np.int64)
    f = p0[r] + (p - rs[r]).astype(np.float64) * (TWO_PI / npr)
    f = np.where(f >= TWO_PI, f - TWO_PI * np.floor(f / TWO_PI), f)
    return tabs["theta_centers"][r], f
