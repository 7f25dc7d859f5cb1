"""Spin-1 transforms (reference: ``spin1.py:5-69``): the spin = 1 case of
``hellblast_amd.spin``, returning / taking the pair (1_a_lm, -1_a_lm) with
1_a = -(E + iB), -1_a = (E - iB), m-major.
"""
from .spin import alm2map_spin_gpu, map2alm_spin_gpu


def map2alm_spin1_gpu(X_rings, Y_rings, thetas, nphi_per_ring, weights_dz, lmax, mmax=None, **kw):
    """``spin1.map2alm_spin1_gpu`` (spin1.py:5-38)."""
    return map2alm_spin_gpu(X_rings, Y_rings, thetas, nphi_per_ring, weights_dz, lmax, mmax, spin=1, **kw)


def alm2map_spin1_gpu(a1, am1, thetas, nphi_per_ring, lmax, mmax=None, **kw):
    """``spin1.alm2map_spin1_gpu`` (spin1.py:40-69)."""
    return alm2map_spin_gpu(a1, am1, thetas, nphi_per_ring, lmax, mmax, spin=1, **kw)
