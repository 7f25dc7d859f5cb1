"""Spin-weighted transforms, Q/U <-> E/B (reference: ``spin.py:32-115``).

Convention (healpy; the reference's analysis ``spin.py:73-78`` implies the
same, its synthesis ``spin.py:84-85`` flips the sign of B):
    Q + iU = sum  s_a_lm  sY_lm,    s_a_lm = -(E_lm + i B_lm)
    Q - iU = sum -s_a_lm -sY_lm,   -s_a_lm = -(-1)^s (E_lm - i B_lm)
with sY_lm of Goldberg et al. (1967).  All alm are m-major
(``alm_ops.to_reference_layout`` converts to the reference's l(l+1)+m).
The l-recursion replaces the reference's per-element Wigner sum
(``wigner_d_ms.cu:20-50``), whose exponents are wrong.
"""
import numpy as np
import torch

from . import _util
from .plan import get_plan, nalm
from .sht import make_geometry


def _pair_maps(Q, U, npix):
    q = _util.flat_maps(Q, npix)
    u = _util.flat_maps(U, npix, dtype=q.dtype)
    if q.shape != u.shape:
        raise ValueError("Q and U must have the same shape")
    return torch.stack([q, u], dim=-2)


def eb_to_spin_pair(E, B, spin):
    """(s_a, -s_a) = (-(E + iB), -(-1)^s (E - iB))."""
    sg = -1.0 if spin % 2 else 1.0
    return -(E + 1j * B), -sg * (E - 1j * B)


def spin_pair_to_eb(a_s, a_ms, spin):
    """Inverse of eb_to_spin_pair: E = -(s_a + (-1)^s -s_a)/2, B = i (s_a - (-1)^s -s_a)/2."""
    sg = -1.0 if spin % 2 else 1.0
    return -(a_s + sg * a_ms) * 0.5, 1j * (a_s - sg * a_ms) * 0.5


def eb_from_qu_gpu(Q_rings, U_rings, thetas, nphi_per_ring, weights_dz, lmax, mmax=None, *, spin=2,
                   phi0=None, alias="reference", precision="f64"):
    """(E, B) m-major from Q, U rings (``spin.py:73-78``), 2 pi sum_r w_r convention.

    Q_rings/U_rings: per-ring lists (reference form) or flat [npix] /
    [nmaps, npix] tensors.  Returns E, B of shape [nalm] or [nmaps, nalm].
    """
    mmax = lmax if mmax is None else int(mmax)
    geom = make_geometry(thetas, nphi_per_ring, weights_dz, phi0, alias, mmax)
    plan = get_plan(geom, int(lmax), mmax, precision, spin=int(spin))
    eb = plan.map2alm_spin(_pair_maps(Q_rings, U_rings, geom.npix))
    return eb[..., 0, :], eb[..., 1, :]


def map2alm_spin_gpu(Q_rings, U_rings, thetas, nphi_per_ring, weights_dz, lmax, mmax=None, *, spin=2, **kw):
    """(s_a_lm, -s_a_lm) m-major (``spin.py:32-71`` returns this pair, l-major)."""
    E, B = eb_from_qu_gpu(Q_rings, U_rings, thetas, nphi_per_ring, weights_dz, lmax, mmax, spin=spin, **kw)
    return eb_to_spin_pair(E, B, spin)


def qu_from_eb_gpu(Ealm, Balm, thetas, nphi_per_ring, lmax, mmax=None, *, spin=2, dtype=torch.float64,
                   phi0=None, alias="reference", precision="f64", flat=False):
    """Q, U from m-major (E, B) (``spin.py:80-115``).

    Returns per-ring views of Q and U (reference form) or, with
    ``flat=True``, flat [npix] / [nmaps, npix] tensors.
    """
    mmax = lmax if mmax is None else int(mmax)
    geom = make_geometry(thetas, nphi_per_ring, None, phi0, alias, mmax)
    plan = get_plan(geom, int(lmax), mmax, precision, spin=int(spin))
    E = torch.as_tensor(Ealm)
    B = torch.as_tensor(Balm)
    if E.shape != B.shape or E.shape[-1] != nalm(int(lmax), mmax):
        raise ValueError("E and B must both be m-major alm of lmax/mmax")
    dt = dtype if isinstance(dtype, torch.dtype) else torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
    qu = plan.alm2map_spin(torch.stack([E.to(plan.device), B.to(plan.device)], dim=-2), dtype=dt)
    Q, U = qu[..., 0, :], qu[..., 1, :]
    if flat:
        return Q, U
    sizes = geom.nphi.tolist()
    return list(torch.split(Q, sizes, dim=-1)), list(torch.split(U, sizes, dim=-1))


def alm2map_spin_gpu(a_s, a_ms, thetas, nphi_per_ring, lmax, mmax=None, *, spin=2, **kw):
    """Q, U from the spin pair (s_a_lm, -s_a_lm), m-major."""
    E, B = spin_pair_to_eb(torch.as_tensor(a_s), torch.as_tensor(a_ms), spin)
    return qu_from_eb_gpu(E, B, thetas, nphi_per_ring, lmax, mmax, spin=spin, **kw)
