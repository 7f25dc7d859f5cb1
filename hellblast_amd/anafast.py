"""anafast on ring lists (drop-in for ``anafast.anafast_ring``, anafast.py:5-7)."""
from .alm_ops import alm2cl
from .sht import map2alm_gpu


def anafast_ring(rings, thetas, nphi_per_ring, weights_dz, lmax, *, precision="f64", alias="reference",
                 phi0=None):
    """C_l of map2alm_gpu(rings, ...), computed on the GPU."""
    alm = map2alm_gpu(rings, thetas, nphi_per_ring, weights_dz, lmax, precision=precision, alias=alias,
                      phi0=phi0)
    return alm2cl(alm, lmax=lmax)
