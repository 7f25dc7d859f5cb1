"""hellblast_amd - MI355X-native HEALPix spherical-harmonic transforms.

Drop-in for the SHT hot path of hellblast: ``map2alm_gpu`` (sht.py),
``alm2map_gpu`` (isht.py), ``anafast_ring`` (anafast.py), the ``hp_compat``
SHT and spin wrappers, ``spin`` / ``spin1`` (Q/U <-> E/B) and
``HealpixRingTables``; next to the path, beam / pixel-window smoothing,
``synfast_tqu``, RING ``ang2pix`` / ``pix2ang`` and FITS map / alm files;
running hand-written gfx950 HIP kernels (libhbx.so, C ABI in include/hbx.h)
on PyTorch-ROCm tensors.

Package aliases as in the reference (``__init__.py:125-127``):
``map2alm``, ``alm2map``, ``anafast``.
"""
from .ring_tables import HealpixRingTables, pack_rings_to_flat, unpack_flat_to_rings
from .sht import map2alm_gpu
from .isht import alm2map_gpu
from .anafast import anafast_ring
from .alm_ops import alm2cl, almxfl, synalm, alm_index, getidx, getsize, to_reference_layout, from_reference_layout
from .spin import map2alm_spin_gpu, eb_from_qu_gpu, qu_from_eb_gpu
from .spin1 import map2alm_spin1_gpu, alm2map_spin1_gpu
from .smoothing import gaussian_beam_fl, smooth_alm_gaussian
from .beams import pixel_window_fl, smooth_alm_with_beam, smooth_map
from .synfast_tqu import synfast_tqu
from .coords_ring import RingMapperGPU, vec2pix_ring_gpu
from .io_fits import read_map_fits, write_map_fits, read_alm_fits, write_alm_fits
from .io_fits_multi import read_tqu_fits, write_tqu_fits
from . import hp_compat

map2alm = map2alm_gpu
alm2map = alm2map_gpu
anafast = anafast_ring

__all__ = [
    "HealpixRingTables", "map2alm_gpu", "alm2map_gpu", "anafast_ring", "alm2cl", "almxfl", "synalm",
    "alm_index", "getidx", "getsize", "to_reference_layout", "from_reference_layout", "hp_compat",
    "map2alm", "alm2map", "anafast", "map2alm_spin_gpu", "eb_from_qu_gpu", "qu_from_eb_gpu",
    "map2alm_spin1_gpu", "alm2map_spin1_gpu", "gaussian_beam_fl", "smooth_alm_gaussian", "pixel_window_fl",
    "smooth_alm_with_beam", "smooth_map", "synfast_tqu", "RingMapperGPU", "vec2pix_ring_gpu",
    "read_map_fits", "write_map_fits", "read_alm_fits", "write_alm_fits", "read_tqu_fits", "write_tqu_fits",
    "pack_rings_to_flat", "unpack_flat_to_rings",
]
