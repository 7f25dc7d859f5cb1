"""ctypes binding of libhbx.so (declared in include/hbx.h).

The library must be built in-tree (``python -m hellblast_amd.build``); there
is no fallback: importing the transform modules without it raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhbx.so")

HBX_OK = 0
HBX_EINVAL = -1
HBX_ENOMEM = -2
HBX_EDEVICE = -3
HBX_EWORKSPACE = -4
HBX_EUNSUPPORTED = -5
HBX_EPLAN = -6
HBX_MAX_SPIN = 8
HBX_STATIC_SCRATCH_BYTES = 64 << 20
HBX_MFMA_F32 = 1
HBX_MFMA_F16X2 = 2
HBX_TIMER_LEGENDRE_ANALYSIS = 1
HBX_TIMER_LEGENDRE_SYNTHESIS = 2
HBX_TIMER_RING_ANALYSIS = 3
HBX_TIMER_RING_SYNTHESIS = 4
HBX_TIMER_SPIN_ANALYSIS = 5
HBX_TIMER_SPIN_SYNTHESIS = 6

HBX_F64 = 0
HBX_F32 = 1
HBX_DT_F64 = 0
HBX_DT_F32 = 1


class HbxError(RuntimeError):
    """A libhbx call failed; ``code`` is the hbx_status value."""

    def __init__(self, code, msg):
        super().__init__(f"hbx error {code}: {msg}")
        self.code = code


class HbxValueError(HbxError, ValueError):
    pass


class HbxUnsupported(HbxError, NotImplementedError):
    pass


class hbx_geom(ctypes.Structure):
    _fields_ = [
        ("nrings", ctypes.c_int32),
        ("ring_start", ctypes.POINTER(ctypes.c_int64)),
        ("nphi", ctypes.POINTER(ctypes.c_int32)),
        ("cos_theta", ctypes.POINTER(ctypes.c_double)),
        ("sin_theta", ctypes.POINTER(ctypes.c_double)),
        ("phi0", ctypes.POINTER(ctypes.c_double)),
        ("weight", ctypes.POINTER(ctypes.c_double)),
        ("mcut", ctypes.POINTER(ctypes.c_int32)),
    ]


class hbx_plan_info(ctypes.Structure):
    _fields_ = [
        ("nrings", ctypes.c_int32), ("npairs", ctypes.c_int32), ("lmax", ctypes.c_int32),
        ("mmax", ctypes.c_int32), ("precision", ctypes.c_int32), ("spin", ctypes.c_int32),
        ("npix", ctypes.c_int64), ("nalm", ctypes.c_int64), ("device_bytes", ctypes.c_uint64),
        ("fp64_band_m", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


class hbx_tiles(ctypes.Structure):
    _fields_ = [("valu_slots", ctypes.c_int32), ("mfma_waves", ctypes.c_int32),
                ("mfma_min_maps", ctypes.c_int32), ("spin_slots", ctypes.c_int32),
                ("mfma_kind", ctypes.c_int32), ("mfma_groups", ctypes.c_int32)]


class hbx_multi_tiles(ctypes.Structure):
    _fields_ = [("valu_slots", ctypes.c_int32), ("syn_maps", ctypes.c_int32),
                ("tiles_per_wave", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class hbx_multi_info(ctypes.Structure):
    _fields_ = [("ngroups", ctypes.c_int32), ("precision", ctypes.c_int32), ("valu_slots", ctypes.c_int32),
                ("syn_maps", ctypes.c_int32), ("tiles_per_wave", ctypes.c_int32),
                ("analysis_tiles", ctypes.c_int32), ("synthesis_tiles", ctypes.c_int32),
                ("analysis_workgroups", ctypes.c_int32), ("synthesis_workgroups", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("workspace_bytes", ctypes.c_uint64)]


HBX_MULTI_MAX_GROUPS = 16

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64

# name -> (restype, argtypes); every symbol declared in include/hbx.h
SIGNATURES = {
    "hbx_abi_version": (ctypes.c_int, []),
    "hbx_last_error": (ctypes.c_char_p, []),
    "hbx_validate_geom": (ctypes.c_int, [ctypes.POINTER(hbx_geom), _i32, _i32]),
    "hbx_plan_create": (ctypes.c_int, [ctypes.POINTER(hbx_geom), _i32, _i32, _i32, _i32, ctypes.c_uint32,
                                       ctypes.POINTER(_vp)]),
    "hbx_plan_destroy": (ctypes.c_int, [_vp]),
    "hbx_plan_get_info": (ctypes.c_int, [_vp, ctypes.POINTER(hbx_plan_info)]),
    "hbx_workspace_bytes": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_u64)]),
    "hbx_plan_derive": (ctypes.c_int, [_vp, ctypes.POINTER(hbx_tiles), ctypes.POINTER(_vp)]),
    "hbx_plan_get_tiles": (ctypes.c_int, [_vp, ctypes.POINTER(hbx_tiles)]),
    "hbx_map2alm": (ctypes.c_int, [_vp, _i32, _vp, _i32, _i64, _vp, _i64, _vp, _u64, _vp]),
    "hbx_alm2map": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _i32, _i64, _vp, _u64, _vp]),
    "hbx_map2alm_static": (ctypes.c_int, [_vp, _i32, _vp, _i32, _i64, _vp, _i64, _vp]),
    "hbx_alm2map_adjoint": (ctypes.c_int, [_vp, _i32, _vp, _i32, _i64, _vp, _i64, _vp, _u64, _vp]),
    "hbx_map2alm_adjoint": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _i32, _i64, _vp, _u64, _vp]),
    "hbx_map2alm_spin": (ctypes.c_int, [_vp, _i32, _vp, _i32, _i64, _vp, _i64, _vp, _u64, _vp]),
    "hbx_alm2map_spin": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _i32, _i64, _vp, _u64, _vp]),
    "hbx_alm2map_spin_adjoint": (ctypes.c_int, [_vp, _i32, _vp, _i32, _i64, _vp, _i64, _vp, _u64, _vp]),
    "hbx_map2alm_spin_adjoint": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _i32, _i64, _vp, _u64, _vp]),
    "hbx_timer_arm": (ctypes.c_int, [ctypes.c_uint32]),
    "hbx_timer_read": (ctypes.c_int, [_i32, ctypes.POINTER(_i32), ctypes.POINTER(ctypes.c_double)]),
    "hbx_alm2cl": (ctypes.c_int, [_i32, _i32, _i32, _vp, _i32, _i64, _vp, _i64, _vp]),
    "hbx_ang2pix_ring": (ctypes.c_int, [_i64, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "hbx_pix2ang_ring": (ctypes.c_int, [_i64, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "hbx_phase_bytes": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_u64)]),
    "hbx_ring_analysis": (ctypes.c_int, [_vp, _i32, _vp, _i32, _i64, _vp, _u64, _vp]),
    "hbx_legendre_analysis": (ctypes.c_int, [_vp, _i32, _vp, _u64, _vp, _i64, _vp, _u64, _vp]),
    "hbx_legendre_synthesis": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _u64, _vp, _u64, _vp]),
    "hbx_ring_synthesis": (ctypes.c_int, [_vp, _i32, _vp, _u64, _vp, _i32, _i64, _vp]),
    "hbx_multi_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(_i32), _i32,
                                        ctypes.POINTER(hbx_multi_tiles), ctypes.POINTER(_vp)]),
    "hbx_multi_destroy": (ctypes.c_int, [_vp]),
    "hbx_multi_workspace_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(_u64)]),
    "hbx_multi_get_info": (ctypes.c_int, [_vp, ctypes.POINTER(hbx_multi_info)]),
    "hbx_multi_map2alm": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_i64), _i32, ctypes.POINTER(_vp),
                                         ctypes.POINTER(_i64), _vp, _u64, _vp]),
    "hbx_multi_alm2map": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_i64), ctypes.POINTER(_vp),
                                         ctypes.POINTER(_i64), _i32, _vp, _u64, _vp]),
}

_lib = None


def lib():
    """Load libhbx.so once; raise ImportError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not found: build it with `python -m hellblast_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    msg = lib().hbx_last_error()
    return msg.decode() if msg else ""


def check(rc):
    """Raise the matching HbxError subclass for a non-zero status."""
    if rc == HBX_OK:
        return
    msg = last_error()
    if rc == HBX_EINVAL or rc == HBX_EWORKSPACE or rc == HBX_EPLAN:
        raise HbxValueError(rc, msg)
    if rc == HBX_EUNSUPPORTED:
        raise HbxUnsupported(rc, msg)
    raise HbxError(rc, msg)
