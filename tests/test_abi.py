"""C-ABI checks that need no GPU: symbols, version, host-side validation."""
import ctypes
import os
import re

import numpy as np
import pytest

from hellblast_amd import _abi
from hellblast_amd.hp_compat import healpix_geom
from hellblast_amd.plan import Geometry

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hbx.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(hbx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _abi.lib()
    names = _declared()
    assert names, "no declarations parsed"
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_abi.SIGNATURES), "ctypes signatures out of sync with include/hbx.h"


def test_abi_version_matches_header():
    v = int(re.search(r"#define HBX_ABI_VERSION (\d+)", open(HEADER).read()).group(1))
    assert _abi.lib().hbx_abi_version() == v


def _geom(nside=4):
    # a private copy: healpix_geom() results are cached and must not be mutated
    c = healpix_geom(nside, 3 * nside)
    return Geometry(c.nphi.copy(), cos_theta=c.cos_theta.copy(), sin_theta=c.sin_theta.copy(),
                    weight=c.weight.copy(), ring_start=c.ring_start.copy())


def _validate(g, lmax=8, mmax=8):
    return _abi.lib().hbx_validate_geom(ctypes.byref(g.as_struct()), lmax, mmax)


def test_valid_healpix_geometry():
    assert _validate(_geom()) == _abi.HBX_OK


@pytest.mark.parametrize("lmax,mmax", [(-1, 0), (4, 5), (4, -1)])
def test_bad_lmax_mmax(lmax, mmax):
    assert _validate(_geom(), lmax, mmax) == _abi.HBX_EINVAL
    assert "mmax" in _abi.last_error() or "lmax" in _abi.last_error()


def test_odd_nphi_rejected():
    g = _geom()
    nphi = g.nphi.copy()
    nphi[3] += 1
    g = Geometry(nphi, cos_theta=g.cos_theta, sin_theta=g.sin_theta)
    assert _validate(g) == _abi.HBX_EINVAL and "even" in _abi.last_error()


def test_ring_start_mismatch_rejected():
    g = _geom()
    rs = g.ring_start.copy()
    rs[5] += 2
    g.ring_start = rs
    assert _validate(g) == _abi.HBX_EINVAL and "ring_start" in _abi.last_error()


def test_non_mirror_geometry_rejected():
    g = _geom()
    c = g.cos_theta.copy()
    s = g.sin_theta.copy()
    c[2] = np.cos(0.5)
    s[2] = np.sin(0.5)
    g.cos_theta, g.sin_theta = c, s
    assert _validate(g) == _abi.HBX_EUNSUPPORTED and "mirror" in _abi.last_error()


def test_bad_trig_rejected():
    g = _geom()
    c = g.cos_theta.copy()
    c[0] = 2.0
    g.cos_theta = c
    assert _validate(g) == _abi.HBX_EINVAL


def test_null_geometry_and_arrays():
    L = _abi.lib()
    assert L.hbx_validate_geom(None, 4, 4) == _abi.HBX_EINVAL
    g = _abi.hbx_geom()
    g.nrings = 3
    assert L.hbx_validate_geom(ctypes.byref(g), 4, 4) == _abi.HBX_EINVAL


def test_plan_create_argument_errors():
    L = _abi.lib()
    g = _geom()
    h = ctypes.c_void_p()
    assert L.hbx_plan_create(ctypes.byref(g.as_struct()), 8, 8, 0, 0, 0, None) == _abi.HBX_EINVAL
    assert L.hbx_plan_create(ctypes.byref(g.as_struct()), 8, 8, 0, 7, 0, ctypes.byref(h)) == _abi.HBX_EINVAL
    assert L.hbx_plan_create(ctypes.byref(g.as_struct()), 8, 8, 0, 0, 1, ctypes.byref(h)) == _abi.HBX_EINVAL
    assert L.hbx_plan_create(ctypes.byref(g.as_struct()), 8, 9, 0, 0, 0, ctypes.byref(h)) == _abi.HBX_EINVAL
    assert not h.value


def test_plan_create_without_device_reports_edevice():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    g = _geom()
    h = ctypes.c_void_p()
    rc = _abi.lib().hbx_plan_create(ctypes.byref(g.as_struct()), 8, 8, 0, 0, 0, ctypes.byref(h))
    assert rc == _abi.HBX_EDEVICE and not h.value


def test_execute_argument_errors():
    L = _abi.lib()
    assert L.hbx_map2alm(None, 1, None, 0, 0, None, 0, None, 0, None) == _abi.HBX_EINVAL
    assert L.hbx_alm2map(None, 1, None, 0, None, 0, 0, None, 0, None) == _abi.HBX_EINVAL
    assert L.hbx_alm2cl(4, 5, 1, None, 0, 15, None, 5, None) == _abi.HBX_EINVAL
    assert L.hbx_alm2cl(4, 4, 1, None, 0, 15, None, 5, None) == _abi.HBX_EINVAL
    assert L.hbx_alm2cl(4, 4, 0, None, 0, 15, None, 5, None) == _abi.HBX_OK
    assert L.hbx_workspace_bytes(None, 1, None) == _abi.HBX_EINVAL


def test_python_errors_map_to_exceptions():
    with pytest.raises(ValueError):
        _abi.check(_abi.HBX_EINVAL)
    with pytest.raises(NotImplementedError):
        _abi.check(_abi.HBX_EUNSUPPORTED)
    with pytest.raises(RuntimeError):
        _abi.check(_abi.HBX_EDEVICE)
