"""hbx_map2alm_static (the no-allocation variant of hbx_map2alm, workspace in
a static device scratch array) called through the C ABI: its alm equal
hbx_map2alm's bit for bit (same kernels, same summation order), on two calls
in a row with the second on a smaller input, for FP64 plans (VALU kernels)
and an FP32 batch (split-fp16 matrix-core kernels and the FP64 band), and
sizes above the scratch are refused with a clear error."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu


def _plan(nside, lmax, precision):
    from hellblast_amd.hp_compat import _plan as hp_plan
    return hp_plan(nside, lmax, lmax, precision, "dz", "reference", "reference")


def _static(plan, maps):
    from hellblast_amd import _abi
    from hellblast_amd.plan import _map_dtype_code
    nb = maps.shape[0]
    out = torch.empty((nb, plan.nalm), dtype=plan.alm_dtype, device=plan.device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(plan.device).cuda_stream)
    rc = _abi.lib().hbx_map2alm_static(plan._h, nb, ctypes.c_void_p(maps.data_ptr()), _map_dtype_code(maps.dtype),
                                       maps.shape[1], ctypes.c_void_p(out.data_ptr()), plan.nalm, stream)
    return rc, out


def _maps(plan, nb, dtype, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.randn((nb, plan.npix), generator=g, device="cuda", dtype=dtype)


@pytest.mark.parametrize("precision,first,second", [
    ("f64", (256, 511, 2), (64, 128, 1)),
    ("f32", (128, 256, 6), (32, 64, 3)),
])
def test_static_map2alm_equals_map2alm(gpu, precision, first, second):
    from hellblast_amd import _abi
    dtype = torch.float64 if precision == "f64" else torch.float32
    for i, (nside, lmax, nb) in enumerate((first, second)):
        print(f"[case] hbx_map2alm_static {precision} call {i + 1}: nside={nside} lmax={lmax} nmaps={nb}",
              flush=True)
        plan = _plan(nside, lmax, precision)
        assert plan.workspace_bytes(nb) <= _abi.HBX_STATIC_SCRATCH_BYTES
        maps = _maps(plan, nb, dtype, 10 * i + nb)
        ref = plan._map2alm_raw(maps)
        rc, got = _static(plan, maps)
        assert rc == _abi.HBX_OK, _abi.last_error()
        torch.cuda.synchronize()
        assert torch.equal(got, ref)


def test_static_map2alm_refuses_sizes_above_the_scratch(gpu):
    from hellblast_amd import _abi
    plan = _plan(512, 1023, "f64")
    nb = 4
    assert plan.workspace_bytes(nb) > _abi.HBX_STATIC_SCRATCH_BYTES
    maps = _maps(plan, nb, torch.float64, 3)
    rc, _ = _static(plan, maps)
    assert rc == _abi.HBX_EUNSUPPORTED
    assert "static scratch" in _abi.last_error()
