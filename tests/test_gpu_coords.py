"""GPU parity of RING ang2pix / pix2ang (hbx_ang2pix_ring, hbx_pix2ang_ring
through coords_ring.RingMapperGPU) against the oracle (oracle/coords.py):
pixel indices bit-exact, angles equal to the last bit."""
import math

import numpy as np
import pytest
import torch

from oracle import coords as OC
from oracle import geometry as OG

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nside", [1, 4, 64, 2048])
@pytest.mark.parametrize("phi0", ["reference", "healpix"])
def test_ring_coords_match_oracle(gpu, nside, phi0):
    print(f"[case] ring coords nside={nside} phi0={phi0}", flush=True)
    from hellblast_amd.coords_ring import RingMapperGPU, vec2pix_ring_gpu
    t = OG.ring_tables(nside)
    p0 = t["phi_center0"] if phi0 == "reference" else t["phi0_healpix"]
    mp = RingMapperGPU(nside, phi0=phi0)
    rng = np.random.default_rng(nside)
    n = 200000
    th = np.arccos(rng.uniform(-1, 1, n))
    th[:4] = [0.0, math.pi, 1e-300, math.pi / 2]
    ph = rng.uniform(-4 * math.pi, 6 * math.pi, n)
    got = mp.ang2pix(torch.as_tensor(th), torch.as_tensor(ph)).cpu().numpy()
    assert np.array_equal(got, OC.ang2pix(th, ph, t, p0))
    pix = rng.integers(0, t["npix"], n)
    pix[:2] = [0, t["npix"] - 1]
    gt, gp = mp.pix2ang(torch.as_tensor(pix))
    et, ep = OC.pix2ang(pix, t, p0)
    assert np.array_equal(gt.cpu().numpy(), et) and np.array_equal(gp.cpu().numpy(), ep)
    # round trip over pixel centres, and out-of-range indices give NaN
    assert torch.equal(mp.ang2pix(gt, gp).cpu(), torch.as_tensor(pix))
    bt, bp = mp.pix2ang(torch.as_tensor([-1, t["npix"]]))
    assert torch.isnan(bt).all() and torch.isnan(bp).all()
    x, y, z = np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)
    v = vec2pix_ring_gpu(x, y, z, mp).cpu().numpy()
    ang = np.arctan2(y, x)
    ang = np.where(ang < 0, ang + 2 * math.pi, ang)
    assert np.array_equal(v, OC.ang2pix(np.arccos(np.clip(z, -1, 1)), ang, t, p0))


def test_hp_compat_geometry_wrappers(gpu):
    """hp_compat.ang2pix / pix2ang / pix2vec (hp_compat.py:131-149) on the
    RING kernels: the reference's phi centres, unit vectors, and NESTED
    refused."""
    print("[case] hp_compat ang2pix/pix2ang/pix2vec nside=32", flush=True)
    from hellblast_amd import hp_compat
    nside = 32
    t = OG.ring_tables(nside)
    pix = np.arange(t["npix"])
    th, ph = hp_compat.pix2ang(nside, torch.as_tensor(pix))
    et, ep = OC.pix2ang(pix, t, t["phi_center0"])
    assert np.array_equal(th.cpu().numpy(), et) and np.array_equal(ph.cpu().numpy(), ep)
    assert torch.equal(hp_compat.ang2pix(nside, th, ph).cpu(), torch.as_tensor(pix))
    x, y, z = hp_compat.pix2vec(nside, torch.as_tensor(pix))
    assert float((x * x + y * y + z * z - 1).abs().max()) <= 1e-15
    assert np.allclose(z.cpu().numpy(), np.cos(et), rtol=0, atol=1e-15)
    with pytest.raises(NotImplementedError):
        hp_compat.ang2pix(nside, th, ph, nest=True)
