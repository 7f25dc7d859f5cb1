"""HEALPix RING geometry tables (drop-in for ``ring_tables.HealpixRingTables``).

Reference: ``ring_tables.py:17-53`` + kernel ``ring_tables_healpix.cu:5-52``.
The reference builds the tables in a CUDA kernel whose thread 0 prefix-sums
values written by other blocks without synchronisation
(``ring_tables_healpix.cu:40-51``).  The tables are O(Nr) integer geometry,
so they are computed exactly on the host here (integer ring indices, one
rounding per value) and handed to the device once as torch tensors.
"""
import numpy as np
import torch


def _default_device():
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


def healpix_geometry(nside):
    """Exact NumPy tables for one nside (ring i = 1..4nside-1 stored at i-1)."""
    nside = int(nside)
    if nside < 1 or (nside & (nside - 1)) != 0:
        raise ValueError("nside must be a power of 2")
    nr = 4 * nside - 1
    i = np.arange(1, nr + 1, dtype=np.int64)
    cap_n = i < nside
    cap_s = i > 3 * nside
    cap = cap_n | cap_s
    ip = np.where(cap_s, 4 * nside - i, i)
    # 1 - |z| on the caps is i'^2/(3 nside^2) exactly representable before one division
    one_minus = (ip * ip).astype(np.float64) / (3.0 * float(nside) * float(nside))
    z = np.where(cap, 1.0 - one_minus, (2 * nside - i).astype(np.float64) * (2.0 / (3.0 * nside)))
    z = np.where(cap_s, -z, z)
    nphi = np.where(cap, 4 * ip, 4 * nside).astype(np.int32)
    dphi = 2.0 * np.pi / nphi
    parity = ((i - nside) & 1).astype(np.float64)
    phi0_ref = np.where(cap, 0.5 * dphi, (1.0 - 0.5 * parity) * dphi)   # ring_tables_healpix.cu:20-34
    phi0_hpx = np.where(cap, 0.5 * dphi, np.where(parity == 0, 0.5 * dphi, 0.0))
    theta = np.arccos(np.clip(z, -1.0, 1.0))
    ring_start = np.zeros(nr + 1, dtype=np.int64)
    ring_start[1:] = np.cumsum(nphi.astype(np.int64))
    z_edges = np.empty(nr + 1)
    z_edges[0] = 1.0
    z_edges[1:nr] = 0.5 * (z[:-1] + z[1:])
    z_edges[nr] = -1.0
    return dict(nside=nside, Nr=nr, npix=int(ring_start[-1]), z_centers=z, theta_centers=theta, nphi=nphi,
                phi_center0=phi0_ref, phi0_healpix=phi0_hpx, ring_start=ring_start, z_edges=z_edges)


class HealpixRingTables:
    """RING-ordered ring tables with the reference's attribute names.

    Attributes are torch tensors on ``device`` (the GPU by default):
    ``z_centers``, ``theta_centers``, ``nphi`` (int32), ``phi_center0``,
    ``ring_start`` (int64, Nr+1), ``z_edges`` (Nr+1); plus ``nside``, ``Nr``,
    ``npix``.  ``phi_center0`` keeps the reference's values
    (``ring_tables_healpix.cu:20-34``); ``phi0_healpix`` holds HEALPix's.
    """

    def __init__(self, nside: int, device=None):
        g = healpix_geometry(nside)
        assert g["npix"] == 12 * g["nside"] ** 2, "npix mismatch"     # ring_tables.py:37
        self._host = g
        self.nside = g["nside"]
        self.Nr = g["Nr"]
        self.npix = g["npix"]
        dev = _default_device() if device is None else torch.device(device)
        self.device = dev
        self.z_centers = torch.as_tensor(g["z_centers"], device=dev)
        self.theta_centers = torch.as_tensor(g["theta_centers"], device=dev)
        self.nphi = torch.as_tensor(g["nphi"], device=dev)
        self.phi_center0 = torch.as_tensor(g["phi_center0"], device=dev)
        self.phi0_healpix = torch.as_tensor(g["phi0_healpix"], device=dev)
        self.ring_start = torch.as_tensor(g["ring_start"], device=dev)
        self.z_edges = torch.as_tensor(g["z_edges"], device=dev)

    def weights_dz(self):
        """Midpoint-Delta-z quadrature weights (``ring_tables.py:39-40``); sum 2."""
        return self.z_edges[:-1] - self.z_edges[1:]

    def weights_area(self):
        """Exact HEALPix equal-area weights 2 nphi_r / npix; sum 2."""
        return 2.0 * self.nphi.to(torch.float64) / float(self.npix)

    def host(self, name):
        return self._host[name]

    def dump_host(self):
        """Host copies of the tables (``ring_tables.py:42-53``)."""
        h = self._host
        return {
            "nside": self.nside, "Nr": self.Nr, "npix": self.npix,
            "z_centers": h["z_centers"].copy(), "theta_centers": h["theta_centers"].copy(),
            "nphi": h["nphi"].copy(), "phi_center0": h["phi_center0"].copy(),
            "ring_start": h["ring_start"].copy(), "z_edges": h["z_edges"].copy(),
        }


def unpack_flat_to_rings(flat, tables):
    """Per-ring views of a flat RING map (``udgrade.py:96-102``): ring r is
    ``flat[..., ring_start[r] : ring_start[r] + nphi[r]]``.  One split on the
    host-side ring lengths; no copies, no per-ring device reads."""
    t = torch.as_tensor(flat)
    if t.shape[-1] != tables.npix:
        raise ValueError(f"map has {t.shape[-1]} pixels but the ring tables describe {tables.npix}")
    return list(torch.split(t, [int(n) for n in tables._host["nphi"]], dim=-1))


def pack_rings_to_flat(rings, tables):
    """Flat RING map from per-ring arrays (``udgrade.py:86-94``), one
    concatenation; ring r must hold nphi[r] pixels."""
    nphi = tables._host["nphi"]
    if len(rings) != tables.Nr:
        raise ValueError(f"{len(rings)} rings given for {tables.Nr} ring tables")
    parts = [torch.as_tensor(r) for r in rings]
    for r, (p, n) in enumerate(zip(parts, nphi)):
        if p.shape[-1] != int(n):
            raise ValueError(f"ring {r} has {p.shape[-1]} pixels, expected {int(n)}")
    dev = parts[0].device
    return torch.cat([p.to(dev) for p in parts], dim=-1)
