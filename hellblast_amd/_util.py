"""Host-side argument normalisation shared by the reference-signature API."""
import numpy as np
import torch


def host_array(x, dtype):
    """Host NumPy copy of a NumPy array, torch tensor or sequence."""
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy().astype(dtype, copy=False)
    return np.asarray(x, dtype=dtype)


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("hellblast_amd transforms need a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def flat_maps(rings, npix, dtype=None):
    """A device tensor [npix] or [nmaps, npix] from a flat map or a ring list.

    A flat tensor already on the device is used without a copy; a list of
    per-ring arrays (the reference's calling form) is concatenated once.
    """
    dev = device()
    if isinstance(rings, (list, tuple)):
        parts = [torch.as_tensor(r) for r in rings]
        t = torch.cat([p.reshape(-1).to(dev) for p in parts])
    else:
        t = torch.as_tensor(rings)
        if t.device != dev:
            t = t.to(dev)
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    elif t.dtype not in (torch.float64, torch.float32):
        t = t.to(torch.float64)
    if t.shape[-1] != npix:
        raise ValueError(f"map has {t.shape[-1]} pixels but the ring tables describe {npix}")
    return t


def alias_mcut(alias, nphi, mmax):
    """Per-ring m cut: 'reference' -> nphi (sht_raw.py:51), 'fold' -> mmax+1."""
    if alias in (None, "reference", "ref", "zero"):
        return None
    if alias in ("fold", "periodic", "healpy"):
        return np.full(len(nphi), mmax + 1, dtype=np.int32)
    raise ValueError("alias must be 'reference' or 'fold'")
