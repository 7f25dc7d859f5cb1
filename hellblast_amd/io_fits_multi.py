"""T, Q, U maps in one FITS file (reference: ``io_fits_multi.py:24-63``).

Layout as the reference writes it: an empty primary HDU, then one IMAGE
extension per map given, named T, Q or U, each with NSIDE (when given),
ORDERING and COORDSYS cards and any extra cards.  Reading returns
((T, Q, U), meta) with meta = {NSIDE, ORDERING, COORDSYS} from the
extensions that carry them; a map not in the file is None.
"""
import numpy as np

from .io_fits import _to_device, _to_host, image_hdu_bytes, list_hdus, read_image, write_file

def write_tqu_fits(path, T=None, Q=None, U=None, nside=None, ordering="RING", coordsys="G", overwrite=True,
                   extra_header=None):
    """(``io_fits_multi.py:24-41``)"""
    hdus = [image_hdu_bytes(None, True)]
    for arr, name in ((T, "T"), (Q, "Q"), (U, "U")):
        if arr is None:
            continue
        hdr = {}
        if nside is not None:
            hdr["NSIDE"] = int(nside)
        hdr["ORDERING"] = str(ordering).upper()
        hdr["COORDSYS"] = coordsys
        for k, v in (extra_header or {}).items():
            hdr[str(k).upper()] = v
        hdus.append(image_hdu_bytes(_to_host(arr), False, hdr, extname=name))
    return write_file(path, hdus, overwrite)


def read_tqu_fits(path, to_gpu=True):
    """(``io_fits_multi.py:43-63``)"""
    maps = {"T": None, "Q": None, "U": None}
    meta = {}
    for h in list_hdus(path)[1:]:
        if h.name in maps:
            maps[h.name] = read_image(path, h)
        for k in ("NSIDE", "ORDERING", "COORDSYS"):
            if k in h.header:
                meta[k] = h.header[k]
    out = tuple(None if maps[k] is None else _to_device(np.asarray(maps[k]), to_gpu) for k in ("T", "Q", "U"))
    return out, meta
