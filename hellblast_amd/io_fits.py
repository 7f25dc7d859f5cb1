"""FITS files for maps and alm: the on-disk format on either side of the SHT.

Reference: ``io_fits.py:11-63`` (read_map_fits, write_map_fits,
read_alm_fits, write_alm_fits) on astropy, which is not a dependency here:
the FITS structures these files use are read and written by a small codec
of this module that follows the FITS Standard 4.0 (IAU FITS Working Group,
2016): 2880-byte blocks, 80-character header cards, big-endian data,
primary / IMAGE / BINTABLE HDUs.

Files written here:
  map  - the map as the primary image (BITPIX -64 / -32 / integer), as the
         reference writes it (``io_fits.py:20-31``);
  alm  - a BINTABLE in HDU 1 with columns INDEX (J, or K past 2^31),
         REAL (D), IMAG (D), INDEX = l^2 + l + m + 1 (the healpy
         convention).  The reference writes REAL and IMAG only, in its own
         l(l+1)+m order (``io_fits.py:46-63``); the INDEX column makes the
         file self-describing and readable by healpy.
Files read here:
  map  - an image HDU, or a BINTABLE (healpy map files: one column with a
         repeat count, rows in pixel order), flattened;
  alm  - INDEX/REAL/IMAG tables by INDEX; REAL/IMAG tables without INDEX
         are the reference's l(l+1)+m layout, (lmax+1)^2 rows.
alm are returned in this package's m-major layout (index
m(2 lmax + 1 - m)/2 + l), the layout map2alm returns and alm2map reads.
"""
import math
import os

import numpy as np
import torch

BLOCK = 2880
CARD = 80

# BITPIX <-> big-endian image element type
_BITPIX = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}
# BINTABLE TFORM type code -> (element dtype, bytes per element)
_TFORM = {"L": ("S1", 1), "B": (">u1", 1), "I": (">i2", 2), "J": (">i4", 4), "K": (">i8", 8),
          "A": ("S1", 1), "E": (">f4", 4), "D": (">f8", 8), "C": (">c8", 8), "M": (">c16", 16)}
_COMMENTARY = ("COMMENT", "HISTORY", "")


# ---------------------------------------------------------------------------
# header cards
# ---------------------------------------------------------------------------
def _format_value(v):
    """Fixed-format value field (FITS 4.0 section 4.2): logicals and numbers
    right-justified to column 30, strings quoted from column 11."""
    if isinstance(v, (bool, np.bool_)):
        return f"{'T' if v else 'F':>20}"
    if isinstance(v, (int, np.integer)):
        return f"{int(v):>20}"
    if isinstance(v, (float, np.floating)):
        f = float(v)
        if not math.isfinite(f):
            raise ValueError("FITS header values must be finite")
        s = repr(f).upper()
        if "." not in s and "E" not in s:
            s += ".0"
        return f"{s:>20}"
    s = str(v).replace("'", "''")
    return "'" + f"{s:<8}" + "'"


def _card(key, value=None, comment=None):
    key = str(key).upper()
    if key in _COMMENTARY:
        text = "" if value is None else str(value)
        c = f"{key:<8}{text}"
    elif len(key) > 8 or " " in key:
        name = key[len("HIERARCH "):] if key.startswith("HIERARCH ") else key
        c = f"HIERARCH {name} = {_format_value(value).strip()}"
    else:
        c = f"{key:<8}= {_format_value(value)}"
        if comment:
            c += f" / {comment}"
    if len(c) > CARD:
        raise ValueError(f"header card longer than {CARD} characters: {c!r}")
    if not c.isascii():
        raise ValueError("FITS header cards are ASCII")
    return f"{c:<{CARD}}"


def _parse_value(text):
    t = text.strip()
    if t.startswith("'"):
        out, i = [], 1
        while i < len(t):
            if t[i] == "'":
                if i + 1 < len(t) and t[i + 1] == "'":
                    out.append("'")
                    i += 2
                    continue
                break
            out.append(t[i])
            i += 1
        return "".join(out).rstrip()
    t = t.split("/", 1)[0].strip()
    if t == "T":
        return True
    if t == "F":
        return False
    if t == "":
        return None
    try:
        return int(t)
    except ValueError:
        pass
    try:
        return float(t.replace("D", "E"))
    except ValueError:
        return t


def _parse_card(c):
    key = c[:8].strip()
    if key.startswith("HIERARCH") and "=" in c:
        name, val = c[9:].split("=", 1)
        return name.strip(), _parse_value(val)
    if c[8:10] == "= " and key not in _COMMENTARY:
        return key, _parse_value(c[10:])
    return key, c[8:].rstrip()


class Header(dict):
    """Header keywords of one HDU (commentary cards gathered in lists)."""

    def add(self, key, value):
        if key in ("COMMENT", "HISTORY"):
            self.setdefault(key, []).append(value)
        elif key:
            self[key] = value


def _pad(n):
    return (-n) % BLOCK


def _header_bytes(cards):
    b = "".join(cards + [f"{'END':<{CARD}}"]).encode("ascii")
    return b + b" " * _pad(len(b))


def _user_cards(header, reserved):
    cards = []
    for k, v in (header or {}).items():
        ku = str(k).upper()
        if ku in reserved:
            raise ValueError(f"header keyword {ku} is set by the writer")
        for item in (v if ku in ("COMMENT", "HISTORY") and isinstance(v, (list, tuple)) else [v]):
            cards.append(_card(ku, item))
    return cards


# ---------------------------------------------------------------------------
# HDU reading
# ---------------------------------------------------------------------------
class HDU:
    """One header-data unit: its header and the byte offset / size of its data."""

    def __init__(self, header, offset, nbytes):
        self.header, self.offset, self.nbytes = header, offset, nbytes

    @property
    def kind(self):
        x = self.header.get("XTENSION")
        if x is None:
            return "PRIMARY"
        return str(x).strip().upper()

    @property
    def name(self):
        return str(self.header.get("EXTNAME", "")).strip().upper()


def _data_bytes(h):
    naxis = int(h.get("NAXIS", 0))
    if naxis == 0:
        return 0
    n = 1
    for i in range(1, naxis + 1):
        n *= int(h[f"NAXIS{i}"])
    return abs(int(h["BITPIX"])) // 8 * int(h.get("GCOUNT", 1)) * (int(h.get("PCOUNT", 0)) + n)


def list_hdus(path):
    """Every HDU of a FITS file, in order."""
    hdus = []
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        pos = 0
        while pos < size:
            h = Header()
            ended = False
            while not ended:
                blk = f.read(BLOCK)
                if len(blk) < BLOCK:
                    raise ValueError(f"{path}: truncated FITS header")
                pos += BLOCK
                text = blk.decode("ascii", errors="replace")
                for i in range(0, BLOCK, CARD):
                    c = text[i:i + CARD]
                    if c[:8] == "END     ":
                        ended = True
                        break
                    h.add(*_parse_card(c))
            if not hdus and h.get("SIMPLE") is not True:
                raise ValueError(f"{path}: not a FITS file (SIMPLE = T missing)")
            n = _data_bytes(h)
            hdus.append(HDU(h, pos, n))
            pos += n + _pad(n)
            f.seek(pos)
    return hdus


def read_image(path, hdu):
    """Image data of an HDU as a NumPy array (native byte order), shape
    (NAXISn, ..., NAXIS1); BSCALE / BZERO applied when present."""
    h = hdu.header
    naxis = int(h.get("NAXIS", 0))
    if naxis == 0:
        return None
    shape = tuple(int(h[f"NAXIS{i}"]) for i in range(naxis, 0, -1))
    dt = np.dtype(_BITPIX[int(h["BITPIX"])])
    a = np.fromfile(path, dtype=dt, count=int(np.prod(shape)), offset=hdu.offset).reshape(shape)
    a = a.astype(dt.newbyteorder("="))
    bscale, bzero = h.get("BSCALE", 1), h.get("BZERO", 0)
    if bscale != 1 or bzero != 0:
        a = a * float(bscale) + float(bzero)
    return a


def _tform(s):
    s = str(s).strip().upper()
    i = 0
    while i < len(s) and s[i].isdigit():
        i += 1
    rep = int(s[:i]) if i else 1
    code = s[i:i + 1]
    if code not in _TFORM:
        raise ValueError(f"BINTABLE column format {s!r} is not supported")
    return rep, code


def _table_dtype(h):
    names, formats, offsets, off = [], [], [], 0
    for j in range(1, int(h["TFIELDS"]) + 1):
        rep, code = _tform(h[f"TFORM{j}"])
        edt, esz = _TFORM[code]
        name = str(h.get(f"TTYPE{j}", f"COL{j}")).strip().upper()
        if code == "A":
            fmt = f"S{rep}"
        else:
            fmt = edt if rep == 1 else (edt, (rep,))
        if rep > 0:
            names.append(name)
            formats.append(fmt)
            offsets.append(off)
        off += rep * esz
    if off != int(h["NAXIS1"]):
        raise ValueError("BINTABLE column formats do not add up to NAXIS1")
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": off})


def read_table(path, hdu):
    """Columns of a BINTABLE HDU: {TTYPE: NumPy array in native byte order}."""
    h = hdu.header
    dt = _table_dtype(h)
    rows = np.fromfile(path, dtype=dt, count=int(h["NAXIS2"]), offset=hdu.offset)
    out = {}
    for name in dt.names:
        col = rows[name]
        if col.dtype.kind in "iufc":
            col = col.astype(col.dtype.newbyteorder("="))
        out[name] = col
    return out


# ---------------------------------------------------------------------------
# HDU writing
# ---------------------------------------------------------------------------
def _bitpix_of(dt):
    dt = np.dtype(dt)
    for bp, s in _BITPIX.items():
        if np.dtype(s).newbyteorder("=") == dt.newbyteorder("="):
            return bp
    raise ValueError(f"no FITS image type for dtype {dt}")


def image_hdu_bytes(data, primary, header=None, extname=None):
    """Header + data bytes of a primary or IMAGE HDU holding `data` (or none)."""
    reserved = {"SIMPLE", "XTENSION", "BITPIX", "NAXIS", "EXTEND", "PCOUNT", "GCOUNT", "END"}
    if data is None:
        a, bitpix, shape = None, 8, ()
    else:
        a = np.asarray(data)
        if a.dtype == np.bool_:
            a = a.astype(np.uint8)
        bitpix = _bitpix_of(a.dtype)
        shape = a.shape
    cards = [_card("SIMPLE", True, "conforms to FITS standard")] if primary else [_card("XTENSION", "IMAGE")]
    cards.append(_card("BITPIX", bitpix))
    cards.append(_card("NAXIS", len(shape)))
    for i, n in enumerate(reversed(shape)):
        cards.append(_card(f"NAXIS{i + 1}", int(n)))
        reserved.add(f"NAXIS{i + 1}")
    if primary:
        cards.append(_card("EXTEND", True))
    else:
        cards += [_card("PCOUNT", 0), _card("GCOUNT", 1)]
    if extname is not None:
        cards.append(_card("EXTNAME", extname))
        reserved.add("EXTNAME")
    cards += _user_cards(header, reserved)
    out = [_header_bytes(cards)]
    if a is not None and a.size:
        d = np.ascontiguousarray(a, dtype=np.dtype(_BITPIX[bitpix])).tobytes()
        out += [d, b"\0" * _pad(len(d))]
    return b"".join(out)


def table_hdu_bytes(columns, header=None, extname=None):
    """Header + data bytes of a BINTABLE HDU; columns = [(name, TFORM, 1-d array)]."""
    nrows = len(columns[0][2])
    fields, cards_cols = [], []
    for j, (name, tform, arr) in enumerate(columns, 1):
        rep, code = _tform(tform)
        if rep != 1 or code not in "BIJKEDCM":
            raise ValueError(f"writer supports scalar numeric columns only, got {tform!r}")
        if len(arr) != nrows:
            raise ValueError("BINTABLE columns differ in length")
        fields.append((name, _TFORM[code][0]))
        cards_cols += [_card(f"TTYPE{j}", name), _card(f"TFORM{j}", tform)]
    rec = np.empty(nrows, dtype=np.dtype(fields))
    for name, _, arr in columns:
        rec[name] = arr
    cards = [_card("XTENSION", "BINTABLE", "binary table extension"), _card("BITPIX", 8), _card("NAXIS", 2),
             _card("NAXIS1", rec.dtype.itemsize), _card("NAXIS2", nrows), _card("PCOUNT", 0), _card("GCOUNT", 1),
             _card("TFIELDS", len(columns))] + cards_cols
    reserved = {"XTENSION", "BITPIX", "NAXIS", "NAXIS1", "NAXIS2", "PCOUNT", "GCOUNT", "TFIELDS", "END"}
    reserved |= {f"TTYPE{j}" for j in range(1, len(columns) + 1)} | {f"TFORM{j}" for j in range(1, len(columns) + 1)}
    if extname is not None:
        cards.append(_card("EXTNAME", extname))
        reserved.add("EXTNAME")
    cards += _user_cards(header, reserved)
    d = rec.tobytes()
    return b"".join([_header_bytes(cards), d, b"\0" * _pad(len(d))])


def write_file(path, hdu_bytes, overwrite):
    if os.path.exists(path) and not overwrite:
        raise OSError(f"File {path!r} already exists.")
    tmp = f"{path}.tmp{os.getpid()}"
    with open(tmp, "wb") as f:
        for b in hdu_bytes:
            f.write(b)
    os.replace(tmp, path)
    return path


# ---------------------------------------------------------------------------
# the reference's entry points
# ---------------------------------------------------------------------------
def _to_host(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def _to_device(a, to_gpu):
    if not to_gpu:
        return a
    return torch.as_tensor(np.ascontiguousarray(a), device=torch.device("cuda", torch.cuda.current_device()))


def _select(path, hdu):
    hdus = list_hdus(path)
    if not 0 <= hdu < len(hdus):
        raise IndexError(f"{path} has {len(hdus)} HDU(s); HDU {hdu} does not exist "
                         f"(a map written by write_map_fits is the primary image, hdu=0)")
    return hdus[hdu]


def read_map_fits(path, hdu=1, to_gpu=True, dtype=np.float64, *, field=0):
    """Map from HDU `hdu` (``io_fits.py:11-18``): an image HDU's data, or
    column `field` of a BINTABLE flattened in row order (healpy map files).
    Returns a CUDA tensor (to_gpu) or a NumPy array of `dtype`."""
    h = _select(path, hdu)
    if h.kind == "BINTABLE":
        cols = read_table(path, h)
        key = list(cols)[field] if isinstance(field, int) else str(field).upper()
        a = cols[key].reshape(-1)
    else:
        a = read_image(path, h)
        if a is None:
            raise ValueError(f"{path}: HDU {hdu} holds no data")
    return _to_device(a.astype(dtype, copy=False), to_gpu)


def write_map_fits(path, m, header=None, overwrite=True):
    """Map (any shape, float or integer) as the primary image
    (``io_fits.py:20-31``); `header` = {keyword: value} cards added to it."""
    return write_file(path, [image_hdu_bytes(_to_host(m), True, header)], overwrite)


def _healpy_index(lmax, mmax):
    from .alm_ops import _lm_arrays
    l, m = _lm_arrays(lmax, mmax)
    return l, m, l * l + l + m + 1


def write_alm_fits(path, alm, header=None, overwrite=True, *, lmax=None, mmax=None):
    """m-major alm [nalm] as a BINTABLE INDEX / REAL / IMAG in HDU 1
    (``io_fits.py:46-63``, plus the healpy INDEX = l^2 + l + m + 1 column).
    lmax defaults to the one of a triangular array (mmax = lmax)."""
    from .plan import lmax_from_size, nalm
    a = _to_host(alm).astype(np.complex128)
    if a.ndim != 1:
        raise ValueError("write_alm_fits writes one alm array; got shape " + str(a.shape))
    if lmax is None:
        if mmax is None:
            lmax = lmax_from_size(a.shape[0])
        else:     # size = (mmax + 1)(lmax + 1) - mmax(mmax + 1)/2
            lmax = (a.shape[0] + mmax * (mmax + 1) // 2) // (mmax + 1) - 1
    mmax = lmax if mmax is None else mmax
    if not 0 <= mmax <= lmax or a.shape[0] != nalm(lmax, mmax):
        raise ValueError(f"alm of size {a.shape[0]} do not match lmax {lmax}, mmax {mmax}")
    l, m, idx = _healpy_index(lmax, mmax)      # rows in m-major order, as the array
    form = "J" if int(idx.max(initial=0)) < 2 ** 31 else "K"
    cols = [("INDEX", form, idx), ("REAL", "D", a.real), ("IMAG", "D", a.imag)]
    hdr = {"MAX-LPOL": int(lmax), "MAX-MPOL": int(mmax)}
    for k, v in (header or {}).items():
        hdr[str(k).upper()] = v
    return write_file(path, [image_hdu_bytes(None, True), table_hdu_bytes(cols, hdr)], overwrite)


def read_alm_fits(path, hdu=1, to_gpu=True, dtype=np.complex128):
    """alm from HDU `hdu` (``io_fits.py:33-44``) in m-major layout:
    INDEX / REAL / IMAG tables (healpy) are placed by INDEX; REAL / IMAG
    tables without INDEX are the reference's l(l+1)+m layout."""
    from .alm_ops import getidx
    from .plan import nalm
    h = _select(path, hdu)
    if h.kind != "BINTABLE":
        a = read_image(path, h)
        return _to_device(np.asarray(a, dtype=dtype), to_gpu)
    cols = read_table(path, h)
    if "REAL" not in cols or "IMAG" not in cols:
        first = cols[list(cols)[0]].reshape(-1)
        return _to_device(first.astype(dtype), to_gpu)
    v = cols["REAL"].reshape(-1).astype(np.float64) + 1j * cols["IMAG"].reshape(-1).astype(np.float64)
    if "INDEX" in cols:
        idx = cols["INDEX"].reshape(-1).astype(np.int64) - 1
        if v.size and idx.min() < 0:
            raise ValueError(f"{path}: INDEX values start at 1")
        l = np.floor(np.sqrt(idx.astype(np.float64))).astype(np.int64)
        l -= (l * l > idx)
        l += ((l + 1) * (l + 1) <= idx)
        m = idx - l * l - l
        if v.size and (m.min() < 0 or np.any(m > l)):
            raise ValueError(f"{path}: INDEX values do not decode to 0 <= m <= l")
        lmax = int(l.max(initial=0))
        mmax = int(m.max(initial=0))
    else:
        n = v.size
        lmax = math.isqrt(n) - 1
        if (lmax + 1) ** 2 != n:
            raise ValueError(f"{path}: {n} rows without INDEX are not an l(l+1)+m array of (lmax+1)^2 entries")
        mmax = lmax
        l = np.repeat(np.arange(lmax + 1), 2 * np.arange(lmax + 1) + 1)
        m = np.arange(n) - l * l - l
        keep = m >= 0
        l, m, v = l[keep], m[keep], v[keep]
    out = np.zeros(nalm(lmax, mmax), dtype=np.complex128)
    out[getidx(lmax, l, m)] = v
    return _to_device(out.astype(dtype), to_gpu)
