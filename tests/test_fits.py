"""FITS map / alm / TQU files (reference: io_fits.py:11-63,
io_fits_multi.py:24-63; hp_compat.py:204-225).

astropy, which the reference calls, is absent here, so the codec is pinned
against the FITS Standard 4.0 directly: header cards checked column by
column, big-endian data at the 2880-byte block after the header, and files
assembled byte by byte in this test (a healpy-style map table with a repeat
count, a reference-layout alm table) read back through the module.
"""
import struct

import numpy as np
import pytest
import torch

from hellblast_amd import hp_compat, io_fits
from hellblast_amd.alm_ops import from_reference_layout, getidx


def _cards(raw, start=0):
    """80-character cards of the header starting at byte `start`, up to END."""
    out = []
    for i in range(start, len(raw), 80):
        c = raw[i:i + 80].decode("ascii")
        out.append(c)
        if c.startswith("END     "):
            return out
    raise AssertionError("no END card")


def _hand_header(cards):
    """Header block from literal card strings (each padded to 80)."""
    b = "".join(f"{c:<80}" for c in cards + ["END"]).encode("ascii")
    return b + b" " * ((-len(b)) % 2880)


def test_primary_image_bytes_follow_the_standard(tmp_path):
    p = str(tmp_path / "m.fits")
    m = np.array([1.5, -2.0, 3.25])
    hp_compat.write_map(p, m, header={"NSIDE": 1, "ORDERING": "RING", "TEMP": 2.5, "FLAG": False})
    raw = open(p, "rb").read()
    assert len(raw) == 2 * 2880
    cards = _cards(raw)
    # fixed format: keyword in columns 1-8, "= " in 9-10, logicals and
    # numbers right-justified to column 30, strings quoted from column 11
    assert cards[0][:30] == "SIMPLE  =                    T"
    assert cards[1][:30] == "BITPIX  =                  -64"
    assert cards[2][:30] == "NAXIS   =                    1"
    assert cards[3][:30] == "NAXIS1  =                    3"
    assert cards[4][:30] == "EXTEND  =                    T"
    assert cards[5][:30] == "NSIDE   =                    1"
    assert cards[6][:20] == "ORDERING= 'RING    '"
    assert cards[7][:30] == "TEMP    =                  2.5"
    assert cards[8][:30] == "FLAG    =                    F"
    assert cards[9].rstrip() == "END"
    assert all(len(c) == 80 for c in cards)
    assert raw[2880:2880 + 24] == struct.pack(">3d", 1.5, -2.0, 3.25)
    assert raw[2880 + 24:] == b"\0" * (2880 - 24)
    # the map is the primary image (HDU 0), as the reference writes it
    back = hp_compat.read_map(p, hdu=0, to_gpu=False)
    assert back.dtype == np.float64 and np.array_equal(back, m)
    with pytest.raises(IndexError):
        hp_compat.read_map(p, to_gpu=False)


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int16, np.int32, np.int64, np.uint8])
def test_map_round_trip_dtypes(tmp_path, dtype):
    p = str(tmp_path / "m.fits")
    rng = np.random.default_rng(1)
    m = (rng.standard_normal((3, 1000)) * 100).astype(dtype)
    io_fits.write_map_fits(p, m)
    h = io_fits.list_hdus(p)[0]
    assert h.header["BITPIX"] == {np.float32: -32, np.float64: -64, np.int16: 16, np.int32: 32, np.int64: 64,
                                  np.uint8: 8}[dtype]
    assert (h.header["NAXIS1"], h.header["NAXIS2"]) == (1000, 3)
    back = io_fits.read_map_fits(p, hdu=0, to_gpu=False, dtype=dtype)
    assert back.shape == (3, 1000) and np.array_equal(back, m)


def test_read_healpy_style_map_table(tmp_path):
    """A BINTABLE map as healpy writes it (one column TEMPERATURE with a
    repeat count of 1024, npix/1024 rows), assembled here byte by byte."""
    nside = 16
    npix = 12 * nside * nside
    m = np.random.default_rng(2).standard_normal(npix).astype(np.float32)
    prim = _hand_header(["SIMPLE  =                    T", "BITPIX  =                    8",
                         "NAXIS   =                    0", "EXTEND  =                    T"])
    tab = _hand_header(["XTENSION= 'BINTABLE'", "BITPIX  =                    8", "NAXIS   =                    2",
                        "NAXIS1  =                 4096", f"NAXIS2  = {npix // 1024:>20}",
                        "PCOUNT  =                    0", "GCOUNT  =                    1",
                        "TFIELDS =                    1", "TTYPE1  = 'TEMPERATURE'", "TFORM1  = '1024E   '",
                        "PIXTYPE = 'HEALPIX '", "ORDERING= 'RING    '", f"NSIDE   = {nside:>20}",
                        "COMMENT written for a test", "HISTORY one", "HISTORY two"])
    data = m.astype(">f4").tobytes()
    p = tmp_path / "hp.fits"
    p.write_bytes(prim + tab + data + b"\0" * ((-len(data)) % 2880))
    hdus = io_fits.list_hdus(str(p))
    assert [h.kind for h in hdus] == ["PRIMARY", "BINTABLE"]
    assert hdus[1].header["NSIDE"] == nside and hdus[1].header["ORDERING"] == "RING"
    assert hdus[1].header["HISTORY"] == ["one", "two"]
    back = hp_compat.read_map(str(p), to_gpu=False)
    assert back.dtype == np.float64 and np.array_equal(back, m.astype(np.float64))
    back32 = io_fits.read_map_fits(str(p), to_gpu=False, dtype=np.float32, field="TEMPERATURE")
    assert np.array_equal(back32, m)


def test_alm_table_has_healpy_index(tmp_path):
    """INDEX = l^2 + l + m + 1, rows in the array's m-major order:
    (l, m) = (0,0) (1,0) (2,0) (1,1) (2,1) (2,2) -> 1 3 7 4 8 9."""
    p = str(tmp_path / "a.fits")
    alm = np.array([1, 2, 3, 4 + 1j, 5 - 2j, 6 + 3j], dtype=np.complex128)
    hp_compat.write_alm(p, alm)
    hdus = io_fits.list_hdus(p)
    assert [h.kind for h in hdus] == ["PRIMARY", "BINTABLE"]
    h = hdus[1].header
    assert (h["TTYPE1"], h["TFORM1"], h["TTYPE2"], h["TFORM2"], h["TTYPE3"], h["TFORM3"]) == \
        ("INDEX", "J", "REAL", "D", "IMAG", "D")
    assert (h["NAXIS1"], h["NAXIS2"], h["MAX-LPOL"], h["MAX-MPOL"]) == (20, 6, 2, 2)
    raw = open(p, "rb").read()
    rows = np.frombuffer(raw[hdus[1].offset:hdus[1].offset + 120],
                         dtype=np.dtype([("i", ">i4"), ("r", ">f8"), ("m", ">f8")]))
    assert rows["i"].tolist() == [1, 3, 7, 4, 8, 9]
    assert rows["r"].tolist() == [1, 2, 3, 4, 5, 6] and rows["m"].tolist() == [0, 0, 0, 1, -2, 3]
    back = hp_compat.read_alm(p, to_gpu=False)
    assert back.dtype == np.complex128 and np.array_equal(back, alm)


def test_alm_round_trip_large_and_truncated(tmp_path):
    rng = np.random.default_rng(3)
    for lmax, mmax in ((700, 700), (300, 120)):
        n = (mmax + 1) * (lmax + 1) - mmax * (mmax + 1) // 2
        alm = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        p = str(tmp_path / f"a{lmax}_{mmax}.fits")
        io_fits.write_alm_fits(p, torch.as_tensor(alm), mmax=None if mmax == lmax else mmax)
        back = io_fits.read_alm_fits(p, to_gpu=False)
        assert np.array_equal(back, alm)
        # rows are decoded from INDEX alone: a shuffled table reads the same
        h = io_fits.list_hdus(p)[1]
        cols = io_fits.read_table(p, h)
        perm = rng.permutation(n)
        q = str(tmp_path / "shuffled.fits")
        io_fits.write_file(q, [io_fits.image_hdu_bytes(None, True), io_fits.table_hdu_bytes(
            [("INDEX", "J", cols["INDEX"][perm]), ("REAL", "D", cols["REAL"][perm]),
             ("IMAG", "D", cols["IMAG"][perm])])], True)
        assert np.array_equal(io_fits.read_alm_fits(q, to_gpu=False), alm)


def test_read_reference_layout_alm(tmp_path):
    """REAL / IMAG without INDEX: the reference's l(l+1)+m array of
    (lmax+1)^2 entries (negative-m slots unused) -> m-major."""
    lmax = 5
    rng = np.random.default_rng(4)
    ref = rng.standard_normal((lmax + 1) ** 2) + 1j * rng.standard_normal((lmax + 1) ** 2)
    p = str(tmp_path / "r.fits")
    io_fits.write_file(p, [io_fits.image_hdu_bytes(None, True),
                           io_fits.table_hdu_bytes([("REAL", "D", ref.real), ("IMAG", "D", ref.imag)])], True)
    back = io_fits.read_alm_fits(p, to_gpu=False)
    assert np.array_equal(back, from_reference_layout(torch.as_tensor(ref), lmax).numpy())
    # (l, m) = (3, 2): reference slot 3*4 + 2, m-major slot getidx(5, 3, 2)
    assert back[getidx(lmax, 3, 2)] == ref[14]
    q = str(tmp_path / "bad.fits")
    io_fits.write_file(q, [io_fits.image_hdu_bytes(None, True),
                           io_fits.table_hdu_bytes([("REAL", "D", ref.real[:7]), ("IMAG", "D", ref.imag[:7])])], True)
    with pytest.raises(ValueError):
        io_fits.read_alm_fits(q, to_gpu=False)


def test_tqu_round_trip(tmp_path):
    p = str(tmp_path / "tqu.fits")
    rng = np.random.default_rng(5)
    T, U = rng.standard_normal(192), rng.standard_normal(192).astype(np.float32)
    hp_compat.write_tqu(p, T=T, U=torch.as_tensor(U), nside=4, extra_header={"BUNIT": "K", "HISTORY": "made"})
    hdus = io_fits.list_hdus(p)
    assert [(h.kind, h.name) for h in hdus] == [("PRIMARY", ""), ("IMAGE", "T"), ("IMAGE", "U")]
    assert hdus[1].header["BUNIT"] == "K" and hdus[1].header["HISTORY"] == ["made"]
    (t, q, u), meta = hp_compat.read_tqu(p, to_gpu=False)
    assert q is None and np.array_equal(t, T) and u.dtype == np.float32 and np.array_equal(u, U)
    assert meta == {"NSIDE": 4, "ORDERING": "RING", "COORDSYS": "G"}


def test_header_values_and_errors(tmp_path):
    p = str(tmp_path / "h.fits")
    hdr = {"STR": "it's", "BIG": 2 ** 40, "SMALL": -1.25e-300, "LONG-KEY-NAME": 7, "ONE": 1.0}
    io_fits.write_map_fits(p, np.zeros(4), header=hdr)
    h = io_fits.list_hdus(p)[0].header
    assert h["STR"] == "it's" and h["BIG"] == 2 ** 40 and h["SMALL"] == -1.25e-300 and h["ONE"] == 1.0
    assert isinstance(h["ONE"], float) and h["LONG-KEY-NAME"] == 7
    with pytest.raises(OSError):
        io_fits.write_map_fits(p, np.zeros(4), overwrite=False)
    with pytest.raises(ValueError):
        io_fits.write_map_fits(p, np.zeros(4), header={"NAXIS": 2})
    with pytest.raises(ValueError):
        io_fits.write_map_fits(p, np.zeros(4), header={"X": float("nan")})
    bad = tmp_path / "bad.fits"
    bad.write_bytes(b"not a fits file".ljust(2880))
    with pytest.raises(ValueError):
        io_fits.list_hdus(str(bad))
    with pytest.raises(ValueError):
        io_fits.write_alm_fits(p, np.zeros(5, dtype=complex))


@pytest.mark.gpu
def test_fits_files_through_the_gpu_transforms(gpu, tmp_path):
    """map file -> read to the GPU -> map2alm -> alm file -> read to the GPU
    -> alm2map: the same numbers as the transforms on the arrays."""
    print("[case] FITS through map2alm/alm2map nside=32", flush=True)
    nside, lmax = 32, 64
    m = np.random.default_rng(6).standard_normal(12 * nside * nside)
    p, q = str(tmp_path / "m.fits"), str(tmp_path / "a.fits")
    hp_compat.write_map(p, m)
    mg = hp_compat.read_map(p, hdu=0)
    assert mg.is_cuda and mg.dtype == torch.float64
    alm = hp_compat.map2alm(mg, lmax=lmax)
    hp_compat.write_alm(q, alm)
    ag = hp_compat.read_alm(q)
    assert ag.is_cuda and torch.equal(ag, alm)
    direct = hp_compat.alm2map(hp_compat.map2alm(torch.as_tensor(m, device="cuda"), lmax=lmax), nside, lmax=lmax)
    assert torch.equal(hp_compat.alm2map(ag, nside, lmax=lmax), direct)
