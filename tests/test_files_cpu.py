"""Input stage and file formats on the CPU: TIFF loading, the slicepos and deformed
list parsers, the command-line driver's argument handling, and argument errors of
the file pipeline (no GPU work).

The TIFF reader is checked against the reference's own fixture
(test/res/INITIAL_x0y0.tif) in two independent ways: Pillow's decoder must give the
same pixels, and the reference's load_tiff test (test/load_tiff_test.cpp:7-48,
libtiff's RGBA decoding, grey == raster >> 8) must hold where libtiff is present.
"""
import ctypes
import ctypes.util
import hashlib
import os
import subprocess

import numpy as np
import pytest

import emida_amd
from emida_amd import _capi as C
from emida_amd import files as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
CLI = os.path.join(ROOT, "emida_amd", "emida")
INITIAL = os.path.join(GOLD, "INITIAL_x0y0.tif")


def test_file_fixtures_unchanged():
    for line in open(os.path.join(GOLD, "fixtures.sha256")):
        digest, name = line.split()[:2]
        assert hashlib.sha256(open(os.path.join(GOLD, name), "rb").read()).hexdigest() == digest, name


# ------------------------------------------------------------------ TIFF
def test_fixture_size():
    assert F.tiff_size(INITIAL) == (873, 873)


def test_fixture_equals_pillow_decode():
    from PIL import Image
    a = F.load_tiff(INITIAL, (873, 873))
    b = np.array(Image.open(INITIAL))
    assert b.dtype == np.uint16
    np.testing.assert_array_equal(a, b)


def _libtiff():
    for name in ("libtiff.so.6", "libtiff.so.5", ctypes.util.find_library("tiff")):
        if not name:
            continue
        try:
            return ctypes.CDLL(name)
        except OSError:
            continue
    return None


def test_reference_load_tiff_known_answer():  # test/load_tiff_test.cpp:7-48
    L = _libtiff()
    if L is None:
        pytest.skip("libtiff not installed")
    L.TIFFOpen.restype = ctypes.c_void_p
    L.TIFFOpen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    L.TIFFClose.argtypes = [ctypes.c_void_p]
    L.TIFFReadRGBAImage.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int]
    raster = F.load_tiff(INITIAL, (873, 873))
    t = L.TIFFOpen(INITIAL.encode(), b"r")
    assert t
    try:
        rgba = np.zeros(873 * 873, np.uint32)
        assert L.TIFFReadRGBAImage(t, 873, 873, rgba.ctypes.data, 0) == 1
    finally:
        L.TIFFClose(t)
    img = rgba.reshape(873, 873)[::-1]  # RGBA rows start at the bottom
    r, g, b = img & 0xFF, (img >> 8) & 0xFF, (img >> 16) & 0xFF
    np.testing.assert_array_equal(r, g)
    np.testing.assert_array_equal(g, b)
    np.testing.assert_array_equal(r, raster >> 8)


@pytest.mark.parametrize("rows_per_strip", [None, 1, 7, 36, 1000])
def test_written_tiffs_read_back(tmp_path, rows_per_strip):
    from PIL import Image
    rng = np.random.default_rng(rows_per_strip or 0)
    r = rng.integers(0, 65536, (37, 53), dtype=np.uint16)
    p = str(tmp_path / "t.tif")
    F.write_tiff(p, r, rows_per_strip=rows_per_strip)
    np.testing.assert_array_equal(np.array(Image.open(p)), r)  # the writer, checked by Pillow
    np.testing.assert_array_equal(F.load_tiff(p, (53, 37)), r)


def test_big_endian_file_is_copied_raw(tmp_path):
    # raw strip reads do not swap bytes (TIFFReadRawStrip): the loader returns the
    # file's bytes as native uint16
    r = np.arange(12 * 10, dtype=np.uint16).reshape(10, 12) * 257 + 1
    p = str(tmp_path / "be.tif")
    F.write_tiff(p, r, rows_per_strip=3, big_endian=True)
    assert F.tiff_size(p) == (12, 10)
    np.testing.assert_array_equal(F.load_tiff(p), r.byteswap())


def _err(fn, *args):
    with pytest.raises(C.EmidaError) as e:
        fn(*args)
    return e.value


def test_tiff_errors(tmp_path):
    from PIL import Image
    r = np.full((8, 6), 1000, np.uint16)
    good = str(tmp_path / "g.tif")
    F.write_tiff(good, r)
    e = _err(F.load_tiff, good, (6, 9))
    assert e.code == C.ERR_ARG and "different than specified size" in str(e)
    e = _err(F.load_tiff, str(tmp_path / "missing.tif"), (6, 8))
    assert e.code == C.ERR_ARG and "cannot open" in str(e)
    e = _err(F.tiff_size, str(tmp_path / "missing.tif"))
    assert e.code == C.ERR_ARG
    notiff = str(tmp_path / "x.tif")
    open(notiff, "wb").write(b"hello world, not a tiff")
    assert _err(F.load_tiff, notiff, (6, 8)).code == C.ERR_ARG
    # truncated pixel data
    data = open(good, "rb").read()
    trunc = str(tmp_path / "trunc.tif")
    open(trunc, "wb").write(data[:20])
    assert _err(F.load_tiff, trunc, (6, 8)).code == C.ERR_ARG
    # compressed and 8-bit images: the reference would copy their raw bytes
    lzw = str(tmp_path / "lzw.tif")
    Image.fromarray(r).save(lzw, compression="tiff_lzw")
    assert _err(F.load_tiff, lzw, (6, 8)).code == C.ERR_UNSUPPORTED
    u8 = str(tmp_path / "u8.tif")
    Image.fromarray(r.astype(np.uint8)).save(u8)
    assert _err(F.load_tiff, u8, (6, 8)).code == C.ERR_UNSUPPORTED


def _patch_short_tag(path, tag, value):
    """Rewrite the inline SHORT value of one IFD entry of a little-endian file."""
    import struct
    b = bytearray(open(path, "rb").read())
    ifd = struct.unpack_from("<I", b, 4)[0]
    n = struct.unpack_from("<H", b, ifd)[0]
    for i in range(n):
        e = ifd + 2 + 12 * i
        if struct.unpack_from("<H", b, e)[0] == tag:
            struct.pack_into("<H", b, e + 8, value)
            open(path, "wb").write(bytes(b))
            return
    raise AssertionError("tag not found")


def test_orientation_and_bad_strip_tables_rejected(tmp_path):
    import struct
    r = np.full((8, 6), 7, np.uint16)
    p = str(tmp_path / "o.tif")
    F.write_tiff(p, r)
    _patch_short_tag(p, 274, 4)  # bottom-left
    e = _err(F.load_tiff, p, (6, 8))
    assert e.code == C.ERR_ARG and "orientation" in str(e)
    # strip byte count larger than the raster: rejected instead of overflowing it
    q = str(tmp_path / "big.tif")
    F.write_tiff(q, r)
    b = bytearray(open(q, "rb").read())
    ifd = struct.unpack_from("<I", b, 4)[0]
    for i in range(struct.unpack_from("<H", b, ifd)[0]):
        e_ = ifd + 2 + 12 * i
        if struct.unpack_from("<H", b, e_)[0] == 279:
            struct.pack_into("<I", b, e_ + 8, 6 * 8 * 2 + 2)
    open(q, "wb").write(bytes(b))
    assert _err(F.load_tiff, q, (6, 8)).code == C.ERR_ARG
    # strip data shorter than the raster: rejected instead of leaving the buffer's tail
    for i in range(struct.unpack_from("<H", b, ifd)[0]):
        e_ = ifd + 2 + 12 * i
        if struct.unpack_from("<H", b, e_)[0] == 279:
            struct.pack_into("<I", b, e_ + 8, 6 * 8 * 2 - 2)
    open(q, "wb").write(bytes(b))
    e = _err(F.load_tiff, q, (6, 8))
    assert e.code == C.ERR_ARG and "shorter than the image" in str(e)
    # a field with no values (ImageWidth, count 0)
    z = str(tmp_path / "zero.tif")
    F.write_tiff(z, r)
    b = bytearray(open(z, "rb").read())
    ifd = struct.unpack_from("<I", b, 4)[0]
    for i in range(struct.unpack_from("<H", b, ifd)[0]):
        e_ = ifd + 2 + 12 * i
        if struct.unpack_from("<H", b, e_)[0] == 256:
            struct.pack_into("<I", b, e_ + 4, 0)
    open(z, "wb").write(bytes(b))
    e = _err(F.tiff_size, z)
    assert e.code == C.ERR_ARG and "empty TIFF field" in str(e)


# ------------------------------------------------------------------ slicepos / deformed list
def test_slicepos_fixture():
    radius, mids = F.read_slicepos(os.path.join(GOLD, "roi-cryst.txt"))
    assert radius == 48
    exp = np.loadtxt(os.path.join(GOLD, "roi-cryst.txt"), skiprows=1, dtype=np.int64)
    assert len(mids) == 89
    np.testing.assert_array_equal(np.array(mids), exp)


def test_slicepos_edge_lines(tmp_path, capfd):
    # std::stoul rules (src/args_parser.cpp:84-105): leading digits count, the rest
    # of a token is ignored, a token without digits skips the line with a warning
    p = tmp_path / "pos.txt"
    p.write_bytes(b"32\n10 20\nnospace\n 5 6\n12abc 7x\n40 50\r\nabc 1\n-1 5\n\n70 80")
    radius, mids = F.read_slicepos(str(p))
    assert radius == 32
    assert mids == [(10, 20), (12, 7), (40, 50), (4294967295, 5), (70, 80)]
    err = capfd.readouterr().err
    assert err.count("Warning: could not parse line") == 4  # nospace, " 5 6", "abc 1", ""


def test_slicepos_bad_size_line(tmp_path, capfd):
    p = tmp_path / "pos.txt"
    p.write_text("size\n1 2\n")
    radius, mids = F.read_slicepos(str(p))
    assert radius is None and mids == [(1, 2)]
    assert "could not parse the first line" in capfd.readouterr().err


def test_deformed_list_fixture():
    work = F.read_deformed_list(os.path.join(GOLD, "def.txt"))
    lines = open(os.path.join(GOLD, "def.txt")).read().splitlines()
    assert len(work) == len(lines) == 492
    for (x, y, name), line in zip(work, lines):
        sx, sy, sn = line.split(" ", 2)
        assert (x, y, name) == (float(sx), float(sy), sn)
    assert work[0] == (0.0, 0.0, "../test_data/deformed/DEFORMED_x0y0.tif")


def test_deformed_list_edge_lines(tmp_path):
    # src/process_files.hpp:61-70: x and y by operator>>, one separator character
    # skipped, the rest of the line is the name (spaces and a CR kept)
    p = tmp_path / "def.txt"
    p.write_bytes(b"1.5 -2 a b.tif\n\n3e1 4\tt.tif\n7 8 win.tif\r\nx 1 bad.tif\n9 10  two.tif")
    work = F.read_deformed_list(str(p))
    assert work == [(1.5, -2.0, "a b.tif"), (0.0, 0.0, ""), (30.0, 4.0, "t.tif"), (7.0, 8.0, "win.tif\r"),
                    (0.0, 0.0, ""), (9.0, 10.0, " two.tif")]


def test_format_output_matches_printf():
    mids = [(10, 20), (30, 40)]
    work = [(0.0, 519.6152422706632, "a.tif"), (60.0, 0.0, "b.tif")]
    res = F.FilesResult(np.array([[[0.25, -1.5], [np.nan, 2.0]], [[0, 0], [0, 0]]], float),
                        np.zeros((2, 2, 6)), np.array([C.OK, C.ERR_ARG], np.int32))
    out = F.format_output(work, mids, res)
    assert out == ("0.000000 519.615242 2 a.tif\n10 20 0.250000 -1.500000\n30 40 nan 2.000000\n")


# ------------------------------------------------------------------ command-line driver
def _cli(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


def _files(tmp_path, radius=16, mids=((40, 40),)):
    pos = tmp_path / "pos.txt"
    pos.write_text(f"{radius}\n" + "".join(f"{x} {y}\n" for x, y in mids))
    init = tmp_path / "init.tif"
    F.write_tiff(str(init), np.full((96, 96), 100, np.uint16))
    lst = tmp_path / "def.txt"
    lst.write_text(f"0.0 0.0 {init}\n")
    return str(init), str(lst), str(pos)


def test_cli_built():
    assert os.access(CLI, os.X_OK), "emida_amd/emida is not built"


def test_cli_help_and_missing_mandatory():
    r = _cli("-h")
    assert r.returncode == 0 and "--slicepos" in r.stderr
    r = _cli("-i", "x.tif")
    assert r.returncode == 1 and "Argument parsing error." in r.stderr


@pytest.mark.parametrize("extra,msg", [
    (["--slicestep", "4,4"], "cannot use slicestep when slicepos specified"),
    (["-c", "65,63"], "cross size must be lesser or equal to slice_size * 2 - 1"),
    (["--precision", "half"], "Invalid precision argument."),
    (["-f", "4"], "the fitsize 4 is not supported"),
    (["--crosspolicy", "fast"], "Invalid crosspolicy argument."),
    (["--batchsize", "0"], "batchsize must be greater than 0"),
    (["--loadworkers=-2"], "loadworkers must be in 1..64"),
    (["--loadworkers", "65"], "loadworkers must be in 1..64"),
    (["-c", "3"], "Argument parsing error."),
])
def test_cli_argument_errors(tmp_path, extra, msg):
    init, lst, pos = _files(tmp_path)
    r = _cli("-i", init, "-d", lst, "-b", pos, *extra)
    assert r.returncode == 1, r.stderr
    assert msg in r.stderr


def test_cli_slice_out_of_bounds_and_bad_size(tmp_path):
    init, lst, pos = _files(tmp_path, radius=16, mids=((8, 40),))
    r = _cli("-i", init, "-d", lst, "-b", pos)
    assert r.returncode == 1 and "A slice out of bounds" in r.stderr
    bad = tmp_path / "bad.txt"
    bad.write_text("r\n40 40\n")
    r = _cli("-i", init, "-d", lst, "-b", str(bad))
    assert r.returncode == 1 and "does not give the slice size" in r.stderr
    good = tmp_path / "good.txt"
    good.write_text("16\n40 40\n")
    r = _cli("-i", str(tmp_path / "none.tif"), "-d", lst, "-b", str(good))
    assert r.returncode == 1 and "cannot open" in r.stderr


# ------------------------------------------------------------------ file pipeline arguments
def test_process_files_needs_allocated_u16_handle(tmp_path):
    g = emida_amd.GpuOffset(np.float64, np.uint16, (96, 96), [(0, 0)], (32, 32), (63, 63), 2)
    e = _err(F.process_files, g, [str(tmp_path / "a.tif")])
    assert e.code == C.ERR_STATE
    g.close()
    g = emida_amd.GpuOffset(np.float64, np.float64, (96, 96), [(0, 0)], (32, 32), (63, 63), 2)
    assert _err(F.process_files, g, ["a"]).code in (C.ERR_STATE, C.ERR_ARG)
    g.close()


# ------------------------------------------------------------------ draw_tiff
def test_draw_tiff_known_answers(tmp_path):
    # src/draw_tiff.hpp:13-59: grey = pixel >> 8; a vector is 20 points at
    # d = 0, 0.05, ..., 0.95 (double accumulation), rounded half away from zero
    from PIL import Image
    pic = (np.arange(20 * 12, dtype=np.uint32).reshape(12, 20) * 271 % 65536).astype(np.uint16)
    p = str(tmp_path / "v.tif")
    F.draw_tiff(pic, p, [(10.0, 0.0), (-10.0, 0.0), (0.0, 3.0), (np.nan, 1.0)], [(5, 5), (3, 8), (18, 0), (1, 1)])
    img = np.array(Image.open(p).convert("RGBA"))
    assert img.shape == (12, 20, 4)
    red = np.zeros((12, 20), bool)
    red[5, 5:16] = True        # (10, 0) from (5, 5): x offsets 0..10, all covered
    red[8, 0:4] = True         # (-10, 0) from (3, 8): x = 3..0, then negative (skipped)
    red[0:4, 18] = True        # (0, 3) from (18, 0): y offsets round(0.05 k * 3) = 0..3
    exp = np.empty((12, 20, 4), np.uint8)
    g = (pic >> 8).astype(np.uint8)
    exp[..., 0], exp[..., 1], exp[..., 2], exp[..., 3] = g, g, g, 255
    exp[red] = (255, 0, 0, 255)
    np.testing.assert_array_equal(img, exp)


def test_cpp_load_tiff_header(tmp_path):
    # include/emida/load_tiff.hpp: the reference's signatures over the C ABI
    from PIL import Image
    exe = os.path.join(ROOT, "tests", "cpp", "test_tiff")
    assert os.access(exe, os.X_OK), "tests/cpp/test_tiff is not built"
    out = str(tmp_path / "drawn.tif")
    r = subprocess.run([exe, INITIAL, out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    a = np.array(Image.open(INITIAL)).astype(np.uint64).ravel()
    w = (np.arange(a.size, dtype=np.uint64) % 9973)
    assert f"size 873 873 sum {int(a.sum())} weighted {int((a * w).sum())}" in r.stdout
    img = np.array(Image.open(out).convert("RGBA"))
    assert tuple(img[100, 100]) == (255, 0, 0, 255) and img.shape == (873, 873, 4)
