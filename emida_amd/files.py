"""Input stage and file formats of the pattern-offset driver.

Mirrors the reference's file side: load_tiff / get_size (src/load_tiff.hpp:14-69),
the slicepos and deformed-list formats (src/args_parser.cpp:62-108,
src/process_files.hpp:54-73, README.md:47-69), the threaded file pipeline
(file_processor, src/process_files.hpp:98-241) and its standard output
(finalize_offsets, src/process_files.hpp:235-264).  Reading, parsing and the
pipeline run in libemida_amd.so; this module converts to numpy.
"""
from __future__ import annotations

import ctypes
import math
import struct
from dataclasses import dataclass, field

import numpy as np

from . import _capi as C
from .gpu_offset import GpuOffset


def _path(p) -> bytes:
    return p if isinstance(p, bytes) else str(p).encode()


def tiff_size(path):
    """get_size (src/load_tiff.hpp:57-69) -> (width, height)."""
    s = C.Size2()
    C.check(C.lib().emida_tiff_size(_path(path), ctypes.byref(s)))
    return s.x, s.y


def load_tiff(path, size=None) -> np.ndarray:
    """load_tiff (src/load_tiff.hpp:14-53): the raw strips of an uncompressed 16-bit
    grey TIFF as a (height, width) uint16 array.  size = (width, height) must match
    the file (default: the file's own size)."""
    if size is None:
        size = tiff_size(path)
    out = np.empty((int(size[1]), int(size[0])), np.uint16)
    C.check(C.lib().emida_tiff_load(_path(path), C.ptr(out), C.size2(size)))
    return out


def write_tiff(path, raster, rows_per_strip=None, big_endian=False):
    """Writes a (height, width) uint16 array as an uncompressed single-sample TIFF
    with top-left orientation, the layout the loader reads (no counterpart in the
    reference, which only reads TIFFs).  Pixel bytes are stored in the file's byte
    order, so the loader's raw copy returns `raster` for little-endian files."""
    a = np.ascontiguousarray(raster, dtype=np.uint16)
    if a.ndim != 2:
        raise ValueError("raster must be 2-D")
    h, w = a.shape
    rps = h if rows_per_strip is None else max(1, min(int(rows_per_strip), h))
    bo = ">" if big_endian else "<"
    data = a.astype(bo + "u2").tobytes()
    strips = [data[r * w * 2:min(h, r + rps) * w * 2] for r in range(0, h, rps)]
    n = len(strips)
    header = 8
    pos = header
    body = b""
    offsets = []
    for s in strips:
        offsets.append(pos + len(body))
        body += s
    if len(body) % 2:
        body += b"\0"
    arrays_at = header + len(body)
    extra = b""

    def long_array(vals):
        nonlocal extra
        if len(vals) == 1:
            return vals[0], None
        off = arrays_at + len(extra)
        extra += struct.pack(bo + f"{len(vals)}I", *vals)
        return None, off

    so_val, so_off = long_array(offsets)
    sc_val, sc_off = long_array([len(s) for s in strips])
    entries = [(256, 4, 1, w), (257, 4, 1, h), (258, 3, 1, 16), (259, 3, 1, 1), (262, 3, 1, 1),
               (273, 4, n, so_val if so_off is None else so_off), (274, 3, 1, 1), (277, 3, 1, 1),
               (278, 4, 1, rps), (279, 4, n, sc_val if sc_off is None else sc_off)]
    ifd_at = arrays_at + len(extra)
    ifd = struct.pack(bo + "H", len(entries))
    for tag, typ, cnt, val in entries:
        if typ == 3 and cnt == 1:
            ifd += struct.pack(bo + "HHIHH", tag, typ, cnt, val, 0)
        else:
            ifd += struct.pack(bo + "HHII", tag, typ, cnt, val)
    ifd += struct.pack(bo + "I", 0)
    hdr = (b"MM" if big_endian else b"II") + struct.pack(bo + "HI", 42, ifd_at)
    with open(path, "wb") as f:
        f.write(hdr + body + extra + ifd)


def draw_tiff(pic, path, vectors, positions):
    """draw_tiff (src/draw_tiff.hpp:26-59): the (height, width) uint16 pattern as an
    RGBA TIFF (grey = pixel >> 8) with each vector (dx, dy) drawn in red from its
    (x, y) position."""
    a = np.ascontiguousarray(pic, np.uint16)
    v = np.ascontiguousarray(vectors, float).reshape(-1, 2)
    pos = list(positions)
    if len(pos) != len(v):
        raise ValueError("one position per vector")
    arr = (C.Size2 * max(1, len(pos)))(*[C.Size2(int(x), int(y)) for x, y in pos])
    C.check(C.lib().emida_draw_tiff(C.ptr(a), C.Size2(a.shape[1], a.shape[0]), _path(path),
                                    v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), arr, len(v)))


def read_slicepos(path):
    """load_slice_begins (src/args_parser.cpp:62-108) -> (radius or None, mids as a
    list of (x, y))."""
    L = C.lib()
    radius = ctypes.c_uint64()
    ok = ctypes.c_int()
    n = ctypes.c_uint32()
    C.check(L.emida_read_slicepos(_path(path), ctypes.byref(radius), ctypes.byref(ok), None, 0, ctypes.byref(n)))
    arr = (C.Size2 * max(1, n.value))()
    C.check(L.emida_read_slicepos(_path(path), ctypes.byref(radius), ctypes.byref(ok), arr, n.value,
                                  ctypes.byref(n)))
    return (radius.value if ok.value else None), [(arr[i].x, arr[i].y) for i in range(n.value)]


def read_deformed_list(path):
    """load_work (src/process_files.hpp:54-73) -> list of (x, y, file name)."""
    L = C.lib()
    n = ctypes.c_uint32()
    nb = ctypes.c_size_t()
    C.check(L.emida_read_deformed_list(_path(path), None, 0, ctypes.byref(n), None, 0, ctypes.byref(nb)))
    xy = np.zeros(2 * max(1, n.value))
    names = ctypes.create_string_buffer(max(1, nb.value))
    C.check(L.emida_read_deformed_list(_path(path), xy.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n.value,
                                       ctypes.byref(n), names, nb.value, ctypes.byref(nb)))
    raw = names.raw[:nb.value].split(b"\0")[:n.value]
    return [(float(xy[2 * i]), float(xy[2 * i + 1]), raw[i].decode(errors="surrogateescape"))
            for i in range(n.value)]


def slice_begins(mids, slice_size):
    """Top-left corners of the ROIs: middle - slice_size / 2 (src/args_parser.cpp:164-166)."""
    hx, hy = int(slice_size[0]) // 2, int(slice_size[1]) // 2
    out = []
    for x, y in mids:
        if x < hx or y < hy:
            raise ValueError("A slice out of bounds")
        out.append((x - hx, y - hy))
    return out


@dataclass
class FilesResult:
    """Per-file results of process_files: offsets (n_files, n_rois, 2), coefs
    (n_files, n_rois, 6), status (n_files,) EMIDA_OK or an error code (the file's
    batch failed; its rows are NaN), errors {first file index of a failed batch:
    message}, order (first file index of each batch in delivery order) and the
    pipeline's timings."""
    offsets: np.ndarray
    coefs: np.ndarray
    status: np.ndarray
    errors: dict = field(default_factory=dict)
    order: list = field(default_factory=list)
    stats: dict = field(default_factory=dict)


def process_files(offs: GpuOffset, paths, load_workers=1, ordered=True) -> FilesResult:
    """file_processor::process_files (src/process_files.hpp:135-241) over an allocated
    GpuOffset with IN = uint16: load_workers threads read batch_size files each and
    upload them, one thread computes, one finalizes.  Returns every file's results."""
    paths = [_path(p) for p in paths]
    n_rois = offs.total_slices // offs.batch_size
    nf = len(paths)
    res = FilesResult(np.full((nf, n_rois, 2), np.nan), np.full((nf, n_rois, 6), np.nan),
                      np.full(nf, -1, np.int32))

    def cb(user, first, count, status, error, o, cf):
        res.order.append(first)
        res.status[first:first + count] = status
        if status != C.OK:
            res.errors[first] = (error or b"").decode(errors="replace")
            return
        k = count * n_rois
        res.offsets[first:first + count] = np.ctypeslib.as_array(o, shape=(2 * k,)).reshape(count, n_rois, 2)
        res.coefs[first:first + count] = np.ctypeslib.as_array(cf, shape=(6 * k,)).reshape(count, n_rois, 6)

    fn = C.FILES_CALLBACK(cb)
    arr = (ctypes.c_char_p * max(1, nf))(*paths)
    st = C.FilesStats()
    C.check(C.lib().emida_process_files(offs._h, arr, nf, int(load_workers), 1 if ordered else 0, fn, None,
                                        ctypes.byref(st)))
    res.stats = {"batches": st.batches, "failed_batches": st.failed_batches, "wall_ms": st.wall_ms,
                 "load_ms": st.load_ms, "upload_wait_ms": st.upload_wait_ms}
    return res


def _c_float(v: float) -> str:
    """printf("%f") of a double; NaN printed as "nan" (src/process_files.hpp:248-252)."""
    if math.isnan(v):
        return "nan"
    return "%f" % v


def format_output(work, mids, result: FilesResult, write_coefs=False) -> str:
    """The standard output of finalize_offsets (src/process_files.hpp:235-264) for
    the files whose batch succeeded, in list order: per file "%f %f %llu %s", then
    per ROI "%lu %lu %f %f" (+ six "%f" coefficients with write_coefs)."""
    lines = []
    for f, (x, y, name) in enumerate(work):
        if result.status[f] != C.OK:
            continue
        lines.append(f"{_c_float(x)} {_c_float(y)} {len(mids)} {name}")
        for r, (mx, my) in enumerate(mids):
            vals = [result.offsets[f, r, 0], result.offsets[f, r, 1]]
            if write_coefs:
                vals += list(result.coefs[f, r])
            lines.append(f"{mx} {my} " + " ".join(_c_float(float(v)) for v in vals))
    return "".join(line + "\n" for line in lines)
