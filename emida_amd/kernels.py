"""Kernel-level seam: the reference's run_* launchers (src/kernels.cuh:56-125) and its
GPU test wrappers algorithm_cross_corr / algorithm_maxarg (src/cross_corr_host.hpp,
src/maxarg_host.hpp), taking and returning numpy arrays.  Each call uploads its
inputs, launches one library kernel through the C ABI and downloads the result."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C
from .gpu_offset import DeviceBuffer, data_index_dtype


def _sums_dtype(in_dtype):
    return np.dtype(np.uint32) if np.dtype(in_dtype) == np.uint16 else np.dtype(in_dtype)


def _begins(begins):
    return np.ascontiguousarray(np.asarray(begins, dtype=np.uint32).reshape(-1, 2))


def run_sum(data, begins, src_size, slice_size, batch_size=1):
    """run_sum(data, sums, begins, src, slice, begins_size, batch) (src/sum.cu:160-167)."""
    data = np.ascontiguousarray(data)
    b = _begins(begins)
    out_t = _sums_dtype(data.dtype)
    d, db = DeviceBuffer.from_array(data), DeviceBuffer.from_array(b)
    out = DeviceBuffer(len(b) * batch_size * out_t.itemsize)
    C.check(C.lib().emida_run_sum(C.type_code(data.dtype), d.ptr, out.ptr, db.ptr, C.size2(src_size),
                                  C.size2(slice_size), len(b), batch_size, None))
    return out.download(out_t)


def run_prepare_pics(pic, hanning_x, hanning_y, sums, begins, src_size, slice_size, out_size=None, batch_size=1,
                     out_init=None):
    """run_prepare_pics (src/prepare_pics.cu:46-63). Padding of out_size slices keeps
    the contents of out_init (zeros by default)."""
    pic = np.ascontiguousarray(pic)
    hx = np.ascontiguousarray(hanning_x)
    hy = np.ascontiguousarray(hanning_y, dtype=hx.dtype)
    sums = np.ascontiguousarray(sums)
    b = _begins(begins)
    out_size = tuple(out_size or slice_size)
    n = len(b) * batch_size * out_size[0] * out_size[1]
    out = DeviceBuffer.from_array(np.zeros(n, hx.dtype) if out_init is None else np.asarray(out_init, hx.dtype))
    bufs = [DeviceBuffer.from_array(a) for a in (pic, hx, hy, sums, b)]
    s_code = C.U16 if sums.dtype == np.uint32 else C.type_code(sums.dtype)
    C.check(C.lib().emida_run_prepare_pics(C.type_code(pic.dtype), C.type_code(hx.dtype), s_code, bufs[0].ptr,
                                           out.ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, bufs[4].ptr,
                                           C.size2(src_size), C.size2(slice_size), C.size2(out_size), len(b),
                                           batch_size, None))
    return out.download(hx.dtype)


def run_cross_corr(pics, ref, size, res_size, ref_slices=1, batch_size=1):
    """run_cross_corr (src/cross_corr.cu:68-74); algorithm_cross_corr
    (src/cross_corr_host.hpp:16-64) calls it with ref_slices = #slices, batch 1."""
    pics = np.ascontiguousarray(pics)
    ref = np.ascontiguousarray(ref, dtype=pics.dtype)
    dp, dr = DeviceBuffer.from_array(pics), DeviceBuffer.from_array(ref)
    n = res_size[0] * res_size[1] * ref_slices * batch_size
    out = DeviceBuffer(n * pics.dtype.itemsize)
    C.check(C.lib().emida_run_cross_corr(C.type_code(pics.dtype), dp.ptr, dr.ptr, out.ptr, C.size2(size),
                                         C.size2(res_size), ref_slices, batch_size, None))
    return out.download(pics.dtype)


def run_cross_corr_mfma(pics, ref, size, res_size, ref_slices=1, batch_size=1, mode=C.MFMA_F16_SPLIT):
    """Matrix-core brute correlation of float32 prepared slices: same layout and
    meaning as run_cross_corr, products of bf16/fp16 halves accumulated in fp32
    (mode MFMA_*_SPLIT: hi*hi + hi*lo + lo*hi)."""
    pics = np.ascontiguousarray(pics, dtype=np.float32)
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    dp, dr = DeviceBuffer.from_array(pics), DeviceBuffer.from_array(ref)
    n = res_size[0] * res_size[1] * ref_slices * batch_size
    out = DeviceBuffer(n * 4)
    C.check(C.lib().emida_run_cross_corr_mfma(int(mode), dp.ptr, dr.ptr, out.ptr, C.size2(size), C.size2(res_size),
                                              ref_slices, batch_size, None))
    return out.download(np.float32)


def run_cross_corr_fft(pics, ref, size, ref_slices=1, batch_size=1):
    """cross_corr_fft (src/gpu_offset.hpp:303-316) on prepared slices: returns the
    (2s)^2 circular correlation buffers, unnormalised."""
    pics = np.ascontiguousarray(pics)
    ref = np.ascontiguousarray(ref, dtype=pics.dtype)
    dp, dr = DeviceBuffer.from_array(pics), DeviceBuffer.from_array(ref)
    n = 4 * size[0] * size[1] * ref_slices * batch_size
    out = DeviceBuffer(n * pics.dtype.itemsize)
    C.check(C.lib().emida_run_cross_corr_fft(C.type_code(pics.dtype), dp.ptr, dr.ptr, out.ptr, C.size2(size),
                                             ref_slices, batch_size, None))
    return out.download(pics.dtype)


def run_cross_corr_fft_device(d_pics, d_ref, d_res, dtype, size, ref_slices=1, batch_size=1, noalloc=False,
                              stream=None):
    """cross_corr_fft on device buffers already in place (no upload or download);
    noalloc: the allocation-free entry point, working in the library's static scratch.
    Buffers with a known size (DeviceBuffer) are checked against what the call reads
    and writes before anything runs."""
    sx, sy = (size, size) if np.isscalar(size) else size
    item = np.dtype(dtype).itemsize
    n_ref = int(sx) * int(sy) * int(ref_slices) * item
    for buf, need, what in ((d_pics, n_ref * batch_size, "pattern slices"), (d_ref, n_ref, "reference slices"),
                            (d_res, 4 * n_ref * batch_size, "result buffers")):
        have = getattr(buf, "nbytes", None)
        if have is not None and have < need:
            raise ValueError(f"cross_corr_fft: the {what} need {need} bytes, the buffer holds {have}")
    fn = C.lib().emida_run_cross_corr_fft_noalloc if noalloc else C.lib().emida_run_cross_corr_fft
    C.check(fn(C.type_code(np.dtype(dtype)), d_pics.ptr, d_ref.ptr, d_res.ptr, C.size2(size), ref_slices, batch_size,
               None if stream is None else ctypes.c_void_p(int(stream))))


def fft_noalloc_scratch_bytes():
    v = ctypes.c_size_t()
    C.check(C.lib().emida_fft_noalloc_scratch_bytes(ctypes.byref(v)))
    return v.value


def run_hadamard(pics, ref, one_size, ref_slices=1, batch_size=1):
    """run_hadamard (src/hadamard.cu:43-49): complex arrays, pics * conj(ref)."""
    pics = np.ascontiguousarray(pics)
    ref = np.ascontiguousarray(ref, dtype=pics.dtype)
    real = np.float32 if pics.dtype == np.complex64 else np.float64
    dp, dr = DeviceBuffer.from_array(pics), DeviceBuffer.from_array(ref)
    C.check(C.lib().emida_run_hadamard(C.type_code(real), dp.ptr, dr.ptr, C.size2(one_size), ref_slices,
                                       batch_size, None))
    return dp.download(pics.dtype)


def run_maxarg_reduce(data, size, batch_size=1):
    """run_maxarg_reduce (src/maxarg.cu:180-187) -> data_index<T>[batch]."""
    data = np.ascontiguousarray(data)
    dd = DeviceBuffer.from_array(data)
    dt = data_index_dtype(data.dtype)
    out = DeviceBuffer(batch_size * dt.itemsize)
    C.check(C.lib().emida_run_maxarg_reduce(C.type_code(data.dtype), dd.ptr, out.ptr, C.size2(size), batch_size,
                                            None))
    return out.download(dt)


def run_extract_neighbors(data, max_i, s, src_size, batch_size=1, pos_policy=C.CROSS_POLICY_BRUTE):
    """run_extract_neighbors<T, pos_policy> (src/extract_neighbors.cu:43-49)."""
    data = np.ascontiguousarray(data)
    max_i = np.ascontiguousarray(max_i)
    dd, dm = DeviceBuffer.from_array(data), DeviceBuffer.from_array(max_i)
    out = DeviceBuffer(batch_size * s * s * data.dtype.itemsize)
    C.check(C.lib().emida_run_extract_neighbors(C.type_code(data.dtype), pos_policy, dd.ptr, data.size, dm.ptr,
                                                out.ptr, s, C.size2(src_size), batch_size, None))
    return out.download(data.dtype)


def run_subpixel_max(neighbors, s):
    """subpixel_max_serial (src/subpixel_max.hpp:76-97) on the GPU -> (offsets, coefs)."""
    nb = np.ascontiguousarray(neighbors)
    n = nb.size // (s * s)
    dn = DeviceBuffer.from_array(nb)
    do, dc = DeviceBuffer(n * 16), DeviceBuffer(n * 48)
    C.check(C.lib().emida_run_subpixel_max(C.type_code(nb.dtype), dn.ptr, s, n, do.ptr, dc.ptr, None))
    return do.download(np.float64).reshape(n, 2), dc.download(np.float64).reshape(n, 6)


def lstsq_matrix(s):
    out = np.zeros(6 * s * s)
    C.check(C.lib().emida_lstsq_matrix(s, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out.reshape(6, s * s)


def generate_hanning(size, dtype=np.float64):
    out = np.zeros(size, dtype)
    C.check(C.lib().emida_generate_hanning(C.type_code(dtype), size, C.ptr(out)))
    return out
