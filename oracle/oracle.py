"""numpy front end of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to verify
its result and by bench.py's cpu_baseline leg. The product library never uses it.

What pins each oracle function (besides the reference file:line it restates):

  hann        closed form at n=5 (test/prepare_test.cpp:18-23); numpy.hanning
              fixtures test/res/hann/{5,128}_{A,res}.txt (window applied to data).
  sums        closed forms (test/sum_test.cpp:44-149); exact uint32 arithmetic;
              brute-force math.fsum on float inputs (order-free bound).
  prepare     exact vectors (test/prepare_test.cpp:47-210).
  cross_corr  integer known answers (test/cross_corr_test.cpp:17-217 including the
              64x64 fixture); numpy brute force on tiny inputs.
  fft_corr    numpy FFT (pocketfft) of the zero-padded grid: C2R(R2C(p)*conj(R2C(r)))
              unnormalised, the identity the reference computes with cuFFT.
  maxarg      known answers (test/maxarg_test.cpp:63-230); numpy.argmax.
  extract     known answers (test/maxarg_test.cpp:8-61); slicing on tiny grids.
  lstsq       exact rational (A^T A)^-1 A^T; the reference table values
              (src/subpixel_max.cpp:22-58) when the reference tree is present.
  subpixel    test/subpixel_max_test.cpp:8-46; exact recovery of sampled quadratics
              for fitsize 3/5/7/9; numpy.linalg.lstsq.
  get_offset  49 + 1 end-to-end goldens (test/get_offset_test.cpp:10-27,108-199),
              BRUTE and FFT, batch 1 and 5, tolerance 7e-13 / 2e-14.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

T_F32, T_F64, T_I32, T_U16 = 0, 1, 2, 3
POLICY_BRUTE, POLICY_FFT = 0, 1

_TCODE = {np.dtype(np.float32): T_F32, np.dtype(np.float64): T_F64,
          np.dtype(np.int32): T_I32, np.dtype(np.uint16): T_U16}


def data_index_dtype(t):
    """numpy layout of emida's data_index<T> (src/common.hpp:26-31)."""
    return np.dtype([("data", np.dtype(t)), ("index", np.uint32)], align=True)


def sums_dtype(in_dtype):
    """sums_trait<IN> (src/cufft_helpers.hpp:157-179)."""
    in_dtype = np.dtype(in_dtype)
    return np.dtype(np.uint32) if in_dtype == np.uint16 else in_dtype


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        vp, u32, i32, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_size_t
        dp = ctypes.POINTER(ctypes.c_double)
        L.or_num_threads.argtypes = [i32]
        L.or_hann.argtypes = [i32, sz, vp]
        L.or_sums.argtypes = [i32, vp, vp, u32, u32, u32, u32, u32, u32, vp]
        L.or_prepare.argtypes = [i32, i32, i32, vp, vp, vp, vp, vp, vp, u32, u32, u32, u32, u32, u32, u32, u32]
        L.or_cross_corr.argtypes = [i32, vp, vp, vp, u32, u32, u32, u32, u32, u32]
        L.or_fft_corr.argtypes = [i32, vp, vp, vp, u32, u32, u32, u32]
        L.or_maxarg.argtypes = [i32, vp, u32, u32, vp]
        L.or_extract.argtypes = [i32, i32, vp, sz, vp, vp, i32, u32, u32, u32]
        L.or_lstsq_matrix.argtypes = [i32, dp]
        L.or_subpixel.argtypes = [i32, vp, i32, sz, vp, vp]
        L.or_get_offset.argtypes = [i32, i32, i32, vp, vp, u32, u32, vp, u32, u32, u32, u32, u32, u32, i32,
                                    vp, vp, vp, vp, vp]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _check(rc, what):
    if rc != 0:
        raise ValueError(f"oracle {what}: unsupported arguments (rc={rc})")


def set_threads(n):
    return lib().or_num_threads(int(n))


def hann(n, dtype=np.float64):
    out = np.empty(n, dtype)
    _check(lib().or_hann(_TCODE[np.dtype(dtype)], n, _p(out)), "hann")
    return out


def _begins(begins):
    b = np.ascontiguousarray(np.asarray(begins, dtype=np.uint32).reshape(-1, 2))
    return b


def sums(pics, begins, src, slice_size, batch=1):
    pics = np.ascontiguousarray(pics)
    b = _begins(begins)
    out = np.zeros(len(b) * batch, sums_dtype(pics.dtype))
    _check(lib().or_sums(_TCODE[pics.dtype], _p(pics), _p(b), src[0], src[1], slice_size[0], slice_size[1],
                         len(b), batch, _p(out)), "sums")
    return out


def prepare(pics, hx, hy, sumv, begins, src, slice_size, out_size=None, batch=1):
    pics = np.ascontiguousarray(pics)
    hx = np.ascontiguousarray(hx)
    hy = np.ascontiguousarray(hy)
    sumv = np.ascontiguousarray(sumv)
    b = _begins(begins)
    out_size = tuple(out_size or slice_size)
    out = np.zeros(len(b) * batch * out_size[0] * out_size[1], hx.dtype)
    s_code = T_U16 if sumv.dtype == np.uint32 else _TCODE[sumv.dtype]
    _check(lib().or_prepare(_TCODE[pics.dtype], _TCODE[hx.dtype], s_code, _p(pics), _p(out), _p(hx), _p(hy),
                            _p(sumv), _p(b), src[0], src[1], slice_size[0], slice_size[1], out_size[0],
                            out_size[1], len(b), batch), "prepare")
    return out


def cross_corr(pics, ref, size, res_size, ref_slices=1, batch=1):
    pics = np.ascontiguousarray(pics)
    ref = np.ascontiguousarray(ref, dtype=pics.dtype)
    out = np.zeros(res_size[0] * res_size[1] * ref_slices * batch, pics.dtype)
    _check(lib().or_cross_corr(_TCODE[pics.dtype], _p(pics), _p(ref), _p(out), size[0], size[1], res_size[0],
                               res_size[1], ref_slices, batch), "cross_corr")
    return out


def fft_corr(pics, ref, size, ref_slices=1, batch=1):
    pics = np.ascontiguousarray(pics)
    ref = np.ascontiguousarray(ref, dtype=pics.dtype)
    out = np.zeros(4 * size[0] * size[1] * ref_slices * batch, pics.dtype)
    _check(lib().or_fft_corr(_TCODE[pics.dtype], _p(pics), _p(ref), _p(out), size[0], size[1], ref_slices,
                             batch), "fft_corr")
    return out


def maxarg(data, area, batch=1):
    data = np.ascontiguousarray(data)
    out = np.zeros(batch, data_index_dtype(data.dtype))
    _check(lib().or_maxarg(_TCODE[data.dtype], _p(data), area, batch, _p(out)), "maxarg")
    return out


def extract(data, max_i, s, size, batch=1, policy=POLICY_BRUTE):
    data = np.ascontiguousarray(data)
    max_i = np.ascontiguousarray(max_i)
    out = np.zeros(batch * s * s, data.dtype)
    _check(lib().or_extract(_TCODE[data.dtype], policy, _p(data), data.size, _p(max_i), _p(out), s, size[0],
                            size[1], batch), "extract")
    return out


def lstsq_matrix(s):
    out = np.zeros(6 * s * s, np.float64)
    _check(lib().or_lstsq_matrix(s, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "lstsq_matrix")
    return out.reshape(6, s * s)


def subpixel(neigh, s, n=None):
    neigh = np.ascontiguousarray(neigh)
    n = neigh.size // (s * s) if n is None else n
    offs = np.zeros((n, 2))
    coefs = np.zeros((n, 6))
    _check(lib().or_subpixel(_TCODE[neigh.dtype], _p(neigh), s, n, _p(offs), _p(coefs)), "subpixel")
    return offs, coefs


def get_offset(pics, temp, src, begins, slice_size, cross_size, batch, fitsize=3, policy=POLICY_BRUTE,
               dtype=np.float64, return_intermediates=False, return_stats=False):
    """Whole-path oracle: gpu_offset<T,IN> construct + allocate_memory(temp) +
    get_offset(pics) (src/gpu_offset.hpp:100-301). Returns offsets (n,2) and
    coefs (n,6) in double; optionally the arg-max records and neighbourhoods, and
    (return_stats) per slice the largest correlation value, the second largest and
    the largest magnitude, (n,3)."""
    pics = np.ascontiguousarray(pics)
    temp = np.ascontiguousarray(temp, dtype=pics.dtype)
    b = _begins(begins)
    total = len(b) * batch
    offs = np.zeros((total, 2))
    coefs = np.zeros((total, 6))
    maxes = np.zeros(total, data_index_dtype(dtype))
    neigh = np.zeros(total * fitsize * fitsize, dtype)
    stats = np.zeros((total, 3)) if return_stats else None
    rc = lib().or_get_offset(_TCODE[np.dtype(dtype)], _TCODE[pics.dtype], policy, _p(pics), _p(temp), src[0],
                             src[1], _p(b), len(b), slice_size[0], slice_size[1], cross_size[0], cross_size[1],
                             batch, fitsize, _p(offs), _p(coefs), _p(maxes), _p(neigh), _p(stats))
    _check(rc, "get_offset")
    out = (offs, coefs)
    if return_intermediates:
        out += (maxes, neigh)
    if return_stats:
        out += (stats,)
    return out


# Relative peak margin above which float rounding cannot move the integer peak (the
# float FFT's error stays far below 1e-4 of the largest magnitude).
FLOAT_MARGIN = 1e-4


def well_conditioned(coefs, fit, stats, rel_margin=FLOAT_MARGIN):
    """The ROIs on which an offset error against the FP64 oracle is measured: the FP64
    peak beats the runner-up by more than rel_margin x the largest magnitude (the
    integer peak is fixed) and the FP64 quadratic fit has its maximum inside the
    F x F neighbourhood (elsewhere the fitted position is an extrapolation whose
    error grows without bound as the curvature vanishes, in any precision)."""
    b, c, d, e, f = (coefs[:, i] for i in range(1, 6))
    det = 4 * f * d - e * e
    with np.errstate(divide="ignore", invalid="ignore"):
        py = (b * e - 2 * c * d) / det
        px = (-e * py - b) / (2 * d)
    posed = (d < 0) & (det > 0) & (px >= 0) & (px <= fit - 1) & (py >= 0) & (py <= fit - 1)
    margin = (stats[:, 0] - stats[:, 1]) > rel_margin * stats[:, 2]
    return posed & margin
