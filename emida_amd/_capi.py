"""ctypes binding of include/emida_amd.h (libemida_amd.so, built in-tree)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libemida_amd.so")

F32, F64, I32, U16 = 0, 1, 2, 3
CROSS_POLICY_BRUTE, CROSS_POLICY_FFT, CROSS_POLICY_BRUTE_MFMA = 0, 1, 2
MFMA_AUTO, MFMA_BF16, MFMA_BF16_SPLIT, MFMA_F16, MFMA_F16_SPLIT = -1, 0, 1, 2, 3
OK, ERR_ARG, ERR_HIP, ERR_STATE, ERR_UNSUPPORTED, ERR_NOMEM = range(6)

_TYPE_OF = {np.dtype(np.float32): F32, np.dtype(np.float64): F64, np.dtype(np.int32): I32,
            np.dtype(np.uint16): U16}
_DTYPE_OF = {v: k for k, v in _TYPE_OF.items()}


def type_code(dtype) -> int:
    try:
        return _TYPE_OF[np.dtype(dtype)]
    except KeyError:
        raise ValueError(f"unsupported element type {dtype}") from None


def dtype_of(code: int):
    return _DTYPE_OF[code]


class EmidaError(RuntimeError):
    """A failing library call (the reference throws std::runtime_error)."""

    def __init__(self, code, msg):
        super().__init__(f"emida error {code}: {msg}")
        self.code = code


class Size2(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint32), ("y", ctypes.c_uint32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("max_batch", ctypes.c_uint32), ("streams", ctypes.c_uint32),
                ("max_queued", ctypes.c_uint32), ("use_graphs", ctypes.c_int), ("slots_per_plan", ctypes.c_uint32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("submit_ns", ctypes.c_uint64), ("launch_ns", ctypes.c_uint64), ("done_ns", ctypes.c_uint64),
                ("batch", ctypes.c_uint32), ("plan", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("launches", ctypes.c_uint64),
                ("graph_launches", ctypes.c_uint64), ("graphs_captured", ctypes.c_uint64),
                ("device_bytes_resident", ctypes.c_uint64), ("device_bytes_batches", ctypes.c_uint64),
                ("device_bytes_peak_inflight", ctypes.c_uint64), ("slots", ctypes.c_uint32)]


class FilesStats(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_uint32), ("failed_batches", ctypes.c_uint32), ("wall_ms", ctypes.c_double),
                ("load_ms", ctypes.c_double), ("upload_wait_ms", ctypes.c_double)]


# void (*)(void* user, uint32 first, uint32 count, int status, const char* error,
#          const double* offsets, const double* coefs)
FILES_CALLBACK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                  ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double))


class Calib(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double), ("z", ctypes.c_double)]


class OffsetInfo(ctypes.Structure):
    _fields_ = [("total_slices", ctypes.c_uint32), ("neigh_size", ctypes.c_uint32),
                ("data_index_bytes", ctypes.c_uint32), ("cross_size", Size2), ("corr_buffer", Size2),
                ("fft_size", ctypes.c_uint32), ("device_bytes", ctypes.c_uint64)]


_lib = None


def lib():
    """The loaded library. Fails loudly when it has not been built: there is no
    fallback implementation."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C emida_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        vp, u32, i32, sz, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_size_t, ctypes.c_uint64
        dp = ctypes.POINTER(ctypes.c_double)
        S = Size2
        sigs = {
            "emida_last_error": ([], ctypes.c_char_p),
            "emida_offset_create": ([ctypes.POINTER(vp), i32, i32, S, ctypes.POINTER(S), u32, S, S, u32, i32, i32,
                                     u32], i32),
            "emida_offset_destroy": ([vp], None),
            "emida_offset_get_info": ([vp, ctypes.POINTER(OffsetInfo)], i32),
            "emida_offset_allocate_memory": ([vp, vp], i32),
            "emida_offset_transfer_pic_to_device_async": ([vp, vp, ctypes.POINTER(u32)], i32),
            "emida_offset_wait_uploads": ([vp], i32),
            "emida_offset_get_offset_core": ([vp, vp, vp, u32], i32),
            "emida_offset_finalize": ([vp, vp, vp, dp, dp], i32),
            "emida_offset_get_offset": ([vp, vp, dp, dp], i32),
            "emida_offset_process_device": ([vp, vp, u32, vp, vp, vp, vp, vp], i32),
            "emida_offset_get_cu_pic": ([vp, ctypes.POINTER(vp)], i32),
            "emida_offset_get_cu_temp": ([vp, ctypes.POINTER(vp)], i32),
            "emida_offset_cross_corr_fft": ([vp], i32),
            "emida_offset_get_cu_cross_res": ([vp, ctypes.POINTER(vp)], i32),
            "emida_offset_get_streams": ([vp, ctypes.POINTER(vp), ctypes.POINTER(vp)], i32),
            "emida_offset_set_mfma_mode": ([vp, i32], i32),
            "emida_offset_get_mfma_mode": ([vp, ctypes.POINTER(i32)], i32),
            "emida_offset_set_mfma_refine": ([vp, i32], i32),
            "emida_offset_set_timing": ([vp, i32], i32),
            "emida_offset_get_timing": ([vp, ctypes.POINTER(ctypes.c_char_p), dp, ctypes.POINTER(u64)], i32),
            "emida_host_alloc": ([ctypes.POINTER(vp), sz], i32),
            "emida_host_free": ([vp], i32),
            "emida_device_alloc": ([ctypes.POINTER(vp), sz], i32),
            "emida_device_free": ([vp], i32),
            "emida_memcpy_h2d": ([vp, vp, sz], i32),
            "emida_memcpy_d2h": ([vp, vp, sz], i32),
            "emida_memcpy_d2d": ([vp, vp, sz], i32),
            "emida_memset_device": ([vp, i32, sz], i32),
            "emida_device_synchronize": ([], i32),
            "emida_device_query": ([ctypes.c_char_p, sz, ctypes.POINTER(i32)], i32),
            "emida_device_mem_info": ([ctypes.POINTER(sz), ctypes.POINTER(sz)], i32),
            "emida_run_sum": ([i32, vp, vp, vp, S, S, u32, u32, vp], i32),
            "emida_run_prepare_pics": ([i32, i32, i32, vp, vp, vp, vp, vp, vp, S, S, S, u32, u32, vp], i32),
            "emida_run_cross_corr": ([i32, vp, vp, vp, S, S, u32, u32, vp], i32),
            "emida_run_cross_corr_fft": ([i32, vp, vp, vp, S, u32, u32, vp], i32),
            "emida_run_cross_corr_fft_noalloc": ([i32, vp, vp, vp, S, u32, u32, vp], i32),
            "emida_fft_noalloc_scratch_bytes": ([ctypes.POINTER(sz)], i32),
            "emida_run_cross_corr_mfma": ([i32, vp, vp, vp, S, S, u32, u32, vp], i32),
            "emida_run_hadamard": ([i32, vp, vp, S, u32, u32, vp], i32),
            "emida_run_maxarg_reduce": ([i32, vp, vp, S, u32, vp], i32),
            "emida_run_extract_neighbors": ([i32, i32, vp, sz, vp, vp, i32, S, u32, vp], i32),
            "emida_run_subpixel_max": ([i32, vp, i32, u32, vp, vp, vp], i32),
            "emida_lstsq_matrix": ([i32, dp], i32),
            "emida_fft_max_slice": ([i32, ctypes.POINTER(u32)], i32),
            "emida_server_create": ([ctypes.POINTER(vp), ctypes.POINTER(ServerConfig)], i32),
            "emida_server_destroy": ([vp], None),
            "emida_server_add_plan": ([vp, i32, S, vp, ctypes.POINTER(S), u32, S, S, i32, i32, i32,
                                       ctypes.POINTER(u32)], i32),
            "emida_server_plan_info": ([vp, u32, ctypes.POINTER(OffsetInfo), ctypes.POINTER(u32)], i32),
            "emida_server_submit": ([vp, u32, vp, dp, dp, ctypes.POINTER(u64)], i32),
            "emida_server_wait": ([vp, u64, ctypes.POINTER(RequestInfo)], i32),
            "emida_server_get_stats": ([vp, ctypes.POINTER(ServerStats)], i32),
            "emida_generate_hanning": ([i32, sz, vp], i32),
            "emida_tiff_size": ([ctypes.c_char_p, ctypes.POINTER(S)], i32),
            "emida_tiff_load": ([ctypes.c_char_p, vp, S], i32),
            "emida_process_files": ([vp, ctypes.POINTER(ctypes.c_char_p), u32, u32, i32, FILES_CALLBACK, vp,
                                     ctypes.POINTER(FilesStats)], i32),
            "emida_read_slicepos": ([ctypes.c_char_p, ctypes.POINTER(u64), ctypes.POINTER(i32), ctypes.POINTER(S), u32,
                                     ctypes.POINTER(u32)], i32),
            "emida_read_deformed_list": ([ctypes.c_char_p, dp, u32, ctypes.POINTER(u32), ctypes.c_char_p, sz,
                                          ctypes.POINTER(sz)], i32),
            "emida_draw_tiff": ([vp, S, ctypes.c_char_p, dp, ctypes.POINTER(S), u32], i32),
            "emida_fit_deformation": ([vp, i32, u32, vp, vp, u32, Calib, ctypes.c_double, u32, vp, vp, vp, vp, vp,
                                       vp, vp], i32),
        }
        for name, (args, res) in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        msg = lib().emida_last_error().decode(errors="replace")
        raise EmidaError(rc, msg)


def size2(v) -> Size2:
    x, y = v
    return Size2(int(x), int(y))


def ptr(a):
    """Address of a numpy array (or None)."""
    if a is None:
        return None
    return ctypes.c_void_p(a.ctypes.data)


def exported_symbols():
    """Every function name declared in include/emida_amd.h that the library must export."""
    hdr = os.path.join(os.path.dirname(_HERE), "include", "emida_amd.h")
    import re
    names = re.findall(r"^\s*(?:int|void|const char\*)\s+(emida_\w+)\s*\(", open(hdr).read(), re.M)
    return names
