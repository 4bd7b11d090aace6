"""Python mirror of the reference's gpu_offset<T, IN> and get_offset<T>.

Same names, argument meaning and error behaviour as src/gpu_offset.hpp:48-326 and
src/get_offset.hpp:8-20; every call goes through the C ABI of libemida_amd.so
(include/emida_amd.h).  Errors raise EmidaError (a RuntimeError), as the reference
throws std::runtime_error.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C

CROSS_POLICY_BRUTE = C.CROSS_POLICY_BRUTE
CROSS_POLICY_FFT = C.CROSS_POLICY_FFT
CROSS_POLICY_BRUTE_MFMA = C.CROSS_POLICY_BRUTE_MFMA
MFMA_AUTO, MFMA_BF16, MFMA_BF16_SPLIT, MFMA_F16, MFMA_F16_SPLIT = (C.MFMA_AUTO, C.MFMA_BF16, C.MFMA_BF16_SPLIT,
                                                              C.MFMA_F16, C.MFMA_F16_SPLIT)


def data_index_dtype(t):
    """data_index<T> (src/common.hpp:26-31): {T data; uint32 index}, 8 or 16 bytes."""
    return np.dtype([("data", np.dtype(t)), ("index", np.uint32)], align=True)


class DeviceBuffer:
    """Device memory owned by the library's allocator (for callers without a GPU
    runtime of their own)."""

    def __init__(self, nbytes: int):
        p = ctypes.c_void_p()
        C.check(C.lib().emida_device_alloc(ctypes.byref(p), int(nbytes)))
        self.ptr = p
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(a)
        if offset + a.nbytes > self.nbytes:
            raise ValueError("upload exceeds the buffer")
        C.check(C.lib().emida_memcpy_h2d(ctypes.c_void_p(self.ptr.value + offset), C.ptr(a), a.nbytes))

    def download(self, dtype, count=None, offset: int = 0) -> np.ndarray:
        dtype = np.dtype(dtype)
        count = (self.nbytes - offset) // dtype.itemsize if count is None else count
        out = np.empty(count, dtype)
        if offset + out.nbytes > self.nbytes:
            raise ValueError("download exceeds the buffer")
        C.check(C.lib().emida_memcpy_d2h(C.ptr(out), ctypes.c_void_p(self.ptr.value + offset), out.nbytes))
        return out

    def zero(self):
        C.check(C.lib().emida_memset_device(self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr is not None and self.ptr.value:
            C.check(C.lib().emida_device_free(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedBuffer:
    """cuda_malloc_host (src/host_helpers.hpp:74-80) as a numpy array."""

    def __init__(self, count: int, dtype):
        dtype = np.dtype(dtype)
        p = ctypes.c_void_p()
        C.check(C.lib().emida_host_alloc(ctypes.byref(p), max(1, count) * dtype.itemsize))
        self._p = p
        buf = (ctypes.c_char * (max(1, count) * dtype.itemsize)).from_address(p.value)
        self.array = np.frombuffer(buf, dtype=dtype, count=count)

    def __del__(self):
        try:
            if self._p is not None:
                C.lib().emida_host_free(self._p)
                self._p = None
        except Exception:
            pass


class GpuOffset:
    """gpu_offset<T, IN> (src/gpu_offset.hpp:100-125).

    T: np.float32 or np.float64 (correlation precision); IN: np.uint16, np.float32
    or np.float64 (pixel type); begins: list of (x, y) ROI corners (copied);
    slice_size, cross_size, src_size: (x, y); policy: CROSS_POLICY_BRUTE,
    CROSS_POLICY_FFT or CROSS_POLICY_BRUTE_MFMA (matrix-core brute correlation);
    s: fitsize 3, 5, 7 or 9; mfma_mode: MFMA_BF16, MFMA_BF16_SPLIT, MFMA_F16 or
    MFMA_F16_SPLIT (None or MFMA_AUTO: the plain mode when refining, the split one
    otherwise, fp16 for uint16 pixels and bf16 for float pixels; the fp16 modes need
    uint16 pixels; .mfma_mode is the mode in effect), and mfma_refine: exact peak equal
    to CROSS_POLICY_BRUTE's results (False: float screen values, T must be float32);
    both for CROSS_POLICY_BRUTE_MFMA.
    """

    def __init__(self, T, IN, src_size, begins, slice_size, cross_size, batch_size, policy=CROSS_POLICY_BRUTE,
                 s=3, in_buffers_size=1, mfma_mode=None, mfma_refine=True):
        self.T = np.dtype(T)
        self.IN = np.dtype(IN)
        self.src_size = tuple(int(v) for v in src_size)
        self.begins = [tuple(int(v) for v in b) for b in begins]
        self.slice_size = tuple(int(v) for v in slice_size)
        self.batch_size = int(batch_size)
        self.policy = int(policy)
        self.s = int(s)
        arr = (C.Size2 * len(self.begins))(*[C.Size2(x, y) for x, y in self.begins])
        h = ctypes.c_void_p()
        C.check(C.lib().emida_offset_create(ctypes.byref(h), C.type_code(self.T), C.type_code(self.IN),
                                            C.size2(src_size), arr, len(self.begins), C.size2(slice_size),
                                            C.size2(cross_size), self.batch_size, self.policy, self.s,
                                            int(in_buffers_size)))
        self._h = h
        if self.policy == CROSS_POLICY_BRUTE_MFMA:
            C.check(C.lib().emida_offset_set_mfma_mode(h, int(MFMA_AUTO if mfma_mode is None else mfma_mode)))
            C.check(C.lib().emida_offset_set_mfma_refine(h, 1 if mfma_refine else 0))
        used = ctypes.c_int32()
        C.check(C.lib().emida_offset_get_mfma_mode(h, ctypes.byref(used)))
        self.mfma_mode = used.value
        info = C.OffsetInfo()
        C.check(C.lib().emida_offset_get_info(h, ctypes.byref(info)))
        self.total_slices = info.total_slices
        self.neigh_size = info.neigh_size
        self.cross_size = (info.cross_size.x, info.cross_size.y)
        self.corr_buffer = (info.corr_buffer.x, info.corr_buffer.y)
        self.fft_size = info.fft_size
        self.data_index = data_index_dtype(self.T)

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            C.lib().emida_offset_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def device_bytes(self):
        info = C.OffsetInfo()
        C.check(C.lib().emida_offset_get_info(self._h, ctypes.byref(info)))
        return info.device_bytes

    def _pic(self, pic, count):
        a = np.ascontiguousarray(pic, dtype=self.IN).ravel()
        if a.size != count:
            raise ValueError(f"expected {count} pixels, got {a.size}")
        return a

    # -- reference API
    def allocate_memory(self, temp):
        """allocate_memory(IN* temp) (src/gpu_offset.hpp:127-211); temp may be None."""
        t = None if temp is None else self._pic(temp, self.src_size[0] * self.src_size[1])
        C.check(C.lib().emida_offset_allocate_memory(self._h, C.ptr(t)))

    def transfer_pic_to_device_async(self, pic) -> int:
        """transfer_pic_to_device_async (src/gpu_offset.hpp:213-218).  The array must
        stay alive until the upload finished (wait_uploads)."""
        a = self._pic(pic, self.src_size[0] * self.src_size[1] * self.batch_size)
        self._keep = a
        i = ctypes.c_uint32()
        C.check(C.lib().emida_offset_transfer_pic_to_device_async(self._h, C.ptr(a), ctypes.byref(i)))
        return i.value

    def wait_uploads(self):
        C.check(C.lib().emida_offset_wait_uploads(self._h))

    def new_result_buffers(self, pinned=True):
        """Host buffers for get_offset_core: (maxes_i, neighbors)."""
        if pinned:
            m = PinnedBuffer(self.total_slices, self.data_index)
            n = PinnedBuffer(self.neigh_size, self.T)
            self._pinned = getattr(self, "_pinned", []) + [m, n]
            return m.array, n.array
        return np.zeros(self.total_slices, self.data_index), np.zeros(self.neigh_size, self.T)

    def get_offset_core(self, maxes_i, neighbors, cu_pic_index=0):
        """get_offset_core (src/gpu_offset.hpp:234-278): asynchronous; results are
        valid after finalize()."""
        C.check(C.lib().emida_offset_get_offset_core(self._h, C.ptr(maxes_i), C.ptr(neighbors), int(cu_pic_index)))

    def finalize(self, maxes_i, neighbors):
        """finalize (src/gpu_offset.hpp:280-301) -> (offsets (n,2), coefs (n,6))."""
        offs = np.zeros((self.total_slices, 2))
        coefs = np.zeros((self.total_slices, 6))
        dp = ctypes.POINTER(ctypes.c_double)
        C.check(C.lib().emida_offset_finalize(self._h, C.ptr(maxes_i), C.ptr(neighbors), offs.ctypes.data_as(dp),
                                              coefs.ctypes.data_as(dp)))
        return offs, coefs

    def get_offset(self, pic):
        """get_offset(IN* pic) (src/gpu_offset.hpp:220-231) -> (offsets, coefs)."""
        a = self._pic(pic, self.src_size[0] * self.src_size[1] * self.batch_size)
        offs = np.zeros((self.total_slices, 2))
        coefs = np.zeros((self.total_slices, 6))
        dp = ctypes.POINTER(ctypes.c_double)
        C.check(C.lib().emida_offset_get_offset(self._h, C.ptr(a), offs.ctypes.data_as(dp), coefs.ctypes.data_as(dp)))
        return offs, coefs

    def process_device(self, d_pics, n_patterns, d_maxes=None, d_neighbors=None, d_offsets=None, d_coefs=None,
                       stream=None):
        """Device-resident batch (no reference counterpart): device pointers in and out."""
        p = lambda b: None if b is None else (b.ptr if isinstance(b, DeviceBuffer) else ctypes.c_void_p(int(b)))
        C.check(C.lib().emida_offset_process_device(self._h, p(d_pics), int(n_patterns), p(d_maxes),
                                                    p(d_neighbors), p(d_offsets), p(d_coefs),
                                                    None if stream is None else ctypes.c_void_p(int(stream))))

    # -- the reference's cross_corr_fft seam and its test accessors (src/gpu_offset.hpp:96-97,303-325)
    def _dptr(self, fn):
        p = ctypes.c_void_p()
        C.check(fn(self._h, ctypes.byref(p)))
        return p.value

    def get_cu_pic(self) -> int:
        """Device address of the prepared pattern slices (FFT policy: the (2s.x) x
        (2s.y) buffers cross_corr_fft reads)."""
        return self._dptr(C.lib().emida_offset_get_cu_pic)

    def get_cu_temp(self) -> int:
        """Device address of the reference data (FFT policy: R2C spectra, 2s.y rows of
        s.x + 1 complex per ROI; brute: prepared reference slices)."""
        return self._dptr(C.lib().emida_offset_get_cu_temp)

    def get_cu_cross_res(self) -> int:
        """Device address of the correlation buffers cross_corr_fft writes."""
        return self._dptr(C.lib().emida_offset_get_cu_cross_res)

    def cross_corr_fft(self):
        """cross_corr_fft() (src/gpu_offset.hpp:303-316): R2C of the cu_pic buffers,
        product with conj(cu_temp), unnormalised C2R into the cu_cross_res buffers."""
        C.check(C.lib().emida_offset_cross_corr_fft(self._h))

    @property
    def in_stream(self) -> int:
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        C.check(C.lib().emida_offset_get_streams(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value

    @property
    def out_stream(self) -> int:
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        C.check(C.lib().emida_offset_get_streams(self._h, ctypes.byref(a), ctypes.byref(b)))
        return b.value

    def set_timing(self, enable=True):
        C.check(C.lib().emida_offset_set_timing(self._h, 1 if enable else 0))

    def get_timing(self):
        """(kernel name, summed HIP-event milliseconds, launches) of the dominant kernel."""
        name = ctypes.c_char_p()
        ms = ctypes.c_double()
        n = ctypes.c_uint64()
        C.check(C.lib().emida_offset_get_timing(self._h, ctypes.byref(name), ctypes.byref(ms), ctypes.byref(n)))
        return name.value.decode(), ms.value, n.value


def get_offset(pic, temp, size, cross_size, c_policy=CROSS_POLICY_BRUTE, T=np.float64):
    """get_offset<T>(pic, temp, size, cross_size, policy) (src/get_offset.hpp:8-14):
    one ROI covering a size.x x size.y picture at {0, 0}, batch 1. Returns (1, 2) offsets."""
    g = GpuOffset(T, T, size, [(0, 0)], size, cross_size, 1, c_policy)
    try:
        g.allocate_memory(temp)
        return g.get_offset(pic)[0]
    finally:
        g.close()


def fft_max_slice(T=np.float32) -> int:
    """Largest (power-of-two) slice side the FFT policy accepts for precision T."""
    v = ctypes.c_uint32()
    C.check(C.lib().emida_fft_max_slice(C.type_code(np.dtype(T)), ctypes.byref(v)))
    return v.value


def synchronize():
    C.check(C.lib().emida_device_synchronize())


def device_query():
    buf = ctypes.create_string_buffer(256)
    cu = ctypes.c_int()
    C.check(C.lib().emida_device_query(buf, 256, ctypes.byref(cu)))
    return buf.value.decode(), cu.value


def device_memory():
    """(free, total) bytes of device memory on the current device."""
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    C.check(C.lib().emida_device_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value
