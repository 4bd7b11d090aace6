"""Python side of the request server (emida_server_* in include/emida_amd.h).

The server has no counterpart in the reference; it serves single uint16 patterns
against resident pattern-offset configurations ("plans") with continuous batching.
Results per request are those of GpuOffset.get_offset on the same pattern.
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np

from . import _capi as C
from .gpu_offset import CROSS_POLICY_BRUTE, MFMA_AUTO


class Server:
    """hbm_budget: bytes of device memory for batch slots; max_batch: patterns per
    launch; streams: concurrent device streams; max_queued: pending requests before
    submit blocks; use_graphs: replay hipGraph-captured pipelines; slots_per_plan:
    batch slots one plan may hold (0: as many as the budget allows)."""

    def __init__(self, hbm_budget=8 << 30, max_batch=16, streams=4, max_queued=256, use_graphs=True,
                 slots_per_plan=2):
        cfg = C.ServerConfig(int(hbm_budget), int(max_batch), int(streams), int(max_queued), 1 if use_graphs else 0,
                             int(slots_per_plan))
        h = ctypes.c_void_p()
        C.check(C.lib().emida_server_create(ctypes.byref(h), ctypes.byref(cfg)))
        self._h = h
        self._plans = []
        self._pending = {}  # request id -> output arrays the library writes into
        self._lock = threading.Lock()

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            C.lib().emida_server_destroy(self._h)  # completes every pending request first
            self._h = ctypes.c_void_p()
            self._pending.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_plan(self, T, ref, begins, slice_size, cross_size=None, policy=CROSS_POLICY_BRUTE, s=3,
                 mfma_mode=MFMA_AUTO) -> int:
        """ref: uint16 (rows, cols) reference pattern; begins: (x, y) ROI corners.
        Returns the plan id."""
        ref = np.ascontiguousarray(ref, dtype=np.uint16)
        src = (ref.shape[1], ref.shape[0])
        if cross_size is None:
            cross_size = (2 * slice_size[0] - 1, 2 * slice_size[1] - 1)
        arr = (C.Size2 * len(begins))(*[C.Size2(int(x), int(y)) for x, y in begins])
        pid = ctypes.c_uint32()
        C.check(C.lib().emida_server_add_plan(self._h, C.type_code(np.dtype(T)), C.size2(src), C.ptr(ref), arr,
                                              len(begins), C.size2(slice_size), C.size2(cross_size), int(policy),
                                              int(mfma_mode), int(s), ctypes.byref(pid)))
        info = C.OffsetInfo()
        mb = ctypes.c_uint32()
        C.check(C.lib().emida_server_plan_info(self._h, pid.value, ctypes.byref(info), ctypes.byref(mb)))
        self._plans.append(dict(n_rois=len(begins), src=src, max_batch=mb.value, cross=(info.cross_size.x,
                                                                                         info.cross_size.y)))
        return pid.value

    def plan(self, pid):
        return self._plans[pid]

    def submit(self, pid, pattern, offsets=None, coefs=None):
        """Queues one pattern; returns (request id, offsets (n, 2), coefs (n, 6)).
        The output arrays are filled when the request completes (see wait); the
        server keeps them alive until then."""
        n = self._plans[pid]["n_rois"]
        pattern = np.ascontiguousarray(pattern, dtype=np.uint16)
        if pattern.size != self._plans[pid]["src"][0] * self._plans[pid]["src"][1]:
            raise ValueError("pattern size does not match the plan")
        if offsets is None:
            offsets = np.empty((n, 2))
        if coefs is None:
            coefs = np.empty((n, 6))
        rid = ctypes.c_uint64()
        dp = ctypes.POINTER(ctypes.c_double)
        if offsets.shape != (n, 2) or offsets.dtype != np.float64 or not offsets.flags.c_contiguous:
            raise ValueError("offsets must be a C-contiguous float64 (n_rois, 2) array")
        if coefs.shape != (n, 6) or coefs.dtype != np.float64 or not coefs.flags.c_contiguous:
            raise ValueError("coefs must be a C-contiguous float64 (n_rois, 6) array")
        # the local references keep the arrays alive until they are registered
        C.check(C.lib().emida_server_submit(self._h, pid, C.ptr(pattern), offsets.ctypes.data_as(dp),
                                            coefs.ctypes.data_as(dp), ctypes.byref(rid)))
        with self._lock:
            self._pending[rid.value] = (offsets, coefs)
        return rid.value, offsets, coefs

    def wait(self, rid):
        """Blocks until the request completed; returns its timing record (ns)."""
        info = C.RequestInfo()
        C.check(C.lib().emida_server_wait(self._h, rid, ctypes.byref(info)))
        with self._lock:
            self._pending.pop(rid, None)
        return dict(submit_ns=info.submit_ns, launch_ns=info.launch_ns, done_ns=info.done_ns, batch=info.batch,
                    plan=info.plan, latency_ms=(info.done_ns - info.submit_ns) * 1e-6)

    def stats(self):
        st = C.ServerStats()
        C.check(C.lib().emida_server_get_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in C.ServerStats._fields_}
