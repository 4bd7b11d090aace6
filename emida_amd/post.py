"""Deformation post-step: per pattern, the projective map between ROI middles and
their displaced positions and its strain and rotation, on the GPU.

Mirrors the reference's Python post-processing of emida's output:
python/process_offsets.py:31-97 (algebraic fit, the map divided by its [2][2]
element) and jove/analyze.py:6-95 (algebraic_fit, post_process: ROI selection by the
fitted quadratic, median quality, scaling to determinant 1, polar decomposition).
Every computation runs in libemida_amd.so (emida_fit_deformation); this module moves
numpy arrays to and from the device and parses emida's standard output.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C
from .gpu_offset import DeviceBuffer


def fit_deformation(mids, offsets, coefs=None, calib=(436.0, 436.0, 436.0), det_min=1e-6, min_rois=10):
    """Batched fit.  mids (N, 2) shared or (P, N, 2) per pattern; offsets (P, N, 2);
    coefs (P, N, 6) or None (every ROI used).  Returns a dict of numpy arrays: mask (P,),
    quality (P,), H, F, epsilon, omega (P, 3, 3); NaN where a pattern is not fitted."""
    offsets = np.ascontiguousarray(offsets, float)
    if offsets.ndim == 2:
        offsets = offsets[np.newaxis]
    P, N = offsets.shape[:2]
    mids = np.ascontiguousarray(mids, float)
    per_pattern = mids.ndim == 3
    if mids.shape[-2:] != (N, 2) or (per_pattern and mids.shape[0] != P):
        raise ValueError("mids must be (N, 2) or (P, N, 2)")
    if coefs is not None:
        coefs = np.ascontiguousarray(coefs, float).reshape(P, N, 6)
    d_mids, d_offs = DeviceBuffer.from_array(mids), DeviceBuffer.from_array(offsets)
    d_coefs = DeviceBuffer.from_array(coefs) if coefs is not None else None
    d_mask = DeviceBuffer(4 * P)
    d_q = DeviceBuffer(8 * P)
    outs = {k: DeviceBuffer(72 * P) for k in ("H", "F", "epsilon", "omega")}
    cal = C.Calib(*(float(v) for v in calib))
    C.check(C.lib().emida_fit_deformation(d_mids.ptr, 1 if per_pattern else 0, N, d_offs.ptr,
                                          d_coefs.ptr if d_coefs is not None else None, P, cal, float(det_min),
                                          int(min_rois), d_mask.ptr, d_q.ptr, outs["H"].ptr, outs["F"].ptr,
                                          outs["epsilon"].ptr, outs["omega"].ptr, None))
    res = {"mask": d_mask.download(np.int32).astype(bool), "quality": d_q.download(np.float64)}
    for k, b in outs.items():
        res[k] = b.download(np.float64).reshape(P, 3, 3)
    return res


def algebraic_fit(old, new, calib):
    """algebraic_fit (jove/analyze.py:6-31, rows as python/process_offsets.py:73-88):
    the unit-norm null vector of the 2N x 9 system as a 3 x 3 matrix, sign chosen so
    that its [2][2] element is positive (the reference returns either sign)."""
    old = np.asarray(old, float)
    new = np.asarray(new, float)
    if len(old) != len(new) or len(old) <= 4:
        raise ValueError("algebraic_fit needs matching point sets of more than 4 points")
    r = fit_deformation(old, new - old, None, calib)
    H = r["H"][0]
    return H / np.linalg.norm(H)


def post_process(data, calib, M=None):
    """post_process (jove/analyze.py:37-95): data is a sequence of (N, 10) arrays
    (ROI middle x, y, offset u, v, six fit coefficients).  Returns (mask, F, epsilon,
    omega, quality) with F, epsilon, omega for the fitted patterns only, as the
    reference.  M (camera to sample) is unused there too."""
    data = np.asarray(data, float)
    r = fit_deformation(data[:, :, :2], data[:, :, 2:4], data[:, :, 4:10], calib)
    m = r["mask"]
    return m, r["F"][m], r["epsilon"][m], r["omega"][m], r["quality"]


def read_output(text):
    """Parses emida's standard output (finalize_offsets, src/process_files.hpp:244-260):
    returns a list of (x, y, file name, rows) with rows (N, 4) or (N, 10) per pattern."""
    lines = text.splitlines() if isinstance(text, str) else list(text)
    out, i = [], 0
    while i < len(lines):
        head = lines[i].split(" ", 3)
        if len(head) < 3 or not head[0]:
            i += 1
            continue
        x, y, n = float(head[0]), float(head[1]), int(head[2])
        name = head[3] if len(head) > 3 else ""
        rows = np.array([[float(v) for v in lines[i + 1 + k].split()] for k in range(n)])
        out.append((x, y, name, rows))
        i += 1 + n
    return out
