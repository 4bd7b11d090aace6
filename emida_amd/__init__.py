"""MI355X (gfx950) implementation of emida's EBSD pattern-offset pipeline.

The compute path is libemida_amd.so (hand-written HIP kernels behind the C ABI in
include/emida_amd.h).  This package is its Python host side: GpuOffset mirrors
gpu_offset<T, IN>, get_offset mirrors get_offset<T>, kernels mirrors the
kernel-level launchers.  Importing the package does not load the library; the
first call does, and fails loudly if it was not built.
"""
from ._capi import (CROSS_POLICY_BRUTE, CROSS_POLICY_BRUTE_MFMA, CROSS_POLICY_FFT, LIB_PATH, MFMA_AUTO,  # noqa: F401
                    MFMA_BF16, MFMA_BF16_SPLIT, MFMA_F16, MFMA_F16_SPLIT, EmidaError)
from .gpu_offset import (DeviceBuffer, GpuOffset, PinnedBuffer, data_index_dtype, device_memory,  # noqa: F401
                         device_query, fft_max_slice, get_offset, synchronize)
from .serve import Server  # noqa: F401

__all__ = ["CROSS_POLICY_BRUTE", "CROSS_POLICY_FFT", "CROSS_POLICY_BRUTE_MFMA", "MFMA_AUTO", "MFMA_BF16", "MFMA_BF16_SPLIT",
           "MFMA_F16", "MFMA_F16_SPLIT", "EmidaError", "GpuOffset", "DeviceBuffer", "PinnedBuffer",
           "data_index_dtype", "get_offset", "synchronize", "device_query", "device_memory", "fft_max_slice", "LIB_PATH", "Server"]
