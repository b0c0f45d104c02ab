"""Helpers for the GPU tests (device <-> host copies of raw pointers)."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def d2h(ptr, shape, dtype):
    out = np.empty(shape, dtype)
    h = hip()
    assert h.hipDeviceSynchronize() == 0
    assert h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), out.nbytes, 2) == 0
    return out


def to_dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def pairs_np(t):
    """int32 device pair tensor -> uint32 numpy [P,2]."""
    return t.cpu().numpy().view(np.uint32).reshape(-1, 2)
