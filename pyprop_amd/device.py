"""Device-memory plumbing: wavefunction buffers are torch complex128/complex64 tensors on
the current HIP device; the library receives their raw pointers and the current stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


def require_gpu():
    if torch is None or not torch.cuda.is_available():
        raise RuntimeError("pyprop_amd needs a HIP device (no CPU fallback)")


def stream_handle():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t) -> C.c_void_p:
    if t is None:
        return C.c_void_p(None)
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def hptr(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


def to_device(a, dtype=None):
    require_gpu()
    if isinstance(a, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(a))
    else:
        t = a
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda").contiguous()


def zeros(shape, dtype=None):
    require_gpu()
    return torch.zeros(tuple(shape), dtype=dtype or torch.complex128, device="cuda")


def synchronize():
    torch.cuda.synchronize()
