"""Vendor-library baselines on the same GPU: rocSOLVER's tridiagonal
eigenvalue routines (STERF: Pal-Walker-Kahan QL/QR; STEDC with no vectors),
called through their C API on torch device buffers.  Reported beside the BR
path, not part of it (the analogue of the reference's cuSOLVER column).
"""
from __future__ import annotations

import ctypes
import ctypes.util

_rb = None
_rs = None
_handles = {}

ROCBLAS_EVECT_NONE = 213  # rocblas_evect_none (rocsolver-extra-types.h)


def _libs():
    global _rb, _rs
    if _rs is None:
        _rb = ctypes.CDLL("/opt/rocm/lib/librocblas.so")
        _rs = ctypes.CDLL("/opt/rocm/lib/librocsolver.so")
        _rb.rocblas_create_handle.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _rb.rocblas_set_stream.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _rs.rocsolver_dsterf.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]
        _rs.rocsolver_dstedc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int, ctypes.c_void_p]
    return _rb, _rs


def _handle(stream):
    rb, _ = _libs()
    h = _handles.get(None)
    if h is None:
        h = ctypes.c_void_p()
        if rb.rocblas_create_handle(ctypes.byref(h)) != 0:
            raise RuntimeError("rocblas_create_handle failed")
        _handles[None] = h
    if rb.rocblas_set_stream(h, ctypes.c_void_p(stream.cuda_stream)) != 0:
        raise RuntimeError("rocblas_set_stream failed")
    return h


def rocsolver_eigvals(d, e, method: str = "sterf", stream=None):
    """Eigenvalues of tridiag(e, d, e) by rocSOLVER on the GPU.  d (n) and e
    (n-1) are float64 device tensors; they are copied, not modified.
    Returns (w, info) with info a device int32 tensor (not synchronised)."""
    import torch
    _, rs = _libs()
    st = stream if stream is not None else torch.cuda.current_stream(d.device)
    n = d.numel()
    w = d.detach().clone()
    ew = torch.zeros(max(n, 1), dtype=torch.float64, device=d.device)
    if n > 1:
        ew[: n - 1].copy_(e.detach()[: n - 1])
    info = torch.zeros(1, dtype=torch.int32, device=d.device)
    h = _handle(st)
    if method == "sterf":
        status = rs.rocsolver_dsterf(h, n, w.data_ptr(), ew.data_ptr(), info.data_ptr())
    elif method == "stedc":
        status = rs.rocsolver_dstedc(h, ROCBLAS_EVECT_NONE, n, w.data_ptr(), ew.data_ptr(), None,
                                     max(n, 1), info.data_ptr())
    else:
        raise ValueError("method is 'sterf' or 'stedc'")
    if status != 0:
        raise RuntimeError(f"rocsolver_d{method} returned rocblas_status {status}")
    return w, info
