"""WLS positioning functions (the ``wls_position``/``wls_batch`` part of
``gnss_gpu._gnss_gpu``, python/gnss_gpu/_bindings.cpp:55-91).

Both run on the GPU.  ``wls_batch`` keeps the reference's Sagnac range model
and Gaussian elimination; ``wls_position`` keeps the reference CPU solver's
semantics (no Sagnac, early return on a singular matrix).  ``wls_batch``
additionally accepts ``n_sat`` (per-epoch satellite counts) for ragged
batches, and ``wls_batch_device`` works on device pointers (torch tensors).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, dptr, f64, i32, iptr


def wls_position(sat_ecef, pseudoranges, weights, max_iter=10, tol=1e-4):
    """Single-epoch WLS; returns (ndarray[4], iterations)."""
    pr = f64(pseudoranges).ravel()
    n_sat = pr.size
    if n_sat < 4:
        raise RuntimeError("wls_position requires at least 4 satellites")
    sat = f64(sat_ecef).ravel()
    w = f64(weights).ravel()
    if sat.size < 3 * n_sat or w.size < n_sat:
        raise ValueError("sat_ecef/weights shorter than pseudoranges")
    result = np.empty(4, dtype=np.float64)
    it = C.c_int32()
    call("gg_wls_position", dptr(sat), dptr(pr), dptr(w), n_sat, int(max_iter), float(tol),
         dptr(result), C.byref(it))
    return result, int(it.value)


def wls_batch(sat_ecef, pseudoranges, weights, max_iter=10, tol=1e-4, n_sat=None):
    """Batched WLS over epochs; returns (ndarray[E, 4], int32[E])."""
    sat = f64(sat_ecef)
    if sat.ndim != 3 or sat.shape[2] != 3:
        raise ValueError("sat_ecef must have shape (n_epoch, n_sat, 3)")
    n_epoch, s_max = int(sat.shape[0]), int(sat.shape[1])
    if s_max < 4:
        raise RuntimeError("wls_batch requires at least 4 satellites")
    pr = f64(pseudoranges).reshape(n_epoch, s_max)
    w = f64(weights).reshape(n_epoch, s_max)
    ns = None
    if n_sat is not None:
        ns = i32(n_sat).ravel()
        if ns.size != n_epoch:
            raise ValueError("n_sat must have one count per epoch")
    results = np.empty((n_epoch, 4), dtype=np.float64)
    iters = np.empty(n_epoch, dtype=np.int32)
    call("gg_wls_batch", dptr(sat), dptr(pr), dptr(w), iptr(ns) if ns is not None else None,
         n_epoch, s_max, int(max_iter), float(tol), dptr(results), iptr(iters))
    return results, iters


def wls_batch_device(sat_ecef, pseudoranges, weights, results, iters, n_sat=None,
                     max_iter=10, tol=1e-4, mode=0, stream=None):
    """Device-pointer WLS on torch tensors (float64 sat [E,S,3], pr/w [E,S],
    results [E,4], int32 iters [E], optional int32 n_sat [E]).  Enqueued on
    ``stream`` (default: torch's current stream); no synchronisation."""
    import torch

    for t in (sat_ecef, pseudoranges, weights, results):
        if t.dtype != torch.float64 or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("float64 contiguous device tensors required")
    if iters.dtype != torch.int32 or not iters.is_contiguous():
        raise ValueError("iters must be a contiguous int32 device tensor")
    n_epoch, s_max = int(sat_ecef.shape[0]), int(sat_ecef.shape[1])
    if stream is None:
        stream = torch.cuda.current_stream(sat_ecef.device).cuda_stream
    ns_ptr = None
    if n_sat is not None:
        if n_sat.dtype != torch.int32 or not n_sat.is_contiguous():
            raise ValueError("n_sat must be a contiguous int32 device tensor")
        ns_ptr = C.c_void_p(n_sat.data_ptr())
    call("gg_wls_batch_device", C.c_void_p(sat_ecef.data_ptr()),
         C.c_void_p(pseudoranges.data_ptr()), C.c_void_p(weights.data_ptr()), ns_ptr,
         n_epoch, s_max, int(max_iter), float(tol), C.c_void_p(results.data_ptr()),
         C.c_void_p(iters.data_ptr()), int(mode), C.c_void_p(stream))


__all__ = ["wls_position", "wls_batch", "wls_batch_device"]
_ = _lib
