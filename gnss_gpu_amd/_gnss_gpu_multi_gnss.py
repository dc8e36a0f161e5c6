"""Multi-GNSS WLS with one clock bias per system (the ``gnss_gpu._gnss_gpu_multi_gnss``
module, python/gnss_gpu/_multi_gnss_bindings.cpp:8-67).

Both functions run on the GPU.  ``wls_multi_gnss_batch`` keeps the reference
batch kernel's Sagnac range model (src/positioning/multi_gnss.cu:211-305);
``wls_multi_gnss`` keeps the reference's single-epoch solver
(multi_gnss.cu:50-153), which takes plain ranges.  The state is
``[x, y, z, cb_0, ..., cb_{n_systems-1}]``; ``system_ids`` index the clocks.
``wls_multi_gnss_batch_device`` works on device pointers (torch tensors).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, dptr, f64, i32, iptr

MAX_SYSTEMS = 5


def _n_state(n_systems):
    n_systems = int(n_systems)
    if n_systems < 0 or n_systems > MAX_SYSTEMS:
        raise ValueError(f"n_systems must be in [0, {MAX_SYSTEMS}]")
    return 3 + n_systems


def wls_multi_gnss(sat_ecef, pseudoranges, weights, system_ids, n_systems, max_iter=10,
                   tol=1e-4):
    """Single-epoch solve; returns (ndarray[3 + n_systems], iterations).

    sat_ecef is (n_sat, 3) or flat (3 * n_sat,).  Fewer than 3 + n_systems
    satellites, or a satellite centroid at the origin, give zeros and -1.
    """
    n_state = _n_state(n_systems)
    pr = f64(pseudoranges).ravel()
    n_sat = pr.size
    sat = f64(sat_ecef).ravel()
    w = f64(weights).ravel()
    sid = i32(system_ids).ravel()
    if sat.size < 3 * n_sat or w.size < n_sat or sid.size < n_sat:
        raise ValueError("sat_ecef/weights/system_ids shorter than pseudoranges")
    result = np.empty(n_state, dtype=np.float64)
    it = C.c_int32()
    call("gg_wls_multi_gnss", dptr(sat), dptr(pr), dptr(w), iptr(sid), n_sat, int(n_systems),
         int(max_iter), float(tol), dptr(result), C.byref(it))
    return result, int(it.value)


def wls_multi_gnss_batch(sat_ecef, pseudoranges, weights, system_ids, n_systems, max_iter=10,
                         tol=1e-4):
    """Batched solve over epochs; returns (ndarray[E, 3 + n_systems], int32[E]).

    Raises RuntimeError when n_sat < 3 + n_systems, as the reference binding does.
    """
    n_state = _n_state(n_systems)
    sat = f64(sat_ecef)
    if sat.ndim != 3 or sat.shape[2] != 3:
        raise ValueError("sat_ecef must have shape (n_epoch, n_sat, 3)")
    n_epoch, n_sat = int(sat.shape[0]), int(sat.shape[1])
    if n_sat < n_state:
        raise RuntimeError("wls_multi_gnss_batch requires at least 3 + n_systems satellites")
    pr = f64(pseudoranges).reshape(n_epoch, n_sat)
    w = f64(weights).reshape(n_epoch, n_sat)
    sid = i32(system_ids).reshape(n_epoch, n_sat)
    results = np.empty((n_epoch, n_state), dtype=np.float64)
    iters = np.empty(n_epoch, dtype=np.int32)
    call("gg_wls_multi_gnss_batch", dptr(sat), dptr(pr), dptr(w), iptr(sid), n_epoch, n_sat,
         int(n_systems), int(max_iter), float(tol), dptr(results), iptr(iters))
    return results, iters


def wls_multi_gnss_batch_device(sat_ecef, pseudoranges, weights, system_ids, results, iters,
                                n_systems, max_iter=10, tol=1e-4, mode=0, stream=None):
    """Device-pointer solve on torch tensors (float64 sat [E,S,3], pr/w [E,S],
    int32 system_ids [E,S], results [E, 3 + n_systems], int32 iters [E]).
    mode 0 is the batch kernel's semantics, 1 the single-epoch solver's.
    Enqueued on ``stream`` (default: torch's current stream); no synchronisation."""
    import torch

    n_state = _n_state(n_systems)
    for t in (sat_ecef, pseudoranges, weights, results):
        if t.dtype != torch.float64 or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("float64 contiguous device tensors required")
    for t in (system_ids, iters):
        if t.dtype != torch.int32 or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("system_ids and iters must be contiguous int32 device tensors")
    n_epoch, n_sat = int(sat_ecef.shape[0]), int(sat_ecef.shape[1])
    if (tuple(pseudoranges.shape) != (n_epoch, n_sat) or tuple(weights.shape) != (n_epoch, n_sat)
            or tuple(system_ids.shape) != (n_epoch, n_sat)
            or tuple(results.shape) != (n_epoch, n_state) or iters.numel() != n_epoch):
        raise ValueError("tensor shapes do not match sat_ecef (n_epoch, n_sat, 3)")
    if stream is None:
        stream = torch.cuda.current_stream(sat_ecef.device).cuda_stream
    call("gg_wls_multi_gnss_batch_device", C.c_void_p(sat_ecef.data_ptr()),
         C.c_void_p(pseudoranges.data_ptr()), C.c_void_p(weights.data_ptr()),
         C.c_void_p(system_ids.data_ptr()), n_epoch, n_sat, int(n_systems), int(max_iter),
         float(tol), C.c_void_p(results.data_ptr()), C.c_void_p(iters.data_ptr()), int(mode),
         C.c_void_p(stream))


__all__ = ["wls_multi_gnss", "wls_multi_gnss_batch", "wls_multi_gnss_batch_device"]
_ = _lib
