"""Ray casting against a triangle list (the ``gnss_gpu._raytrace`` module,
python/gnss_gpu/_raytrace_bindings.cpp).  The reference scans every triangle;
here a tree is built over the triangles on each call and the BVH kernels run,
which gives the same answers."""

from __future__ import annotations

import numpy as np

from . import _lib
from ._bvh import _rx
from ._lib import call, dptr, f64, iptr


def raytrace_los_check(rx_ecef, sat_ecef, triangles):
    """is_los bool [n_sat]."""
    rx = _rx(rx_ecef)
    sat = f64(sat_ecef).reshape(-1, 3)
    tris = f64(triangles).reshape(-1, 3, 3)
    out = np.zeros(sat.shape[0], dtype=np.int32)
    call("gg_raytrace_los_check", dptr(rx), dptr(sat), dptr(tris), tris.shape[0], iptr(out),
         sat.shape[0])
    return out != 0


def raytrace_multipath(rx_ecef, sat_ecef, triangles):
    """Returns (reflection_points [n_sat, 3], excess_delays [n_sat])."""
    rx = _rx(rx_ecef)
    sat = f64(sat_ecef).reshape(-1, 3)
    tris = f64(triangles).reshape(-1, 3, 3)
    refl = np.zeros((sat.shape[0], 3), dtype=np.float64)
    delay = np.zeros(sat.shape[0], dtype=np.float64)
    call("gg_raytrace_multipath", dptr(rx), dptr(sat), dptr(tris), tris.shape[0], dptr(refl),
         dptr(delay), sat.shape[0])
    return refl, delay


__all__ = ["raytrace_los_check", "raytrace_multipath"]
_ = _lib
