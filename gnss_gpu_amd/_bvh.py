"""BVH ray casting (the ``gnss_gpu._bvh`` module, python/gnss_gpu/_bvh_bindings.cpp).

Trees use the reference's flat node rows ``[n_nodes, 10]`` (min_xyz, max_xyz,
left, right, tri_start, tri_count; ``left == -1`` marks a leaf) and the
triangles ``[n_tri, 3, 3]`` in the tree's order.  The build runs on the host,
the queries on the GPU (binary64).
"""

from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import call, dptr, f64, iptr


def _tree(nodes_flat, sorted_tris):
    nodes = f64(nodes_flat).reshape(-1, 10)
    tris = f64(sorted_tris).reshape(-1, 3, 3)
    return nodes, tris


def _rx(rx_ecef):
    rx = f64(rx_ecef).ravel()
    if rx.size < 3:
        raise RuntimeError("rx_ecef must have at least 3 elements")
    return rx


def bvh_build(triangles):
    """Returns (nodes_flat [n_nodes, 10] float64, sorted_indices int32 [n_tri])."""
    tris = f64(triangles).reshape(-1, 3, 3)
    n_tri = tris.shape[0]
    nodes = np.zeros((max(2 * n_tri, 1), 10), dtype=np.float64)
    order = np.zeros(max(n_tri, 1), dtype=np.int32)
    n_nodes = np.zeros(1, dtype=np.int32)
    call("gg_bvh_build", dptr(tris), n_tri, dptr(nodes), iptr(n_nodes), iptr(order))
    return nodes[: int(n_nodes[0])].copy(), order[:n_tri].copy()


def raytrace_los_check_bvh(rx_ecef, sat_ecef, nodes_flat, sorted_tris):
    """is_los bool [n_sat] for one receiver."""
    rx = _rx(rx_ecef)
    sat = f64(sat_ecef).reshape(-1, 3)
    nodes, tris = _tree(nodes_flat, sorted_tris)
    out = np.zeros(sat.shape[0], dtype=np.int32)
    call("gg_raytrace_los_check_bvh", dptr(rx), dptr(sat), dptr(nodes), nodes.shape[0],
         dptr(tris), tris.shape[0], iptr(out), sat.shape[0])
    return out != 0


def raytrace_multipath_bvh(rx_ecef, sat_ecef, nodes_flat, sorted_tris):
    """Returns (reflection_points [n_sat, 3], excess_delays [n_sat])."""
    rx = _rx(rx_ecef)
    sat = f64(sat_ecef).reshape(-1, 3)
    nodes, tris = _tree(nodes_flat, sorted_tris)
    refl = np.zeros((sat.shape[0], 3), dtype=np.float64)
    delay = np.zeros(sat.shape[0], dtype=np.float64)
    call("gg_raytrace_multipath_bvh", dptr(rx), dptr(sat), dptr(nodes), nodes.shape[0],
         dptr(tris), tris.shape[0], dptr(refl), dptr(delay), sat.shape[0])
    return refl, delay


def _batch_shapes(rx_ecef, sat_ecef):
    rx = f64(rx_ecef)
    sat = f64(sat_ecef)
    if rx.ndim != 2 or rx.shape[1] != 3:
        raise RuntimeError("rx_ecef must have shape (N, 3)")
    if sat.ndim != 3 or sat.shape[2] != 3:
        raise RuntimeError("sat_ecef must have shape (N, n_sat, 3)")
    if rx.shape[0] != sat.shape[0]:
        raise RuntimeError("rx_ecef and sat_ecef must share the leading N")
    return rx, sat


def raytrace_los_check_bvh_batch(rx_ecef, sat_ecef, nodes_flat, sorted_tris):
    """is_los bool [N, n_sat]; a non-finite coordinate gives False."""
    rx, sat = _batch_shapes(rx_ecef, sat_ecef)
    nodes, tris = _tree(nodes_flat, sorted_tris)
    n_epoch, n_sat = sat.shape[0], sat.shape[1]
    out = np.zeros((n_epoch, n_sat), dtype=np.int32)
    call("gg_raytrace_los_check_bvh_batch", dptr(rx), dptr(sat), dptr(nodes), nodes.shape[0],
         dptr(tris), tris.shape[0], iptr(out), n_epoch, n_sat)
    return out != 0


def raytrace_multipath_bvh_batch(rx_ecef, sat_ecef, nodes_flat, sorted_tris):
    """Returns (reflection_points [N, n_sat, 3], excess_delays [N, n_sat])."""
    rx, sat = _batch_shapes(rx_ecef, sat_ecef)
    nodes, tris = _tree(nodes_flat, sorted_tris)
    n_epoch, n_sat = sat.shape[0], sat.shape[1]
    refl = np.zeros((n_epoch, n_sat, 3), dtype=np.float64)
    delay = np.zeros((n_epoch, n_sat), dtype=np.float64)
    call("gg_raytrace_multipath_bvh_batch", dptr(rx), dptr(sat), dptr(nodes), nodes.shape[0],
         dptr(tris), tris.shape[0], dptr(refl), dptr(delay), n_epoch, n_sat)
    return refl, delay


__all__ = ["bvh_build", "raytrace_los_check_bvh", "raytrace_multipath_bvh",
           "raytrace_los_check_bvh_batch", "raytrace_multipath_bvh_batch"]
_ = _lib
