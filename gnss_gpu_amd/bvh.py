"""BVH-accelerated ray casting for NLOS detection (the ``BVHAccelerator`` of
python/gnss_gpu/bvh.py:4-187): same constructor, attributes and queries."""

from __future__ import annotations

import numpy as np


class BVHAccelerator:
    """Tree over a triangle mesh [N, 3, 3]; ``_nodes_flat`` holds the reference's
    [n_nodes, 10] rows and ``_sorted_tris`` the triangles in the tree's order."""

    def __init__(self, triangles):
        self.triangles = np.asarray(triangles, dtype=np.float64)
        if self.triangles.ndim != 3 or self.triangles.shape[1:] != (3, 3):
            raise ValueError("triangles must have shape [N, 3, 3]")
        from ._bvh import bvh_build

        self._nodes_flat, self._sorted_indices = bvh_build(self.triangles)
        self._sorted_tris = self.triangles[self._sorted_indices].copy()

    def device_handle(self):
        """Device-resident copy of the tree (created on first use, freed with the
        object) for the device filter's ray-traced update."""
        h = getattr(self, "_device", None)
        if h is None:
            import ctypes as C
            import weakref

            from . import _lib

            h = C.c_void_p()
            nodes = np.ascontiguousarray(self._nodes_flat, dtype=np.float64)
            tris = np.ascontiguousarray(self._sorted_tris, dtype=np.float64)
            _lib.call("gg_bvh_create", _lib.dptr(nodes), nodes.shape[0], _lib.dptr(tris),
                      tris.shape[0], C.byref(h))
            self._device = h
            weakref.finalize(self, _lib.lib.gg_bvh_destroy, h)
        return h

    @property
    def n_nodes(self):
        return self._nodes_flat.shape[0]

    @property
    def n_triangles(self):
        return self.triangles.shape[0]

    def check_los(self, rx_ecef, sat_ecef):
        """is_los bool [n_sat] for one receiver."""
        from ._bvh import raytrace_los_check_bvh

        rx = np.asarray(rx_ecef, dtype=np.float64).ravel()
        sat = np.asarray(sat_ecef, dtype=np.float64).reshape(-1, 3)
        return raytrace_los_check_bvh(rx, sat, self._nodes_flat, self._sorted_tris)

    def compute_multipath(self, rx_ecef, sat_ecef):
        """Returns (excess_delays [n_sat], reflection_points [n_sat, 3])."""
        from ._bvh import raytrace_multipath_bvh

        rx = np.asarray(rx_ecef, dtype=np.float64).ravel()
        sat = np.asarray(sat_ecef, dtype=np.float64).reshape(-1, 3)
        refl, delay = raytrace_multipath_bvh(rx, sat, self._nodes_flat, self._sorted_tris)
        return delay, refl

    @staticmethod
    def _batch_args(rx_ecef, sat_ecef):
        rx = np.ascontiguousarray(np.asarray(rx_ecef, dtype=np.float64))
        sat = np.ascontiguousarray(np.asarray(sat_ecef, dtype=np.float64))
        if rx.ndim != 2 or rx.shape[1] != 3:
            raise ValueError("rx_ecef must have shape (N, 3)")
        if sat.ndim != 3 or sat.shape[2] != 3:
            raise ValueError("sat_ecef must have shape (N, n_sat, 3)")
        if rx.shape[0] != sat.shape[0]:
            raise ValueError("rx_ecef and sat_ecef must share the leading N")
        return rx, sat

    def check_los_batch(self, rx_ecef, sat_ecef):
        """is_los bool [N, n_sat]; non-finite entries give False (padding)."""
        from ._bvh import raytrace_los_check_bvh_batch

        rx, sat = self._batch_args(rx_ecef, sat_ecef)
        return raytrace_los_check_bvh_batch(rx, sat, self._nodes_flat, self._sorted_tris)

    def compute_multipath_batch(self, rx_ecef, sat_ecef):
        """Returns (excess_delays [N, n_sat], reflection_points [N, n_sat, 3])."""
        from ._bvh import raytrace_multipath_bvh_batch

        rx, sat = self._batch_args(rx_ecef, sat_ecef)
        refl, delay = raytrace_multipath_bvh_batch(rx, sat, self._nodes_flat,
                                                   self._sorted_tris)
        return delay, refl

    @classmethod
    def from_building_model(cls, building_model):
        return cls(building_model.triangles)

    @classmethod
    def from_obj(cls, filepath):
        from .raytrace import BuildingModel

        return cls(BuildingModel.from_obj(filepath).triangles)

    @classmethod
    def create_box(cls, center, width, depth, height):
        from .raytrace import BuildingModel

        return cls(BuildingModel.create_box(center, width, depth, height).triangles)


__all__ = ["BVHAccelerator"]
