"""3D building model for LOS / multipath queries (the ``BuildingModel`` of
python/gnss_gpu/raytrace.py:4-134): same constructor, OBJ loader, box helper
and queries; the queries run on the GPU."""

from __future__ import annotations

import numpy as np


class BuildingModel:
    """Triangle mesh [N, 3, 3] (ECEF metres) for ray casting."""

    def __init__(self, triangles):
        self.triangles = np.asarray(triangles, dtype=np.float64)
        if self.triangles.ndim != 3 or self.triangles.shape[1:] != (3, 3):
            raise ValueError("triangles must have shape [N, 3, 3]")

    def check_los(self, rx_ecef, sat_ecef):
        """is_los bool [n_sat]: True when no triangle lies between receiver and satellite."""
        from ._raytrace import raytrace_los_check

        rx = np.asarray(rx_ecef, dtype=np.float64).ravel()
        sat = np.asarray(sat_ecef, dtype=np.float64).reshape(-1, 3)
        return raytrace_los_check(rx, sat, self.triangles.reshape(-1, 3, 3))

    def compute_multipath(self, rx_ecef, sat_ecef):
        """Returns (excess_delays [n_sat], reflection_points [n_sat, 3])."""
        from ._raytrace import raytrace_multipath

        rx = np.asarray(rx_ecef, dtype=np.float64).ravel()
        sat = np.asarray(sat_ecef, dtype=np.float64).reshape(-1, 3)
        refl, delay = raytrace_multipath(rx, sat, self.triangles.reshape(-1, 3, 3))
        return delay, refl

    @classmethod
    def from_obj(cls, filepath):
        """Wavefront OBJ ('v' and 'f' records; polygons fan-triangulated from their
        first vertex, raytrace.py:56-93)."""
        vertices, triangles = [], []
        with open(filepath, "r") as fh:
            for line in fh:
                fields = line.split()
                if not fields or fields[0].startswith("#"):
                    continue
                if fields[0] == "v":
                    vertices.append([float(v) for v in fields[1:4]])
                elif fields[0] == "f":
                    idx = [int(tok.split("/")[0]) - 1 for tok in fields[1:]]
                    for k in range(1, len(idx) - 1):
                        triangles.append([vertices[idx[0]], vertices[idx[k]],
                                          vertices[idx[k + 1]]])
        return cls(np.array(triangles, dtype=np.float64).reshape(-1, 3, 3))

    @classmethod
    def create_box(cls, center, width, depth, height):
        """Axis-aligned box as 12 triangles, two per face (raytrace.py:95-134)."""
        c = np.asarray(center, dtype=np.float64).ravel()
        h = np.array([width, depth, height], dtype=np.float64) / 2.0
        # corners 0-3: bottom layer counter-clockwise from (-x, -y); 4-7: top layer
        corners = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                            [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
        v = c + corners * h
        faces = [(0, 1, 2), (0, 2, 3), (4, 6, 5), (4, 7, 6), (0, 5, 1), (0, 4, 5),
                 (2, 7, 3), (2, 6, 7), (0, 3, 7), (0, 7, 4), (1, 5, 6), (1, 6, 2)]
        return cls(np.array([[v[a], v[b], v[d]] for a, b, d in faces], dtype=np.float64))


__all__ = ["BuildingModel"]
