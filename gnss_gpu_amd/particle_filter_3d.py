"""Particle filters with a ray-traced NLOS likelihood (``ParticleFilter3D``,
python/gnss_gpu/particle_filter_3d.py:15-96, and ``ParticleFilter3DBVH``,
particle_filter_3d_bvh.py:27-120).

Both keep the legacy host-array filter (``ParticleFilter``) for predict,
resample and estimate, and replace ``update``: for every particle and
satellite a ray is cast through the building mesh; a blocked ray scores the
residual with sigma_nlos after removing nlos_bias from positive residuals,
mixed with the LOS term by blocked_nlos_prob / clear_nlos_prob.
"""

from __future__ import annotations

import numpy as np

from .bvh import BVHAccelerator
from .particle_filter import ParticleFilter
from .raytrace import BuildingModel


class _RayTracedUpdate:
    def _satellite_rows(self, sat_ecef, pseudoranges, weights):
        if not self._initialized:
            raise RuntimeError(f"{type(self).__name__} not initialized. Call initialize() first.")
        sat = np.asarray(sat_ecef, dtype=np.float64).reshape(-1, 3)
        pr = np.asarray(pseudoranges, dtype=np.float64).ravel()
        n_sat = len(pr)
        if weights is None:
            weights = np.ones(n_sat, dtype=np.float64)
        else:
            weights = np.asarray(weights, dtype=np.float64).ravel()
        return sat, pr, weights, n_sat

    def _maybe_resample(self):
        if self.get_ess() < self.ess_threshold * self.n_particles:
            self._resample()


class ParticleFilter3D(_RayTracedUpdate, ParticleFilter):
    """Ray-traced NLOS likelihood over a ``BuildingModel``."""

    def __init__(self, building_model, sigma_los=3.0, sigma_nlos=30.0, nlos_bias=20.0,
                 blocked_nlos_prob=1.0, clear_nlos_prob=0.0, **kwargs):
        ParticleFilter.__init__(self, **kwargs)
        if not isinstance(building_model, BuildingModel):
            raise TypeError("building_model must be a BuildingModel instance")
        self.building_model = building_model
        self.sigma_los = sigma_los
        self.sigma_nlos = sigma_nlos
        self.nlos_bias = nlos_bias
        self.blocked_nlos_prob = blocked_nlos_prob
        self.clear_nlos_prob = clear_nlos_prob
        from ._gnss_gpu_pf3d import pf_weight_3d

        self._pf_weight_3d = pf_weight_3d

    def update(self, sat_ecef, pseudoranges, weights=None):
        sat, pr, weights, n_sat = self._satellite_rows(sat_ecef, pseudoranges, weights)
        self._pf_weight_3d(self._px, self._py, self._pz, self._pcb, sat.ravel(), pr, weights,
                           self.building_model.triangles.reshape(-1, 3, 3), self._log_weights,
                           self.n_particles, n_sat, float(self.sigma_los),
                           float(self.sigma_nlos), float(self.nlos_bias),
                           float(self.blocked_nlos_prob), float(self.clear_nlos_prob))
        self._maybe_resample()


class ParticleFilter3DBVH(_RayTracedUpdate, ParticleFilter):
    """Ray-traced NLOS likelihood over a ``BVHAccelerator`` (drop-in for
    ``ParticleFilter3D`` with ``bvh`` in place of ``building_model``)."""

    def __init__(self, bvh, sigma_los=3.0, sigma_nlos=30.0, nlos_bias=20.0,
                 blocked_nlos_prob=1.0, clear_nlos_prob=0.0, **kwargs):
        ParticleFilter.__init__(self, **kwargs)
        if not isinstance(bvh, BVHAccelerator):
            raise TypeError("bvh must be a BVHAccelerator instance")
        self.bvh = bvh
        self.sigma_los = sigma_los
        self.sigma_nlos = sigma_nlos
        self.nlos_bias = nlos_bias
        self.blocked_nlos_prob = blocked_nlos_prob
        self.clear_nlos_prob = clear_nlos_prob
        from ._gnss_gpu_pf3d_bvh import pf_weight_3d_bvh

        self._pf_weight_3d_bvh = pf_weight_3d_bvh

    def update(self, sat_ecef, pseudoranges, weights=None):
        sat, pr, weights, n_sat = self._satellite_rows(sat_ecef, pseudoranges, weights)
        self._pf_weight_3d_bvh(self._px, self._py, self._pz, self._pcb, sat.ravel(), pr, weights,
                               self.bvh._nodes_flat, self.bvh._sorted_tris, self._log_weights,
                               self.n_particles, n_sat, float(self.sigma_los),
                               float(self.sigma_nlos), float(self.nlos_bias),
                               float(self.blocked_nlos_prob), float(self.clear_nlos_prob))
        self._maybe_resample()


__all__ = ["ParticleFilter3D", "ParticleFilter3DBVH"]
