"""``gnss_gpu.particle_filter_3d_bvh`` counterpart: ``ParticleFilter3DBVH``."""

from .particle_filter_3d import ParticleFilter3DBVH

__all__ = ["ParticleFilter3DBVH"]
