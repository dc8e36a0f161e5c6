"""The ``gnss_gpu._gnss_gpu_pf3d`` module (python/gnss_gpu/_pf3d_bindings.cpp:10-52):
``pf_weight_3d`` over a triangle list; see ``_gnss_gpu_pf3d_bvh``."""

from ._gnss_gpu_pf3d_bvh import pf_weight_3d

__all__ = ["pf_weight_3d"]
