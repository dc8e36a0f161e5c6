"""gnss_gpu_amd: MI355X-native GNSS particle-filter step and batched WLS.

Public names follow python/gnss_gpu/__init__.py:36-92 for the in-scope part of
that package:

    ParticleFilter, ParticleFilterDevice, wls_position, wls_batch, MultiGNSSSolver,
    BuildingModel, BVHAccelerator, ParticleFilter3D, ParticleFilter3DBVH

plus the batched multi-filter engine ``BatchedParticleFilter``.  The compiled
HIP library (libgnss_gpu_amd.so) is loaded on first use of any of them; there
is no CPU fallback.
"""

__version__ = "0.1.0"

_LAZY = {
    "ParticleFilter": ("particle_filter", "ParticleFilter"),
    "ParticleFilterDevice": ("particle_filter_device", "ParticleFilterDevice"),
    "BatchedParticleFilter": ("batched", "BatchedParticleFilter"),
    "wls_position": ("_gnss_gpu", "wls_position"),
    "wls_batch": ("_gnss_gpu", "wls_batch"),
    "wls_batch_device": ("_gnss_gpu", "wls_batch_device"),
    "MultiGNSSSolver": ("multi_gnss", "MultiGNSSSolver"),
    "BuildingModel": ("raytrace", "BuildingModel"),
    "BVHAccelerator": ("bvh", "BVHAccelerator"),
    "ParticleFilter3D": ("particle_filter_3d", "ParticleFilter3D"),
    "ParticleFilter3DBVH": ("particle_filter_3d", "ParticleFilter3DBVH"),
}

__all__ = list(_LAZY)


def __getattr__(name):
    if name in _LAZY:
        import importlib

        mod, attr = _LAZY[name]
        value = getattr(importlib.import_module(f"{__name__}.{mod}"), attr)
        globals()[name] = value
        return value
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
