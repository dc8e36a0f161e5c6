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

# Reference module names whose contents live in another module here:
# gnss_gpu.particle_filter_3d_bvh (ParticleFilter3DBVH) and gnss_gpu._gnss_gpu_pf3d
# (pf_weight_3d, python/gnss_gpu/_pf3d_bindings.cpp:10-52).  Importing
# gnss_gpu_amd.<alias> returns the target module itself.
_MODULE_ALIASES = {
    "particle_filter_3d_bvh": "particle_filter_3d",
    "_gnss_gpu_pf3d": "_gnss_gpu_pf3d_bvh",
}


class _AliasFinder:
    """Meta-path finder that resolves the aliased module names."""

    @staticmethod
    def find_spec(fullname, path=None, target=None):
        prefix, _, leaf = fullname.rpartition(".")
        if prefix != __name__ or leaf not in _MODULE_ALIASES:
            return None
        import importlib.util

        return importlib.util.spec_from_loader(fullname, _AliasFinder)

    @staticmethod
    def create_module(spec):
        import importlib

        return importlib.import_module(f"{__name__}.{_MODULE_ALIASES[spec.name.rpartition('.')[2]]}")

    @staticmethod
    def exec_module(module):
        return None


def _install_aliases():
    import sys

    if not any(f is _AliasFinder for f in sys.meta_path):
        sys.meta_path.append(_AliasFinder)


_install_aliases()


def __getattr__(name):
    if name in _LAZY:
        import importlib

        mod, attr = _LAZY[name]
        value = getattr(importlib.import_module(f"{__name__}.{mod}"), attr)
        globals()[name] = value
        return value
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
