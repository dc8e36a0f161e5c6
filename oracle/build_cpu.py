"""Build the CPU baseline library liboracle_cpu_pf.so (gcc, OpenMP): the C
restatements of the particle-filter step (cpu_pf.c), of the observation updates
(cpu_updates.c), of the WLS solvers, single-clock and multi-GNSS (cpu_wls.c), and
of the ray-traced NLOS weight over a BVH (cpu_raytrace.c).

Measurement infrastructure only; the output lands in oracle/_build/.
"""

from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "liboracle_cpu_pf.so")


def build() -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = [os.path.join(HERE, "cpu_pf.c"), os.path.join(HERE, "cpu_wls.c"),
            os.path.join(HERE, "cpu_updates.c"), os.path.join(HERE, "cpu_raytrace.c")]
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c11",
           "-D_GNU_SOURCE", *srcs, "-o", tmp, "-lm"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    os.replace(tmp, LIB)
    return LIB


def load():
    import ctypes as C

    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    D, I, U64, P = C.c_double, C.c_int, C.c_uint64, C.c_void_p
    DP = C.POINTER(C.c_double)
    IP = C.POINTER(C.c_int)
    lib.cpu_pf_create.restype = P
    lib.cpu_pf_create.argtypes = [I, I, U64, I]
    lib.cpu_pf_destroy.argtypes = [P]
    lib.cpu_pf_threads.restype = I
    lib.cpu_pf_initialize.argtypes = [P, DP, D, D, D]
    lib.cpu_pf_step.restype = I
    lib.cpu_pf_step.argtypes = [P, DP, D, D, D, D, I, DP, DP, IP, I, D, D, D, DP]
    lib.cpu_wls_threads.restype = I
    lib.cpu_wls_batch.restype = None
    lib.cpu_wls_batch.argtypes = [DP, DP, DP, IP, I, I, I, D, I, I, DP, IP]
    lib.cpu_wls_mg_batch.restype = None
    lib.cpu_wls_mg_batch.argtypes = [DP, DP, DP, IP, I, I, I, I, D, I, I, I, DP, IP]
    lib.cpu_updates_threads.restype = I
    lib.cpu_raytrace_threads.restype = I
    lib.cpu_weight_3d_bvh.restype = I
    lib.cpu_weight_3d_bvh.argtypes = [DP, DP, DP, DP, I, DP, DP, DP, I, DP, I, DP, I, D, D, D,
                                      D, D, DP, I]
    for name, args in {
        "cpu_dd_pr_loglik": [I, DP, I, DP, DP, DP, DP, DP, DP, D, D, I, D, DP, I],
        "cpu_carrier_afv_loglik": [I, DP, DP, I, DP, DP, DP, D, D, DP, I],
        "cpu_dd_carrier_afv_loglik": [I, DP, I, DP, DP, DP, DP, DP, DP, DP, D, D, I, D, DP, I],
        "cpu_doppler_update": [I, DP, DP, I, DP, DP, DP, DP, D, D, D, D, DP, I],
        "cpu_doppler_kf_update": [I, DP, DP, DP, I, DP, DP, DP, DP, D, D, DP, I],
    }.items():
        fn = getattr(lib, name)
        fn.restype = None
        fn.argtypes = args
    return lib


if __name__ == "__main__":
    print(build())
