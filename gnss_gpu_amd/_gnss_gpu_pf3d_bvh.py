"""Ray-traced NLOS pseudorange weights (the ``gnss_gpu._gnss_gpu_pf3d_bvh`` and
``gnss_gpu._gnss_gpu_pf3d`` modules, python/gnss_gpu/_pf3d_bvh_bindings.cpp:18-92,
_pf3d_bindings.cpp:10-52).

Same arguments as the reference: float64 host arrays px, py, pz, pcb [N];
``log_weights`` [N] (float64, C-contiguous) is overwritten in place.
"""

from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import call, dptr, f64


def _out(log_weights, n):
    if (not isinstance(log_weights, np.ndarray) or log_weights.dtype != np.float64
            or not log_weights.flags.c_contiguous or log_weights.size < n):
        raise ValueError("log_weights must be a C-contiguous float64 array of n_particles")
    return log_weights


def _inputs(px, py, pz, pcb, sat_ecef, pseudoranges, weights_sat, n_particles, n_sat):
    n, ns = int(n_particles), int(n_sat)
    parts = [f64(a).ravel() for a in (px, py, pz, pcb)]
    if any(a.size < n for a in parts):
        raise ValueError("particle arrays shorter than n_particles")
    sat = f64(sat_ecef).ravel()
    pr = f64(pseudoranges).ravel()
    ws = f64(weights_sat).ravel()
    if sat.size < 3 * ns or pr.size < ns or ws.size < ns:
        raise ValueError("satellite arrays shorter than n_sat")
    return n, ns, parts, sat, pr, ws


def pf_weight_3d_bvh(px, py, pz, pcb, sat_ecef, pseudoranges, weights_sat, nodes_flat,
                     sorted_tris, log_weights, n_particles, n_sat, sigma_pr_los, sigma_pr_nlos,
                     nlos_bias, blocked_nlos_prob=1.0, clear_nlos_prob=0.0):
    n, ns, parts, sat, pr, ws = _inputs(px, py, pz, pcb, sat_ecef, pseudoranges, weights_sat,
                                        n_particles, n_sat)
    lw = _out(log_weights, n)
    nodes = f64(nodes_flat).reshape(-1, 10)
    tris = f64(sorted_tris).reshape(-1, 3, 3)
    call("gg_pf_weight_3d_bvh", *[dptr(a) for a in parts], dptr(sat), dptr(pr), dptr(ws),
         dptr(nodes), nodes.shape[0], dptr(tris), tris.shape[0], dptr(lw), n, ns,
         float(sigma_pr_los), float(sigma_pr_nlos), float(nlos_bias), float(blocked_nlos_prob),
         float(clear_nlos_prob))


def pf_weight_3d(px, py, pz, pcb, sat_ecef, pseudoranges, weights_sat, triangles, log_weights,
                 n_particles, n_sat, sigma_pr_los, sigma_pr_nlos, nlos_bias,
                 blocked_nlos_prob=1.0, clear_nlos_prob=0.0):
    n, ns, parts, sat, pr, ws = _inputs(px, py, pz, pcb, sat_ecef, pseudoranges, weights_sat,
                                        n_particles, n_sat)
    lw = _out(log_weights, n)
    tris = f64(triangles).reshape(-1, 3, 3)
    call("gg_pf_weight_3d", *[dptr(a) for a in parts], dptr(sat), dptr(pr), dptr(ws),
         dptr(tris), tris.shape[0], dptr(lw), n, ns, float(sigma_pr_los), float(sigma_pr_nlos),
         float(nlos_bias), float(blocked_nlos_prob), float(clear_nlos_prob))


__all__ = ["pf_weight_3d_bvh", "pf_weight_3d"]
_ = _lib
