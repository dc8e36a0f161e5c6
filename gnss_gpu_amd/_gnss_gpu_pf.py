"""Host-array particle-filter functions (the legacy ``gnss_gpu._gnss_gpu_pf`` module).

Same signatures and in-place semantics as the reference bindings
(python/gnss_gpu/_particle_filter_bindings.cpp:10-145): the caller's float64
NumPy arrays px, py, pz, pcb (and log_weights) are read and overwritten.

Each call moves the four arrays into a device filter of the same size, kept
per size for the process, runs the step on the GPU and writes the result back
(gg_pf_device_set_positions / gg_pf_device_get_positions: four binary64
arrays each way, no repacking on the host), so no device memory is allocated
per call.  Between calls the state lives only in the caller's arrays, as in
the reference.
"""

from __future__ import annotations

import numpy as np

from . import _gnss_gpu_pf_device as _dev
from ._lib import call, dptr, f64

_FILTERS: dict[int, _dev.PFDeviceState] = {}


def _filter(n: int, seed: int = 0) -> _dev.PFDeviceState:
    """The cached device filter for n particles (created and given a frame on
    first use)."""
    st = _FILTERS.get(n)
    if st is None or not st.allocated:
        st = _dev.pf_device_create(n)
        _dev.pf_device_initialize(st, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, seed)
        _FILTERS[n] = st
    return st


def _count(n, *arrays):
    n = int(n)
    if n <= 0:
        raise ValueError("n_particles must be positive")
    for a in arrays:
        if np.asarray(a).size < n:
            raise ValueError("array shorter than n_particles")
    return n


def _column(a, n):
    a = np.asarray(a)
    return f64(a.reshape(-1)[:n])


def _upload(px, py, pz, pcb, n, seed=0, log_weights=None):
    st = _filter(n, seed)
    call("gg_pf_device_set_positions", st.handle,
         *(dptr(_column(a, n)) for a in (px, py, pz, pcb)))
    if log_weights is not None:
        _dev.pf_device_set_log_weights(st, _column(log_weights, n))
    return st


def _download(st, px, py, pz, pcb, n):
    outs = [np.empty(n, dtype=np.float64) for _ in range(4)]
    call("gg_pf_device_get_positions", st.handle, *(dptr(o) for o in outs))
    for dst, src in zip((px, py, pz, pcb), outs):
        dst.reshape(-1)[:n] = src


def pf_initialize(px, py, pz, pcb, init_x, init_y, init_z, init_cb, spread_pos, spread_cb,
                  n_particles, seed):
    n = _count(n_particles, px, py, pz, pcb)
    st = _filter(n, seed)
    _dev.pf_device_initialize(st, init_x, init_y, init_z, init_cb, spread_pos, spread_cb, seed)
    _download(st, px, py, pz, pcb, n)


def pf_predict(px, py, pz, pcb, vx, vy, vz, dt, sigma_pos, sigma_cb, n_particles, seed, step):
    n = _count(n_particles, px, py, pz, pcb)
    st = _upload(px, py, pz, pcb, n, seed)
    v = [float(np.asarray(c, dtype=np.float64).reshape(-1)[0]) for c in (vx, vy, vz)]
    _dev.pf_device_predict(st, v[0], v[1], v[2], dt, sigma_pos, sigma_cb, seed, step)
    _download(st, px, py, pz, pcb, n)


def pf_weight(px, py, pz, pcb, sat_ecef, pseudoranges, weights_sat, log_weights, n_particles,
              n_sat, sigma_pr):
    n = _count(n_particles, px, py, pz, pcb, log_weights)
    st = _upload(px, py, pz, pcb, n)
    _dev.pf_device_weight(st, sat_ecef, pseudoranges, weights_sat, n_sat, sigma_pr)
    out = np.empty(n, dtype=np.float64)
    _dev.pf_device_get_log_weights(st, out)
    log_weights.reshape(-1)[:n] = out


def pf_compute_ess(log_weights, n_particles):
    n = _count(n_particles, log_weights)
    st = _filter(n)
    _dev.pf_device_set_log_weights(st, _column(log_weights, n))
    return _dev.pf_device_ess(st)


def pf_resample_systematic(px, py, pz, pcb, log_weights, n_particles, seed):
    n = _count(n_particles, px, py, pz, pcb, log_weights)
    st = _upload(px, py, pz, pcb, n, seed, log_weights)
    _dev.pf_device_resample_systematic(st, seed)
    _download(st, px, py, pz, pcb, n)


def pf_resample_megopolis(px, py, pz, pcb, log_weights, n_particles, n_iterations, seed):
    n = _count(n_particles, px, py, pz, pcb, log_weights)
    st = _upload(px, py, pz, pcb, n, seed, log_weights)
    _dev.pf_device_resample_megopolis(st, n_iterations, seed)
    _download(st, px, py, pz, pcb, n)


def pf_estimate(px, py, pz, pcb, log_weights, result, n_particles):
    n = _count(n_particles, px, py, pz, pcb, log_weights)
    st = _upload(px, py, pz, pcb, n, log_weights=log_weights)
    result.reshape(-1)[:4] = _dev.pf_device_estimate(st)


def pf_get_particles(px, py, pz, pcb, output, n_particles):
    """output[4 i + k] = (px, py, pz, pcb)[k][i]: the reference's kernel only
    interleaves the caller's arrays (src/particle_filter/weight.cu:331-368), so
    this is a host copy of the same values, exact to the bit."""
    n = _count(n_particles, px, py, pz, pcb)
    rows = np.stack([_column(a, n) for a in (px, py, pz, pcb)], axis=1)
    output.reshape(-1)[: 4 * n] = rows.reshape(-1)
