"""Host-array particle-filter functions (the legacy ``gnss_gpu._gnss_gpu_pf`` module).

Same signatures and in-place semantics as the reference bindings
(python/gnss_gpu/_particle_filter_bindings.cpp:10-145): the caller's float64
NumPy arrays px, py, pz, pcb (and log_weights) are read and overwritten.  Each
call uploads the arrays into a cached device filter of the same size, runs the
GPU step and copies the result back, so no device memory is allocated per call.
"""

from __future__ import annotations

import numpy as np

from . import _gnss_gpu_pf_device as _dev

_CACHE: dict[int, _dev.PFDeviceState] = {}


def _state_for(n: int) -> _dev.PFDeviceState:
    st = _CACHE.get(n)
    if st is None or not st.allocated:
        st = _dev.pf_device_create(n)
        _CACHE[n] = st
    return st


def _check(n, *arrays):
    n = int(n)
    for a in arrays:
        if np.asarray(a).size < n:
            raise ValueError("array shorter than n_particles")
    return n


def _load(px, py, pz, pcb, n, seed=0):
    st = _state_for(n)
    if not getattr(st, "_legacy_init", False):
        _dev.pf_device_initialize(st, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, seed)
        st._legacy_init = True
    rows = np.zeros((n, 16), dtype=np.float64)
    rows[:, 0] = np.asarray(px, dtype=np.float64)[:n]
    rows[:, 1] = np.asarray(py, dtype=np.float64)[:n]
    rows[:, 2] = np.asarray(pz, dtype=np.float64)[:n]
    rows[:, 3] = np.asarray(pcb, dtype=np.float64)[:n]
    _dev.pf_device_set_particle_states(st, rows)
    return st


def _store(st, px, py, pz, pcb, n):
    p = _dev.pf_device_get_particles(st)
    px[:n] = p[:, 0]
    py[:n] = p[:, 1]
    pz[:n] = p[:, 2]
    pcb[:n] = p[:, 3]


def _set_lw(st, log_weights, n):
    _dev.pf_device_set_log_weights(st, np.asarray(log_weights, dtype=np.float64)[:n])


def pf_initialize(px, py, pz, pcb, init_x, init_y, init_z, init_cb, spread_pos, spread_cb,
                  n_particles, seed):
    n = _check(n_particles, px, py, pz, pcb)
    st = _state_for(n)
    _dev.pf_device_initialize(st, init_x, init_y, init_z, init_cb, spread_pos, spread_cb, seed)
    st._legacy_init = True
    _store(st, px, py, pz, pcb, n)


def pf_predict(px, py, pz, pcb, vx, vy, vz, dt, sigma_pos, sigma_cb, n_particles, seed, step):
    n = _check(n_particles, px, py, pz, pcb)
    st = _load(px, py, pz, pcb, n, seed)
    v = [float(np.asarray(x, dtype=np.float64).ravel()[0]) for x in (vx, vy, vz)]
    _dev.pf_device_predict(st, v[0], v[1], v[2], dt, sigma_pos, sigma_cb, seed, step)
    _store(st, px, py, pz, pcb, n)


def pf_weight(px, py, pz, pcb, sat_ecef, pseudoranges, weights_sat, log_weights, n_particles,
              n_sat, sigma_pr):
    n = _check(n_particles, px, py, pz, pcb, log_weights)
    st = _load(px, py, pz, pcb, n)
    _dev.pf_device_weight(st, sat_ecef, pseudoranges, weights_sat, n_sat, sigma_pr)
    out = np.empty(n, dtype=np.float64)
    _dev.pf_device_get_log_weights(st, out)
    log_weights[:n] = out


def pf_compute_ess(log_weights, n_particles):
    n = _check(n_particles, log_weights)
    st = _state_for(n)
    if not getattr(st, "_legacy_init", False):
        _dev.pf_device_initialize(st, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0)
        st._legacy_init = True
    _set_lw(st, log_weights, n)
    return _dev.pf_device_ess(st)


def pf_resample_systematic(px, py, pz, pcb, log_weights, n_particles, seed):
    n = _check(n_particles, px, py, pz, pcb, log_weights)
    st = _load(px, py, pz, pcb, n, seed)
    _set_lw(st, log_weights, n)
    _dev.pf_device_resample_systematic(st, seed)
    _store(st, px, py, pz, pcb, n)


def pf_resample_megopolis(px, py, pz, pcb, log_weights, n_particles, n_iterations, seed):
    n = _check(n_particles, px, py, pz, pcb, log_weights)
    st = _load(px, py, pz, pcb, n, seed)
    _set_lw(st, log_weights, n)
    _dev.pf_device_resample_megopolis(st, n_iterations, seed)
    _store(st, px, py, pz, pcb, n)


def pf_estimate(px, py, pz, pcb, log_weights, result, n_particles):
    n = _check(n_particles, px, py, pz, pcb, log_weights)
    st = _load(px, py, pz, pcb, n)
    _set_lw(st, log_weights, n)
    result[:4] = _dev.pf_device_estimate(st)


def pf_get_particles(px, py, pz, pcb, output, n_particles):
    n = _check(n_particles, px, py, pz, pcb)
    st = _load(px, py, pz, pcb, n)
    output.reshape(-1)[: 4 * n] = _dev.pf_device_get_particles(st).reshape(-1)
