"""Device-resident particle filter: the function module behind ParticleFilterDevice.

Same names, positional order and defaults as the reference's pybind11 module
``gnss_gpu._gnss_gpu_pf_device`` (python/gnss_gpu/_pf_device_bindings.cpp:24-398),
implemented over the C ABI of libgnss_gpu_amd.so.  Particle state lives on the
GPU as binary32 offsets in a local ENU frame around a binary64 anchor; every
getter returns binary64 absolute (ECEF) values like the reference.

The double-differenced pseudorange, carrier-phase AFV and Doppler updates are
here; the GMM update raises NotImplementedError.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, dptr, f64, i32, iptr


class PFDeviceState:
    """Opaque handle to device-resident particle state (PFDeviceState,
    include/gnss_gpu/pf_device.h:7-60)."""

    def __init__(self, n_particles: int, n_isb: int = 0):
        n_particles = int(n_particles)
        if n_particles <= 0:
            raise ValueError("n_particles must be positive")
        h = C.c_void_p()
        call("gg_pf_device_create", int(n_particles), int(n_isb), C.byref(h))
        self._h = h
        self.n_particles = n_particles
        self.n_isb = int(n_isb)
        self.allocated = True

    @property
    def handle(self) -> C.c_void_p:
        if not self.allocated:
            raise RuntimeError("PFDeviceState has been destroyed")
        return self._h

    def destroy(self) -> None:
        if self.allocated:
            _lib.lib.gg_pf_destroy(self._h)
            self.allocated = False

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def pf_device_create(n_particles, n_isb=0):
    return PFDeviceState(n_particles, n_isb)


def pf_device_destroy(state):
    if state is not None:
        state.destroy()


def pf_device_initialize(state, init_x, init_y, init_z, init_cb, spread_pos, spread_cb, seed,
                         init_vx=0.0, init_vy=0.0, init_vz=0.0, spread_vel=0.0,
                         init_vel_sigma=0.0, init_isb=None, spread_isb=0.0):
    isb = None
    if state.n_isb:
        isb = f64(np.zeros(state.n_isb) if init_isb is None else init_isb).ravel()
        if isb.size != state.n_isb:
            raise ValueError(f"init_isb must have {state.n_isb} values")
    call("gg_pf_device_initialize", state.handle, float(init_x), float(init_y), float(init_z),
         float(init_cb), float(spread_pos), float(spread_cb), int(seed) & 0xFFFFFFFFFFFFFFFF,
         float(init_vx), float(init_vy), float(init_vz), float(spread_vel),
         float(init_vel_sigma), dptr(isb) if isb is not None else None, float(spread_isb))


def pf_device_predict(state, vx, vy, vz, dt, sigma_pos, sigma_cb, seed, step, sigma_vel=0.0,
                      velocity_guide_alpha=1.0, velocity_kf=False, velocity_process_noise=0.0,
                      sigma_isb=0.0):
    call("gg_pf_device_predict", state.handle, float(vx), float(vy), float(vz), float(dt),
         float(sigma_pos), float(sigma_cb), int(seed) & 0xFFFFFFFFFFFFFFFF, int(step),
         float(sigma_vel), float(velocity_guide_alpha), 1 if velocity_kf else 0,
         float(velocity_process_noise), float(sigma_isb))


def pf_device_weight(state, sat_ecef, pseudoranges, weights_sat, n_sat, sigma_pr, nu=0.0,
                     per_particle_nlos_threshold_m=0.0, per_particle_huber=False,
                     per_particle_huber_k=1.5, system_ids=None):
    n_sat = int(n_sat)
    sat = f64(sat_ecef).ravel()
    pr = f64(pseudoranges).ravel()
    w = f64(weights_sat).ravel()
    if sat.size < 3 * n_sat or pr.size < n_sat or w.size < n_sat:
        raise ValueError("satellite arrays are shorter than n_sat")
    sys = None if system_ids is None else i32(system_ids).ravel()
    if sys is not None and sys.size < n_sat:
        raise ValueError("system_ids is shorter than n_sat")
    call("gg_pf_device_weight", state.handle, dptr(sat), dptr(pr), dptr(w),
         iptr(sys) if sys is not None else None, n_sat, float(sigma_pr), float(nu),
         float(per_particle_nlos_threshold_m), 1 if per_particle_huber else 0,
         float(per_particle_huber_k))


def _dd_arrays(n, sat_k, ref, obs, brk, brr, w, wl=None):
    n = int(n)
    arrs = [f64(sat_k).ravel(), f64(ref).ravel(), f64(obs).ravel(), f64(brk).ravel(),
            f64(brr).ravel(), f64(w).ravel()] + ([] if wl is None else [f64(wl).ravel()])
    need = [3 * n, 3 * n, n, n, n, n] + ([] if wl is None else [n])
    if any(a.size < k for a, k in zip(arrs, need)):
        raise ValueError("double-difference arrays are shorter than n_dd")
    return n, arrs


def pf_device_weight_dd_pseudorange(state, sat_ecef_k, ref_ecef, dd_pseudorange, base_range_k,
                                    base_range_ref, weights_dd, n_dd, sigma_pr,
                                    per_particle_nlos_threshold_m=0.0, per_particle_huber=False,
                                    per_particle_huber_k=1.5):
    """DD pseudorange weight update; overwrites the log-weights
    (_pf_device_bindings.cpp:111-143, pf_device.cu:1576-1620)."""
    n, a = _dd_arrays(n_dd, sat_ecef_k, ref_ecef, dd_pseudorange, base_range_k, base_range_ref,
                      weights_dd)
    call("gg_pf_device_weight_dd_pseudorange", state.handle, *[dptr(x) for x in a], n,
         float(sigma_pr), float(per_particle_nlos_threshold_m), 1 if per_particle_huber else 0,
         float(per_particle_huber_k))


def pf_device_weight_carrier_afv(state, sat_ecef, carrier_phase, weights_sat, n_sat,
                                 wavelength=0.190293673, sigma_cycles=0.05):
    """Undifferenced carrier-phase AFV update; adds to the log-weights
    (_pf_device_bindings.cpp:167-183, pf_device.cu:2062-2116)."""
    n = int(n_sat)
    sat = f64(sat_ecef).ravel()
    cp = f64(carrier_phase).ravel()
    w = f64(weights_sat).ravel()
    if sat.size < 3 * n or cp.size < n or w.size < n:
        raise ValueError("satellite arrays are shorter than n_sat")
    call("gg_pf_device_weight_carrier_afv", state.handle, dptr(sat), dptr(cp), dptr(w), n,
         float(wavelength), float(sigma_cycles))


def pf_device_weight_3d_bvh(state, bvh, sat_ecef, pseudoranges, weights_sat, n_sat,
                            sigma_pr_los=3.0, sigma_pr_nlos=30.0, nlos_bias=20.0,
                            blocked_nlos_prob=1.0, clear_nlos_prob=0.0, system_ids=None):
    """Ray-traced NLOS weight on the device particles (no reference binding: the
    device-filter counterpart of pf_weight_3d_bvh, weight_3d_bvh.cu:127-220);
    overwrites the log-weights.  ``bvh`` is a ``BVHAccelerator``."""
    n = int(n_sat)
    sat = f64(sat_ecef).ravel()
    pr = f64(pseudoranges).ravel()
    w = f64(weights_sat).ravel()
    if sat.size < 3 * n or pr.size < n or w.size < n:
        raise ValueError("satellite arrays are shorter than n_sat")
    sys = None
    if system_ids is not None:
        sys = np.ascontiguousarray(np.asarray(system_ids, dtype=np.int32).ravel())
        if sys.size < n:
            raise ValueError("system_ids shorter than n_sat")
    call("gg_pf_device_weight_3d_bvh", state.handle, bvh.device_handle(), dptr(sat), dptr(pr),
         dptr(w), iptr(sys) if sys is not None else None, n, float(sigma_pr_los),
         float(sigma_pr_nlos), float(nlos_bias), float(blocked_nlos_prob), float(clear_nlos_prob))


def pf_device_weight_dd_carrier_afv(state, sat_ecef_k, ref_ecef, dd_carrier, base_range_k,
                                    base_range_ref, weights_dd, wavelengths_m, n_dd,
                                    sigma_cycles=0.05, per_particle_nlos_threshold_cycles=0.0,
                                    per_particle_huber=False, per_particle_huber_k=1.5):
    """DD carrier-phase AFV update; adds to the log-weights
    (_pf_device_bindings.cpp:185-220, pf_device.cu:2122-2170)."""
    n, a = _dd_arrays(n_dd, sat_ecef_k, ref_ecef, dd_carrier, base_range_k, base_range_ref,
                      weights_dd, wavelengths_m)
    call("gg_pf_device_weight_dd_carrier_afv", state.handle, *[dptr(x) for x in a], n,
         float(sigma_cycles), float(per_particle_nlos_threshold_cycles),
         1 if per_particle_huber else 0, float(per_particle_huber_k))


def _doppler_arrays(sat_ecef, sat_vel, doppler_hz, weights_sat, n_sat):
    n = int(n_sat)
    sat, vel = f64(sat_ecef).ravel(), f64(sat_vel).ravel()
    dop, w = f64(doppler_hz).ravel(), f64(weights_sat).ravel()
    if sat.size < 3 * n or vel.size < 3 * n or dop.size < n or w.size < n:
        raise ValueError("Doppler arrays are shorter than n_sat")
    return n, sat, vel, dop, w


def pf_device_weight_doppler(state, sat_ecef, sat_vel, doppler_hz, weights_sat, n_sat,
                             wavelength_m=0.19029367279836488, sigma_mps=0.5,
                             velocity_update_gain=0.25, max_velocity_update_mps=10.0):
    """Doppler likelihood and per-particle velocity nudge
    (_pf_device_bindings.cpp:225-253, pf_device.cu:2183-2226)."""
    n, sat, vel, dop, w = _doppler_arrays(sat_ecef, sat_vel, doppler_hz, weights_sat, n_sat)
    call("gg_pf_device_weight_doppler", state.handle, dptr(sat), dptr(vel), dptr(dop), dptr(w),
         n, float(wavelength_m), float(sigma_mps), float(velocity_update_gain),
         float(max_velocity_update_mps))


def pf_device_doppler_kf_update(state, sat_ecef, sat_vel, doppler_hz, weights_sat, n_sat,
                                wavelength_m=0.19029367279836488, sigma_mps=0.5):
    """Per-particle velocity Kalman update from Doppler rows (RBPF)
    (_pf_device_bindings.cpp:255-278, pf_device.cu:2228-2270)."""
    n, sat, vel, dop, w = _doppler_arrays(sat_ecef, sat_vel, doppler_hz, weights_sat, n_sat)
    call("gg_pf_device_doppler_kf_update", state.handle, dptr(sat), dptr(vel), dptr(dop),
         dptr(w), n, float(wavelength_m), float(sigma_mps))


def pf_device_weight_learned(state, sat_ecef, pseudoranges, weights_sat, n_sat, cn0, los_flag,
                             system_ids=None, table=False):
    """Weight update with the attached learned likelihood (no reference
    counterpart; replaces the Student-t term of pf_device.cu:262-266)."""
    n_sat = int(n_sat)
    sat = f64(sat_ecef).ravel()
    pr = f64(pseudoranges).ravel()
    w = f64(weights_sat).ravel()
    c = np.ascontiguousarray(np.asarray(cn0, dtype=np.float32).ravel())
    fl = np.ascontiguousarray(np.asarray(los_flag, dtype=np.float32).ravel())
    if sat.size < 3 * n_sat or pr.size < n_sat or w.size < n_sat or c.size < n_sat or fl.size < n_sat:
        raise ValueError("satellite arrays are shorter than n_sat")
    sys = None if system_ids is None else i32(system_ids).ravel()
    F = C.POINTER(C.c_float)
    call("gg_pf_device_weight_learned", state.handle, dptr(sat), dptr(pr), dptr(w),
         iptr(sys) if sys is not None else None, n_sat, c.ctypes.data_as(F),
         fl.ctypes.data_as(F), 1 if table else 0)


def pf_device_set_mlp(state, model):
    call("gg_pf_set_mlp", state.handle, None if model is None else model.handle)


def pf_device_ess(state):
    out = C.c_double()
    call("gg_pf_device_ess", state.handle, C.byref(out))
    return float(out.value)


def pf_device_position_spread(state, center_x, center_y, center_z):
    out = C.c_double()
    call("gg_pf_device_position_spread", state.handle, float(center_x), float(center_y),
         float(center_z), C.byref(out))
    return float(out.value)


def pf_device_resample_systematic(state, seed, u0_mode=0):
    call("gg_pf_device_resample_systematic", state.handle, int(seed) & 0xFFFFFFFFFFFFFFFF,
         int(u0_mode))


def pf_device_resample_megopolis(state, n_iterations, seed):
    call("gg_pf_device_resample_megopolis", state.handle, int(n_iterations),
         int(seed) & 0xFFFFFFFFFFFFFFFF)


def pf_device_resample_megopolis_chesser(state, n_iterations):
    call("gg_pf_device_resample_megopolis_chesser", state.handle, int(n_iterations))


def pf_device_estimate(state):
    out = np.empty(4, dtype=np.float64)
    call("gg_pf_device_estimate", state.handle, dptr(out))
    return out


def pf_device_estimate_full(state):
    out = np.empty(4 + state.n_isb, dtype=np.float64)
    call("gg_pf_device_estimate_full", state.handle, dptr(out))
    return out


def pf_device_get_particles(state):
    out = np.empty((state.n_particles, 4), dtype=np.float64)
    call("gg_pf_device_get_particles", state.handle, dptr(out))
    return out


def pf_device_get_particles_full(state):
    out = np.empty((state.n_particles, 4 + state.n_isb), dtype=np.float64)
    call("gg_pf_device_get_particles_full", state.handle, dptr(out))
    return out


def pf_device_get_particle_states(state):
    out = np.empty((state.n_particles, 16), dtype=np.float64)
    call("gg_pf_device_get_particle_states", state.handle, dptr(out))
    return out


def pf_device_set_particle_states(state, input):
    arr = f64(input)
    if arr.ndim != 2 or arr.shape != (state.n_particles, 16):
        raise RuntimeError("pf_device_set_particle_states: input shape must be (n_particles, 16)")
    call("gg_pf_device_set_particle_states", state.handle, dptr(arr))


def pf_device_get_log_weights(state, out):
    if out.size != state.n_particles:
        raise RuntimeError("pf_device_get_log_weights: output size must match n_particles")
    tmp = np.empty(state.n_particles, dtype=np.float64)
    call("gg_pf_device_get_log_weights", state.handle, dptr(tmp))
    out.reshape(-1)[:] = tmp


def pf_device_set_log_weights(state, input):
    arr = f64(input).reshape(-1)
    if arr.size != state.n_particles:
        raise RuntimeError("pf_device_set_log_weights: input size must match n_particles")
    call("gg_pf_device_set_log_weights", state.handle, dptr(arr))


def pf_device_get_resample_ancestors(state, out):
    if out.size != state.n_particles:
        raise RuntimeError(
            "pf_device_get_resample_ancestors: output size must match n_particles")
    tmp = np.empty(state.n_particles, dtype=np.int32)
    call("gg_pf_device_get_resample_ancestors", state.handle, iptr(tmp))
    out.reshape(-1)[:] = tmp


def pf_device_shift_clock_bias(state, shift):
    call("gg_pf_device_shift_clock_bias", state.handle, float(shift))


def pf_device_shift_position(state, dx, dy, dz):
    call("gg_pf_device_shift_position", state.handle, float(dx), float(dy), float(dz))


def pf_device_position_update(state, ref_x, ref_y, ref_z, sigma_pos):
    call("gg_pf_device_position_update", state.handle, float(ref_x), float(ref_y), float(ref_z),
         float(sigma_pos))


def pf_device_sync(state):
    if state is not None and state.allocated:
        call("gg_pf_device_sync", state.handle)


def stream_handle(stream):
    """hipStream_t of ``stream``: a torch.cuda.Stream, a raw handle (int), or
    None (the filter's own stream)."""
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return C.c_void_p(int(stream.cuda_stream))
    return C.c_void_p(int(stream))


def pf_device_set_stream(state, stream=None):
    """Enqueue this filter's work on ``stream`` from now on (after finishing
    what is queued on its current stream); None returns to a stream of its own.
    Calls that hand results back to the host synchronise that stream."""
    call("gg_pf_set_stream", state.handle, stream_handle(stream))


def pf_device_weight_gmm(state, sat_ecef, pseudoranges, weights_sat, n_sat, sigma_pr,
                         w_los=0.7, mu_nlos=15.0, sigma_nlos=30.0):
    """LOS / NLOS Gaussian-mixture weight update (_pf_device_bindings.cpp:147-165,
    pf_device.cu:373-437, 1622-1680): overwrites the log-weights."""
    n_sat = int(n_sat)
    sat = f64(sat_ecef).ravel()
    pr = f64(pseudoranges).ravel()
    w = f64(weights_sat).ravel()
    if sat.size < 3 * n_sat or pr.size < n_sat or w.size < n_sat:
        raise ValueError("satellite arrays are shorter than n_sat")
    call("gg_pf_device_weight_gmm", state.handle, dptr(sat), dptr(pr), dptr(w), n_sat,
         float(sigma_pr), float(w_los), float(mu_nlos), float(sigma_nlos))
