"""Backward trajectory sampling over a particle history (FFBSi and the
resampling genealogy), the functions of python/gnss_gpu/particle_ffbsi.py with
the same arguments and results, computed on the GPU.

The caller's ``numpy.random.Generator`` is consumed exactly as the reference
consumes it: one ``random()`` per sampled step, from step T-1 down to 0, and
``integers(0, N)`` at a step whose backward weights are degenerate (a NaN or
+inf log-weight, or all -inf); nothing when step T-1 is degenerate (all indices
are then 0).  The uniforms are drawn up front and every step runs in one GPU
launch; when the launch reports a degenerate step, the generator is rewound to
its saved state, the draws are repeated with an integer at that step, and the
trajectory is run again.

``ffbsi_sample_indices_batch`` samples several trajectories (one generator each)
in one launch; trajectory k equals ``ffbsi_sample_indices(..., rngs[k])``.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, dptr, f64, iptr


def transition_logpdf(x_next, x_t, vel, dt, sigma_pos, sigma_cb):
    """log p(x_next | x_t[i]) of one predict step (particle_ffbsi.py:14-29); x_t (N, 4)."""
    xt = f64(x_t).reshape(-1, 4)
    xn = f64(x_next).ravel()[:4]
    v = f64(vel).ravel()[:3]
    out = np.empty(xt.shape[0], dtype=np.float64)
    call("gg_transition_logpdf", dptr(xn), dptr(xt), xt.shape[0], dptr(v), float(dt),
         float(sigma_pos), float(sigma_cb), dptr(out))
    return out


def _checked(log_weights, X=None, vel=None, dt=None, sigma_pos=None, ancestors=None):
    lw = f64(log_weights)
    T, N = lw.shape
    out = {"lw": lw, "T": T, "N": N}
    if X is not None:
        X = f64(X)
        if X.shape != (T, N, 4):
            raise ValueError(f"X shape {X.shape} != ({T},{N},4)")
        vel, dt, sigma_pos = f64(vel), f64(dt).ravel(), f64(sigma_pos).ravel()
        if vel.shape[0] != T or dt.shape[0] != T or sigma_pos.shape[0] != T:
            raise ValueError("vel, dt, sigma_pos must have length T")
        out.update(X=X, vel=np.ascontiguousarray(vel.reshape(T, 3)), dt=dt, sp=sigma_pos)
    if ancestors is not None:
        anc = np.asarray(ancestors)
        if anc.shape != (T, N):
            raise ValueError(f"ancestors shape {anc.shape} != ({T},{N})")
        body = anc[: T - 1]
        if body.size and (body.min() < 0 or body.max() >= N):
            raise ValueError("ancestors must lie in [0, N)")
        out["anc"] = np.ascontiguousarray(anc, dtype=np.int32)
    return out


def _launch(a, sigma_cb, u, forced):
    M, T = u.shape
    idx = np.zeros((M, T), dtype=np.int64)
    status = np.zeros(M, dtype=np.int32)
    anc = a.get("anc")
    ffbsi = anc is None
    call("gg_ffbsi_sample", dptr(a["lw"]), dptr(a["X"]) if ffbsi else None,
         dptr(a["vel"]) if ffbsi else None, dptr(a["dt"]) if ffbsi else None,
         dptr(a["sp"]) if ffbsi else None, float(sigma_cb), iptr(anc) if not ffbsi else None,
         T, a["N"], M, dptr(u), iptr(forced), C.c_void_p(idx.ctypes.data), iptr(status))
    return idx, status


def _sample(launch, T, N, rngs, last_only):
    """Draw the uniforms from each generator in the reference's order, run the
    backward walk, and redraw around degenerate steps (see the module note).
    launch(u [M, T], forced [M, T]) -> (indices [M, T], status [M])."""
    M = len(rngs)
    saved = [r.bit_generator.state for r in rngs]
    degenerate = [set() for _ in range(M)]
    u = np.zeros((M, T))
    forced = np.full((M, T), -1, dtype=np.int32)

    def draw(k):
        rngs[k].bit_generator.state = saved[k]
        forced[k] = -1
        for t in range(T - 1, -1, -1):
            if t in degenerate[k]:
                forced[k, t] = int(rngs[k].integers(0, N))
            else:
                u[k, t] = rngs[k].random()
            if last_only:
                break

    for k in range(M):
        draw(k)
    result = np.zeros((M, T), dtype=np.int64)
    todo = list(range(M))
    while todo:
        idx, status = launch(np.ascontiguousarray(u[todo]), np.ascontiguousarray(forced[todo]))
        again = []
        for j, k in enumerate(todo):
            st = int(status[j])
            if st == -1:
                result[k] = idx[j]
            elif st == -2:  # degenerate last step: zeros, generator untouched
                rngs[k].bit_generator.state = saved[k]
            elif st == -3:
                raise ValueError("a forced draw or genealogy ancestor lies outside [0, N)")
            else:
                degenerate[k].add(st)
                draw(k)
                again.append(k)
        todo = again
    return result


def _host(a, sigma_cb, rngs, last_only):
    return _sample(lambda u, f: _launch(a, sigma_cb, u, f), a["T"], a["N"], rngs, last_only)


def ffbsi_sample_indices(log_weights, X, vel, dt, sigma_pos, sigma_cb, rng):
    """One FFBSi backward trajectory as ancestor indices I[0..T-1] (int64)."""
    a = _checked(log_weights, X, vel, dt, sigma_pos)
    return _host(a, sigma_cb, [rng], last_only=False)[0]


def ffbsi_sample_indices_batch(log_weights, X, vel, dt, sigma_pos, sigma_cb, rngs):
    """len(rngs) trajectories in one launch: int64 [M, T]."""
    a = _checked(log_weights, X, vel, dt, sigma_pos)
    return _host(a, sigma_cb, list(rngs), last_only=False)


def ffbsi_smooth_sample(log_weights, X, vel, dt, sigma_pos, sigma_cb, rng):
    """One smoothed path (T, 3) from FFBSi indices."""
    idx = ffbsi_sample_indices(log_weights, X, vel, dt, sigma_pos, sigma_cb, rng)
    return np.asarray(np.asarray(X)[np.arange(len(idx)), idx, :3], dtype=np.float64)


def genealogy_smooth_indices(log_weights, ancestors, rng):
    """Backward trace along the resampling genealogy (ancestors[t, j]: source of
    slot j after the resample at the end of epoch t; row T-1 is ignored)."""
    a = _checked(log_weights, ancestors=ancestors)
    return _host(a, 1.0, [rng], last_only=True)[0]  # sigma_cb: unused by the trace


def genealogy_smooth_sample(log_weights, X, ancestors, rng):
    """One smoothed path (T, 3) via the genealogy trace."""
    idx = genealogy_smooth_indices(log_weights, ancestors, rng)
    return np.asarray(np.asarray(X, dtype=np.float64)[np.arange(len(idx)), idx, :3],
                      dtype=np.float64)


class ParticleHistory:
    """Device-resident particle history of a ``ParticleFilterDevice`` for FFBSi
    (no reference counterpart: the reference's smoother experiments copy
    get_particles() / get_log_weights() to host lists every epoch,
    experiments/exp_ffbsi_eval.py:189-190, and pass the stacked arrays).

    ``capacity`` epochs of N particles are held in HBM as binary64 ECEF rows
    [T, N, 4] and log-weights [T, N] (40 B per particle and epoch: 1M particles
    x 1000 epochs is 40 GB).  ``record`` stores the filter's current particles
    and log-weights (after the update, before the resample, as the reference
    experiment does) with the predict parameters of that epoch;
    ``ffbsi_sample_indices`` samples trajectories on the stored history and
    returns what ``ffbsi_sample_indices_batch`` returns for the equivalent host
    arrays and generators.
    """

    def __init__(self, pf, capacity):
        import torch

        self.pf = pf
        self.N = int(pf.n_particles)
        self.capacity = int(capacity)
        dev = torch.device("cuda")
        self.X = torch.empty((self.capacity, self.N, 4), dtype=torch.float64, device=dev)
        self.lw = torch.empty((self.capacity, self.N), dtype=torch.float64, device=dev)
        self.vel = np.zeros((self.capacity, 3))
        self.dt = np.zeros(self.capacity)
        self.sigma_pos = np.zeros(self.capacity)
        self.T = 0

    def record(self, velocity=None, dt=1.0, sigma_pos=None):
        """Store the current epoch; velocity / dt / sigma_pos are the predict
        parameters that led into it (row t drives the t-1 -> t transition)."""
        if self.T >= self.capacity:
            raise RuntimeError("ParticleHistory is full")
        t = self.T
        call("gg_pf_device_record", self.pf._state.handle, C.c_void_p(self.X[t].data_ptr()),
             C.c_void_p(self.lw[t].data_ptr()))
        self.vel[t] = (np.zeros(3) if velocity is None
                       else np.asarray(velocity, dtype=np.float64).ravel()[:3])
        self.dt[t] = float(dt)
        self.sigma_pos[t] = float(self.pf.sigma_pos if sigma_pos is None else sigma_pos)
        self.T += 1

    def ffbsi_sample_indices(self, sigma_cb, rngs):
        """len(rngs) FFBSi trajectories over the recorded epochs: int64 [M, T]."""
        import torch

        T, N = self.T, self.N
        if T == 0:
            raise RuntimeError("ParticleHistory is empty")
        self.pf.sync()
        dev = self.X.device
        vel = torch.from_numpy(np.ascontiguousarray(self.vel[:T])).to(dev)
        dt = torch.from_numpy(np.ascontiguousarray(self.dt[:T])).to(dev)
        sp = torch.from_numpy(np.ascontiguousarray(self.sigma_pos[:T])).to(dev)

        def launch(u, forced):
            M = u.shape[0]
            tu = torch.from_numpy(u).to(dev)
            tf = torch.from_numpy(forced).to(dev)
            ti = torch.zeros((M, T), dtype=torch.int64, device=dev)
            ts = torch.zeros(M, dtype=torch.int32, device=dev)
            stream = torch.cuda.current_stream(dev).cuda_stream
            call("gg_ffbsi_sample_device", C.c_void_p(self.lw.data_ptr()),
                 C.c_void_p(self.X.data_ptr()), C.c_void_p(vel.data_ptr()),
                 C.c_void_p(dt.data_ptr()), C.c_void_p(sp.data_ptr()), float(sigma_cb), None,
                 T, N, M, C.c_void_p(tu.data_ptr()), C.c_void_p(tf.data_ptr()),
                 C.c_void_p(ti.data_ptr()), C.c_void_p(ts.data_ptr()), C.c_void_p(stream))
            return ti.cpu().numpy(), ts.cpu().numpy()

        return _sample(launch, T, N, list(rngs), last_only=False)


__all__ = ["ParticleHistory", "transition_logpdf", "ffbsi_sample_indices",
           "ffbsi_sample_indices_batch", "ffbsi_smooth_sample", "genealogy_smooth_indices",
           "genealogy_smooth_sample"]
_ = _lib
