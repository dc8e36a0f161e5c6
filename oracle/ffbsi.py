"""Restatement of the reference's backward trajectory sampling
(python/gnss_gpu/particle_ffbsi.py) in NumPy (TEST INFRASTRUCTURE ONLY).

The draws are written out: numpy's Generator.choice(N, p=w) takes one
``random()`` u and returns searchsorted(cumsum(p) / cumsum(p)[-1], u, 'right');
``choice_index`` restates that and is pinned against Generator.choice itself.

transition_logpdf : particle_ffbsi.py:14-29
ffbsi_indices     : particle_ffbsi.py:32-104
genealogy_indices : particle_ffbsi.py:121-151
"""

from __future__ import annotations

import numpy as np


def choice_index(w, u):
    """The index Generator.choice(len(w), p=w / sum(w)) returns for uniform u."""
    p = np.asarray(w, dtype=np.float64)
    p = p / p.sum()
    cdf = np.cumsum(p)
    cdf /= cdf[-1]
    return int(np.searchsorted(cdf, u, side="right"))


def transition_logpdf(x_next, x_t, vel, dt, sigma_pos, sigma_cb):
    xt = np.asarray(x_t, dtype=np.float64)
    x_next = np.asarray(x_next, dtype=np.float64)
    dp = x_next[:3] - (xt[:, :3] + np.asarray(vel, dtype=np.float64).ravel() * float(dt))
    dcb = x_next[3] - xt[:, 3]
    return (-0.5 * np.sum(dp * dp, axis=1) / float(sigma_pos) ** 2
            - 0.5 * dcb * dcb / float(sigma_cb) ** 2)


def _normalised(logp):
    mx = np.max(logp)
    with np.errstate(invalid="ignore", over="ignore"):
        w = np.exp(logp - mx)
    s = np.sum(w)
    return w, s


def ffbsi_indices(log_weights, X, vel, dt, sigma_pos, sigma_cb, rng):
    """One backward trajectory (int64 [T]); the rng is consumed as the reference
    does: one random() per step from T-1 down, integers(0, N) at a degenerate
    step, nothing when the last step is degenerate (all indices 0)."""
    lw = np.asarray(log_weights, dtype=np.float64)
    X = np.asarray(X, dtype=np.float64)
    T, N = lw.shape
    idx = np.zeros(T, dtype=np.int64)
    w, s = _normalised(lw[T - 1])
    if not (s > 0.0 and np.isfinite(s)):
        return idx
    idx[T - 1] = choice_index(w, rng.random())
    for t in range(T - 2, -1, -1):
        lf = transition_logpdf(X[t + 1, idx[t + 1]], X[t], vel[t + 1], dt[t + 1],
                               sigma_pos[t + 1], sigma_cb)
        mxt = np.max(lw[t])
        with np.errstate(invalid="ignore", over="ignore"):
            log_norm = mxt + np.log(np.sum(np.exp(lw[t] - mxt)))
        w, s = _normalised(lw[t] - log_norm + lf)
        if not (s > 0.0 and np.isfinite(s)):
            idx[t] = int(rng.integers(0, N))
            continue
        idx[t] = choice_index(w, rng.random())
    return idx


def genealogy_indices(log_weights, ancestors, rng):
    lw = np.asarray(log_weights, dtype=np.float64)
    anc = np.asarray(ancestors, dtype=np.int64)
    T = lw.shape[0]
    idx = np.zeros(T, dtype=np.int64)
    w, s = _normalised(lw[T - 1])
    if not (s > 0.0 and np.isfinite(s)):
        return idx
    idx[T - 1] = choice_index(w, rng.random())
    for t in range(T - 2, -1, -1):
        idx[t] = int(anc[t, idx[t + 1]])
    return idx
