"""Binary64 restatement of the reference particle filter, driven by the same
random numbers as the GPU library.

Follows src/particle_filter/pf_device.cu:
  init      :67-104    predict :106-187    weight :189-276 (Huber :19-26)
  ESS       :924-965 + :1722-1762          estimate :968-1023 + :1926-1969
  spread    :1025-1076 + :1764-1807        systematic :1813-1878
  Megopolis :1231-1298 + :1884-1920
  DD pseudorange :282-368   carrier AFV :439-494   DD carrier AFV :502-604
  Doppler :610-740          Doppler KF :742-880
and python/gnss_gpu/particle_filter_device.py:198-303,679-714 for the
orchestration.  State is absolute ECEF binary64 ([x y z cb isb...]), like the
reference.  Random draws: the reference's cuRAND streams cannot be reproduced
(closed library), so the noise comes from the library's Philox streams, drawn
in the library's local ENU frame and rotated to ECEF.  Resampling uses the
bit-exact twins in detmath on binary32-rounded log-weights (the library stores
log-weights in binary32).
"""

from __future__ import annotations

import numpy as np

from . import detmath, philox


def enu_rotation(x):
    x = np.asarray(x, dtype=np.float64).reshape(3)
    r = float(np.linalg.norm(x))
    if not r > 1.0:
        return np.eye(3)
    lon = np.arctan2(x[1], x[0])
    lat = np.arctan2(x[2], np.hypot(x[0], x[1]))
    sl, cl, sp, cp = np.sin(lon), np.cos(lon), np.sin(lat), np.cos(lat)
    return np.array([[-sl, cl, 0.0], [-sp * cl, -sp * sl, cp], [cp * cl, cp * sl, sp]])


def huber_cost(residual, sigma, w, k):
    """pf_device.cu:19-26."""
    z = np.abs(residual) / sigma
    rho = np.where(z < k, 0.5 * z * z, k * z - 0.5 * k * k)
    return w * rho


def log_likelihood(pos, cb, isb, sat, pr, w, sys, sigma_pr, nu=0.0, gate_thr=0.0,
                   huber=False, huber_k=1.5):
    """Per-particle log-weight, pf_device.cu:189-276 (overwrites, not adds).
    pos [N,3], cb [N], isb [N, n_isb] (may be empty), sat [S,3]."""
    N = pos.shape[0]
    S = sat.shape[0]
    if S == 0:
        return np.zeros(N)
    r = np.linalg.norm(pos[:, None, :] - sat[None, :, :], axis=2)
    isb_term = np.zeros((N, S))
    for s in range(S):
        if sys is not None and sys[s] >= 1 and isb.shape[1] >= sys[s]:
            isb_term[:, s] = isb[:, sys[s] - 1]
    resid = pr[None, :] - (r + cb[:, None] + isb_term)
    use_gate = gate_thr > 0.0
    keep = np.ones((N, S), dtype=bool)
    if use_gate:
        keep = np.abs(resid) <= gate_thr
        min_kept = min(S, 4)
        gate_on = keep.sum(axis=1) >= min_kept
        keep = np.where(gate_on[:, None], keep, True)
    if nu > 0.0:
        term = -0.5 * (nu + 1.0) * np.log(1.0 + w[None, :] * resid * resid / (nu * sigma_pr ** 2))
    elif huber and huber_k > 0.0:
        term = -huber_cost(resid, sigma_pr, w[None, :], huber_k)
    else:
        term = -0.5 * w[None, :] * resid * resid / (sigma_pr * sigma_pr)
    return np.sum(np.where(keep, term, 0.0), axis=1)


def log_likelihood_gmm(pos, cb, sat, pr, w, sigma_los, w_los, mu_nlos, sigma_nlos):
    """Per-particle GMM log-weight, pfd_weight_gmm_kernel pf_device.cu:373-437
    (overwrites): sum over satellites of max(a, b) + log(exp(a - max) + exp(b - max))
    with a = log w_los - log sigma_los - w e^2 / (2 sigma_los^2),
    b = log(1 - w_los) - log sigma_nlos - w (e - mu_nlos)^2 / (2 sigma_nlos^2),
    e = pr - |sat - x| - cb.  pos [N,3], cb [N], sat [S,3]."""
    r = np.linalg.norm(pos[:, None, :] - sat[None, :, :], axis=2)
    e = pr[None, :] - (r + cb[:, None])
    a = np.log(w_los) - np.log(sigma_los) - 0.5 * w[None, :] * e * e / (sigma_los * sigma_los)
    en = e - mu_nlos
    b = np.log(1.0 - w_los) - np.log(sigma_nlos) - 0.5 * w[None, :] * en * en / (
        sigma_nlos * sigma_nlos)
    mx = np.maximum(a, b)
    return np.sum(mx + np.log(np.exp(a - mx) + np.exp(b - mx)), axis=1)


def _gated_sum(resid, w, sigma, thr, huber, huber_k):
    """Sum over observations of the Gaussian or weighted-Huber cost with the
    per-particle gate of the DD kernels (min(n, 3) kept, pf_device.cu:317-335)."""
    n = resid.shape[1]
    keep = np.ones(resid.shape, dtype=bool)
    if thr > 0.0:
        keep = np.abs(resid) <= thr
        gate_on = keep.sum(axis=1) >= min(n, 3)
        keep = np.where(gate_on[:, None], keep, True)
    if huber and huber_k > 0.0:
        term = -huber_cost(resid, sigma, w[None, :], huber_k)
    else:
        term = -0.5 * w[None, :] * resid * resid / (sigma * sigma)
    return np.sum(np.where(keep, term, 0.0), axis=1)


def dd_pseudorange_loglik(pos, sat_k, sat_ref, dd, br_k, br_ref, w, sigma, thr=0.0,
                          huber=False, huber_k=1.5):
    """DD pseudorange log-weight, pf_device.cu:282-368 (overwrites):
    residual = dd - (r_k - r_ref - br_k + br_ref)."""
    rk = np.linalg.norm(pos[:, None, :] - np.asarray(sat_k)[None, :, :], axis=2)
    rr = np.linalg.norm(pos[:, None, :] - np.asarray(sat_ref)[None, :, :], axis=2)
    resid = np.asarray(dd)[None, :] - (((rk - rr) - np.asarray(br_k)[None, :]) +
                                       np.asarray(br_ref)[None, :])
    return _gated_sum(resid, np.asarray(w, dtype=np.float64), sigma, thr, huber, huber_k)


def afv(x):
    """Distance to the nearest integer cycle, x - rint(x)."""
    x = np.asarray(x, dtype=np.float64)
    return x - np.rint(x)


def carrier_afv_loglik(pos, cb, sat, cp, w, wavelength, sigma_cycles):
    """Undifferenced carrier AFV term, pf_device.cu:439-494 (added to lw):
    -w afv^2 / (2 sigma^2), afv of cp - (r + cb) / wavelength."""
    r = np.linalg.norm(pos[:, None, :] - np.asarray(sat)[None, :, :], axis=2)
    a = afv(np.asarray(cp)[None, :] - (r + np.asarray(cb)[:, None]) * (1.0 / wavelength))
    return np.sum(-0.5 * np.asarray(w)[None, :] * a * a / (sigma_cycles * sigma_cycles), axis=1)


def dd_carrier_afv_loglik(pos, sat_k, sat_ref, ddc, br_k, br_ref, w, wl, sigma_cycles, thr=0.0,
                          huber=False, huber_k=1.5):
    """DD carrier AFV term, pf_device.cu:502-604 (added to lw), per-pair
    wavelengths; gate and Huber on the AFV in cycles."""
    rk = np.linalg.norm(pos[:, None, :] - np.asarray(sat_k)[None, :, :], axis=2)
    rr = np.linalg.norm(pos[:, None, :] - np.asarray(sat_ref)[None, :, :], axis=2)
    exp_c = (((rk - rr) - np.asarray(br_k)[None, :]) + np.asarray(br_ref)[None, :]) * \
        (1.0 / np.asarray(wl, dtype=np.float64))[None, :]
    a = afv(np.asarray(ddc)[None, :] - exp_c)
    return _gated_sum(a, np.asarray(w, dtype=np.float64), sigma_cycles, thr, huber, huber_k)


def _los(pos, sat):
    d = np.asarray(sat)[None, :, :] - pos[:, None, :]
    r = np.linalg.norm(d, axis=2)
    ok = r >= 1.0
    return d / np.where(ok, r, 1.0)[..., None], ok


def doppler_update(pos, vel, sat, sv, dop, w, wl, sigma, gain, max_dv):
    """pfd_weight_doppler_kernel, pf_device.cu:610-740, binary64, per particle:
    returns (log-weight increment [N], new velocity [N, 3])."""
    pos = np.asarray(pos, dtype=np.float64)
    vel = np.array(vel, dtype=np.float64, copy=True)
    l, ok = _los(pos, sat)
    obs = -np.asarray(dop) * wl
    svr = np.einsum("nsk,sk->ns", l, np.asarray(sv))
    w = np.maximum(np.asarray(w, dtype=np.float64), 0.0)[None, :] * ok
    rxr = np.einsum("nsk,nk->ns", l, vel)
    wsum = w.sum(axis=1)
    anyw = wsum > 0
    cd = np.where(anyw, (w * (obs[None, :] - svr + rxr)).sum(axis=1) / np.where(anyw, wsum, 1), 0)
    pred = svr - rxr + cd[:, None]
    res = obs[None, :] - pred
    add = np.where(anyw, np.sum(-0.5 * w * res * res / (sigma * sigma), axis=1), 0.0)
    g = min(1.0, max(0.0, gain))
    if g > 0.0 and len(obs) >= 4:
        for i in np.nonzero(anyw)[0]:
            H = np.concatenate([-l[i], np.ones((len(obs), 1))], axis=1)
            rhs = obs - svr[i]
            A = np.zeros((4, 5))
            A[:, :4] = (H * w[i][:, None]).T @ H + 1e-9 * np.eye(4)
            A[:, 4] = (H * w[i][:, None]).T @ rhs
            try:
                sol = np.linalg.solve(A[:, :4], A[:, 4])
            except np.linalg.LinAlgError:
                continue
            if not np.all(np.isfinite(sol)):
                continue
            dv = sol[:3] - vel[i]
            nrm = np.linalg.norm(dv)
            if max_dv > 0.0 and nrm > max_dv:
                dv *= max_dv / max(nrm, 1e-9)
            vel[i] = vel[i] + g * dv
    return add, vel


def doppler_kf_update(pos, vel, cov, sat, sv, dop, w, wl, sigma):
    """pfd_doppler_kf_update_kernel, pf_device.cu:742-880, binary64: sequential
    scalar Kalman updates with centred rows; returns (log-weight increment [N],
    velocity [N, 3], covariance [N, 3, 3])."""
    pos = np.asarray(pos, dtype=np.float64)
    mu = np.array(vel, dtype=np.float64, copy=True)
    P = np.array(cov, dtype=np.float64, copy=True).reshape(-1, 3, 3)
    l, ok = _los(pos, sat)
    obs = -np.asarray(dop) * wl
    svr = np.einsum("nsk,sk->ns", l, np.asarray(sv))
    w = np.maximum(np.asarray(w, dtype=np.float64), 0.0)
    use = ok & (w[None, :] > 0)
    wm = np.where(use, w[None, :], 0.0)
    wsum = wm.sum(axis=1)
    add = np.zeros(pos.shape[0])
    for i in range(pos.shape[0]):
        if not wsum[i] > 0:
            continue
        hm = (wm[i][:, None] * -l[i]).sum(axis=0) / wsum[i]
        ym = (wm[i] * (obs - svr[i])).sum() / wsum[i]
        for k in range(len(obs)):
            if not use[i, k]:
                continue
            h = -l[i, k] - hm
            yc = (obs[k] - svr[i, k]) - ym
            ph = P[i] @ h
            S = h @ ph + sigma * sigma / w[k]
            if not (S > 1e-18 and np.isfinite(S)):
                continue
            innov = yc - h @ mu[i]
            K = ph / S
            mu[i] = mu[i] + K * innov
            P[i] = P[i] - np.outer(K, h @ P[i])
            P[i] = 0.5 * (P[i] + P[i].T)
            for d in range(3):
                P[i, d, d] = P[i, d, d] if (np.isfinite(P[i, d, d]) and P[i, d, d] > 1e-12) else 1e-12
            add[i] += -0.5 * (innov * innov / S + np.log(S))
    return add, mu, P


def ess(lw):
    """(sum w)^2 / sum w^2 with w = exp(lw - max), pf_device.cu:1722-1762."""
    lw = np.asarray(lw, dtype=np.float64)
    w = np.exp(lw - np.max(lw))
    return float(np.sum(w) ** 2 / np.sum(w * w))


def weighted_mean(state, lw):
    """pf_device.cu:968-1023 + :1926-1969."""
    lw = np.asarray(lw, dtype=np.float64)
    w = np.exp(lw - np.max(lw))
    return (w[:, None] * state).sum(axis=0) / w.sum()


def position_spread(pos, lw, center):
    """pf_device.cu:1025-1076 + :1764-1807."""
    lw = np.asarray(lw, dtype=np.float64)
    w = np.exp(lw - np.max(lw))
    d2 = np.sum((pos - np.asarray(center)[None, :]) ** 2, axis=1)
    tw = w.sum()
    if not tw > 0.0:
        return 0.0
    return float(np.sqrt(max(np.sum(w * d2) / tw, 0.0)))


class OraclePF:
    """Binary64 particle filter mirroring ParticleFilterDevice call by call."""

    def __init__(self, n_particles, seed=42, n_isb=0):
        self.n = int(n_particles)
        self.seed = int(seed)
        self.n_isb = int(n_isb)
        self.state = None          # [N, 4 + n_isb] absolute
        self.vel = None            # [N, 3] ECEF velocity
        self.vcov = None           # [N, 3, 3] ECEF velocity covariance (RBPF)
        self.lw = np.zeros(self.n)
        self.R = np.eye(3)
        self.resample_count = 0
        self.ancestors = np.arange(self.n, dtype=np.int32)

    # pf_device.cu:67-104 (velocity and its covariance: :84-101)
    def initialize(self, pos, cb=0.0, spread_pos=100.0, spread_cb=1000.0, isb=None,
                   spread_isb=0.0, velocity=None, spread_vel=0.0, velocity_init_sigma=0.0):
        pos = np.asarray(pos, dtype=np.float64).reshape(3)
        self.R = enu_rotation(pos)
        i = np.arange(self.n)
        z = philox.normal4(self.seed, i, 0, philox.PURPOSE_INIT, 0)
        D = 4 + self.n_isb
        st = np.zeros((self.n, D))
        st[:, :3] = pos[None, :] + (spread_pos * z[:, :3]) @ self.R
        st[:, 3] = cb + spread_cb * z[:, 3]
        if self.n_isb:
            zi = philox.normal4(self.seed, i, 0, philox.PURPOSE_INIT, 1)
            isb0 = np.zeros(self.n_isb) if isb is None else np.asarray(isb, dtype=np.float64)
            st[:, 4:] = isb0[None, :] + spread_isb * zi[:, : self.n_isb]
        self.state = st
        v = np.zeros(3) if velocity is None else np.asarray(velocity, dtype=np.float64)
        self.vel = np.tile(v, (self.n, 1))
        if spread_vel > 0.0:
            zv = philox.normal4(self.seed, i, 0, philox.PURPOSE_INIT, 2)
            self.vel = self.vel + (spread_vel * zv[:, :3]) @ self.R
        self.vcov = np.tile(np.eye(3) * float(velocity_init_sigma) ** 2, (self.n, 1, 1))
        self.lw = np.zeros(self.n)
        self.resample_count = 0

    @staticmethod
    def _chol3(a):
        """Lower Cholesky factor of the symmetric 3x3 matrices a [N, 3, 3] with
        the reference's guards (pf_device.cu:153-165): sqrt of max(., 0), the
        column below a pivot of at most 1e-12 left at zero."""
        l00 = np.sqrt(np.maximum(a[:, 0, 0], 0.0))
        ok0 = l00 > 1e-12
        l10 = np.where(ok0, a[:, 0, 1] / np.where(ok0, l00, 1.0), 0.0)
        l20 = np.where(ok0, a[:, 0, 2] / np.where(ok0, l00, 1.0), 0.0)
        l11 = np.sqrt(np.maximum(a[:, 1, 1] - l10 * l10, 0.0))
        ok1 = l11 > 1e-12
        l21 = np.where(ok1, (a[:, 1, 2] - l20 * l10) / np.where(ok1, l11, 1.0), 0.0)
        l22 = np.sqrt(np.maximum(a[:, 2, 2] - l20 * l20 - l21 * l21, 0.0))
        L = np.zeros_like(a)
        L[:, 0, 0], L[:, 1, 0], L[:, 1, 1] = l00, l10, l11
        L[:, 2, 0], L[:, 2, 1], L[:, 2, 2] = l20, l21, l22
        return L

    # pf_device.cu:106-187: velocity guide blend v = (1 - a) v + a v_guide
    # (:121-128), velocity noise N(0, sigma_vel) unless the velocity Kalman
    # filter is on (:129-133), position noise N(0, sigma_pos) per axis or, with
    # the RBPF, the Cholesky draw from sigma_pos^2 I + dt^2 Sigma_v followed by
    # Sigma_v += q dt I (:141-176), then p += v dt + n and the clock walk.
    # The noise is drawn in the library's local ENU frame (R fixed at
    # initialisation) and rotated to ECEF; the Cholesky factor is that of the
    # local-frame covariance R Sigma_v R^T, as the library factors it.
    def predict(self, velocity, dt, sigma_pos, sigma_cb, step, sigma_isb=0.0, alpha=1.0,
                sigma_vel=0.0, velocity_kf=False, velocity_process_noise=0.0):
        z, zv, zi = philox.predict_normals(self.seed, self.n, step, self.n_isb)
        vg = np.asarray(velocity, dtype=np.float64).reshape(3)
        a = min(1.0, max(0.0, alpha))
        self.vel = (1.0 - a) * self.vel + a * vg[None, :]
        if not velocity_kf and sigma_vel > 0.0:
            self.vel = self.vel + (sigma_vel * zv) @ self.R
        if velocity_kf:
            Sl = np.einsum("ij,njk,lk->nil", self.R, self.vcov, self.R)
            A = sigma_pos * sigma_pos * np.eye(3)[None] + dt * dt * Sl
            L = self._chol3(A)
            noise = np.einsum("nij,nj->ni", L, z[:, :3]) @ self.R
            qdt = max(velocity_process_noise, 0.0) * max(dt, 0.0)
            self.vcov = self.vcov + qdt * np.eye(3)[None]
        else:
            noise = (sigma_pos * z[:, :3]) @ self.R
        self.state[:, :3] += self.vel * dt + noise
        self.state[:, 3] += sigma_cb * z[:, 3]
        if self.n_isb:
            self.state[:, 4:] += sigma_isb * zi

    def weight(self, sat, pr, w=None, sys=None, sigma_pr=5.0, nu=0.0, gate_thr=0.0,
               huber=False, huber_k=1.5):
        sat = np.asarray(sat, dtype=np.float64).reshape(-1, 3)
        pr = np.asarray(pr, dtype=np.float64).ravel()
        w = np.ones(pr.size) if w is None else np.asarray(w, dtype=np.float64).ravel()
        isb = self.state[:, 4:]
        self.lw = log_likelihood(self.state[:, :3], self.state[:, 3], isb, sat, pr, w, sys,
                                 sigma_pr, nu, gate_thr, huber, huber_k)

    def ess(self):
        return ess(self.lw)

    def estimate(self):
        return weighted_mean(self.state, self.lw)

    def _apply(self, anc):
        self.ancestors = anc
        self.state = self.state[anc]
        self.vel = self.vel[anc]
        self.vcov = self.vcov[anc]
        self.lw = np.zeros(self.n)
        self.resample_count += 1

    def resample_systematic(self, u0=None, lw32=None):
        """pf_device.cu:1813-1878 with the library's fixed scan order."""
        if u0 is None:
            u0 = detmath.reference_u0(self.seed, self.n)
        lw = self.lw.astype(np.float32) if lw32 is None else lw32
        self._apply(detmath.systematic_ancestors(lw, self.n, u0))

    def resample_megopolis(self, n_iter=15, lw32=None):
        """pf_device.cu:1884-1920 (reference acceptance rule)."""
        lw = self.lw.astype(np.float32) if lw32 is None else lw32
        self._apply(detmath.megopolis_ref_ancestors(lw, self.n, n_iter, self.seed))

    def resample_megopolis_chesser(self, n_iter, lw32=None):
        lw = self.lw.astype(np.float32) if lw32 is None else lw32
        self._apply(detmath.megopolis_ancestors(lw, self.n, n_iter, self.seed,
                                                self.resample_count))
