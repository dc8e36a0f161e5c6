"""Seeded synthetic GNSS scenarios for the particle-filter and WLS workloads.

Satellites sit on a sphere of radius WGS84_A + 20,200 km and are placed along
local ENU directions (elevation/azimuth) from the receiver; pseudoranges are
geometric range + receiver clock bias (+ inter-system bias) + noise, with an
optional NLOS generator (positive exponential bias on a random subset).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

WGS84_A = 6378137.0
SAT_RADIUS = WGS84_A + 20_200_000.0
ANCHOR_CFG1 = np.array([-3.96e6, 3.35e6, 3.70e6])

SYS_GPS, SYS_GAL, SYS_BDS, SYS_GLO = 0, 1, 2, 3


def enu_rotation(x):
    """Rotation ECEF->local ENU at x (geocentric latitude); identity near 0."""
    x = np.asarray(x, dtype=np.float64).reshape(3)
    r = float(np.linalg.norm(x))
    if not r > 1.0:
        return np.eye(3)
    lon = np.arctan2(x[1], x[0])
    lat = np.arctan2(x[2], np.hypot(x[0], x[1]))
    sl, cl, sp, cp = np.sin(lon), np.cos(lon), np.sin(lat), np.cos(lat)
    return np.array([[-sl, cl, 0.0], [-sp * cl, -sp * sl, cp], [cp * cl, cp * sl, sp]])


def place_satellites(rx, el_deg, az_deg, radius=SAT_RADIUS):
    """Satellite ECEF positions at the given elevations/azimuths seen from rx."""
    rx = np.asarray(rx, dtype=np.float64).reshape(3)
    el = np.radians(np.asarray(el_deg, dtype=np.float64))
    az = np.radians(np.asarray(az_deg, dtype=np.float64))
    enu = np.stack([np.cos(el) * np.sin(az), np.cos(el) * np.cos(az), np.sin(el)], axis=1)
    u = enu @ enu_rotation(rx)  # ECEF unit vectors
    b = u @ rx
    rho = -b + np.sqrt(b * b - rx @ rx + radius * radius)
    return rx[None, :] + rho[:, None] * u


@dataclass
class Scenario:
    """A trajectory of epochs with a fixed constellation."""

    sat: np.ndarray          # [S, 3] ECEF
    sys: np.ndarray          # [S] system ids (0 = reference system)
    truth: np.ndarray        # [T, 3] receiver ECEF
    cb: float                # receiver clock bias [m]
    isb: np.ndarray          # [n_isb] inter-system biases [m]
    vel: np.ndarray          # [3] ECEF velocity [m/s]
    pr: np.ndarray           # [T, S] pseudoranges
    nlos: np.ndarray         # [T, S] bool
    el: np.ndarray = field(default_factory=lambda: np.zeros(0))
    az: np.ndarray = field(default_factory=lambda: np.zeros(0))

    @property
    def n_epochs(self):
        return self.pr.shape[0]

    @property
    def n_sat(self):
        return self.sat.shape[0]


def _pseudoranges(rng, sat, sys, truth, cb, isb, los_sigma, nlos_frac, nlos_mean):
    T, S = truth.shape[0], sat.shape[0]
    rng_ = np.linalg.norm(sat[None, :, :] - truth[:, None, :], axis=2)
    bias = np.zeros((T, S))
    nlos = np.zeros((T, S), dtype=bool)
    if nlos_frac > 0.0:
        k = int(round(nlos_frac * S))
        for t in range(T):
            idx = rng.choice(S, size=k, replace=False)
            nlos[t, idx] = True
        bias[nlos] = rng.exponential(nlos_mean, size=int(nlos.sum()))
    isb_row = np.array([0.0 if s == 0 else isb[s - 1] for s in sys])
    noise = rng.normal(0.0, los_sigma, size=(T, S))
    return rng_ + cb + isb_row[None, :] + noise + bias, nlos


def scenario_plain(n_epochs=100, n_sat=8, seed=0, speed=10.0, cb=100.0, los_sigma=3.0,
                   anchor=ANCHOR_CFG1):
    """8 GPS satellites, el U(15,85), az U(0,360); straight 10 m/s trajectory."""
    rng = np.random.default_rng(seed)
    el = rng.uniform(15.0, 85.0, n_sat)
    az = rng.uniform(0.0, 360.0, n_sat)
    sat = place_satellites(anchor, el, az)
    R = enu_rotation(anchor)
    vel = np.array([speed, 0.0, 0.0]) @ R  # due east
    truth = anchor[None, :] + np.arange(n_epochs)[:, None] * vel[None, :]
    sys = np.zeros(n_sat, dtype=np.int32)
    pr, nlos = _pseudoranges(rng, sat, sys, truth, cb, np.zeros(0), los_sigma, 0.0, 0.0)
    return Scenario(sat, sys, truth, cb, np.zeros(0), vel, pr, nlos, el, az)


def scenario_multi_gnss(n_epochs=500, per_sys=3, n_sys=4, seed=0, speed=10.0, cb=100.0,
                        isb=(15.0, -8.0, 25.0), nlos_frac=0.3, nlos_mean=20.0, los_sigma=3.0,
                        anchor=ANCHOR_CFG1):
    """GPS/Galileo/BeiDou/GLONASS, per_sys satellites each, with ISBs and NLOS."""
    rng = np.random.default_rng(seed)
    S = per_sys * n_sys
    el = rng.uniform(15.0, 85.0, S)
    az = rng.uniform(0.0, 360.0, S)
    sat = place_satellites(anchor, el, az)
    sys = np.repeat(np.arange(n_sys, dtype=np.int32), per_sys)
    R = enu_rotation(anchor)
    vel = np.array([speed, 0.0, 0.0]) @ R
    truth = anchor[None, :] + np.arange(n_epochs)[:, None] * vel[None, :]
    isb = np.asarray(isb[: n_sys - 1], dtype=np.float64)
    pr, nlos = _pseudoranges(rng, sat, sys, truth, cb, isb, los_sigma, nlos_frac, nlos_mean)
    return Scenario(sat, sys, truth, cb, isb, vel, pr, nlos, el, az)


def batched_filters(n_filters=256, n_min=1_000, n_max=100_000, s_min=6, s_max=20, seed=0):
    """Particle counts log-uniform in [n_min, n_max], satellite counts U{s_min..s_max}."""
    rng = np.random.default_rng(seed)
    n = np.round(np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_filters))).astype(np.int32)
    s = rng.integers(s_min, s_max + 1, n_filters).astype(np.int32)
    return n, s


def wls_epochs(n_epoch=10_000, s_min=6, s_max=20, seed=0, noise=3.0, cb=3000.0,
               anchor=ANCHOR_CFG1, sagnac=True):
    """Ragged WLS batch: per-epoch satellite counts U{s_min..s_max}; rows beyond an
    epoch's count are zero-padded (weights 0)."""
    from .range_model import geometric_ranges_sagnac

    rng = np.random.default_rng(seed)
    counts = rng.integers(s_min, s_max + 1, n_epoch).astype(np.int32)
    sat = np.zeros((n_epoch, s_max, 3))
    pr = np.zeros((n_epoch, s_max))
    w = np.zeros((n_epoch, s_max))
    truth = anchor[None, :] + rng.normal(0.0, 100.0, (n_epoch, 3))
    for e in range(n_epoch):
        k = counts[e]
        el = rng.uniform(10.0, 85.0, k)
        az = rng.uniform(0.0, 360.0, k)
        s_ecef = place_satellites(truth[e], el, az)
        sat[e, :k] = s_ecef
        if sagnac:
            rr = geometric_ranges_sagnac(truth[e], s_ecef)
        else:
            rr = np.linalg.norm(s_ecef - truth[e], axis=1)
        pr[e, :k] = rr + cb + rng.normal(0.0, noise, k)
        w[e, :k] = 1.0
    return sat, pr, w, counts, truth


def batched_scenarios(n_sat, n_epochs, seed=0, nlos_frac=0.3, nlos_mean=20.0, los_sigma=3.0,
                      speed=10.0, cb=100.0):
    """One GPS-only scenario per filter with n_sat[f] satellites: the anchor of
    configuration 1 moved by up to 5 km, el U(15,85), az U(0,360), a straight
    trajectory at `speed` m/s in a random direction, and the NLOS generator of
    configuration 2 (nlos_frac of the satellites per epoch, bias Exp(nlos_mean))."""
    rng = np.random.default_rng(seed)
    out = []
    for s in np.asarray(n_sat):
        a0 = ANCHOR_CFG1 + rng.uniform(-5000.0, 5000.0, 3)
        el = rng.uniform(15.0, 85.0, int(s))
        az = rng.uniform(0.0, 360.0, int(s))
        sat = place_satellites(a0, el, az)
        R = enu_rotation(a0)
        h = rng.uniform(0.0, 2 * np.pi)
        vel = np.array([speed * np.cos(h), speed * np.sin(h), 0.0]) @ R
        truth = a0[None, :] + np.arange(n_epochs)[:, None] * vel[None, :]
        sys = np.zeros(int(s), dtype=np.int32)
        pr, nlos = _pseudoranges(rng, sat, sys, truth, cb, np.zeros(0), los_sigma, nlos_frac,
                                 nlos_mean)
        out.append(Scenario(sat, sys, truth, cb, np.zeros(0), vel, pr, nlos, el, az))
    return out
