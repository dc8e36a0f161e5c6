"""Urban-canyon satellite-epoch generator with random box buildings.

Synthetic inputs for the learned likelihood and its benchmark (host side,
NumPy).  A receiver drives along a street (the local east axis) lined on both
sides by rectangular buildings 10-150 m tall.  For each satellite the line of
sight from the antenna is tested against every building (slab test); a
blocked satellite is received through a single reflection off the facade
across the street, which lengthens the path by 2 w cos(el) |cos(az)| (w: the
receiver's distance to that facade, az from north, the street running east),
plus a spread.  C/N0 ~ U(20, 50) dB-Hz, 8 dB lower when blocked.  With the
default street widths (half width 18-40 m) and elevations U(15, 85) deg about
40% of the satellite-epochs are blocked.  The
reference's own canyon demo (examples/demo_urban_canyon_sim.py:120-160) masks
low-elevation satellites with a fixed bias; this generator derives the
blocking and the bias from the building geometry instead, in the box-building
model of python/gnss_gpu/raytrace.py (BuildingModel).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

C_LIGHT = 299792458.0


@dataclass
class Canyon:
    lo: np.ndarray      # [B, 3] box minimum corner (east, north, up) [m]
    hi: np.ndarray      # [B, 3] box maximum corner
    half_width: float   # nominal street half width [m]


def make_canyon(rng, length=2000.0, half_width=(18.0, 40.0), depth=(10.0, 40.0),
                height=(10.0, 150.0), block=(15.0, 60.0), gap=(5.0, 40.0)) -> Canyon:
    """Buildings on both sides of a street along +east from -length/2 to length/2."""
    lo, hi = [], []
    hw = float(rng.uniform(*half_width))
    for side in (-1.0, 1.0):
        x = -length / 2
        while x < length / 2:
            bl = rng.uniform(*block)
            setback = hw + rng.uniform(0.0, 4.0)
            d = rng.uniform(*depth)
            h = rng.uniform(*height)
            if side > 0:
                lo.append([x, setback, 0.0])
                hi.append([x + bl, setback + d, h])
            else:
                lo.append([x, -setback - d, 0.0])
                hi.append([x + bl, -setback, h])
            x += bl + rng.uniform(*gap)
    return Canyon(np.array(lo), np.array(hi), hw)


def los_dirs(el, az):
    """Unit ENU vectors toward satellites at elevation el, azimuth az (rad, from north)."""
    ce = np.cos(el)
    return np.stack([ce * np.sin(az), ce * np.cos(az), np.sin(el)], axis=-1)


def blocked(canyon: Canyon, pos, dirs, max_range=2000.0):
    """Ray/box slab test: pos [..., 3] ENU, dirs [..., 3] unit; True where any box
    is hit within max_range.  Broadcasts pos and dirs against each other."""
    pos = np.asarray(pos, dtype=np.float64)[..., None, :]
    d = np.asarray(dirs, dtype=np.float64)[..., None, :]
    inv = 1.0 / np.where(np.abs(d) < 1e-12, 1e-12, d)
    t0 = (canyon.lo - pos) * inv
    t1 = (canyon.hi - pos) * inv
    tmin = np.max(np.minimum(t0, t1), axis=-1)
    tmax = np.min(np.maximum(t0, t1), axis=-1)
    hit = (tmax >= np.maximum(tmin, 0.0)) & (tmin <= max_range)
    return np.any(hit, axis=-1)


def nlos_excess(canyon: Canyon, pos, el, az, rng):
    """Single-reflection excess path off the facade across the street [m]."""
    north = np.asarray(pos)[..., 1]
    # the satellite on the north side (cos az > 0) reflects off the south facade
    w = np.where(np.cos(az) > 0, north + canyon.half_width, canyon.half_width - north)
    base = 2.0 * np.abs(w) * np.cos(el) * np.abs(np.cos(az))
    return base * rng.uniform(0.8, 1.5, np.shape(base)) + rng.exponential(5.0, np.shape(base))


@dataclass
class SatEpochs:
    """Per satellite-epoch records, [E, S] arrays (S fixed per set)."""
    el: np.ndarray
    az: np.ndarray
    u: np.ndarray          # [E, S, 3] unit ENU direction receiver -> satellite
    cn0: np.ndarray        # dB-Hz
    nlos: np.ndarray       # bool, blocked at the true position
    bias: np.ndarray       # NLOS excess path [m] (0 when line of sight)
    noise: np.ndarray      # code noise [m]
    truth: np.ndarray      # [E, 3] receiver ENU position


def sat_epochs(rng, n_epochs, n_sat, canyon=None, sigma_code=3.0, el_range=(15.0, 85.0),
               speed=5.0):
    """Satellite-epochs along a drive through one canyon (a new canyon when None)."""
    canyon = canyon if canyon is not None else make_canyon(rng)
    x = -900.0 + speed * np.arange(n_epochs)
    x = ((x + 900.0) % 1800.0) - 900.0
    truth = np.stack([x, rng.uniform(-2.0, 2.0, n_epochs), np.full(n_epochs, 1.5)], axis=1)
    el = np.deg2rad(rng.uniform(*el_range, (n_epochs, n_sat)))
    az = np.deg2rad(rng.uniform(0.0, 360.0, (n_epochs, n_sat)))
    u = los_dirs(el, az)
    nl = blocked(canyon, truth[:, None, :], u)
    bias = np.where(nl, nlos_excess(canyon, truth[:, None, :], el, az, rng), 0.0)
    cn0 = rng.uniform(20.0, 50.0, (n_epochs, n_sat)) - 8.0 * nl
    noise = rng.normal(0.0, sigma_code, (n_epochs, n_sat))
    return canyon, SatEpochs(el, az, u, cn0, nl, bias, noise, truth)


# --------------------------------------------------------------------------
# Features of the learned likelihood (shared by training, inference and the
# oracle): per pair (particle, satellite)
#   x = [e / 10, el / (pi/2), sin az, cos az, (cn0 - 35) / 10, los_at_mean,
#        n_sat / 20, ess / N]
# The first is the pseudorange residual at the particle [m]; the rest are per
# satellite and epoch.
# --------------------------------------------------------------------------
RES_SCALE = 10.0
N_FEAT = 8


def sat_features(el, az, cn0, los_at_mean, n_sat, ess_frac):
    """Per-satellite features 1..7 (the residual is feature 0), [..., 7] float32."""
    el = np.asarray(el, dtype=np.float64)
    shp = el.shape
    f = np.stack([el / (np.pi / 2), np.sin(az), np.cos(az), (np.asarray(cn0) - 35.0) / 10.0,
                  np.asarray(los_at_mean, dtype=np.float64),
                  np.broadcast_to(np.asarray(n_sat, dtype=np.float64) / 20.0, shp),
                  np.broadcast_to(np.asarray(ess_frac, dtype=np.float64), shp)], axis=-1)
    return f.astype(np.float32)


def training_set(rng, n_epochs, n_sat=16, n_canyons=None, mean_err=3.0, canyon_kw=None):
    """Satellite-epochs for the learned likelihood from fresh canyons.

    Returns (f [E, S, 7] float32 features, r0 [E, S] residual at the true
    position [m], u [E, S, 3] unit vectors to the satellites, n_sat [E] int32).
    canyon_kw goes to make_canyon.
    The line-of-sight feature is tested from a stand-in for the particle mean:
    the true position plus a N(0, mean_err^2) horizontal error; the ESS
    fraction is drawn from U(0.05, 1)."""
    n_canyons = n_canyons or max(1, n_epochs // 500)
    per = -(-n_epochs // n_canyons)
    fs, rs, us = [], [], []
    for _ in range(n_canyons):
        can = make_canyon(rng, **(canyon_kw or {}))
        can, se = sat_epochs(rng, per, n_sat, canyon=can)
        est = se.truth.copy()
        est[:, :2] += rng.normal(0.0, mean_err, (per, 2))
        flag = ~blocked(can, est[:, None, :], se.u)
        ess = rng.uniform(0.05, 1.0, (per, 1))
        fs.append(sat_features(se.el, se.az, se.cn0, flag, n_sat, ess))
        rs.append((se.bias + se.noise).astype(np.float32))
        us.append(se.u.astype(np.float32))
    f = np.concatenate(fs)[:n_epochs]
    r0 = np.concatenate(rs)[:n_epochs]
    u = np.concatenate(us)[:n_epochs]
    return f, r0, u, np.full(n_epochs, n_sat, dtype=np.int32)


@dataclass
class Drive:
    """A canyon drive in ECEF for the particle filter."""
    canyon: Canyon
    anchor: np.ndarray     # ECEF origin of the local frame
    R: np.ndarray          # ECEF -> ENU rotation
    truth_enu: np.ndarray  # [T, 3]
    truth: np.ndarray      # [T, 3] ECEF
    vel: np.ndarray        # [T, 3] ECEF velocity (truth[t] - truth[t-1])
    sat: np.ndarray        # [T, S, 3] ECEF
    pr: np.ndarray         # [T, S]
    cn0: np.ndarray        # [T, S]
    u: np.ndarray          # [T, S, 3] ENU unit vectors to the satellites
    nlos: np.ndarray       # [T, S]
    cb: float
    sys: np.ndarray = None  # [S] system index (0 = reference system)
    isb: np.ndarray = None  # [n_sys - 1] inter-system biases [m] added to pr

    def los_flags(self, est_ecef, t):
        """Line-of-sight flags of epoch t's satellites seen from an ECEF point."""
        e = self.R @ (np.asarray(est_ecef, dtype=np.float64) - self.anchor)
        return (~blocked(self.canyon, e[None, :], self.u[t])).astype(np.float32)


def drive_ecef(rng, n_epochs, n_sat=16, anchor=(-3957199.0, 3310205.0, 3737911.0), cb=30.0,
               speed=5.0, n_sys=1, isb=(12.0, -7.5, 20.0), canyon_kw=None):
    """A straight drive along a fresh canyon; satellites are placed 2.2e7 m from
    the receiver along the simulated lines of sight.  With n_sys > 1 the
    satellites are split evenly over n_sys systems (row order), system k > 0
    carrying the inter-system bias isb[k-1]; canyon_kw goes to make_canyon
    (its length is stretched to cover the drive)."""
    from .sim import enu_rotation

    kw = dict(canyon_kw or {})
    kw["length"] = max(float(kw.get("length", 2000.0)), speed * n_epochs + 400.0)
    can = make_canyon(rng, **kw)
    can, se = sat_epochs(rng, n_epochs, n_sat, canyon=can, speed=speed)
    anchor = np.asarray(anchor, dtype=np.float64)
    R = enu_rotation(anchor)
    # keep the drive continuous: no wrap inside one drive
    truth_enu = se.truth.copy()
    truth_enu[:, 0] = -0.5 * speed * n_epochs + speed * np.arange(n_epochs)
    truth = anchor[None, :] + truth_enu @ R
    vel = np.vstack([np.zeros(3), np.diff(truth, axis=0)])
    nl = blocked(can, truth_enu[:, None, :], se.u)
    bias = np.where(nl, nlos_excess(can, truth_enu[:, None, :], se.el, se.az, rng), 0.0)
    cn0 = rng.uniform(20.0, 50.0, nl.shape) - 8.0 * nl
    sat = truth[:, None, :] + 2.2e7 * (se.u @ R)
    rng_true = np.linalg.norm(sat - truth[:, None, :], axis=2)
    sys_ids = (np.arange(n_sat) * n_sys // n_sat).astype(np.int32)
    isb_v = np.asarray(isb, dtype=np.float64)[: max(n_sys - 1, 0)]
    offs = np.concatenate([[0.0], isb_v])[sys_ids]
    pr = rng_true + cb + offs[None, :] + bias + se.noise
    return Drive(can, anchor, R, truth_enu, truth, vel, sat, pr, cn0.astype(np.float32), se.u,
                 nl, cb, sys_ids, isb_v)
