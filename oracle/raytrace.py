"""Binary64 restatement of the reference's ray casting and ray-traced NLOS weight
(TEST INFRASTRUCTURE ONLY).  Brute force over every triangle, no tree: the
library's BVH queries must give the same answers.

mt_hit        : Moller-Trumbore, src/raytrace/bvh.cu:250-272
los_blocked   : bvh.cu:293-354, 517-585 (LOS), weight_3d_bvh.cu:79-112 (weight)
multipath     : bvh.cu:362-462, 587-691
weight_3d     : src/particle_filter/weight_3d_bvh.cu:127-220
"""

from __future__ import annotations

import numpy as np


def mt_hit(o, d, tris):
    """Rays o, d [R, 3] against triangles [T, 3, 3]: (hit [R, T], t [R, T]).
    hit is the intersection test without the t range (bvh.cu:250-270)."""
    o = np.asarray(o, dtype=np.float64)[:, None, :]
    d = np.asarray(d, dtype=np.float64)[:, None, :]
    v0 = tris[None, :, 0, :]
    e1 = tris[None, :, 1, :] - tris[None, :, 0, :]
    e2 = tris[None, :, 2, :] - tris[None, :, 0, :]
    h = np.cross(d, e2)
    a = np.sum(e1 * h, axis=-1)
    ok = ~(np.abs(a) < 1e-12)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = 1.0 / a
        s = o - v0
        u = f * np.sum(s * h, axis=-1)
        ok &= ~((u < 0.0) | (u > 1.0))
        q = np.cross(s, e1)
        v = f * np.sum(d * q, axis=-1)
        ok &= ~((v < 0.0) | (u + v > 1.0))
        t = f * np.sum(e2 * q, axis=-1)
    return ok, t


def los_blocked(origins, targets, tris, tmin=1e-12, recip=False):
    """A triangle with tmin < t < |target - origin| on the unit ray (blocked [R]).
    The direction is the difference divided by the range (LOS queries) or
    multiplied by its reciprocal (recip=True, the weight kernel)."""
    origins = np.asarray(origins, dtype=np.float64).reshape(-1, 3)
    targets = np.asarray(targets, dtype=np.float64).reshape(-1, 3)
    tris = np.asarray(tris, dtype=np.float64).reshape(-1, 3, 3)
    R = origins.shape[0]
    if tris.shape[0] == 0:
        return np.zeros(R, dtype=bool)
    dvec = targets - origins
    dist = np.sqrt(np.sum(dvec * dvec, axis=1))
    with np.errstate(divide="ignore", invalid="ignore"):
        d = dvec * (1.0 / dist)[:, None] if recip else dvec / dist[:, None]
    out = np.zeros(R, dtype=bool)
    for c in range(0, R, 4096):
        hit, t = mt_hit(origins[c:c + 4096], d[c:c + 4096], tris)
        dd = dist[c:c + 4096, None]
        out[c:c + 4096] = np.any(hit & (t > tmin) & (t < dd), axis=1)
    return out


def los_check(rx, sat, tris, batch=False):
    """is_los for receivers rx [E, 3] and satellites sat [E, S, 3]; a receiver at the
    satellite (range < 1e-15) is LOS; in batch mode a non-finite entry is not LOS."""
    rx = np.asarray(rx, dtype=np.float64).reshape(-1, 3)
    sat = np.asarray(sat, dtype=np.float64).reshape(rx.shape[0], -1, 3)
    E, S = sat.shape[0], sat.shape[1]
    o = np.repeat(rx, S, axis=0)
    s = sat.reshape(-1, 3)
    fin = np.all(np.isfinite(o), axis=1) & np.all(np.isfinite(s), axis=1)
    dist = np.sqrt(np.sum((s - o) ** 2, axis=1))
    good = fin & ~(dist < 1e-15)
    los = np.ones(E * S, dtype=bool)
    if good.any():
        los[good] = ~los_blocked(o[good], s[good], tris)
    if batch:
        los[~fin] = False
    return los.reshape(E, S)


def multipath(rx, sat, tris, batch=False):
    """Smallest positive excess delay of a specular reflection (receiver mirrored in
    each triangle's plane, mirror-to-satellite ray crossing the triangle); returns
    (reflection points [E, S, 3], delays [E, S]); zeros when there is none."""
    rx = np.asarray(rx, dtype=np.float64).reshape(-1, 3)
    sat = np.asarray(sat, dtype=np.float64).reshape(rx.shape[0], -1, 3)
    tris = np.asarray(tris, dtype=np.float64).reshape(-1, 3, 3)
    E, S = sat.shape[0], sat.shape[1]
    refl = np.zeros((E, S, 3))
    delay = np.zeros((E, S))
    for e in range(E):
        p = rx[e]
        for s in range(S):
            q = sat[e, s]
            if batch and not (np.all(np.isfinite(p)) and np.all(np.isfinite(q))):
                continue
            d_rs = np.sqrt(np.sum((p - q) ** 2))
            best, found = 0.0, False
            for tri in tris:
                e1, e2 = tri[1] - tri[0], tri[2] - tri[0]
                n = np.cross(e1, e2)
                nn = np.sqrt(np.sum(n * n))
                if nn < 1e-15:
                    continue
                n = n / nn
                h = np.sum((p - tri[0]) * n)
                m = p - 2.0 * h * n
                dv = q - m
                dl = np.sqrt(np.sum(dv * dv))
                if dl < 1e-15:
                    continue
                dv = dv / dl
                hit, t = mt_hit(m[None], dv[None], tri[None])
                t = t[0, 0]
                if not hit[0, 0] or not t > 1e-12:
                    continue
                r = m + t * dv
                ex = (np.sqrt(np.sum((p - r) ** 2)) + np.sqrt(np.sum((r - q) ** 2))) - d_rs
                if ex > 0.0 and (not found or ex < best):
                    best, found = ex, True
                    refl[e, s] = r
            delay[e, s] = best
    return refl, delay


def mixture(los, nlos, p):
    """mixed_log_likelihood (weight_3d_bvh.cu:59-72)."""
    if p <= 0.0:
        return los
    if p >= 1.0:
        return nlos
    return np.logaddexp(np.log(1.0 - p) + los, np.log(p) + nlos)


def weight_3d(px, py, pz, pcb, sat, pr, ws, tris, sigma_los, sigma_nlos, nlos_bias,
              blocked_nlos_prob=1.0, clear_nlos_prob=0.0):
    """Log-weights [N] (weight_3d_bvh.cu:157-219), satellites summed in order."""
    pos = np.stack([np.asarray(a, dtype=np.float64) for a in (px, py, pz)], axis=1)
    cb = np.asarray(pcb, dtype=np.float64)
    sat = np.asarray(sat, dtype=np.float64).reshape(-1, 3)
    lw = np.zeros(pos.shape[0])
    for s in range(sat.shape[0]):
        dv = sat[s][None, :] - pos
        dist = np.sqrt(np.sum(dv * dv, axis=1))
        blocked = los_blocked(pos, np.broadcast_to(sat[s], pos.shape), tris, tmin=1e-6,
                              recip=True)
        res = pr[s] - (dist + cb)
        los = -0.5 * ws[s] * res * res / (sigma_los * sigma_los)
        rn = np.where(res > 0.0, res - nlos_bias, res)
        nlos = -0.5 * ws[s] * rn * rn / (sigma_nlos * sigma_nlos)
        lw += np.where(blocked, mixture(los, nlos, blocked_nlos_prob),
                       mixture(los, nlos, clear_nlos_prob))
    return lw
