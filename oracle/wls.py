"""Binary64 restatement of the reference WLS solvers (src/positioning/wls.cu).

wls_position_ref : wls.cu:10-123   (CPU, no Sagnac, singular -> current estimate)
wls_batch_ref    : wls.cu:126-241  (GPU kernel, Sagnac, pivot < 1e-15 skipped)
wls_multi_gnss_ref       : multi_gnss.cu:50-153  (CPU, no Sagnac, 3 + n_systems states)
wls_multi_gnss_batch_ref : multi_gnss.cu:211-305 (GPU kernel, Sagnac)
All keep the reference's sequential summation over satellites.  The batch
version is vectorised over epochs; the loop over satellites and the Gaussian
elimination run in the reference's order.
"""

from __future__ import annotations

import numpy as np

WGS84_A = 6378137.0
OMEGA_E = 7.2921151467e-5
C_LIGHT = 299792458.0


def _los(sx, sy, sz, x, y, z, sagnac):
    if sagnac:
        dx0, dy0, dz0 = x - sx, y - sy, z - sz
        ra = np.sqrt(dx0 * dx0 + dy0 * dy0 + dz0 * dz0)
        th = OMEGA_E * (ra / C_LIGHT)
        ct, st = np.cos(th), np.sin(th)
        sxr = sx * ct + sy * st
        syr = -sx * st + sy * ct
        dx, dy, dz = x - sxr, y - syr, z - sz
        r = np.sqrt(dx * dx + dy * dy + dz * dz)
        r = np.maximum(r, 1e-12)
    else:
        dx, dy, dz = x - sx, y - sy, z - sz
        r = np.sqrt(dx * dx + dy * dy + dz * dz)
    return dx, dy, dz, r


def _solve4(A, cpu_rule):
    """Gaussian elimination with partial pivoting, A [E,4,5]; returns (x [E,4], ok [E])."""
    A = A.copy()
    E = A.shape[0]
    ok = np.ones(E, dtype=bool)
    rows = np.arange(E)
    for col in range(4):
        piv = col + np.argmax(np.abs(A[:, col:, col]), axis=1)
        tmp = A[rows, col, :].copy()
        A[rows, col, :] = A[rows, piv, :]
        A[rows, piv, :] = tmp
        tiny = np.abs(A[:, col, col]) < 1e-15
        if cpu_rule:
            ok &= ~tiny
        for row in range(col + 1, 4):
            with np.errstate(divide="ignore", invalid="ignore"):
                fac = A[:, row, col] / A[:, col, col]
            upd = A[:, row, col:] - fac[:, None] * A[:, col, col:]
            apply = ~tiny if not cpu_rule else ok
            A[:, row, col:] = np.where(apply[:, None], upd, A[:, row, col:])
    x = np.zeros((E, 4))
    for row in range(3, -1, -1):
        v = A[:, row, 4].copy()
        for c in range(row + 1, 4):
            v = v - A[:, row, c] * x[:, c]
        with np.errstate(divide="ignore", invalid="ignore"):
            x[:, row] = v / A[:, row, row]
    return x, ok


def wls_solve(sat, pr, w, n_sat=None, max_iter=10, tol=1e-4, sagnac=True, cpu_rule=False):
    """sat [E,S,3], pr/w [E,S]; returns (results [E,4], iters [E])."""
    sat = np.asarray(sat, dtype=np.float64)
    pr = np.asarray(pr, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    E, S = pr.shape
    ns = np.full(E, S, dtype=np.int64) if n_sat is None else np.asarray(n_sat, dtype=np.int64)
    res = np.zeros((E, 4))
    iters = np.full(E, -1, dtype=np.int32)
    valid = ns >= 4
    live = valid.copy()
    mask = np.arange(S)[None, :] < ns[:, None]
    cx = np.zeros(E); cy = np.zeros(E); cz = np.zeros(E)
    for s in range(S):
        m = mask[:, s]
        cx = np.where(m, cx + sat[:, s, 0], cx)
        cy = np.where(m, cy + sat[:, s, 1], cy)
        cz = np.where(m, cz + sat[:, s, 2], cz)
    nsf = np.maximum(ns, 1).astype(np.float64)
    cx, cy, cz = cx / nsf, cy / nsf, cz / nsf
    cn = np.sqrt(cx * cx + cy * cy + cz * cz)
    if cpu_rule:
        live &= cn >= 1e-6
    with np.errstate(divide="ignore", invalid="ignore"):
        scale = WGS84_A / cn
    x, y, z = cx * scale, cy * scale, cz * scale
    cb = np.zeros(E)
    for s in range(S):
        dx, dy, dz, r = _los(sat[:, s, 0], sat[:, s, 1], sat[:, s, 2], x, y, z, sagnac)
        cb = np.where(mask[:, s], cb + (pr[:, s] - r), cb)
    cb = cb / nsf
    it = np.zeros(E, dtype=np.int32)
    active = live.copy()
    for k in range(max_iter):
        if not active.any():
            break
        HtWH = np.zeros((E, 4, 4))
        HtWy = np.zeros((E, 4))
        for s in range(S):
            m = mask[:, s]
            dx, dy, dz, r = _los(sat[:, s, 0], sat[:, s, 1], sat[:, s, 2], x, y, z, sagnac)
            resid = pr[:, s] - (r + cb)
            ws = w[:, s]
            H = np.stack([dx / r, dy / r, dz / r, np.ones(E)], axis=1)
            for a in range(4):
                HtWy[:, a] = np.where(m, HtWy[:, a] + H[:, a] * ws * resid, HtWy[:, a])
                for b in range(4):
                    HtWH[:, a, b] = np.where(m, HtWH[:, a, b] + H[:, a] * ws * H[:, b],
                                             HtWH[:, a, b])
        A = np.concatenate([HtWH, HtWy[:, :, None]], axis=2)
        d, ok = _solve4(A, cpu_rule)
        stop_singular = active & ~ok
        upd = active & ok
        x = np.where(upd, x + d[:, 0], x)
        y = np.where(upd, y + d[:, 1], y)
        z = np.where(upd, z + d[:, 2], z)
        cb = np.where(upd, cb + d[:, 3], cb)
        nrm = np.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2 + d[:, 2] ** 2 + d[:, 3] ** 2)
        conv = upd & (nrm < tol)
        it = np.where(conv, k + 1, it)
        it = np.where(stop_singular, k, it)
        it = np.where(upd & ~conv, k + 1, it)
        active = upd & ~conv
    res[:, 0], res[:, 1], res[:, 2], res[:, 3] = x, y, z, cb
    iters[live] = it[live]
    if cpu_rule:
        res[~live] = 0.0
    else:
        res[~valid] = np.nan
    return res, iters


def wls_position_ref(sat_ecef, pr, w, max_iter=10, tol=1e-4):
    """wls.cu:10-123: returns (result[4], iters) with -1 for < 4 satellites."""
    pr = np.asarray(pr, dtype=np.float64).ravel()
    S = pr.size
    sat = np.asarray(sat_ecef, dtype=np.float64).reshape(1, S, 3)
    res, it = wls_solve(sat, pr[None, :], np.asarray(w, dtype=np.float64).reshape(1, S),
                        None, max_iter, tol, sagnac=False, cpu_rule=True)
    return res[0], int(it[0])


def wls_batch_ref(sat, pr, w, max_iter=10, tol=1e-4, n_sat=None):
    """wls.cu:126-241."""
    return wls_solve(sat, pr, w, n_sat, max_iter, tol, sagnac=True, cpu_rule=False)


# ----------------------------------------------------------- multi-GNSS ----
MAX_SYSTEMS = 5  # include/gnss_gpu/multi_gnss.h:5


def _solve_n(A):
    """solve_NxN_cpu / solve_NxN_device (multi_gnss.cu:15-47, 157-187) on
    A [E, n, n+1]: partial pivoting, a pivot below 1e-15 skips its column,
    back substitution sets the unknown of a pivot <= 1e-15 to zero."""
    A = A.copy()
    E, n = A.shape[0], A.shape[1]
    rows = np.arange(E)
    for col in range(n):
        piv = np.full(E, col)
        for row in range(col + 1, n):
            piv = np.where(np.abs(A[rows, row, col]) > np.abs(A[rows, piv, col]), row, piv)
        tmp = A[rows, col, :].copy()
        A[rows, col, :] = A[rows, piv, :]
        A[rows, piv, :] = tmp
        keep = ~(np.abs(A[:, col, col]) < 1e-15)
        for row in range(col + 1, n):
            with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
                fac = A[:, row, col] / A[:, col, col]
                upd = A[:, row, col:] - fac[:, None] * A[:, col, col:]
            A[:, row, col:] = np.where(keep[:, None], upd, A[:, row, col:])
    x = np.zeros((E, n))
    for row in range(n - 1, -1, -1):
        v = A[:, row, n].copy()
        for c in range(row + 1, n):
            v = v - A[:, row, c] * x[:, c]
        p = A[:, row, row]
        with np.errstate(divide="ignore", invalid="ignore"):
            x[:, row] = np.where(np.abs(p) > 1e-15, v / p, 0.0)
    return x


def wls_multi_gnss_solve(sat, pr, w, sys_ids, n_systems, max_iter=10, tol=1e-4,
                         sagnac=True, cpu_rule=False):
    """Multi-GNSS Gauss-Newton WLS with one clock per system.

    sat [E,S,3], pr/w [E,S], sys_ids int [E,S]; returns (results [E, 3+n_systems],
    iters [E]).  sagnac/cpu_rule=(True, False) is wls_multi_gnss_batch_kernel
    (multi_gnss.cu:211-305): Sagnac ranges with r >= 1e-12 (:189-209), no
    guards.  (False, True) is the CPU wls_multi_gnss (multi_gnss.cu:50-153):
    plain ranges, zeros and -1 for S < 3 + n_systems or a centroid below 1e-6 m.
    A system id outside [0, n_systems) adds no clock column and is predicted
    with clock 0's bias (:111, :120); a system with no satellites keeps a
    zero column, whose unknown the solver sets to 0.
    """
    sat = np.asarray(sat, dtype=np.float64)
    pr = np.asarray(pr, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    sid = np.asarray(sys_ids, dtype=np.int64)
    E, S = pr.shape
    ns = int(n_systems)
    n = 3 + ns
    res = np.zeros((E, n))
    iters = np.full(E, -1, dtype=np.int32)
    if cpu_rule and S < n:
        return res, iters
    valid = (sid >= 0) & (sid < ns)
    cx = np.zeros(E); cy = np.zeros(E); cz = np.zeros(E)
    for s in range(S):
        cx = cx + sat[:, s, 0]
        cy = cy + sat[:, s, 1]
        cz = cz + sat[:, s, 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        cx, cy, cz = cx / S, cy / S, cz / S
        cn = np.sqrt(cx * cx + cy * cy + cz * cz)
        scale = WGS84_A / cn
        x, y, z = cx * scale, cy * scale, cz * scale
    live = cn >= 1e-6 if cpu_rule else np.ones(E, dtype=bool)
    cb = np.zeros((E, max(ns, 1)))
    cnt = np.zeros((E, max(ns, 1)), dtype=np.int64)
    for s in range(S):
        dx, dy, dz, r = _los(sat[:, s, 0], sat[:, s, 1], sat[:, s, 2], x, y, z, sagnac)
        for k in range(ns):
            m = valid[:, s] & (sid[:, s] == k)
            cb[:, k] = np.where(m, cb[:, k] + (pr[:, s] - r), cb[:, k])
            cnt[:, k] += m
    cb = np.where(cnt > 0, cb / np.maximum(cnt, 1), cb)
    it = np.full(E, max_iter if max_iter > 0 else 0, dtype=np.int32)
    active = live.copy()
    for k in range(max_iter):
        if not active.any():
            break
        HtWH = np.zeros((E, n, n))
        HtWy = np.zeros((E, n))
        for s in range(S):
            dx, dy, dz, r = _los(sat[:, s, 0], sat[:, s, 1], sat[:, s, 2], x, y, z, sagnac)
            ss = sid[:, s]
            vs = valid[:, s]
            cbs = np.where(vs, cb[np.arange(E), np.clip(ss, 0, max(ns - 1, 0))],
                           cb[:, 0] if ns > 0 else 0.0)
            resid = pr[:, s] - (r + cbs)
            ws = w[:, s]
            H = np.zeros((E, n))
            H[:, 0], H[:, 1], H[:, 2] = dx / r, dy / r, dz / r
            for j in range(ns):
                H[:, 3 + j] = np.where(vs & (ss == j), 1.0, 0.0)
            for a in range(n):
                HtWy[:, a] += H[:, a] * ws * resid
                for b in range(n):
                    HtWH[:, a, b] += H[:, a] * ws * H[:, b]
        d = _solve_n(np.concatenate([HtWH, HtWy[:, :, None]], axis=2))
        upd = active
        x = np.where(upd, x + d[:, 0], x)
        y = np.where(upd, y + d[:, 1], y)
        z = np.where(upd, z + d[:, 2], z)
        for j in range(ns):
            cb[:, j] = np.where(upd, cb[:, j] + d[:, 3 + j], cb[:, j])
        nrm = np.sqrt(np.sum(d * d, axis=1))
        conv = upd & (nrm < tol)
        it = np.where(conv, k + 1, it)
        active = upd & ~conv
    res[:, 0], res[:, 1], res[:, 2] = x, y, z
    res[:, 3:] = cb[:, :ns]
    iters[live] = it[live]
    res[~live] = 0.0
    return res, iters


def wls_multi_gnss_ref(sat_ecef, pr, w, sys_ids, n_systems, max_iter=10, tol=1e-4):
    """multi_gnss.cu:50-153 (CPU single epoch): returns (result[3+n_systems], iters)."""
    pr = np.asarray(pr, dtype=np.float64).ravel()
    S = pr.size
    sat = np.asarray(sat_ecef, dtype=np.float64).reshape(1, S, 3)
    res, it = wls_multi_gnss_solve(sat, pr[None, :], np.asarray(w, dtype=np.float64).reshape(1, S),
                                   np.asarray(sys_ids).reshape(1, S), n_systems, max_iter, tol,
                                   sagnac=False, cpu_rule=True)
    return res[0], int(it[0])


def wls_multi_gnss_batch_ref(sat, pr, w, sys_ids, n_systems, max_iter=10, tol=1e-4):
    """multi_gnss.cu:211-305 (GPU batch kernel)."""
    return wls_multi_gnss_solve(sat, pr, w, sys_ids, n_systems, max_iter, tol,
                                sagnac=True, cpu_rule=False)
