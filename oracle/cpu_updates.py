"""ctypes wrappers of the C restatement of the observation updates
(cpu_updates.c): the CPU baseline of the update kernels.  TEST/MEASUREMENT
INFRASTRUCTURE ONLY; the product never imports it.  Arguments and results are
those of the NumPy oracle functions of the same names in oracle/pf.py."""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import build_cpu

_lib = None


def _l():
    global _lib
    if _lib is None:
        _lib = build_cpu.load()
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def threads() -> int:
    return int(_l().cpu_updates_threads())


def _check(m):
    if m > 256:
        raise ValueError("at most 256 observations")


def dd_pseudorange_loglik(pos, sat_k, sat_ref, dd, br_k, br_ref, w, sigma, thr=0.0, huber=False,
                          huber_k=1.5, n_threads=0):
    pos, sat_k, sat_ref = _d(pos), _d(sat_k), _d(sat_ref)
    dd, br_k, br_ref, w = _d(dd), _d(br_k), _d(br_ref), _d(w)
    n, m = pos.shape[0], dd.size
    _check(m)
    out = np.empty(n)
    _l().cpu_dd_pr_loglik(n, _p(pos), m, _p(sat_k), _p(sat_ref), _p(dd), _p(br_k), _p(br_ref),
                          _p(w), sigma, thr, int(huber), huber_k, _p(out), n_threads)
    return out


def carrier_afv_loglik(pos, cb, sat, cp, w, wavelength, sigma_cycles, n_threads=0):
    pos, cb, sat, cp, w = _d(pos), _d(cb), _d(sat), _d(cp), _d(w)
    n, m = pos.shape[0], cp.size
    _check(m)
    out = np.empty(n)
    _l().cpu_carrier_afv_loglik(n, _p(pos), _p(cb), m, _p(sat), _p(cp), _p(w), wavelength,
                                sigma_cycles, _p(out), n_threads)
    return out


def dd_carrier_afv_loglik(pos, sat_k, sat_ref, ddc, br_k, br_ref, w, wl, sigma_cycles, thr=0.0,
                          huber=False, huber_k=1.5, n_threads=0):
    pos, sat_k, sat_ref = _d(pos), _d(sat_k), _d(sat_ref)
    ddc, br_k, br_ref, w, wl = _d(ddc), _d(br_k), _d(br_ref), _d(w), _d(wl)
    n, m = pos.shape[0], ddc.size
    _check(m)
    out = np.empty(n)
    _l().cpu_dd_carrier_afv_loglik(n, _p(pos), m, _p(sat_k), _p(sat_ref), _p(ddc), _p(br_k),
                                   _p(br_ref), _p(w), _p(wl), sigma_cycles, thr, int(huber),
                                   huber_k, _p(out), n_threads)
    return out


def doppler_update(pos, vel, sat, sv, dop, w, wl, sigma, gain, max_dv, n_threads=0):
    pos, sat, sv, dop, w = _d(pos), _d(sat), _d(sv), _d(dop), _d(w)
    v = _d(vel).copy()
    n, m = pos.shape[0], dop.size
    _check(m)
    out = np.empty(n)
    _l().cpu_doppler_update(n, _p(pos), _p(v), m, _p(sat), _p(sv), _p(dop), _p(w), wl, sigma,
                            gain, max_dv, _p(out), n_threads)
    return out, v


def doppler_kf_update(pos, vel, cov, sat, sv, dop, w, wl, sigma, n_threads=0):
    pos, sat, sv, dop, w = _d(pos), _d(sat), _d(sv), _d(dop), _d(w)
    v = _d(vel).copy()
    P = _d(np.asarray(cov).reshape(-1, 9)).copy()
    n, m = pos.shape[0], dop.size
    _check(m)
    out = np.empty(n)
    _l().cpu_doppler_kf_update(n, _p(pos), _p(v), _p(P), m, _p(sat), _p(sv), _p(dop), _p(w), wl,
                               sigma, _p(out), n_threads)
    return out, v, P.reshape(-1, 3, 3)
