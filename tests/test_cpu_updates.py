"""The C restatement of the observation updates (oracle/cpu_updates.c, the
timed CPU baseline of the update kernels) agrees with the NumPy oracle
(oracle/pf.py) on random particle clouds, including the gate and Huber
variants of the DD kinds.  CPU only."""

import numpy as np
import pytest

from oracle import cpu_updates as cu
from oracle import pf as opf
from gnss_gpu_amd import sim


def _setup(seed=5, n=400, S=9):
    rng = np.random.default_rng(seed)
    sc = sim.scenario_plain(n_epochs=1, n_sat=S, seed=seed)
    sat, tp = sc.sat, sc.truth[0]
    pos = tp + rng.normal(0, 30.0, (n, 3))
    cb = 50.0 + rng.normal(0, 3.0, n)
    base = tp + np.array([400.0, -300.0, 80.0])
    rr = np.linalg.norm(sat - tp, axis=1)
    br = np.linalg.norm(sat - base, axis=1)
    ks = np.arange(1, S)
    dd = (rr[ks] - rr[0]) - (br[ks] - br[0]) + rng.normal(0, 2.0, S - 1)
    dd[2] += 60.0  # one pair beyond the gate
    wl = np.full(S - 1, 0.190293673)
    ddc = dd / wl + rng.normal(0, 0.1, S - 1)
    w = rng.uniform(0.5, 1.5, S - 1)
    return rng, sat, tp, pos, cb, br, ks, dd, ddc, wl, w, rr


@pytest.mark.parametrize("thr,huber", [(0.0, False), (20.0, False), (0.0, True), (20.0, True)])
def test_dd_kinds_match_oracle(thr, huber):
    rng, sat, tp, pos, cb, br, ks, dd, ddc, wl, w, rr = _setup()
    sk, sref = sat[ks], np.repeat(sat[[0]], len(ks), axis=0)
    a = cu.dd_pseudorange_loglik(pos, sk, sref, dd, br[ks], np.repeat(br[0], len(ks)), w, 0.8,
                                 thr, huber, 1.5)
    b = opf.dd_pseudorange_loglik(pos, sk, sref, dd, br[ks], np.repeat(br[0], len(ks)), w, 0.8,
                                  thr, huber, 1.5)
    np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-9)
    thr_c = 0.3 if thr else 0.0
    a = cu.dd_carrier_afv_loglik(pos, sk, sref, ddc, br[ks], np.repeat(br[0], len(ks)), w, wl,
                                 0.05, thr_c, huber, 1.5)
    b = opf.dd_carrier_afv_loglik(pos, sk, sref, ddc, br[ks], np.repeat(br[0], len(ks)), w, wl,
                                  0.05, thr_c, huber, 1.5)
    np.testing.assert_allclose(a, b, rtol=1e-7, atol=1e-6)


def test_carrier_afv_matches_oracle():
    rng, sat, tp, pos, cb, br, ks, dd, ddc, wl, w, rr = _setup(seed=6)
    cp = (rr + 50.0) / 0.190293673 + rng.normal(0, 0.05, len(rr)) + 1234.0
    ws = rng.uniform(0.5, 1.5, len(rr))
    a = cu.carrier_afv_loglik(pos, cb, sat, cp, ws, 0.190293673, 0.05)
    b = opf.carrier_afv_loglik(pos, cb, sat, cp, ws, 0.190293673, 0.05)
    np.testing.assert_allclose(a, b, rtol=1e-7, atol=1e-6)


def _doppler_setup(seed=7, n=300, S=9):
    rng = np.random.default_rng(seed)
    sc = sim.scenario_plain(n_epochs=1, n_sat=S, seed=seed)
    sat, tp = sc.sat, sc.truth[0]
    pos = tp + rng.normal(0, 20.0, (n, 3))
    R = sim.enu_rotation(tp)
    v0 = np.array([1.0, 0.5, -0.2]) @ R
    vel = v0 + rng.normal(0, 0.7, (n, 3))
    sv = np.cross(sat, R[2])
    sv = 3000.0 * sv / np.linalg.norm(sv, axis=1)[:, None]
    rr = np.linalg.norm(sat - tp, axis=1)
    losv = (sat - tp) / rr[:, None]
    dop = -np.sum((sv - v0) * losv, axis=1) / 0.19029367279836488 + rng.normal(0, 0.3, S)
    w = rng.uniform(0.5, 1.5, S)
    A = rng.normal(0, 1.0, (n, 3, 3))
    cov = np.einsum("nij,nkj->nik", A, A) + 0.5 * np.eye(3)
    return pos, vel, cov, sat, sv, dop, w


def test_doppler_update_matches_oracle():
    pos, vel, cov, sat, sv, dop, w = _doppler_setup()
    a, va = cu.doppler_update(pos, vel, sat, sv, dop, w, 0.19029367279836488, 0.5, 0.25, 10.0)
    b, vb = opf.doppler_update(pos, vel, sat, sv, dop, w, 0.19029367279836488, 0.5, 0.25, 10.0)
    np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-8)
    np.testing.assert_allclose(va, vb, rtol=0, atol=1e-8)


def test_doppler_kf_update_matches_oracle():
    pos, vel, cov, sat, sv, dop, w = _doppler_setup(seed=8, n=120)
    a, va, Pa = cu.doppler_kf_update(pos, vel, cov, sat, sv, dop, w, 0.19029367279836488, 0.5)
    b, vb, Pb = opf.doppler_kf_update(pos, vel, cov, sat, sv, dop, w, 0.19029367279836488, 0.5)
    np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(va, vb, rtol=0, atol=1e-9)
    np.testing.assert_allclose(Pa, Pb, rtol=1e-9, atol=1e-12)
