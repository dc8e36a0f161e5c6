"""The legacy host-array path on the GPU: ``_gnss_gpu_pf`` (in-place functions
over the caller's float64 arrays, python/gnss_gpu/_particle_filter_bindings.cpp:
10-145) and ``ParticleFilter`` (python/gnss_gpu/particle_filter.py:31-222)
against the binary64 oracle.

Bars: the caller's arrays are the ones written (in place); positions within
1 mm of the oracle (the arrays cross into binary32 offsets on every call);
log-weights 1e-5 relative; ESS 1e-6 relative; resampled arrays equal the
oracle's particles at its ancestors (systematic with the reference offset
u0 = (seed mod 1e6) / (1e6 N); Megopolis with the reference rule, 15
iterations, the same keys every call), identified exactly because the
particles lie metres apart.
"""

import json
import os

import numpy as np
import pytest

from oracle import detmath
from oracle import pf as opf

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_cases.json")))
SEED = 42


def _tokyo():
    g = GOLD["tokyo_constellation"]
    return np.array(g["true_pos"]), np.array(g["sat_ecef"])


def _arrays(n):
    return [np.full(n, np.nan) for _ in range(4)]


def test_functions_write_the_callers_arrays(gpu):
    from gnss_gpu_amd import _gnss_gpu_pf as lp

    n = 5000
    truth, sat = _tokyo()
    px, py, pz, pcb = arrs = _arrays(n)
    ids = [id(a) for a in arrs]
    assert lp.pf_initialize(px, py, pz, pcb, *truth, 3000.0, 30.0, 10.0, n, SEED) is None
    assert [id(a) for a in (px, py, pz, pcb)] == ids
    assert np.all(np.isfinite(np.stack(arrs)))
    orc = opf.OraclePF(n, seed=SEED)
    orc.initialize(truth, 3000.0, 30.0, 10.0)
    np.testing.assert_allclose(np.stack(arrs, 1), orc.state, rtol=0, atol=1e-4)
    before = np.stack(arrs, 1).copy()
    vx, vy, vz = np.array([2.0]), np.array([-1.0]), np.array([0.5])
    lp.pf_predict(px, py, pz, pcb, vx, vy, vz, 1.0, 0.0, 0.0, n, SEED, 1)
    np.testing.assert_allclose(np.stack(arrs, 1)[:, :3] - before[:, :3],
                               np.tile([2.0, -1.0, 0.5], (n, 1)), rtol=0, atol=1e-4)
    out = np.empty(4 * n)
    lp.pf_get_particles(px, py, pz, pcb, out, n)
    np.testing.assert_array_equal(out.reshape(n, 4), np.stack(arrs, 1))


def test_functions_match_oracle(gpu):
    """initialize, three predicts, weight, ESS, estimate, both resamplers."""
    from gnss_gpu_amd import _gnss_gpu_pf as lp

    n = 20_000
    truth, sat = _tokyo()
    rng = np.random.default_rng(0)
    pr = np.linalg.norm(sat - truth, axis=1) + 3000.0 + rng.normal(0, 3, len(sat))
    px, py, pz, pcb = arrs = _arrays(n)
    lp.pf_initialize(px, py, pz, pcb, *truth, 3000.0, 30.0, 10.0, n, SEED)
    orc = opf.OraclePF(n, seed=SEED)
    orc.initialize(truth, 3000.0, 30.0, 10.0)
    v = np.array([1.0, -2.0, 0.3])
    for step in (1, 2, 3):
        lp.pf_predict(px, py, pz, pcb, v[:1], v[1:2], v[2:], 1.0, 1.0, 3.0, n, SEED, step)
        orc.predict(v, 1.0, 1.0, 3.0, step)
        np.testing.assert_allclose(np.stack(arrs, 1), orc.state, rtol=0, atol=1e-3)
    lw = np.full(n, np.nan)
    lp.pf_weight(px, py, pz, pcb, sat.ravel(), pr, np.ones(len(pr)), lw, n, len(pr), 5.0)
    state = np.stack(arrs, 1)
    ref = opf.log_likelihood(state[:, :3], state[:, 3], np.zeros((n, 0)), sat, pr,
                             np.ones(len(pr)), None, 5.0)
    np.testing.assert_allclose(lw, ref, rtol=1e-5, atol=1e-5)
    assert lp.pf_compute_ess(lw, n) == pytest.approx(opf.ess(lw), rel=1e-6)
    res = np.full(4, np.nan)
    lp.pf_estimate(px, py, pz, pcb, lw, res, n)
    np.testing.assert_allclose(res, opf.weighted_mean(state, lw), rtol=0, atol=1e-3)
    lw32 = lw.astype(np.float32)
    # systematic: the reference's deterministic offset, the same every call
    anc = detmath.systematic_ancestors(lw32, n, detmath.reference_u0(SEED, n))
    s_arrs = [a.copy() for a in arrs]
    lp.pf_resample_systematic(*s_arrs, lw, n, SEED)
    np.testing.assert_allclose(np.stack(s_arrs, 1), state[anc], rtol=0, atol=1e-3)
    # Megopolis with the reference rule and the wrapper's 15 iterations
    anc = detmath.megopolis_ref_ancestors(lw32, n, 15, SEED)
    m_arrs = [a.copy() for a in arrs]
    lp.pf_resample_megopolis(*m_arrs, lw, n, 15, SEED)
    np.testing.assert_allclose(np.stack(m_arrs, 1), state[anc], rtol=0, atol=1e-3)
    # the log-weights are the caller's and stay as they were
    np.testing.assert_array_equal(lw.astype(np.float32), lw32)


@pytest.mark.parametrize("resampling", ["systematic", "megopolis"])
def test_particle_filter_class_matches_oracle(gpu, resampling):
    """ParticleFilter over 25 epochs of the configuration-1 trajectory
    (10,000 particles, 8 GPS satellites, Gaussian 5 m, resampling at ESS < N/2)
    beside OraclePF fed the filter's own binary32 log-weights."""
    from gnss_gpu_amd import ParticleFilter, sim

    sc = sim.scenario_plain(n_epochs=25, seed=0)
    n = 10_000
    pf = ParticleFilter(n_particles=n, sigma_pos=1.0, sigma_cb=3.0, sigma_pr=5.0,
                        resampling=resampling, ess_threshold=0.5, seed=SEED)
    pf.initialize(sc.truth[0], clock_bias=sc.cb, spread_pos=10.0, spread_cb=10.0)
    orc = opf.OraclePF(n, seed=SEED)
    orc.initialize(sc.truth[0], sc.cb, 10.0, 10.0)
    captured = []
    plain_resample = pf._resample

    def resample_capturing():  # the log-weights the resampler consumes
        captured.append(pf._log_weights.copy())
        plain_resample()

    pf._resample = resample_capturing
    resampled = 0
    for t in range(sc.n_epochs):
        if t > 0:
            pf.predict(velocity=sc.vel, dt=1.0)
            orc.predict(sc.vel, 1.0, 1.0, 3.0, t)
        np.testing.assert_allclose(pf.get_particles(), orc.state, rtol=0, atol=1e-3)
        state = pf.get_particles()
        n_cap = len(captured)
        pf.update(sc.sat, sc.pr[t])
        orc.weight(sc.sat, sc.pr[t], sigma_pr=5.0)
        ref = opf.log_likelihood(state[:, :3], state[:, 3], np.zeros((n, 0)), sc.sat, sc.pr[t],
                                 np.ones(sc.n_sat), None, 5.0)
        lw = pf._log_weights
        if len(captured) > n_cap:  # resampled inside update
            resampled += 1
            np.testing.assert_allclose(captured[-1], ref, rtol=1e-5, atol=1e-5)
            assert np.all(lw == 0.0)
            lw32 = captured[-1].astype(np.float32)
            if resampling == "systematic":
                orc.resample_systematic(lw32=lw32)
            else:
                orc.resample_megopolis(15, lw32=lw32)
            np.testing.assert_allclose(pf.get_particles(), orc.state, rtol=0, atol=1e-3)
        else:
            np.testing.assert_allclose(lw, ref, rtol=1e-5, atol=1e-5)
            assert pf.get_ess() == pytest.approx(opf.ess(lw), rel=1e-6)
        np.testing.assert_allclose(pf.estimate(), orc.estimate(), rtol=0, atol=1e-3)
    assert resampled >= 3


def test_particle_filter_class_resampling_decision(gpu):
    """update resamples exactly when ESS < threshold N (particle_filter.py:158-161)."""
    from gnss_gpu_amd import ParticleFilter

    truth, sat = _tokyo()
    pr = np.linalg.norm(sat - truth, axis=1) + 3000.0
    pf = ParticleFilter(n_particles=4096, sigma_pr=5.0, resampling="systematic", seed=SEED)
    pf.initialize(truth, clock_bias=3000.0, spread_pos=0.0, spread_cb=0.0)
    pf.update(sat, pr)  # identical particles: ESS = N, no resampling
    assert pf.get_ess() == pytest.approx(4096.0, rel=1e-9)
    assert np.all(pf._log_weights == pf._log_weights[0])
    pf2 = ParticleFilter(n_particles=4096, sigma_pr=5.0, resampling="systematic", seed=SEED)
    pf2.initialize(truth, clock_bias=3000.0, spread_pos=50.0, spread_cb=50.0)
    pf2.update(sat, pr)  # spread far beyond sigma: ESS << N/2, resampled, weights reset
    assert np.all(pf2._log_weights == 0.0)
    assert pf2.get_ess() == pytest.approx(4096.0, rel=1e-9)
