"""GPU parity of the particle-filter step against the oracle.

Bar: resampling ancestors bit-exact for identical log-weights and uniforms;
log-weights within 1e-5 relative (binary32 path); positions and estimates
within 1 mm of the binary64 oracle.
"""

import json
import os

import numpy as np
import pytest

from oracle import detmath, philox
from oracle import pf as opf

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_cases.json")))
SEED = 42


def _dev():
    from gnss_gpu_amd import _gnss_gpu_pf_device as dev

    return dev


def _stream_scenario(n_sat=8):
    rng = np.random.RandomState(123)
    d = rng.randn(n_sat, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    sat = d * 20_000_000.0
    pr = np.linalg.norm(sat, axis=1) + 100.0
    return sat, pr, np.ones(n_sat)


# ------------------------------------------------ device building blocks ----
def test_device_exp_and_philox_bit_exact(gpu):
    import ctypes as C

    from gnss_gpu_amd import _lib

    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-745, 0, 5000), rng.uniform(-30, 5, 5000),
                        np.array([0.0, -1e-300, -744.9, -708.5, -745.2])])
    out = np.empty_like(x)
    _lib.call("gg_debug_math", 0, _lib.dptr(x), _lib.dptr(out), x.size)
    np.testing.assert_array_equal(out.view(np.uint64), detmath.exp64(x).view(np.uint64))
    xf = rng.uniform(-100, 10, 5000).astype(np.float32).astype(np.float64)
    out = np.empty_like(xf)
    _lib.call("gg_debug_math", 1, _lib.dptr(xf), _lib.dptr(out), xf.size)
    np.testing.assert_array_equal(out.astype(np.float32).view(np.uint32),
                                  detmath.exp32(xf.astype(np.float32)).view(np.uint32))
    idx = np.arange(1000, dtype=np.float64)
    out = np.empty(4000)
    _lib.call("gg_debug_math", 2, _lib.dptr(idx), _lib.dptr(out), 1000)
    ref = philox.draw(0x123456789ABCDEF0, np.arange(1000), 7, 3, 1)
    np.testing.assert_array_equal(out.reshape(1000, 4).astype(np.uint64), np.stack(ref, axis=1))
    out = np.empty(8000)
    _lib.call("gg_debug_math", 3, _lib.dptr(idx), _lib.dptr(out), 1000)
    st = philox.xo_seed(0x123456789ABCDEF0, np.arange(1000), 7, 3, 1)
    ref = np.stack([philox.xo_next(st) for _ in range(8)], axis=1)
    np.testing.assert_array_equal(out.reshape(1000, 8).astype(np.uint64), ref)
    _ = C


# -------------------------------------------------------- closed forms ------
def _pfd(**kw):
    from gnss_gpu_amd import ParticleFilterDevice

    args = dict(n_particles=1024, sigma_pos=1.0, sigma_cb=300.0, sigma_pr=5.0,
                resampling="systematic", seed=SEED)
    args.update(kw)
    return ParticleFilterDevice(**args)


def test_closed_form_gaussian_and_huber(gpu):
    g = GOLD["stream_closed_forms"]
    sat, pr, w = _stream_scenario(5)
    pr_out = pr.copy()
    pr_out[0] += 100.0
    vals = []
    for huber in (False, True):
        pf = _pfd(per_particle_huber=huber, per_particle_huber_undiff_pr_k=1.5)
        pf.initialize(np.zeros(3), clock_bias=100.0, spread_pos=0.0, spread_cb=0.0)
        pf.update(sat, pr_out, w, resample=False)
        lw = pf.get_log_weights()
        assert np.all(lw == lw[0])
        vals.append(float(lw[0]))
    assert vals[0] == pytest.approx(g["gaussian_log_weight"], rel=1e-6)
    assert vals[1] == pytest.approx(g["huber_log_weight"], rel=1e-6)


def test_nlos_gate(gpu):
    sat, pr, w = _stream_scenario(5)
    out = pr.copy()
    out[0] += 100.0
    pf = _pfd(per_particle_nlos_gate=True, per_particle_nlos_undiff_pr_threshold_m=30.0)
    pf.initialize(np.zeros(3), clock_bias=100.0, spread_pos=0.0, spread_cb=0.0)
    pf.update(sat, out, w, resample=False)
    assert float(np.max(np.abs(pf.get_log_weights()))) < 1e-6
    pf2 = _pfd(per_particle_nlos_gate=True, per_particle_nlos_undiff_pr_threshold_m=30.0)
    pf2.initialize(np.zeros(3), clock_bias=100.0, spread_pos=0.0, spread_cb=0.0)
    pf2.update(sat, pr + 100.0, w, resample=False)
    assert float(np.max(pf2.get_log_weights())) < -900.0


def test_predict_deterministic_velocity(gpu):
    """sigma = 0: every particle moves by v * dt (test_cuda_streams.py:239-255)."""
    dev = _dev()
    n = 2048
    st = dev.pf_device_create(n)
    dev.pf_device_initialize(st, 0.0, 0.0, 0.0, 100.0, 0.0, 0.0, SEED)
    dev.pf_device_predict(st, 2.0, -1.0, 0.5, 0.25, 0.0, 0.0, SEED, 1)
    s = dev.pf_device_get_particle_states(st)
    np.testing.assert_allclose(s[:, :3], np.tile([0.5, -0.25, 0.125], (n, 1)), atol=1e-12)
    np.testing.assert_allclose(s[:, 3], 100.0)
    np.testing.assert_allclose(s[:, 4:7], np.tile([2.0, -1.0, 0.5], (n, 1)))


# ------------------------------------------------------- oracle parity ------
def _setup(n, anchor=None, spread=30.0, spread_cb=20.0, seed=SEED, n_isb=0):
    from gnss_gpu_amd import ParticleFilterDevice

    anchor = np.array(GOLD["tokyo_constellation"]["true_pos"]) if anchor is None else anchor
    pf = ParticleFilterDevice(n_particles=n, sigma_pos=1.0, sigma_cb=3.0, sigma_pr=5.0,
                              resampling="systematic", seed=seed, n_isb=n_isb)
    pf.initialize(anchor, clock_bias=3000.0, spread_pos=spread, spread_cb=spread_cb)
    orc = opf.OraclePF(n, seed=seed, n_isb=n_isb)
    orc.initialize(anchor, 3000.0, spread, spread_cb)
    return pf, orc


def test_initialize_matches_oracle(gpu):
    pf, orc = _setup(5000)
    p = pf.get_particles()
    np.testing.assert_allclose(p, orc.state, rtol=0, atol=1e-4)
    # moments (test_particle_filter.py:51-79)
    assert abs(p[:, 0].mean() - orc.state[:, 0].mean()) < 1e-4
    assert abs(p[:, :3].std(axis=0).mean() - 30.0) < 3.0


def test_predict_matches_oracle(gpu):
    pf, orc = _setup(5000)
    v = np.array([3.0, -2.0, 1.0])
    for step in range(1, 4):
        pf.predict(velocity=v, dt=1.0)
        orc.predict(v, 1.0, 1.0, 3.0, step)
    np.testing.assert_allclose(pf.get_particles(), orc.state, rtol=0, atol=1e-3)


@pytest.mark.parametrize("kind", ["gauss", "cauchy", "huber"])
def test_weight_ess_estimate_match_oracle(gpu, kind):
    g = GOLD["tokyo_constellation"]
    sat = np.array(g["sat_ecef"])
    tp = np.array(g["true_pos"])
    pr = np.linalg.norm(sat - tp, axis=1) + g["true_cb"] + np.random.default_rng(1).normal(0, 3, 8)
    pf, orc = _setup(20000)
    if kind == "cauchy":
        pf.nu = 1.0
    if kind == "huber":
        pf.per_particle_huber = True
    pf.update(sat, pr, resample=False)
    lw = pf.get_log_weights()
    st = pf.get_particles()
    ref = opf.log_likelihood(st[:, :3], st[:, 3], np.zeros((len(st), 0)), sat, pr, np.ones(8), None,
                             5.0, nu=1.0 if kind == "cauchy" else 0.0, huber=kind == "huber")
    np.testing.assert_allclose(lw, ref, rtol=1e-5, atol=1e-5)
    # the per-particle exponentials are binary32 (like lw), summed in binary64
    assert pf.get_ess() == pytest.approx(opf.ess(lw), rel=1e-6)
    est = pf.estimate()
    np.testing.assert_allclose(est, opf.weighted_mean(st, lw), rtol=0, atol=1e-3)
    c = tp + 5.0
    assert pf.get_position_spread(c) == pytest.approx(opf.position_spread(st[:, :3], lw, c), rel=1e-6)


@pytest.mark.parametrize("w_los,mu,sig_n", [(0.7, 15.0, 30.0), (0.95, 40.0, 10.0), (1.0, 15.0, 30.0)])
def test_gmm_update_matches_oracle(gpu, w_los, mu, sig_n):
    """update_gmm (particle_filter_device.py:305-351, pf_device.cu:373-437) on the
    Tokyo constellation with two NLOS-biased satellites and per-satellite
    weights: log-weights within 1e-5 relative of the binary64 oracle at the
    device's particles, then the ESS-driven resampling of the caller's path."""
    g = GOLD["tokyo_constellation"]
    sat = np.array(g["sat_ecef"])
    tp = np.array(g["true_pos"])
    rng = np.random.default_rng(5)
    pr = np.linalg.norm(sat - tp, axis=1) + g["true_cb"] + rng.normal(0, 3, 8)
    pr[[2, 6]] += [30.0, 55.0]
    w = rng.uniform(0.5, 2.0, 8)
    pf, orc = _setup(20000)
    pf.update_gmm(sat, pr, weights=w, w_los=w_los, mu_nlos=mu, sigma_nlos=sig_n, resample=False)
    lw = pf.get_log_weights()
    st = pf.get_particles()
    ref = opf.log_likelihood_gmm(st[:, :3], st[:, 3], sat, pr, w, 5.0, w_los, mu, sig_n)
    np.testing.assert_allclose(lw, ref, rtol=1e-5, atol=1e-5)
    assert pf.get_ess() == pytest.approx(opf.ess(lw), rel=1e-6)
    np.testing.assert_allclose(pf.estimate(), opf.weighted_mean(st, lw), rtol=0, atol=1e-3)
    # the drop-in module function, then the caller's resampling
    dev = _dev()
    dev.pf_device_weight_gmm(pf._state, sat.ravel(), pr, w, 8, 5.0, w_los, mu, sig_n)
    np.testing.assert_array_equal(pf.get_log_weights(), lw)
    pf.update_gmm(sat, pr, weights=w, w_los=w_los, mu_nlos=mu, sigma_nlos=sig_n)
    assert np.all(pf.get_log_weights() == 0.0) or pf.get_ess() >= 0.5 * 20000


@pytest.mark.parametrize("n", [1, 5, 1000, 1024, 1025, 4099, 100_000, 1_000_000])
def test_systematic_ancestors_bit_exact(gpu, n):
    dev = _dev()
    st = dev.pf_device_create(n)
    dev.pf_device_initialize(st, -3957199.0, 3310205.0, 3737911.0, 3000.0, 40.0, 10.0, SEED)
    lw = np.random.default_rng(n).normal(0.0, 4.0, n)
    dev.pf_device_set_log_weights(st, lw)
    before = dev.pf_device_get_particles(st)
    dev.pf_device_resample_systematic(st, SEED)
    anc = np.empty(n, np.int32)
    dev.pf_device_get_resample_ancestors(st, anc)
    ref = detmath.systematic_ancestors(lw.astype(np.float32), n, detmath.reference_u0(SEED, n))
    np.testing.assert_array_equal(anc, ref)
    after = dev.pf_device_get_particles(st)
    np.testing.assert_allclose(after, before[anc], rtol=0, atol=1e-3)
    assert np.all(dev.pf_device_get_log_weights is not None)
    out = np.empty(n)
    dev.pf_device_get_log_weights(st, out)
    assert np.all(out == 0.0)
    assert dev.pf_device_ess(st) == pytest.approx(n)


@pytest.mark.parametrize("far", [0.0, 5.0e8])
def test_resampled_estimate_is_the_resampled_mean(gpu, far):
    """The estimate after a systematic resample is the plain mean of the
    resampled particles.  The last work item adds the items' sums as 2^-24
    fixed point; particles 5e8 m from the anchor give sums beyond that range,
    which must take the binary64 chunk sums instead, to the same mean (the
    states are re-anchored on their mean, so the two halves go to +far and
    -far)."""
    dev = _dev()
    n = 100_000
    st = dev.pf_device_create(n)
    dev.pf_device_initialize(st, -3957199.0, 3310205.0, 3737911.0, 3000.0, 40.0, 10.0, SEED)
    s = dev.pf_device_get_particle_states(st)
    s[: n // 2, 0] += far
    s[n // 2:, 0] -= far
    dev.pf_device_set_particle_states(st, s)
    lw = np.random.default_rng(3).normal(0.0, 3.0, n)
    dev.pf_device_set_log_weights(st, lw)
    dev.pf_device_resample_systematic(st, SEED)
    p = dev.pf_device_get_particles(st)
    est = dev.pf_device_estimate(st)
    mean = p.mean(axis=0)
    tol = 1e-6 + 1e-12 * np.abs(mean)  # the particles' own binary32 offsets, summed exactly on the host
    assert np.all(np.abs(est - mean) <= tol), (est - mean)


def test_resampled_estimate_propagates_nan(gpu):
    """Non-finite states make the resampled estimate non-finite: the fixed-point
    sums flag them and the binary64 chunk sums carry the NaN (one NaN state
    makes the anchor, and so every offset of that axis, NaN)."""
    dev = _dev()
    n = 20_000
    st = dev.pf_device_create(n)
    dev.pf_device_initialize(st, -3957199.0, 3310205.0, 3737911.0, 3000.0, 40.0, 10.0, SEED)
    s = dev.pf_device_get_particle_states(st)
    s[123, 0] = np.nan
    dev.pf_device_set_particle_states(st, s)
    lw = np.full(n, -50.0)
    lw[123] = 0.0
    dev.pf_device_set_log_weights(st, lw)
    dev.pf_device_resample_systematic(st, SEED)
    est = dev.pf_device_estimate(st)
    assert np.isnan(est[0]) and np.all(np.isfinite(est[3:]))


@pytest.mark.parametrize("n", [1000, 4099, 1_000_000])
@pytest.mark.parametrize("seed", [2_000_000, 2_000_001])
def test_systematic_counts_on_the_targets(gpu, n, seed):
    """Equal weights put every CDF value on a systematic target (u0 = 0 for a
    seed that is a multiple of 10^6) or within rounding of one, so every
    output count goes through the exact target comparisons instead of the
    floor(c n - u0) + 1 fast path; ancestors must still equal the oracle's."""
    dev = _dev()
    st = dev.pf_device_create(n)
    dev.pf_device_initialize(st, 0.0, 0.0, 0.0, 0.0, 10.0, 10.0, 5)
    lw = np.zeros(n)
    dev.pf_device_set_log_weights(st, lw)
    dev.pf_device_resample_systematic(st, seed)
    anc = np.empty(n, np.int32)
    dev.pf_device_get_resample_ancestors(st, anc)
    ref = detmath.systematic_ancestors(lw.astype(np.float32), n, detmath.reference_u0(seed, n))
    np.testing.assert_array_equal(anc, ref)


def test_systematic_fresh_u0_and_degenerate(gpu):
    dev = _dev()
    n = 3000
    st = dev.pf_device_create(n)
    dev.pf_device_initialize(st, 0.0, 0.0, 0.0, 0.0, 10.0, 10.0, 7)
    lw = np.full(n, -1e4)
    lw[1234] = 0.0
    dev.pf_device_set_log_weights(st, lw)
    dev.pf_device_resample_systematic(st, 7, 1)
    anc = np.empty(n, np.int32)
    dev.pf_device_get_resample_ancestors(st, anc)
    assert np.all(anc == 1234)
    lw = np.random.default_rng(3).normal(0, 2, n)
    dev.pf_device_set_log_weights(st, lw)
    dev.pf_device_resample_systematic(st, 7, 1)
    dev.pf_device_get_resample_ancestors(st, anc)
    ref = detmath.systematic_ancestors(lw.astype(np.float32), n, detmath.fresh_u0(7, 1))
    np.testing.assert_array_equal(anc, ref)


@pytest.mark.parametrize("n,iters", [(1000, 15), (70_000, 32), (5, 3)])
def test_megopolis_reference_rule_bit_exact(gpu, n, iters):
    dev = _dev()
    st = dev.pf_device_create(n)
    dev.pf_device_initialize(st, 0.0, 0.0, 0.0, 0.0, 10.0, 10.0, 11)
    lw = np.random.default_rng(n).normal(0.0, 2.0, n)
    dev.pf_device_set_log_weights(st, lw)
    before = dev.pf_device_get_particles(st)
    dev.pf_device_resample_megopolis(st, iters, 11)
    anc = np.empty(n, np.int32)
    dev.pf_device_get_resample_ancestors(st, anc)
    ref = detmath.megopolis_ref_ancestors(lw.astype(np.float32), n, iters, 11)
    np.testing.assert_array_equal(anc, ref)
    np.testing.assert_allclose(dev.pf_device_get_particles(st), before[anc], rtol=0, atol=1e-3)


@pytest.mark.parametrize("n,iters", [(1000, 15), (70_000, 32), (300_000, 64)])
def test_megopolis_chesser_bit_exact(gpu, n, iters):
    dev = _dev()
    st = dev.pf_device_create(n)
    dev.pf_device_initialize(st, 0.0, 0.0, 0.0, 0.0, 10.0, 10.0, 13)
    lw = np.random.default_rng(n + 1).normal(0.0, 2.0, n)
    dev.pf_device_set_log_weights(st, lw)
    dev.pf_device_resample_megopolis_chesser(st, iters)
    anc = np.empty(n, np.int32)
    dev.pf_device_get_resample_ancestors(st, anc)
    ref = detmath.megopolis_ancestors(lw.astype(np.float32), n, iters, 13, 0)
    np.testing.assert_array_equal(anc, ref)


def test_state_roundtrip_and_frame_shifts(gpu):
    dev = _dev()
    n = 3000
    st = dev.pf_device_create(n)
    dev.pf_device_initialize(st, -3957199.0, 3310205.0, 3737911.0, 3000.0, 50.0, 20.0, 3)
    s = dev.pf_device_get_particle_states(st)
    rng = np.random.default_rng(4)
    s[:, :3] += rng.normal(0, 5.0, (n, 3))
    s[:, 4:7] = [1.0, 2.0, 3.0]
    dev.pf_device_set_particle_states(st, s)
    s2 = dev.pf_device_get_particle_states(st)
    np.testing.assert_allclose(s2[:, :4], s[:, :4], rtol=0, atol=1e-4)
    np.testing.assert_allclose(s2[:, 4:7], s[:, 4:7], atol=1e-9)
    lw = rng.normal(0, 1, n)
    dev.pf_device_set_log_weights(st, lw)
    out = np.empty(n)
    dev.pf_device_get_log_weights(st, out)
    np.testing.assert_allclose(out, lw.astype(np.float32).astype(np.float64))
    e0 = dev.pf_device_estimate(st)
    dev.pf_device_shift_clock_bias(st, 12.5)
    dev.pf_device_shift_position(st, 1.0, -2.0, 3.0)
    e1 = dev.pf_device_estimate(st)
    np.testing.assert_allclose(e1 - e0, [1.0, -2.0, 3.0, 12.5], atol=1e-6)
    p = dev.pf_device_get_particles(st)
    np.testing.assert_allclose(p[:, 3], s[:, 3] + 12.5, atol=1e-4)


@pytest.mark.parametrize("spread", [3000.0, 50000.0])
@pytest.mark.parametrize("kind", ["gauss", "cauchy"])
def test_weight_far_particles_match_oracle(gpu, kind, spread):
    """Particles kilometres from the anchor: the range is evaluated in its
    exact form (3 km mixes series and exact threads in one wavefront, 50 km is
    exact everywhere); both against the binary64 likelihood."""
    g = GOLD["tokyo_constellation"]
    sat = np.array(g["sat_ecef"])
    tp = np.array(g["true_pos"])
    pr = np.linalg.norm(sat - tp, axis=1) + g["true_cb"] + np.random.default_rng(2).normal(0, 3, 8)
    pf, orc = _setup(20000, spread=spread)
    if kind == "cauchy":
        pf.nu = 1.0
    pf.update(sat, pr, resample=False)
    lw = pf.get_log_weights()
    st = pf.get_particles()
    far = np.linalg.norm(st[:, :3] - tp, axis=1) > 3500.0
    assert far.any() and (spread > 10000.0 or (~far).any())
    ref = opf.log_likelihood(st[:, :3], st[:, 3], np.zeros((len(st), 0)), sat, pr, np.ones(8), None,
                             5.0, nu=1.0 if kind == "cauchy" else 0.0)
    np.testing.assert_allclose(lw, ref, rtol=1e-5, atol=1e-4)


def test_many_satellites_grow_capacity(gpu):
    sat, pr, w = _stream_scenario(80)
    dev = _dev()
    st = dev.pf_device_create(10_000)
    dev.pf_device_initialize(st, 0.0, 0.0, 0.0, 100.0, 50.0, 200.0, SEED)
    dev.pf_device_weight(st, sat.ravel(), pr, w, 80, 5.0)
    est = dev.pf_device_estimate(st)
    assert np.all(np.isfinite(est))
    lw = np.empty(10_000)
    dev.pf_device_get_log_weights(st, lw)
    s = dev.pf_device_get_particles(st)
    ref = opf.log_likelihood(s[:, :3], s[:, 3], np.zeros((10_000, 0)), sat, pr, w, None, 5.0)
    np.testing.assert_allclose(lw, ref, rtol=1e-5, atol=1e-4)


def test_ancestors_api_semantics(gpu):
    from gnss_gpu_amd import ParticleFilterDevice

    pf = ParticleFilterDevice(n_particles=4096, resampling="megopolis", seed=1)
    pf.initialize(np.zeros(3))
    with pytest.raises(RuntimeError):
        pf.get_resample_ancestors()
    pf2 = ParticleFilterDevice(n_particles=4096, resampling="systematic", seed=1)
    pf2.initialize(np.zeros(3))
    sat, pr, w = _stream_scenario(8)
    pf2.update(sat, pr, w)
    a = pf2.get_resample_ancestors()
    assert a.dtype == np.int32 and a.shape == (4096,)
    assert a.min() >= 0 and a.max() < 4096


def test_uninitialized_raises(gpu):
    from gnss_gpu_amd import ParticleFilterDevice

    pf = ParticleFilterDevice(n_particles=100)
    with pytest.raises(RuntimeError, match="not initialized"):
        pf.estimate()
    with pytest.raises(RuntimeError, match="not initialized"):
        pf.predict()


def test_isb_states_match_oracle(gpu):
    sc_anchor = np.array([-3.96e6, 3.35e6, 3.70e6])
    from gnss_gpu_amd import sim

    sc = sim.scenario_multi_gnss(n_epochs=2, seed=2, anchor=sc_anchor)
    from gnss_gpu_amd import ParticleFilterDevice

    n = 20_000
    pf = ParticleFilterDevice(n_particles=n, sigma_pos=1.0, sigma_cb=3.0, sigma_pr=5.0, nu=1.0,
                              resampling="systematic", seed=5, n_isb=3, sigma_isb=0.5)
    pf.initialize(sc.truth[0], clock_bias=sc.cb, spread_pos=20.0, spread_cb=10.0,
                  isb=sc.isb, spread_isb=5.0)
    orc = opf.OraclePF(n, seed=5, n_isb=3)
    orc.initialize(sc.truth[0], sc.cb, 20.0, 10.0, isb=sc.isb, spread_isb=5.0)
    from gnss_gpu_amd import _gnss_gpu_pf_device as dev

    np.testing.assert_allclose(dev.pf_device_get_particles_full(pf._state), orc.state, atol=1e-4)
    pf.predict(velocity=sc.vel, dt=1.0)
    orc.predict(sc.vel, 1.0, 1.0, 3.0, 1, sigma_isb=0.5)
    full = dev.pf_device_get_particles_full(pf._state)
    np.testing.assert_allclose(full, orc.state, atol=1e-3)
    pf.update(sc.sat, sc.pr[1], system_ids=sc.sys, resample=False)
    lw = pf.get_log_weights()
    ref = opf.log_likelihood(full[:, :3], full[:, 3], full[:, 4:], sc.sat, sc.pr[1], np.ones(12),
                             sc.sys, 5.0, nu=1.0)
    np.testing.assert_allclose(lw, ref, rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(pf.estimate_full(), opf.weighted_mean(full, lw), atol=1e-3)


def test_isb_closed_forms_through_device(gpu):
    """test_isb_closed_forms on the device: particles at the truth with the true
    inter-system biases weigh 0; ISB j off by delta weighs
    -n_{j+1} delta^2 / (2 sigma^2) (1e-5 relative)."""
    from gnss_gpu_amd import ParticleFilterDevice
    from tests.test_oracle_pinning import _isb_case

    truth, sat, pr, sys_ids, cb, isb, counts = _isb_case()
    for j, delta in ((None, 0.0), (0, 1.5), (1, -2.0), (2, 3.0)):
        off = isb.copy()
        if j is not None:
            off[j] += delta
        pf = ParticleFilterDevice(n_particles=2048, sigma_pr=5.0, resampling="systematic",
                                  seed=SEED, n_isb=3)
        pf.initialize(truth, clock_bias=cb, spread_pos=0.0, spread_cb=0.0, isb=off,
                      spread_isb=0.0)
        pf.update(sat, pr, system_ids=sys_ids, resample=False)
        lw = pf.get_log_weights()
        want = 0.0 if j is None else -counts[j + 1] * delta * delta / 50.0
        np.testing.assert_allclose(lw, want, rtol=1e-5, atol=1e-5)


def test_end_to_end_plain_filter_matches_oracle(gpu):
    """Config-1 trajectory (N=10,000, 8 GPS, Gaussian 5 m, systematic at
    ESS<N/2), all 100 epochs.  Every epoch: the log-weights within 1e-5
    relative of the oracle's likelihood at the GPU's particles; against the
    independently propagated oracle within 1e-5 relative plus the Gaussian
    slope bound sum_s (|e_s| delta + delta^2 / 2) / sigma^2 for the measured
    state drift delta; ancestors equal when the oracle is fed the GPU's
    binary32 log-weights; estimates within 1 mm; drift below 1 mm."""
    from gnss_gpu_amd import ParticleFilterDevice, sim

    sc = sim.scenario_plain(n_epochs=100, seed=0)
    n = 10_000
    sig = 5.0
    pf = ParticleFilterDevice(n_particles=n, sigma_pos=1.0, sigma_cb=3.0, sigma_pr=sig,
                              resampling="systematic", ess_threshold=0.5, seed=0)
    pf.initialize(sc.truth[0], clock_bias=sc.cb, spread_pos=10.0, spread_cb=10.0)
    orc = opf.OraclePF(n, seed=0)
    orc.initialize(sc.truth[0], sc.cb, 10.0, 10.0)
    w = np.ones(sc.n_sat)
    errs, diffs, mism, resampled, worst = [], [], 0, 0, 0.0
    for t in range(sc.n_epochs):
        if t > 0:
            pf.predict(velocity=sc.vel, dt=1.0)
            orc.predict(sc.vel, 1.0, 1.0, 3.0, t)
        pf.update(sc.sat, sc.pr[t], resample=False)
        orc.weight(sc.sat, sc.pr[t], sigma_pr=sig)
        lw_gpu = pf.get_log_weights()
        p = pf.get_particles()
        given = opf.log_likelihood(p[:, :3], p[:, 3], np.zeros((n, 0)), sc.sat, sc.pr[t], w, None,
                                   sig)
        np.testing.assert_allclose(lw_gpu, given, rtol=1e-5, atol=1e-5)
        delta = np.linalg.norm(p[:, :3] - orc.state[:, :3], axis=1) + np.abs(p[:, 3] - orc.state[:, 3])
        assert delta.max() < 1e-3, (t, delta.max())
        e = sc.pr[t][None, :] - (np.linalg.norm(orc.state[:, None, :3] - sc.sat[None], axis=2) +
                                 orc.state[:, 3:4])
        bound = (np.abs(e).sum(axis=1) * delta + 0.5 * sc.n_sat * delta * delta) / sig ** 2
        err = np.abs(lw_gpu - orc.lw)
        worst = max(worst, float(np.max(err / np.maximum(np.abs(orc.lw), 1.0))))
        assert np.all(err <= 1e-5 * np.abs(orc.lw) + bound + 1e-5), t
        if pf.resample_if_needed():
            resampled += 1
            assert orc.ess() < 0.5 * n * 1.001
            orc.resample_systematic(lw32=lw_gpu.astype(np.float32))
            mism += int(np.sum(pf.get_resample_ancestors() != orc.ancestors))
        est = pf.estimate()
        diffs.append(np.abs(est - orc.estimate()).max())
        errs.append(np.linalg.norm(est[:3] - sc.truth[t]))
    print(f"\n[config 1, 100 epochs] resampled {resampled}, ancestor mismatches {mism}, "
          f"max estimate diff {max(diffs):.3e} m, max log-weight rel diff vs independent "
          f"oracle {worst:.3e}", flush=True)
    assert resampled > 5
    assert mism == 0
    assert max(diffs) < 1e-3
    assert np.median(errs) < 10.0


def _dd_case():
    d = GOLD["dd_closed_forms"]
    rng = np.random.RandomState(123)
    dirs = rng.randn(5, 3)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    sat = dirs * 20_000_000.0
    base = np.array(d["base_pos"])
    rr = np.linalg.norm(sat, axis=1)
    br = np.linalg.norm(sat - base, axis=1)
    dd = (rr[1:] - rr[0]) - (br[1:] - br[0])
    n = len(dd)
    from types import SimpleNamespace

    def res(**kw):
        return SimpleNamespace(sat_ecef_k=sat[1:], sat_ecef_ref=np.repeat(sat[[0]], n, axis=0),
                               base_range_k=br[1:], base_range_ref=np.repeat(br[0], n),
                               dd_weights=np.ones(n), n_dd=n, **kw)
    return d, dd, res


@pytest.mark.parametrize("huber", [False, True])
def test_dd_closed_forms_through_device(gpu, huber):
    """The reference's DD pseudorange and DD carrier AFV constants
    (test_cuda_streams.py:604-771) through ParticleFilterDevice; log-weights
    are binary32 here, so the comparison is at 1e-6 relative."""
    from gnss_gpu_amd import ParticleFilterDevice

    d, dd, res = _dd_case()

    def lw(update, gate=False, **kw):
        pf = ParticleFilterDevice(n_particles=1024, sigma_pos=1.0, sigma_cb=300.0, sigma_pr=5.0,
                                  resampling="systematic", seed=SEED, per_particle_huber=huber,
                                  per_particle_nlos_gate=gate, **kw)
        pf.initialize(np.zeros(3), clock_bias=0.0, spread_pos=0.0, spread_cb=0.0)
        update(pf)
        return pf.get_log_weights()

    ddo = dd.copy()
    ddo[0] += d["dd_pr_outlier_m"]
    out = lw(lambda pf: pf.update_dd_pseudorange(res(dd_pseudorange_m=ddo),
                                                 sigma_pr=d["dd_pr_sigma_m"], resample=False))
    want = d["dd_pr_huber"] if huber else d["dd_pr_gaussian"]
    np.testing.assert_allclose(out, want, rtol=1e-6)
    gated = lw(lambda pf: pf.update_dd_pseudorange(res(dd_pseudorange_m=ddo),
                                                   sigma_pr=d["dd_pr_sigma_m"], resample=False),
               gate=True, per_particle_nlos_dd_pr_threshold_m=d["dd_pr_gate_m"])
    assert float(np.max(np.abs(gated))) < d["gated_abs_max"]
    wl = np.full(len(dd), d["wavelength_m"])
    ddc = dd / wl
    ddc[0] += d["dd_cp_outlier_cycles"]
    out = lw(lambda pf: pf.update_dd_carrier_afv(res(dd_carrier_cycles=ddc, wavelengths_m=wl),
                                                 sigma_cycles=d["dd_cp_sigma_cycles"],
                                                 resample=False))
    want = d["dd_cp_huber"] if huber else d["dd_cp_gaussian"]
    np.testing.assert_allclose(out, want, rtol=1e-6)
    gated = lw(lambda pf: pf.update_dd_carrier_afv(res(dd_carrier_cycles=ddc, wavelengths_m=wl),
                                                   sigma_cycles=d["dd_cp_sigma_cycles"],
                                                   resample=False),
               gate=True, per_particle_nlos_dd_carrier_threshold_cycles=d["dd_cp_gate_cycles"])
    assert float(np.max(np.abs(gated))) < d["gated_abs_max"]


def test_dd_and_carrier_afv_match_oracle(gpu):
    """A particle cloud around the Tokyo receiver with a base 800 m away:
    DD pseudorange (Gaussian, Huber, gate), undifferenced carrier AFV and DD
    carrier AFV against the binary64 terms of the oracle.  The carrier terms
    add to the log-weights, so they are checked as differences."""
    from gnss_gpu_amd import ParticleFilterDevice

    g = GOLD["tokyo_constellation"]
    sat = np.array(g["sat_ecef"])
    tp = np.array(g["true_pos"])
    base = tp + np.array([500.0, -600.0, 120.0])
    rng = np.random.default_rng(9)
    S = len(sat)
    rr = np.linalg.norm(sat - tp, axis=1)
    br = np.linalg.norm(sat - base, axis=1)
    ref = 3
    ks = [k for k in range(S) if k != ref]
    dd = (rr[ks] - rr[ref]) - (br[ks] - br[ref]) + rng.normal(0, 0.3, len(ks))
    dd[1] += 25.0
    wl = np.full(len(ks), 0.190293673)
    wl[::2] = 0.2442102134
    ddc = dd / wl + rng.normal(0, 0.02, len(ks))
    w = rng.uniform(0.5, 1.5, len(ks))
    sk, sr = sat[ks], np.repeat(sat[[ref]], len(ks), axis=0)
    brk, brr = br[ks], np.repeat(br[ref], len(ks))
    from types import SimpleNamespace

    def res(**kw):
        return SimpleNamespace(sat_ecef_k=sk, sat_ecef_ref=sr, base_range_k=brk,
                               base_range_ref=brr, dd_weights=w, n_dd=len(ks), **kw)

    for huber, gate in ((False, False), (True, False), (False, True)):
        pf = ParticleFilterDevice(n_particles=20_000, resampling="systematic", seed=SEED,
                                  per_particle_huber=huber, per_particle_nlos_gate=gate,
                                  per_particle_nlos_dd_pr_threshold_m=3.0)
        pf.initialize(tp, clock_bias=50.0, spread_pos=4.0, spread_cb=2.0)
        pf.update_dd_pseudorange(res(dd_pseudorange_m=dd), sigma_pr=0.8, resample=False)
        st = pf.get_particles()
        ref_lw = opf.dd_pseudorange_loglik(st[:, :3], sk, sr, dd, brk, brr, w, 0.8,
                                           thr=3.0 if gate else 0.0, huber=huber)
        np.testing.assert_allclose(pf.get_log_weights(), ref_lw, rtol=1e-5, atol=2e-3)
    # carrier terms on top of a pseudorange update (and after a resample)
    pf = ParticleFilterDevice(n_particles=20_000, resampling="systematic", seed=SEED,
                              sigma_pr=5.0)
    pf.initialize(tp, clock_bias=50.0, spread_pos=1.5, spread_cb=0.3)
    cp = (rr + 50.0) / 0.190293673 + rng.normal(0, 0.02, S) + 1234.0
    pf.update(sat, rr + 50.0 + rng.normal(0, 2.0, S), resample=False)
    lw0 = pf.get_log_weights()
    pf.update_carrier_afv(sat, cp, sigma_cycles=0.05, resample=False)
    st = pf.get_particles()
    add = opf.carrier_afv_loglik(st[:, :3], st[:, 3], sat, cp, np.ones(S), 0.190293673, 0.05)
    np.testing.assert_allclose(pf.get_log_weights() - lw0, add, rtol=1e-4, atol=0.05)
    lw1 = pf.get_log_weights()
    pf.update_dd_carrier_afv(res(dd_carrier_cycles=ddc, wavelengths_m=wl), sigma_cycles=0.05,
                             resample=False)
    add2 = opf.dd_carrier_afv_loglik(st[:, :3], sk, sr, ddc, brk, brr, w, wl, 0.05)
    np.testing.assert_allclose(pf.get_log_weights() - lw1, add2, rtol=1e-4, atol=0.05)
    # after a resample the weights restart from zero
    from gnss_gpu_amd import _gnss_gpu_pf_device as dev

    dev.pf_device_resample_systematic(pf._state, 7)
    pf.update_carrier_afv(sat, cp, sigma_cycles=0.05, resample=False)
    st = pf.get_particles()
    add3 = opf.carrier_afv_loglik(st[:, :3], st[:, 3], sat, cp, np.ones(S), 0.190293673, 0.05)
    np.testing.assert_allclose(pf.get_log_weights(), add3, rtol=1e-4, atol=0.05)


def _doppler_rows():
    d = GOLD["doppler_rows"]
    rx = np.array(d["rx_pos"])
    sat = np.array(d["sat_ecef"])
    sv = np.array(d["sat_vel"])
    tv = np.array(d["true_vel"])
    los = (sat - rx) / np.linalg.norm(sat - rx, axis=1)[:, None]
    dop = -(np.einsum("sk,sk->s", sv - tv[None, :], los) + d["true_cd"]) / d["wavelength_m"]
    return d, rx, sat, sv, tv, dop


def test_doppler_updates_recover_velocity(gpu):
    """test_cuda_streams.py:285-335 and :429-484: the Doppler nudge (gain 1)
    and the per-particle Kalman update recover the true velocity."""
    from gnss_gpu_amd import ParticleFilterDevice

    d, rx, sat, sv, tv, dop = _doppler_rows()
    n = 2048
    pf = ParticleFilterDevice(n_particles=n, seed=SEED, resampling="systematic")
    pf.initialize(rx, clock_bias=100.0, spread_pos=0.0, spread_cb=0.0, velocity=np.zeros(3),
                  spread_vel=0.0)
    pf.update_doppler(sat, sv, dop, weights=np.ones(len(dop)), sigma_mps=d["sigma_mps"],
                      velocity_update_gain=1.0, max_velocity_update_mps=100.0, resample=False)
    st = pf.get_particle_states()
    np.testing.assert_allclose(st[:, 4:7], np.tile(tv, (n, 1)), atol=d["nudge_atol"])
    pf = ParticleFilterDevice(n_particles=n, seed=SEED, resampling="systematic")
    pf.initialize(rx, clock_bias=100.0, spread_pos=0.0, spread_cb=0.0, velocity=np.zeros(3),
                  velocity_init_sigma=d["kf_init_sigma"])
    pf.update_doppler_kf(sat, sv, dop, weights=np.ones(len(dop)), sigma_mps=d["sigma_mps"],
                         resample=False)
    st = pf.get_particle_states()
    np.testing.assert_allclose(st[:, 4:7], np.tile(tv, (n, 1)), atol=d["kf_atol"])
    assert np.all(st[:, [7, 11, 15]] < 100.0) and np.all(st[:, [7, 11, 15]] > 0.0)


def test_doppler_updates_match_oracle(gpu):
    """Particle clouds with per-particle velocities and covariances against the
    binary64 oracle: log-weight increments, velocities, covariances (the
    device works in the filter's ENU frame, the oracle in ECEF)."""
    from gnss_gpu_amd import ParticleFilterDevice

    d, rx, sat, sv, tv, dop = _doppler_rows()
    rng = np.random.default_rng(4)
    dop_n = dop + rng.normal(0.0, 0.3, len(dop))
    w = rng.uniform(0.5, 2.0, len(dop))
    wl, sig = d["wavelength_m"], d["sigma_mps"]
    n = 8192
    for gain, cap in ((0.0, 10.0), (0.6, 0.4)):
        pf = ParticleFilterDevice(n_particles=n, seed=SEED, resampling="systematic")
        pf.initialize(rx, clock_bias=100.0, spread_pos=40.0, spread_cb=3.0,
                      velocity=np.array([1.0, 0.5, -0.2]), spread_vel=0.7)
        before = pf.get_particle_states()
        lw0 = np.zeros(n)
        pf.update_doppler(sat, sv, dop_n, weights=w, wavelength=wl, sigma_mps=sig,
                          velocity_update_gain=gain, max_velocity_update_mps=cap, resample=False)
        after = pf.get_particle_states()
        add, v = opf.doppler_update(before[:, :3], before[:, 4:7], sat, sv, dop_n, w, wl, sig,
                                    gain, cap)
        np.testing.assert_allclose(pf.get_log_weights() - lw0, add, rtol=1e-5, atol=1e-3)
        np.testing.assert_allclose(after[:, 4:7], v, rtol=0, atol=2e-5)
    pf = ParticleFilterDevice(n_particles=n, seed=SEED, resampling="systematic")
    pf.initialize(rx, clock_bias=100.0, spread_pos=40.0, spread_cb=3.0,
                  velocity=np.array([1.0, 0.5, -0.2]), spread_vel=0.7, velocity_init_sigma=3.0)
    before = pf.get_particle_states()
    pf.update_doppler_kf(sat, sv, dop_n, weights=w, wavelength=wl, sigma_mps=sig, resample=False)
    after = pf.get_particle_states()
    add, mu, P = opf.doppler_kf_update(before[:, :3], before[:, 4:7], before[:, 7:16], sat, sv,
                                       dop_n, w, wl, sig)
    np.testing.assert_allclose(pf.get_log_weights(), add, rtol=1e-5, atol=1e-3)
    np.testing.assert_allclose(after[:, 4:7], mu, rtol=0, atol=2e-5)
    np.testing.assert_allclose(after[:, 7:16], P.reshape(n, 9), rtol=1e-5, atol=1e-6)
