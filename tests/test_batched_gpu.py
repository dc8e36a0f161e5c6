"""Batched, graph-driven filters against the host-driven drop-in path and
against themselves run eagerly."""

import numpy as np
import pytest

from oracle import pf as opf

pytestmark = pytest.mark.gpu


def _scenario(seed, n_sat, n_epochs):
    from gnss_gpu_amd import sim

    rng = np.random.default_rng(seed)
    anchor = sim.ANCHOR_CFG1 + rng.normal(0, 1000.0, 3)
    el = rng.uniform(15, 85, n_sat)
    az = rng.uniform(0, 360, n_sat)
    sat = sim.place_satellites(anchor, el, az)
    vel = rng.normal(0, 5.0, 3)
    truth = anchor[None] + np.arange(n_epochs)[:, None] * vel[None]
    pr = np.linalg.norm(sat[None] - truth[:, None], axis=2) + 50.0 + rng.normal(0, 3, (n_epochs, n_sat))
    return sat, pr, truth, vel


def _run_single(n, seed, sat, pr, truth, vel, steps, lik_nu=0.0):
    from gnss_gpu_amd import ParticleFilterDevice

    pf = ParticleFilterDevice(n_particles=n, sigma_pos=1.0, sigma_cb=2.0, sigma_pr=5.0, nu=lik_nu,
                              resampling="systematic", seed=seed)
    pf.initialize(truth[0] - vel, clock_bias=50.0, spread_pos=8.0, spread_cb=8.0)
    ests, res = [], []
    for t in range(steps):
        pf.predict(velocity=vel, dt=1.0)
        pf.update(sat, pr[t], resample=False)
        res.append(pf.resample_if_needed())
        ests.append(pf.estimate())
    return np.array(ests), res, pf


def test_batch_of_one_matches_drop_in_path(gpu):
    from gnss_gpu_amd.batched import BatchedParticleFilter

    n, steps = 20_000, 25
    sat, pr, truth, vel = _scenario(1, 8, steps)
    ref, ref_res, _ = _run_single(n, 7, sat, pr, truth, vel, steps)
    b = BatchedParticleFilter([n], sat_cap=8, n_slots=steps)
    b.load_epochs([[sat]] * steps, [[pr[t]] for t in range(steps)], [[8]] * steps,
                  vel=np.tile(vel, (steps, 1, 1)))
    b.initialize(np.concatenate([truth[0] - vel, [50.0]])[None], [7], 8.0, 8.0)
    st = b.step_args(sigma_pos=1.0, sigma_cb=2.0, sigma_pr=5.0)
    got, got_res = [], []
    for t in range(steps):
        b.step(st)
        d = b.read_dyn()[0]
        got.append([d.est[k] for k in range(4)])
        got_res.append(bool(d.did_resample))
    assert got_res == ref_res
    np.testing.assert_array_equal(np.array(got), ref)


def test_ragged_batch_matches_independent_filters(gpu):
    from gnss_gpu_amd.batched import BatchedParticleFilter

    ns = [1000, 3333, 1024, 50_000, 7, 20_000]
    sats = [6, 20, 9, 12, 5, 15]
    seeds = [11, 12, 13, 14, 15, 16]
    steps = 12
    scen = [_scenario(100 + f, sats[f], steps) for f in range(len(ns))]
    b = BatchedParticleFilter(ns, sat_cap=sats, n_slots=steps)
    b.load_epochs([[scen[f][0] for f in range(len(ns))] for _ in range(steps)],
                  [[scen[f][1][t] for f in range(len(ns))] for t in range(steps)],
                  [sats] * steps,
                  vel=np.array([[scen[f][3] for f in range(len(ns))]] * steps))
    init = np.array([np.concatenate([scen[f][2][0] - scen[f][3], [50.0]]) for f in range(len(ns))])
    b.initialize(init, seeds, 8.0, 8.0)
    st = b.step_args(sigma_pos=1.0, sigma_cb=2.0, sigma_pr=5.0, nu=1.0, lik=1)
    est = []
    for t in range(steps):
        b.step(st)
        est.append(b.estimates())
    est = np.array(est)
    for f in range(len(ns)):
        ref, _, _ = _run_single(ns[f], seeds[f], *scen[f], steps, lik_nu=1.0)
        np.testing.assert_array_equal(est[:, f, :], ref)


def test_graph_replays_match_eager_steps(gpu):
    from gnss_gpu_amd.batched import BatchedParticleFilter

    ns = [5000, 12345]
    steps = 10
    scen = [_scenario(200 + f, 10, steps) for f in range(2)]

    def make():
        b = BatchedParticleFilter(ns, sat_cap=10, n_slots=steps)
        b.load_epochs([[scen[f][0] for f in range(2)] for _ in range(steps)],
                      [[scen[f][1][t] for f in range(2)] for t in range(steps)], [[10, 10]] * steps,
                      vel=np.array([[scen[f][3] for f in range(2)]] * steps))
        init = np.array([np.concatenate([scen[f][2][0] - scen[f][3], [50.0]]) for f in range(2)])
        b.initialize(init, [3, 4], 8.0, 8.0)
        return b

    eager = make()
    st = eager.step_args(sigma_pos=1.0, sigma_cb=2.0, sigma_pr=5.0)
    ref = []
    for t in range(steps):
        eager.step(st)
        ref.append(eager.estimates())
    g = make()
    g.capture(st, steps=1)
    for t in range(steps):
        g.replay(1)
        _ = g.get_log_weights(0)          # eager work between replays
        np.testing.assert_array_equal(g.estimates(), ref[t])


@pytest.mark.parametrize("resampler", ["megopolis", "megopolis_ref"])
def test_batched_megopolis_steps_match_oracle_ancestors(gpu, resampler):
    from gnss_gpu_amd.batched import BatchedParticleFilter
    from oracle import detmath

    n, steps = 30_000, 4
    sat, pr, truth, vel = _scenario(5, 10, steps)
    b = BatchedParticleFilter([n], sat_cap=10, n_slots=steps)
    b.load_epochs([[sat]] * steps, [[pr[t]] for t in range(steps)], [[10]] * steps,
                  vel=np.tile(vel, (steps, 1, 1)))
    b.initialize(np.concatenate([truth[0] - vel, [50.0]])[None], [21], 8.0, 8.0)
    st = b.step_args(sigma_pos=1.0, sigma_cb=2.0, sigma_pr=5.0, ess_frac=1.1)  # always resample
    for t in range(steps):
        b.step(st, resampler=resampler, n_iter=16)
        d = b.read_dyn()[0]
        assert d.did_resample == 1
    p = b.get_particles(0)
    assert np.all(np.isfinite(p))
    assert np.linalg.norm(b.estimates()[0, :3] - truth[steps - 1]) < 30.0
    _ = detmath
    _ = opf


@pytest.mark.parametrize("n_isb", [0, 3])
def test_megopolis_whole_steps_equal_two_part_steps(gpu, n_isb):
    """A whole Megopolis step keeps the predicted state in packed records (only,
    after a step that resampled; a step that then does not resample copies them
    back), the two-part step in the state arrays: the same particles, weights,
    estimates and records after every step of a sequence that covers each pair
    of (previous step resampled, this step resamples)."""
    from gnss_gpu_amd.batched import BatchedParticleFilter

    n = 30_000 + 517
    fracs = [1.1, 0.0, 0.0, 1.1, 1.1, 0.0, 1.1, 0.0]
    steps = len(fracs)
    sat, pr, truth, vel = _scenario(9, 10, steps)
    sys_ids = (np.arange(10) % 4) if n_isb else None
    init = np.concatenate([truth[0] - vel, [50.0], np.zeros(n_isb)])[None]

    def make():
        b = BatchedParticleFilter([n], n_isb=n_isb, sat_cap=10, n_slots=steps)
        b.load_epochs([[sat]] * steps, [[pr[t]] for t in range(steps)], [[10]] * steps,
                      sys=None if sys_ids is None else [[sys_ids]] * steps,
                      vel=np.tile(vel, (steps, 1, 1)))
        b.initialize(init, [33], 8.0, 8.0, 2.0 if n_isb else 0.0)
        return b

    whole, parts = make(), make()
    for t, fr in enumerate(fracs):
        st = whole.step_args(sigma_pos=1.0, sigma_cb=2.0, sigma_isb=0.1, sigma_pr=5.0,
                             ess_frac=fr)
        whole.step(st, resampler="megopolis", n_iter=16)
        parts.step_part(st, 0, resampler="megopolis", n_iter=16)
        parts.step_part(st, 1, resampler="megopolis", n_iter=16)
        dw, dp = whole.read_dyn()[0], parts.read_dyn()[0]
        assert dw.did_resample == dp.did_resample == (1 if fr > 1.0 else 0), t
        assert dw.cur == dp.cur, t
        np.testing.assert_array_equal(whole.get_particles(0), parts.get_particles(0))
        np.testing.assert_array_equal(whole.get_log_weights_raw(0), parts.get_log_weights_raw(0))
        np.testing.assert_array_equal(whole.estimates(), parts.estimates())
    assert np.linalg.norm(whole.estimates()[0, :3] - truth[steps - 1]) < 30.0
    whole.close()
    parts.close()

def test_multi_step_graph_matches_eager_steps(gpu):
    """A graph holding several consecutive steps advances the device epoch
    counter inside the graph; its replays equal the eager steps."""
    from gnss_gpu_amd.batched import BatchedParticleFilter

    ns = [7000, 3000]
    steps, per_graph = 12, 4
    scen = [_scenario(300 + f, 9, steps) for f in range(2)]

    def make():
        b = BatchedParticleFilter(ns, sat_cap=9, n_slots=steps)
        b.load_epochs([[scen[f][0] for f in range(2)] for _ in range(steps)],
                      [[scen[f][1][t] for f in range(2)] for t in range(steps)], [[9, 9]] * steps,
                      vel=np.array([[scen[f][3] for f in range(2)]] * steps))
        init = np.array([np.concatenate([scen[f][2][0] - scen[f][3], [50.0]]) for f in range(2)])
        b.initialize(init, [5, 6], 8.0, 8.0)
        return b

    eager = make()
    st = eager.step_args(sigma_pos=1.0, sigma_cb=2.0, sigma_pr=5.0)
    ref = []
    for t in range(steps):
        eager.step(st)
        if (t + 1) % per_graph == 0:
            ref.append(eager.estimates())
    g = make()
    g.capture(st, steps=per_graph)
    for r in range(steps // per_graph):
        g.replay(1)
        np.testing.assert_array_equal(g.estimates(), ref[r])
