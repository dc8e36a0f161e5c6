"""Launch shapes and the autotuner of the device-driven step.

Every shape gives the same bits: the likelihood kernel with 1, 2 or 4 tiles per
workgroup, and the Megopolis chain kernel with 1, 2, 4 or 8 iterations of
proposal loads in flight, produce identical estimates, particles and ancestors
(the Megopolis shapes are also the ones the oracle-ancestor tests of the batched
step run with by default: depth 1).  The autotuner times every candidate from a
snapshot and restores it, so a filter that was autotuned continues exactly as
its untuned twin.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [50_000, 7_000, 120_000]


def _make():
    from gnss_gpu_amd import sim
    from gnss_gpu_amd._lib import LIK_STUDENT
    from gnss_gpu_amd.batched import BatchedParticleFilter

    T = 12
    sc = sim.scenario_multi_gnss(n_epochs=T, seed=5)
    S = sc.n_sat
    F = len(SIZES)
    pf = BatchedParticleFilter(SIZES, n_isb=3, sat_cap=S, n_slots=T)
    pf.load_epochs([[sc.sat] * F] * T, [[sc.pr[t]] * F for t in range(T)], [[S] * F] * T,
                   sys=[[sc.sys] * F] * T, vel=np.tile(sc.vel, (T, F, 1)))
    init = np.tile(np.concatenate([sc.truth[0] - sc.vel, [sc.cb], sc.isb]), (F, 1))
    pf.initialize(init, [3, 4, 5], 10.0, 10.0, 5.0)
    st = pf.step_args(sigma_pos=1.0, sigma_cb=1.0, sigma_isb=0.1, sigma_pr=5.0, nu=1.0,
                      lik=LIK_STUDENT, n_slots=T, ess_frac=0.9)
    return pf, st


def _run(pf, st, resampler, steps=6):
    est, anc = [], []
    for _ in range(steps):
        pf.step(st, resampler, 32)
        est.append(pf.estimates())
        anc.append(np.concatenate([pf.get_ancestors(f) for f in range(len(SIZES))]))
    parts = [pf.get_particles(f) for f in range(len(SIZES))]
    return np.stack(est), anc, parts


def _same(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    for x, y in zip(a[1], b[1]):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a[2], b[2]):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("depth", [2, 4, 8])
def test_megopolis_load_depth_gives_the_same_bits(gpu, depth):
    base_pf, st = _make()
    base = _run(base_pf, st, "megopolis")
    assert base_pf.read_dyn()[0].resample_count >= 3
    base_pf.close()
    pf, st = _make()
    pf.set_tuning(mego_depth=depth)
    assert pf.get_tuning()["mego_depth"] == depth
    _same(_run(pf, st, "megopolis"), base)
    pf.close()


@pytest.mark.parametrize("tpw", [2, 4])
def test_weight_tiles_per_workgroup_give_the_same_bits(gpu, tpw):
    base_pf, st = _make()
    base = _run(base_pf, st, "systematic")
    base_pf.close()
    pf, st = _make()
    pf.set_tuning(weight_tpw=tpw)
    _same(_run(pf, st, "systematic"), base)
    pf.close()


@pytest.mark.parametrize("resampler", ["systematic", "megopolis"])
def test_autotune_keeps_the_trajectory(gpu, resampler):
    twin, st = _make()
    ref = _run(twin, st, resampler)
    twin.close()
    pf, st = _make()
    epoch0 = pf.read_dyn()[0].epoch
    rep = pf.autotune(st, resampler, 32, steps=4, reps=2)
    n_expect = 3 * (4 if resampler == "megopolis" else 1)
    assert len(rep["candidates"]) == n_expect
    assert all(ms > 0.0 for _, ms in rep["candidates"])
    best = min(rep["candidates"], key=lambda c: c[1])
    assert rep["chosen"] == best[0]
    # the snapshot restored the filters: the tuned run continues as the twin
    assert pf.read_dyn()[0].epoch == epoch0
    _same(_run(pf, st, resampler), ref)
    pf.close()


@pytest.mark.parametrize("resampler", ["systematic", "megopolis"])
def test_profiled_steps_are_the_steps(gpu, resampler):
    """Stage profiling (every kernel bracketed by its own start/stop events)
    runs exactly the step's kernels: the trajectory equals plain steps."""
    twin, st = _make()
    ref = _run(twin, st, resampler, steps=5)
    twin.close()
    pf, st = _make()
    res, oth, cnt = pf.profile_steps(st, 5, resampler, 32)
    assert sum(cnt) == 5 and cnt[0] >= 1
    names = ({"predict_weight", "resample"} if resampler == "systematic"
             else {"predict_weight", "decide", "megopolis_gather"})
    assert set(res) == names
    assert all(v > 0.0 for v in res.values())
    np.testing.assert_array_equal(pf.estimates(), ref[0][-1])
    for f in range(len(SIZES)):
        np.testing.assert_array_equal(pf.get_particles(f), ref[2][f])
    pf.close()
