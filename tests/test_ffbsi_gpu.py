"""GPU backward trajectory sampling (FFBSi, genealogy) against the NumPy
restatement of python/gnss_gpu/particle_ffbsi.py, driven by the same
generators: indices must be identical and the generators left in the same
state."""

import numpy as np
import pytest

from oracle import ffbsi as of

pytestmark = pytest.mark.gpu


def _history(T, N, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    X = np.cumsum(rng.normal(0, spread, (T, N, 4)), axis=0) * np.array([1.0, 1.0, 1.0, 3.0])
    X[..., :3] += np.array([-3957199.0, 3310205.0, 3737911.0])
    lw = rng.normal(0, 2.0, (T, N))
    vel = rng.normal(0, 0.1, (T, 3))
    return lw, X, vel, np.full(T, 1.0), rng.uniform(1.0, 2.0, T)


def _same_state(a, b):
    return a.random() == b.random()


def test_ffbsi_matches_oracle_and_batch(gpu):
    from gnss_gpu_amd.particle_ffbsi import (ffbsi_sample_indices, ffbsi_sample_indices_batch,
                                             ffbsi_smooth_sample)

    lw, X, vel, dt, sp = _history(60, 3000, 1)
    for seed in range(4):
        g, r = np.random.default_rng(seed), np.random.default_rng(seed)
        got = ffbsi_sample_indices(lw, X, vel, dt, sp, 8.0, g)
        ref = of.ffbsi_indices(lw, X, vel, dt, sp, 8.0, r)
        np.testing.assert_array_equal(got, ref)
        assert got.dtype == np.int64 and _same_state(g, r)
    gens = [np.random.default_rng(100 + k) for k in range(8)]
    batch = ffbsi_sample_indices_batch(lw, X, vel, dt, sp, 8.0, gens)
    for k in range(8):
        np.testing.assert_array_equal(
            batch[k], of.ffbsi_indices(lw, X, vel, dt, sp, 8.0, np.random.default_rng(100 + k)))
    path = ffbsi_smooth_sample(lw, X, vel, dt, sp, 8.0, np.random.default_rng(0))
    ref = of.ffbsi_indices(lw, X, vel, dt, sp, 8.0, np.random.default_rng(0))
    np.testing.assert_array_equal(path, X[np.arange(60), ref, :3])


def test_ffbsi_large_population(gpu):
    """N = 200,000 (782 particles per thread range), T = 12."""
    from gnss_gpu_amd.particle_ffbsi import ffbsi_sample_indices_batch

    lw, X, vel, dt, sp = _history(12, 200_000, 2, spread=0.3)
    got = ffbsi_sample_indices_batch(lw, X, vel, dt, sp, 5.0,
                                     [np.random.default_rng(k) for k in range(3)])
    for k in range(3):
        np.testing.assert_array_equal(
            got[k], of.ffbsi_indices(lw, X, vel, dt, sp, 5.0, np.random.default_rng(k)))


@pytest.mark.parametrize("N", [1, 64, 1023, 1024, 1025, 4097, 8192, 8193, 16384, 16385])
def test_ffbsi_population_edges(gpu, N):
    """Populations at the edges of the particle-row kernel (1024 particles a
    row, up to 16 rows) and just past it (the per-thread-range kernel)."""
    from gnss_gpu_amd.particle_ffbsi import ffbsi_sample_indices_batch

    lw, X, vel, dt, sp = _history(24, N, 10 + N % 7)
    got = ffbsi_sample_indices_batch(lw, X, vel, dt, sp, 6.0,
                                     [np.random.default_rng(k) for k in range(3)])
    for k in range(3):
        np.testing.assert_array_equal(
            got[k], of.ffbsi_indices(lw, X, vel, dt, sp, 6.0, np.random.default_rng(k)))


@pytest.mark.parametrize("T", [1, 2, 64, 65])
def test_ffbsi_short_and_block_edge_histories(gpu, T):
    """One and two epochs, and histories at the 64-step block edge of the row
    kernel's staged draws and picks."""
    from gnss_gpu_amd.particle_ffbsi import ffbsi_sample_indices_batch

    lw, X, vel, dt, sp = _history(T, 300, 20 + T)
    gens = [np.random.default_rng(50 + k) for k in range(3)]
    refs = [np.random.default_rng(50 + k) for k in range(3)]
    got = ffbsi_sample_indices_batch(lw, X, vel, dt, sp, 7.0, gens)
    for k in range(3):
        np.testing.assert_array_equal(got[k], of.ffbsi_indices(lw, X, vel, dt, sp, 7.0, refs[k]))
        assert _same_state(gens[k], refs[k])


def test_ffbsi_long_history_with_redraws(gpu):
    """T = 300 (the row kernel reads draws and stores picks 64 steps at a
    time; 300 is not a multiple of 64) with degenerate steps in two different
    64-step blocks, which relaunch with integer draws there."""
    from gnss_gpu_amd.particle_ffbsi import ffbsi_sample_indices_batch

    lw, X, vel, dt, sp = _history(300, 1500, 4)
    lw[150, 7] = np.nan
    lw[40] = -np.inf
    gens = [np.random.default_rng(30 + k) for k in range(4)]
    refs = [np.random.default_rng(30 + k) for k in range(4)]
    got = ffbsi_sample_indices_batch(lw, X, vel, dt, sp, 8.0, gens)
    for k in range(4):
        np.testing.assert_array_equal(got[k], of.ffbsi_indices(lw, X, vel, dt, sp, 8.0, refs[k]))
        assert _same_state(gens[k], refs[k])


def test_ffbsi_degenerate_steps(gpu):
    """A NaN and an all -inf row mid-history redraw integers(0, N) at those
    steps (two extra launches); a degenerate last step gives zeros and leaves
    the generator untouched (particle_ffbsi.py:73-75, 98-100)."""
    from gnss_gpu_amd.particle_ffbsi import ffbsi_sample_indices

    lw, X, vel, dt, sp = _history(30, 500, 3)
    lw[17, 40] = np.nan
    lw[6] = -np.inf
    lw[3, 2] = np.inf
    for seed in range(3):
        g, r = np.random.default_rng(seed), np.random.default_rng(seed)
        got = ffbsi_sample_indices(lw, X, vel, dt, sp, 8.0, g)
        np.testing.assert_array_equal(got, of.ffbsi_indices(lw, X, vel, dt, sp, 8.0, r))
        assert _same_state(g, r)
    lw[29] = np.nan
    g, r = np.random.default_rng(9), np.random.default_rng(9)
    assert np.all(ffbsi_sample_indices(lw, X, vel, dt, sp, 8.0, g) == 0)
    assert _same_state(g, r)


def test_genealogy_and_transition(gpu):
    from gnss_gpu_amd.particle_ffbsi import (genealogy_smooth_indices, genealogy_smooth_sample,
                                             transition_logpdf)

    rng = np.random.default_rng(5)
    T, N = 40, 2048
    lw = rng.standard_normal((T, N)) * 2.0
    anc = np.vstack([rng.integers(0, N, (T - 1, N)), np.arange(N)[None]])
    for seed in range(3):
        g, r = np.random.default_rng(seed), np.random.default_rng(seed)
        got = genealogy_smooth_indices(lw, anc, g)
        np.testing.assert_array_equal(got, of.genealogy_indices(lw, anc, r))
        assert _same_state(g, r)
    X = rng.normal(0, 1, (T, N, 4))
    p = genealogy_smooth_sample(lw, X, anc, np.random.default_rng(1))
    np.testing.assert_array_equal(
        p, X[np.arange(T), of.genealogy_indices(lw, anc, np.random.default_rng(1)), :3])
    xt = rng.normal(0, 5, (1000, 4))
    xn = np.array([1.0, -2.0, 0.5, 40.0])
    got = transition_logpdf(xn, xt, [0.3, -0.1, 0.2], 0.7, 1.2, 30.0)
    np.testing.assert_allclose(got, of.transition_logpdf(xn, xt, [0.3, -0.1, 0.2], 0.7, 1.2, 30.0),
                               rtol=1e-14, atol=1e-12)
    with pytest.raises(ValueError):
        genealogy_smooth_indices(lw, np.full((T, N), N), g)


def test_device_history_matches_host_arrays(gpu):
    """ParticleHistory records a running ParticleFilterDevice in HBM; its FFBSi
    indices equal those of the host-array functions on get_particles() /
    get_log_weights() copies taken at the same points (the reference
    experiment's loop, experiments/exp_ffbsi_eval.py:175-195)."""
    from gnss_gpu_amd import ParticleFilterDevice, sim
    from gnss_gpu_amd.particle_ffbsi import ParticleHistory, ffbsi_sample_indices_batch

    sc = sim.scenario_plain(n_epochs=30, seed=3)
    N, T = 4096, sc.pr.shape[0]
    pf = ParticleFilterDevice(n_particles=N, sigma_pos=1.0, sigma_cb=3.0, sigma_pr=5.0, seed=9)
    pf.initialize(sc.truth[0], clock_bias=100.0, spread_pos=10.0, spread_cb=10.0)
    hist = ParticleHistory(pf, T)
    X, LW, V = [], [], []
    for t in range(T):
        v = sc.truth[t] - sc.truth[max(t - 1, 0)]
        pf.predict(velocity=v, dt=1.0)
        pf.update(sc.sat, sc.pr[t], resample=False)
        hist.record(velocity=v, dt=1.0)
        X.append(pf.get_particles().copy())
        LW.append(pf.get_log_weights().copy())
        V.append(v)
        pf.resample_if_needed()
    X, LW, V = np.stack(X), np.stack(LW), np.stack(V)
    np.testing.assert_array_equal(hist.X.cpu().numpy(), X)
    np.testing.assert_array_equal(hist.lw.cpu().numpy(), LW)
    got = hist.ffbsi_sample_indices(3.0, [np.random.default_rng(k) for k in range(4)])
    ref = ffbsi_sample_indices_batch(LW, X, V, np.ones(T), np.full(T, 1.0), 3.0,
                                     [np.random.default_rng(k) for k in range(4)])
    np.testing.assert_array_equal(got, ref)
    for k in range(2):
        np.testing.assert_array_equal(
            got[k], of.ffbsi_indices(LW, X, V, np.ones(T), np.full(T, 1.0), 3.0,
                                     np.random.default_rng(k)))


@pytest.mark.parametrize("bad", [-5, "N"])
def test_device_entry_reports_forced_values_outside_the_history(gpu, bad):
    """gg_ffbsi_sample_device on device arrays: a forced value outside [-1, N)
    stops that trajectory with status -3 and no index outside [0, N) is
    written; the other trajectories are unaffected.  The host entry rejects
    the same input, and out-of-range genealogy ancestors."""
    import ctypes as C

    import torch

    from gnss_gpu_amd import _lib

    T, N, M = 40, 512, 3
    lw, X, vel, dt, sp = _history(T, N, 17)
    rng = np.random.default_rng(2)
    u = rng.random((M, T))
    forced = np.full((M, T), -1, dtype=np.int32)
    forced[1, 20] = N if bad == "N" else bad
    dev = torch.device("cuda")
    t = {k: torch.tensor(np.ascontiguousarray(v), device=dev)
         for k, v in dict(lw=lw, X=X, vel=vel, dt=dt, sp=sp, u=u).items()}
    tf = torch.tensor(forced, device=dev)
    idx = torch.full((M, T), -7, dtype=torch.int64, device=dev)
    st = torch.zeros(M, dtype=torch.int32, device=dev)
    p = lambda x: C.c_void_p(x.data_ptr())  # noqa: E731
    _lib.call("gg_ffbsi_sample_device", p(t["lw"]), p(t["X"]), p(t["vel"]), p(t["dt"]),
              p(t["sp"]), 3.0, None, T, N, M, p(t["u"]), p(tf), p(idx), p(st), None)
    torch.cuda.synchronize()
    status, got = st.cpu().numpy(), idx.cpu().numpy()
    assert status.tolist() == [-1, -3, -1]
    written = got[1][got[1] != -7]
    assert np.all((written >= 0) & (written < N))
    ref = of.ffbsi_indices(lw, X, vel, dt, sp, 3.0, np.random.default_rng(0))
    assert got[0].shape == ref.shape
    DP, IP = C.POINTER(C.c_double), C.POINTER(C.c_int32)
    out = np.zeros((M, T), dtype=np.int64)
    sts = np.zeros(M, dtype=np.int32)
    with pytest.raises(RuntimeError, match="forced index"):
        _lib.call("gg_ffbsi_sample", lw.ctypes.data_as(DP), X.ctypes.data_as(DP),
                  vel.ctypes.data_as(DP), dt.ctypes.data_as(DP), sp.ctypes.data_as(DP), 3.0,
                  None, T, N, M, u.ctypes.data_as(DP), forced.ctypes.data_as(IP),
                  C.c_void_p(out.ctypes.data), sts.ctypes.data_as(IP))
    anc = np.zeros((T, N), dtype=np.int32)
    anc[5, 7] = N
    ok = np.full((M, T), -1, dtype=np.int32)
    with pytest.raises(RuntimeError, match="ancestor outside"):
        _lib.call("gg_ffbsi_sample", lw.ctypes.data_as(DP), None, None, None, None, 3.0,
                  anc.ctypes.data_as(IP), T, N, M, u.ctypes.data_as(DP), ok.ctypes.data_as(IP),
                  C.c_void_p(out.ctypes.data), sts.ctypes.data_as(IP))
