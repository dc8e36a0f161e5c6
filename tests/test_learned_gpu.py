"""Learned likelihood on the GPU against the CPU oracle (bf16 operands,
binary32 accumulation: tolerances are stated per test)."""

import numpy as np
import pytest

from oracle import learned as ol

pytestmark = pytest.mark.gpu


def _params(seed):
    p = ol.init_params(seed).astype(np.float64)
    # give the output layer weight so that mu and l move with the inputs
    p[ol.OFF_W3:ol.OFF_B3] *= 10.0
    return p.astype(np.float32)


def test_mlp_eval_matches_oracle(gpu):
    from gnss_gpu_amd.learned import LearnedLikelihood

    rng = np.random.default_rng(0)
    G = 96
    p = _params(3)
    e = rng.normal(0.0, 25.0, (G, 16)).astype(np.float32)
    f = rng.normal(0.0, 1.0, (G, 7)).astype(np.float32)
    m = LearnedLikelihood(p)
    ll, z3 = m.eval_groups(e, f)
    fw = ol.mlp_forward(p, e, np.broadcast_to(f[:, None, :], (G, 16, 7)), mode="bf16")
    # binary32 accumulation in a different order than the oracle's binary64:
    # z3 within 2e-3 absolute (the bf16 roundings of h1/h2 can flip by one ulp)
    np.testing.assert_allclose(z3, fw["z3"], rtol=0, atol=2e-3)
    np.testing.assert_allclose(ll, fw["ll"], rtol=5e-3, atol=5e-3)
    # and the bf16 network stays close to the exact binary64 one
    ex = ol.mlp_forward(p, e, np.broadcast_to(f[:, None, :], (G, 16, 7)), mode="exact")
    assert np.max(np.abs(z3 - ex["z3"])) < 0.1


def test_mlp_params_roundtrip(gpu):
    from gnss_gpu_amd.learned import LearnedLikelihood

    p = _params(5)
    m = LearnedLikelihood(p)
    np.testing.assert_array_equal(m.get_params(), p)


def test_mlp_gradient_matches_oracle(gpu):
    """Loss and gradient of the contrastive objective on the GPU (bf16 matrix
    cores, binary32 sums) against the oracle's bf16-mode binary64 gradient."""
    from gnss_gpu_amd import urban
    from gnss_gpu_amd.learned import LearnedLikelihood

    rng = np.random.default_rng(4)
    E, S = 6, 12
    f, r0, u, ns = urban.training_set(rng, E, n_sat=S, n_canyons=2)
    cand = np.zeros((E, 64, 4), np.float32)
    cand[:, 1:, :2] = rng.normal(0, 5, (E, 63, 2))
    cand[:, 1:, 2] = rng.normal(0, 3, (E, 63))
    cand[:, 1:, 3] = rng.normal(0, 5, (E, 63))
    p = _params(7)
    m = LearnedLikelihood(p)
    m.load_data(f, r0, u, ns)
    loss, grad = m.loss_grad(cand)
    ref_g = np.zeros(ol.N_PARAMS)
    for e in range(E):
        lo, g, _ = ol.contrastive_loss_and_grad(p, r0[e], u[e], f[e], cand[e, :, :3], cand[e, :, 3],
                                                mode="bf16")
        assert loss[e] == pytest.approx(lo, rel=2e-3, abs=2e-3)
        ref_g += g
    # per block of parameters: relative error of the gradient vector
    for a, b in [(ol.OFF_W1, ol.OFF_B1), (ol.OFF_B1, ol.OFF_W2), (ol.OFF_W2, ol.OFF_B2),
                 (ol.OFF_B2, ol.OFF_W3), (ol.OFF_W3, ol.OFF_B3), (ol.OFF_B3, ol.N_PARAMS)]:
        rel = np.linalg.norm(grad[a:b] - ref_g[a:b]) / max(np.linalg.norm(ref_g[a:b]), 1e-12)
        assert rel < 3e-2, (a, b, rel)


def test_mlp_training_reduces_loss(gpu):
    from gnss_gpu_amd import urban
    from gnss_gpu_amd.learned import LearnedLikelihood

    rng = np.random.default_rng(5)
    f, r0, u, ns = urban.training_set(rng, 4000, n_sat=16)
    m = LearnedLikelihood(seed=1)
    m.load_data(f, r0, u, ns)
    loss = m.train(300, batch=128, lr=2e-3, seed=3)
    assert np.all(np.isfinite(loss))
    assert loss[-50:].mean() < 0.5 * loss[:10].mean()


def _device_features(anchor, sat, cn0, flag, ess_frac=1.0):
    """Satellite features the weight kernel derives: elevation / azimuth from
    the unit vector to the satellite in the filter's ENU frame at its anchor."""
    from gnss_gpu_amd import sim

    R = sim.enu_rotation(anchor)
    ul = (np.asarray(sat) - anchor) @ R.T
    ul /= np.linalg.norm(ul, axis=1, keepdims=True)
    ch = np.hypot(ul[:, 0], ul[:, 1])
    S = len(sat)
    return np.stack([np.arcsin(ul[:, 2]) / (np.pi / 2), ul[:, 0] / ch, ul[:, 1] / ch,
                     (np.asarray(cn0) - 35.0) / 10.0, np.asarray(flag, np.float64),
                     np.full(S, S / 20.0), np.full(S, ess_frac)], axis=1).astype(np.float32)


def test_learned_weight_update_matches_oracle(gpu):
    """Per-particle sum over satellites of the learned pair log-likelihoods
    (ParticleFilterDevice.update with cn0/los_flag) against the oracle; bf16
    network, binary32 residuals on the device: 0.05 + 1e-3 |lw| absolute."""
    from gnss_gpu_amd import ParticleFilterDevice, urban
    from gnss_gpu_amd.learned import LearnedLikelihood

    rng = np.random.default_rng(8)
    d = urban.drive_ecef(rng, 3, n_sat=12)
    p = _params(9)
    m = LearnedLikelihood(p)
    n = 3000
    pf = ParticleFilterDevice(n_particles=n, seed=3)
    pf.set_learned_likelihood(m)
    pf.initialize(d.truth[1], clock_bias=d.cb, spread_pos=5.0, spread_cb=5.0)
    st = pf.get_particles()
    flag = d.los_flags(d.truth[1] + np.array([3.0, 0.0, 0.0]), 1)
    pf.update(d.sat[1], d.pr[1], cn0=d.cn0[1], los_flag=flag, resample=False)
    lw = pf.get_log_weights()
    f = _device_features(d.truth[1], d.sat[1], d.cn0[1], flag)
    e = d.pr[1][None, :] - (np.linalg.norm(d.sat[1][None] - st[:, None, :3], axis=2) + st[:, 3:4])
    ref = np.zeros(n)
    for s in range(12):
        ref += ol.mlp_forward(p, e[:, s].astype(np.float32), np.broadcast_to(f[s], (n, 7)),
                              mode="bf16")["ll"]
    np.testing.assert_allclose(lw, ref, rtol=1e-3, atol=0.05)


def test_learned_filter_beats_cauchy_in_canyons(gpu):
    """A network trained on the GPU (this test) against the Cauchy likelihood
    of the plain filter on fresh canyon drives: median horizontal error."""
    from gnss_gpu_amd import ParticleFilterDevice, urban
    from gnss_gpu_amd.learned import LearnedLikelihood

    rng = np.random.default_rng(11)
    f, r0, u, ns = urban.training_set(rng, 30000, n_sat=16)
    m = LearnedLikelihood(seed=2)
    m.load_data(f, r0, u, ns)
    loss = m.train(1500, batch=256, lr=2e-3, seed=5)
    assert loss[-100:].mean() < loss[:10].mean()
    T = 80
    p50 = {"cauchy": [], "learned": []}
    for k in range(3):
        d = urban.drive_ecef(np.random.default_rng(100 + k), T, n_sat=16)
        for kind in ("cauchy", "learned"):
            pf = ParticleFilterDevice(n_particles=50_000, sigma_pos=1.0, sigma_cb=1.0,
                                      sigma_pr=5.0, nu=1.0, seed=7)
            if kind == "learned":
                pf.set_learned_likelihood(m)
            pf.initialize(d.truth[0], clock_bias=d.cb, spread_pos=5.0, spread_cb=5.0)
            est = d.truth[0].copy()
            errs = []
            for t in range(1, T):
                pf.predict(velocity=d.vel[t], dt=1.0)
                if kind == "learned":
                    pf.update(d.sat[t], d.pr[t], cn0=d.cn0[t], los_flag=d.los_flags(est, t))
                else:
                    pf.update(d.sat[t], d.pr[t])
                est = pf.estimate()[:3]
                errs.append(float(np.linalg.norm((d.R @ (est - d.truth[t]))[:2])))
            p50[kind].append(float(np.median(errs[10:])))
    print("P50 horizontal error [m]", p50)
    assert np.mean(p50["learned"]) <= np.mean(p50["cauchy"])


def test_learned_code_tables_equal_per_pair_network(gpu):
    """The per-satellite code tables return exactly the per-pair network's
    outputs: identical log-weights (bit for bit) from both modes."""
    from gnss_gpu_amd import ParticleFilterDevice, urban
    from gnss_gpu_amd.learned import LearnedLikelihood

    rng = np.random.default_rng(21)
    d = urban.drive_ecef(rng, 4, n_sat=14)
    m = LearnedLikelihood(_params(13))
    out = []
    for table in (False, True):
        pf = ParticleFilterDevice(n_particles=70_000, seed=5)
        pf.set_learned_likelihood(m, table=table)
        pf.initialize(d.truth[1], clock_bias=d.cb, spread_pos=30.0, spread_cb=30.0)
        pf.predict(velocity=d.vel[2], dt=1.0)
        pf.update(d.sat[2], d.pr[2], cn0=d.cn0[2], los_flag=d.los_flags(d.truth[2], 2),
                  resample=False)
        out.append(pf.get_log_weights())
    np.testing.assert_array_equal(out[0], out[1])


def test_batched_learned_modes_match(gpu):
    """Graph-free batched steps with the learned likelihood: the direct and the
    table mode advance two ragged filters identically."""
    from gnss_gpu_amd import urban
    from gnss_gpu_amd._lib import LIK_LEARNED
    from gnss_gpu_amd.batched import BatchedParticleFilter
    from gnss_gpu_amd.learned import LearnedLikelihood

    m = LearnedLikelihood(_params(17))
    T = 6
    drives = [urban.drive_ecef(np.random.default_rng(40 + f), T, n_sat=s)
              for f, s in enumerate((12, 9))]
    ns = [20_000, 5_000]
    est = []
    for mode in (0, 1):
        b = BatchedParticleFilter(ns, sat_cap=[12, 9], n_slots=T)
        b.load_epochs([[dr.sat[t] for dr in drives] for t in range(T)],
                      [[dr.pr[t] for dr in drives] for t in range(T)], [[12, 9]] * T,
                      vel=np.array([[dr.vel[t] for dr in drives] for t in range(T)]))
        b.set_learned(m)
        b.load_features([[dr.cn0[t] for dr in drives] for t in range(T)],
                        [[dr.los_flags(dr.truth[t], t) for dr in drives] for t in range(T)])
        b.initialize(np.array([np.concatenate([dr.truth[0], [dr.cb]]) for dr in drives]), [3, 4],
                     5.0, 5.0)
        st = b.step_args(lik=LIK_LEARNED, learned_mode=mode, sigma_pos=1.0, sigma_cb=1.0)
        e = []
        for t in range(T):
            b.step(st)
            e.append(b.estimates())
        est.append(np.array(e))
    np.testing.assert_array_equal(est[0], est[1])


def test_fp8_weights_eval_matches_oracle(gpu):
    """FP8 E4M3 weight mode: the device network against the oracle's fp8 mode
    (same per-tensor scales and roundings), tolerances of the bf16 test; the
    quantised network differs from the bf16 one."""
    from gnss_gpu_amd.learned import LearnedLikelihood

    rng = np.random.default_rng(8)
    G = 64
    p = _params(11)
    e = rng.normal(0.0, 25.0, (G, 16)).astype(np.float32)
    f = rng.normal(0.0, 1.0, (G, 7)).astype(np.float32)
    m = LearnedLikelihood(p)
    _, z_bf16 = m.eval_groups(e, f)
    m.set_fp8(True)
    ll, z3 = m.eval_groups(e, f)
    fb = np.broadcast_to(f[:, None, :], (G, 16, 7))
    fw = ol.mlp_forward(p, e, fb, mode="fp8")
    np.testing.assert_allclose(z3, fw["z3"], rtol=0, atol=2e-3)
    np.testing.assert_allclose(ll, fw["ll"], rtol=5e-3, atol=5e-3)
    assert np.max(np.abs(z3 - z_bf16)) > 1e-3
    # the master parameters are untouched by the mode
    np.testing.assert_array_equal(m.get_params(), p)
    m.set_fp8(False)
    _, z_back = m.eval_groups(e, f)
    np.testing.assert_array_equal(z_back, z_bf16)


def test_fp8_code_tables_equal_per_pair_network(gpu):
    """Code tables built from the FP8 network equal its per-pair evaluation."""
    from gnss_gpu_amd import ParticleFilterDevice, urban
    from gnss_gpu_amd.learned import LearnedLikelihood

    d = urban.drive_ecef(np.random.default_rng(23), 4, n_sat=12)
    m = LearnedLikelihood(_params(19))
    m.set_fp8(True)
    out = []
    for table in (False, True):
        pf = ParticleFilterDevice(n_particles=50_000, seed=6)
        pf.set_learned_likelihood(m, table=table)
        pf.initialize(d.truth[1], clock_bias=d.cb, spread_pos=30.0, spread_cb=30.0)
        pf.predict(velocity=d.vel[2], dt=1.0)
        pf.update(d.sat[2], d.pr[2], cn0=d.cn0[2], los_flag=d.los_flags(d.truth[2], 2),
                  resample=False)
        out.append(pf.get_log_weights())
    np.testing.assert_array_equal(out[0], out[1])
