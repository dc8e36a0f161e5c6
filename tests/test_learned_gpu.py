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
