"""GPU batched WLS against the binary64 restatement of src/positioning/wls.cu."""

import json
import os

import numpy as np
import pytest

from oracle import wls as owls

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_cases.json")))


def _tokyo():
    g = GOLD["tokyo_constellation"]
    return np.array(g["sat_ecef"]), np.array(g["true_pos"]), g["true_cb"]


def test_wls_position_non_sagnac_truth(gpu):
    from gnss_gpu_amd import wls_position

    sat, tp, cb = _tokyo()
    pr = np.linalg.norm(sat - tp, axis=1) + cb
    res, it = wls_position(sat.flatten(), pr, np.ones(8))
    assert np.linalg.norm(res[:3] - tp) < GOLD["wls_bounds"]["single_epoch_error_m"]
    assert abs(res[3] - cb) < 0.01
    ref, rit = owls.wls_position_ref(sat, pr, np.ones(8))
    np.testing.assert_allclose(res, ref, rtol=0, atol=1e-6)
    assert it == rit and it <= 10


def test_wls_batch_sagnac_truth_and_noise(gpu):
    from gnss_gpu_amd import wls_batch
    from gnss_gpu_amd.range_model import geometric_ranges_sagnac

    sat, tp, cb = _tokyo()
    pr = geometric_ranges_sagnac(tp, sat) + cb
    E = 100
    rng = np.random.default_rng(42)
    sb = np.tile(sat, (E, 1, 1))
    pb = np.tile(pr, (E, 1)) + rng.normal(0, 3.0, (E, 8))
    pb[0] = pr
    wb = np.ones((E, 8))
    res, it = wls_batch(sb, pb, wb)
    assert res.shape == (E, 4) and it.dtype == np.int32
    assert np.linalg.norm(res[0, :3] - tp) < 1e-3
    err = np.linalg.norm(res[:, :3] - tp, axis=1)
    assert np.all(err < GOLD["wls_bounds"]["batch_error_m_with_3m_noise"])
    ref, rit = owls.wls_batch_ref(sb, pb, wb)
    np.testing.assert_allclose(res, ref, rtol=0, atol=1e-6)
    np.testing.assert_array_equal(it, rit)


def test_wls_batch_ragged_counts_large(gpu):
    from gnss_gpu_amd import sim, wls_batch

    sat, pr, w, counts, truth = sim.wls_epochs(n_epoch=10_000, seed=7)
    res, it = wls_batch(sat, pr, w, n_sat=counts)
    ref, rit = owls.wls_batch_ref(sat, pr, w, n_sat=counts)
    np.testing.assert_allclose(res, ref, rtol=0, atol=1e-5)
    np.testing.assert_array_equal(it, rit)
    err = np.linalg.norm(res[:, :3] - truth, axis=1)
    assert np.median(err) < 10.0


def test_wls_batch_short_epoch_and_weights(gpu):
    from gnss_gpu_amd import wls_batch

    sat, tp, cb = _tokyo()
    pr = np.linalg.norm(sat - tp, axis=1) + cb
    sb = np.tile(sat, (3, 1, 1))
    pb = np.tile(pr, (3, 1))
    w = np.ones((3, 8))
    w[1] = np.linspace(0.5, 2.0, 8)
    res, it = wls_batch(sb, pb, w, n_sat=[8, 8, 3])
    assert it[2] == -1 and np.all(np.isnan(res[2]))
    ref, rit = owls.wls_batch_ref(sb[:2], pb[:2], w[:2])
    np.testing.assert_allclose(res[:2], ref, rtol=0, atol=1e-6)


def test_wls_position_singular_geometry(gpu):
    from gnss_gpu_amd import wls_position

    s1 = np.tile(np.array([[1.5e7, 1.0e7, 2.0e7]]), (5, 1))
    res, it = wls_position(s1.ravel(), np.full(5, 2.2e7), np.ones(5))
    ref, rit = owls.wls_position_ref(s1, np.full(5, 2.2e7), np.ones(5))
    assert it == rit == 0
    np.testing.assert_allclose(res, ref, rtol=1e-12)


def test_wls_batch_device_tensors(gpu):
    import torch

    from gnss_gpu_amd import sim, wls_batch_device

    sat, pr, w, counts, truth = sim.wls_epochs(n_epoch=2000, seed=9)
    dev = torch.device("cuda:0")
    ts = torch.from_numpy(sat).to(dev)
    tp = torch.from_numpy(pr).to(dev)
    tw = torch.from_numpy(w).to(dev)
    tn = torch.from_numpy(counts).to(dev)
    res = torch.empty((2000, 4), dtype=torch.float64, device=dev)
    it = torch.empty(2000, dtype=torch.int32, device=dev)
    wls_batch_device(ts, tp, tw, res, it, n_sat=tn)
    torch.cuda.synchronize()
    ref, rit = owls.wls_batch_ref(sat, pr, w, n_sat=counts)
    np.testing.assert_allclose(res.cpu().numpy(), ref, rtol=0, atol=1e-5)
    np.testing.assert_array_equal(it.cpu().numpy(), rit)
