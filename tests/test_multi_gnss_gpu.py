"""GPU multi-GNSS WLS against the binary64 restatement of
src/positioning/multi_gnss.cu, on the scenarios of tests/test_multi_gnss.py."""

import json
import os

import numpy as np
import pytest

from oracle import wls as owls

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_cases.json")))
MG = GOLD["multi_gnss_scenario"]


def _scenario(isb_gal=0.0, isb_glo=0.0, sagnac=True):
    from gnss_gpu_amd.range_model import geometric_ranges_sagnac

    sat = np.vstack([MG["gps_sats"], MG["galileo_sats"], MG["glonass_sats"]])
    tp = np.array(MG["true_pos"])
    sid = np.array([0] * 4 + [2] * 4 + [1] * 4, dtype=np.int32)
    rng = geometric_ranges_sagnac(tp, sat) if sagnac else np.linalg.norm(sat - tp, axis=1)
    pr = rng + MG["true_cb_gps"] + np.repeat([0.0, isb_gal, isb_glo], 4)
    return sat, pr, sid, tp


def test_solver_recovers_reference_isbs(gpu):
    """test_multi_gnss.py:109-151 through MultiGNSSSolver.solve_batch (Sagnac)."""
    from gnss_gpu_amd.multi_gnss import MultiGNSSSolver

    for case in MG["isb_cases"]:
        sat, pr, sid, tp = _scenario(case["isb_galileo"], case["isb_glonass"])
        n = case["n_sat"]
        solver = MultiGNSSSolver(systems=case["systems"])
        pos, cbs, its = solver.solve_batch(sat[None, :n], pr[None, :n], sid[None, :n])
        assert np.linalg.norm(pos[0] - tp) < MG["position_error_m"]
        cb = dict(zip(solver.systems, cbs[0]))
        assert abs(cb[2] - cb[0] - case["isb_galileo"]) < MG["isb_error_m"]
        if 1 in cb:
            assert abs(cb[1] - cb[0] - case["isb_glonass"]) < MG["isb_error_m"]
        mapped = solver._map_ids(sid[None, :n])
        ref, rit = owls.wls_multi_gnss_batch_ref(sat[None, :n], pr[None, :n], np.ones((1, n)),
                                                 mapped, solver.n_systems)
        np.testing.assert_allclose(np.hstack([pos, cbs]), ref, rtol=0, atol=1e-6)
        np.testing.assert_array_equal(its, rit)


def test_single_epoch_plain_ranges(gpu):
    """wls_multi_gnss keeps the reference CPU solver's plain-range model
    (multi_gnss.cu:50-153): exact on plain ranges, equal to the oracle on
    Sagnac ranges (where it is ~30 m off, as the reference is)."""
    from gnss_gpu_amd.multi_gnss import MultiGNSSSolver

    solver = MultiGNSSSolver(systems=[0, 1, 2])
    for sagnac in (False, True):
        sat, pr, sid, tp = _scenario(15.0, -8.0, sagnac=sagnac)
        pos, biases, it = solver.solve(sat, pr, sid)
        ref, rit = owls.wls_multi_gnss_ref(sat, pr, np.ones(12), solver._map_ids(sid), 3)
        np.testing.assert_allclose(np.concatenate([pos, [biases[0], biases[1], biases[2]]]), ref,
                                   rtol=0, atol=1e-6)
        assert it == rit
        if not sagnac:
            assert np.linalg.norm(pos - tp) < MG["gps_only_error_m"]
            assert abs(biases[2] - biases[0] - 15.0) < 1e-6


def test_raw_binding_gps_and_dual_system(gpu):
    """test_multi_gnss.py:207-232 through the module functions (plain ranges for
    the single-epoch solver, Sagnac for the batch)."""
    from gnss_gpu_amd._gnss_gpu_multi_gnss import wls_multi_gnss, wls_multi_gnss_batch

    sat = np.array(GOLD["tokyo_constellation"]["sat_ecef"])
    tp = np.array(MG["true_pos"])
    pr = np.linalg.norm(sat - tp, axis=1) + 3000.0
    res, it = wls_multi_gnss(sat.ravel(), pr, np.ones(8), np.zeros(8, np.int32), 1, 10, 1e-4)
    assert res.shape == (4,) and np.linalg.norm(res[:3] - tp) < MG["gps_only_error_m"]
    sat, pr, sid, tp = _scenario(10.0, sagnac=False)
    mapped = np.where(sid == 2, 1, 0).astype(np.int32)
    res, it = wls_multi_gnss(sat[:8].ravel(), pr[:8], np.ones(8), mapped[:8], 2, 10, 1e-4)
    assert np.linalg.norm(res[:3] - tp) < MG["position_error_m"]
    assert abs(res[4] - res[3] - 10.0) < MG["isb_error_m"]
    sat, pr, sid, tp = _scenario(10.0)
    res, it = wls_multi_gnss_batch(sat[None, :8], pr[None, :8], np.ones((1, 8)),
                                   mapped[None, :8], 2)
    assert res.shape == (1, 5) and it.dtype == np.int32
    assert abs(res[0, 4] - res[0, 3] - 10.0) < MG["isb_error_m"]


def test_batch_noise_matches_oracle(gpu):
    """test_multi_gnss.py:234-260: 50 epochs with 2 m noise, GPS/GLONASS/Galileo."""
    from gnss_gpu_amd._gnss_gpu_multi_gnss import wls_multi_gnss_batch

    sat, pr, sid, tp = _scenario(5.0)
    E = 50
    rng = np.random.default_rng(42)
    sb = np.tile(sat, (E, 1, 1))
    pb = np.tile(pr, (E, 1)) + rng.normal(0, MG["batch_noise_sigma_m"], (E, 12))
    mapped = np.array([{0: 0, 1: 1, 2: 2}[s] for s in sid], dtype=np.int32)
    ib = np.tile(mapped, (E, 1))
    wb = rng.uniform(0.5, 2.0, (E, 12))
    res, it = wls_multi_gnss_batch(sb, pb, wb, ib, 3)
    assert np.all(np.linalg.norm(res[:, :3] - tp, axis=1) < MG["batch_error_m"])
    ref, rit = owls.wls_multi_gnss_batch_ref(sb, pb, wb, ib, 3)
    np.testing.assert_allclose(res, ref, rtol=0, atol=1e-6)
    np.testing.assert_array_equal(it, rit)


@pytest.mark.parametrize("n_systems", [0, 1, 2, 3, 4, 5])
def test_every_system_count_matches_oracle(gpu, n_systems):
    """Random constellations, 40 satellites (two passes of the 32-lane group).
    Every third epoch has one satellite with a foreign id (-1 or 8; its row has
    no clock column, so convergence turns linear), and every fifth epoch has no
    satellite of the last system (its clock column stays zero)."""
    from gnss_gpu_amd import sim
    from gnss_gpu_amd._gnss_gpu_multi_gnss import wls_multi_gnss_batch

    E = 300
    sat, pr, w, ns, truth = sim.wls_epochs(n_epoch=E, s_min=40, s_max=40, seed=20 + n_systems)
    rng = np.random.default_rng(n_systems)
    ids = rng.integers(0, max(n_systems, 1), pr.shape).astype(np.int32)
    ids[1::3, 7] = rng.choice([-1, 8], len(ids[1::3]))
    if n_systems > 1:
        ids[::5][ids[::5] == n_systems - 1] = 0
    isb = np.concatenate([[0.0], rng.normal(0, 30.0, 5)])
    pr = pr + np.where((ids >= 0) & (ids < n_systems), isb[np.clip(ids, 0, 5)], 0.0)
    w = rng.uniform(0.2, 3.0, pr.shape)
    res, it = wls_multi_gnss_batch(sat, pr, w, ids, n_systems, max_iter=60)
    ref, rit = owls.wls_multi_gnss_batch_ref(sat, pr, w, ids, n_systems, max_iter=60)
    assert np.all(rit < 60)
    np.testing.assert_allclose(res, ref, rtol=0, atol=1e-6)
    np.testing.assert_array_equal(it, rit)


def test_device_tensors_and_large_batch(gpu):
    import torch

    from gnss_gpu_amd import sim
    from gnss_gpu_amd._gnss_gpu_multi_gnss import wls_multi_gnss_batch_device

    E, S, NS = 10_000, 24, 4
    sat, pr, w, ns, truth = sim.wls_epochs(n_epoch=E, s_min=S, s_max=S, seed=31)
    ids = np.tile(np.arange(S, dtype=np.int32) % NS, (E, 1))
    pr = pr + np.array([0.0, 12.0, -7.5, 20.0])[ids]
    dev = torch.device("cuda:0")
    t = [torch.from_numpy(a).to(dev) for a in (sat, pr, w, ids)]
    res = torch.empty((E, 3 + NS), dtype=torch.float64, device=dev)
    it = torch.empty(E, dtype=torch.int32, device=dev)
    wls_multi_gnss_batch_device(*t, res, it, NS)
    torch.cuda.synchronize()
    ref, rit = owls.wls_multi_gnss_batch_ref(sat, pr, w, ids, NS)
    np.testing.assert_allclose(res.cpu().numpy(), ref, rtol=0, atol=1e-5)
    np.testing.assert_array_equal(it.cpu().numpy(), rit)
    # the single-epoch rule on the device: too few satellites and zero centroid
    res1 = torch.empty((2, 3 + NS), dtype=torch.float64, device=dev)
    it1 = torch.empty(2, dtype=torch.int32, device=dev)
    z = [torch.zeros((2, 6, 3), dtype=torch.float64, device=dev),
         torch.full((2, 6), 2e7, dtype=torch.float64, device=dev),
         torch.ones((2, 6), dtype=torch.float64, device=dev),
         torch.zeros((2, 6), dtype=torch.int32, device=dev)]
    wls_multi_gnss_batch_device(*z, res1, it1, NS, mode=1)
    torch.cuda.synchronize()
    assert np.all(it1.cpu().numpy() == -1) and np.all(res1.cpu().numpy() == 0)
    wls_multi_gnss_batch_device(*z, res1[:, :6].contiguous(), it1, 3, mode=1)
    torch.cuda.synchronize()
    assert np.all(it1.cpu().numpy() == -1)
