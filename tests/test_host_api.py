"""Host-side behaviour that needs no GPU: argument checks that run before any
device work, and the synthetic scenario generators."""

import numpy as np
import pytest

from gnss_gpu_amd import sim


def test_wls_rejects_fewer_than_four_satellites():
    from gnss_gpu_amd import wls_batch, wls_position

    with pytest.raises(RuntimeError, match="at least 4 satellites"):
        wls_position(np.zeros(9), np.zeros(3), np.ones(3))
    with pytest.raises(RuntimeError, match="at least 4 satellites"):
        wls_batch(np.zeros((2, 3, 3)), np.zeros((2, 3)), np.ones((2, 3)))


def test_wls_batch_shape_checks():
    from gnss_gpu_amd import wls_batch

    with pytest.raises(ValueError):
        wls_batch(np.zeros((2, 5)), np.zeros((2, 5)), np.ones((2, 5)))
    with pytest.raises(ValueError):
        wls_batch(np.zeros((2, 5, 3)), np.zeros((2, 5)), np.ones((2, 5)), n_sat=[5])


def test_particle_filter_device_validates_modes_before_allocating():
    from gnss_gpu_amd import ParticleFilterDevice

    with pytest.raises(ValueError):
        ParticleFilterDevice(n_particles=10, megopolis_mode="bogus")
    with pytest.raises(ValueError):
        ParticleFilterDevice(n_particles=10, systematic_u0="bogus")


def test_legacy_filter_requires_initialize():
    from gnss_gpu_amd import ParticleFilter

    pf = ParticleFilter(n_particles=16)
    for call in (lambda: pf.predict(), lambda: pf.estimate(), lambda: pf.get_ess(),
                 lambda: pf.get_particles(), lambda: pf.update(np.zeros((4, 3)), np.zeros(4))):
        with pytest.raises(RuntimeError, match="not initialized"):
            call()


def test_scenario_geometry():
    sc = sim.scenario_plain(n_epochs=5, seed=0)
    rho = np.linalg.norm(sc.sat, axis=1)
    np.testing.assert_allclose(rho, sim.SAT_RADIUS, rtol=1e-12)
    # elevations reproduce the requested ones
    R = sim.enu_rotation(sc.truth[0])
    enu = (sc.sat - sc.truth[0]) @ R.T
    el = np.degrees(np.arcsin(enu[:, 2] / np.linalg.norm(enu, axis=1)))
    np.testing.assert_allclose(el, sc.el, atol=1e-9)
    assert sc.pr.shape == (5, 8)
    assert np.linalg.norm(sc.truth[1] - sc.truth[0]) == pytest.approx(10.0)


def test_multi_gnss_nlos_fraction_and_isb():
    sc = sim.scenario_multi_gnss(n_epochs=50, seed=1)
    assert sc.n_sat == 12
    assert np.all(sc.nlos.sum(axis=1) == round(0.3 * 12))
    resid = sc.pr - np.linalg.norm(sc.sat[None] - sc.truth[:, None], axis=2) - sc.cb
    los = ~sc.nlos
    for s_id in (1, 2, 3):
        m = los & (sc.sys[None, :] == s_id)
        assert abs(resid[m].mean() - sc.isb[s_id - 1]) < 1.5


def test_wls_epochs_ragged():
    sat, pr, w, counts, truth = sim.wls_epochs(n_epoch=50, seed=3)
    assert counts.min() >= 6 and counts.max() <= 20
    for e in range(50):
        assert np.all(w[e, counts[e]:] == 0) and np.all(w[e, :counts[e]] == 1)


def test_multi_gnss_solver_host_logic():
    """multi_gnss.py:38-53, 218-240 and test_multi_gnss.py:174-197: PRN parsing,
    names, the sorted system numbering and the early -1 for too few satellites
    (no device call)."""
    import json
    import os

    from gnss_gpu_amd.multi_gnss import (SYSTEM_BEIDOU, SYSTEM_GALILEO, SYSTEM_GPS,
                                         MultiGNSSSolver)

    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                    "reference_cases.json")))["multi_gnss_scenario"]
    for prn, sys_id, num in g["prn_to_system"]:
        assert MultiGNSSSolver.prn_to_system(prn) == (sys_id, num)
    assert MultiGNSSSolver.prn_to_system("X09") == (SYSTEM_GPS, 9)
    for sys_id, name in g["system_names"]:
        assert MultiGNSSSolver.system_name(sys_id) == name
    assert MultiGNSSSolver.system_name(9) == "Unknown(9)"
    solver = MultiGNSSSolver(systems=[SYSTEM_BEIDOU, SYSTEM_GPS, SYSTEM_GALILEO])
    assert solver.systems == [SYSTEM_GPS, SYSTEM_GALILEO, SYSTEM_BEIDOU] and solver.n_systems == 3
    np.testing.assert_array_equal(solver._map_ids([3, 0, 2, 1, 4, 3]), [2, 0, 1, 0, 0, 2])
    ins = g["insufficient"]
    pos, biases, it = MultiGNSSSolver(systems=ins["systems"]).solve(
        np.array(ins["sat_ecef"]), np.array(ins["pr"]), np.array(ins["systems"]))
    assert it == -1 and np.all(pos == 0) and biases == {0: 0.0, 2: 0.0}
    pos, cbs, its = MultiGNSSSolver(systems=[0, 1, 2]).solve_batch(
        np.zeros((3, 5, 3)), np.zeros((3, 5)), np.zeros((3, 5), dtype=np.int32))
    assert pos.shape == (3, 3) and cbs.shape == (3, 3) and np.all(its == -1)


def test_multi_gnss_argument_checks():
    from gnss_gpu_amd._gnss_gpu_multi_gnss import wls_multi_gnss, wls_multi_gnss_batch

    with pytest.raises(RuntimeError, match="3 \\+ n_systems"):
        wls_multi_gnss_batch(np.zeros((2, 5, 3)), np.zeros((2, 5)), np.ones((2, 5)),
                             np.zeros((2, 5), dtype=np.int32), 3)
    with pytest.raises(ValueError, match="n_systems"):
        wls_multi_gnss(np.zeros(30), np.zeros(10), np.ones(10), np.zeros(10, np.int32), 6)
    with pytest.raises(ValueError):
        wls_multi_gnss_batch(np.zeros((2, 9)), np.zeros((2, 9)), np.ones((2, 9)),
                             np.zeros((2, 9), dtype=np.int32), 1)


def _random_boxes(n_boxes, seed):
    """test_bvh.py:64-77: a mesh of random boxes, 12 triangles each."""
    from gnss_gpu_amd.raytrace import BuildingModel

    rng = np.random.RandomState(seed)
    out = []
    for _ in range(n_boxes):
        c = [rng.uniform(-500, 500), rng.uniform(-500, 500), rng.uniform(0, 100)]
        out.append(BuildingModel.create_box(c, rng.uniform(5, 30), rng.uniform(5, 30),
                                            rng.uniform(10, 60)).triangles)
    return np.concatenate(out, axis=0)


def test_bvh_build_is_a_valid_tree():
    """Host build (no GPU): every triangle in exactly one leaf, leaves of at most
    4, each box contains its children's boxes or its triangles, depth <= 62."""
    from gnss_gpu_amd.bvh import BVHAccelerator

    for tris in (_random_boxes(100, 42), _random_boxes(1, 0), np.zeros((0, 3, 3))):
        b = BVHAccelerator(tris)
        assert b.n_triangles == tris.shape[0]
        if tris.shape[0] == 0:
            assert b.n_nodes == 0
            continue
        nodes = b._nodes_flat
        assert sorted(b._sorted_indices.tolist()) == list(range(tris.shape[0]))
        np.testing.assert_array_equal(b._sorted_tris, tris[b._sorted_indices])
        covered = np.zeros(tris.shape[0], dtype=int)
        stack = [(0, 0)]
        while stack:
            i, depth = stack.pop()
            assert depth <= 62
            lo, hi = nodes[i, :3], nodes[i, 3:6]
            left, right, start, count = (int(v) for v in nodes[i, 6:])
            if left == -1:
                assert 1 <= count <= 4
                covered[start:start + count] += 1
                pts = b._sorted_tris[start:start + count].reshape(-1, 3)
                assert np.all(pts >= lo - 1e-9) and np.all(pts <= hi + 1e-9)
            else:
                for ch in (left, right):
                    assert np.all(nodes[ch, :3] >= lo) and np.all(nodes[ch, 3:6] <= hi)
                    stack.append((ch, depth + 1))
        assert np.all(covered == 1)


def test_malformed_trees_are_rejected_before_any_device_work():
    from gnss_gpu_amd._bvh import raytrace_los_check_bvh

    tris = _random_boxes(2, 1)
    from gnss_gpu_amd.bvh import BVHAccelerator

    b = BVHAccelerator(tris)
    sat = np.array([[0.0, 0.0, 2e7]])
    bad = b._nodes_flat.copy()
    bad[0, 6] = 0  # root is its own child
    with pytest.raises(RuntimeError, match="child index"):
        raytrace_los_check_bvh(np.zeros(3), sat, bad, b._sorted_tris)
    bad = b._nodes_flat.copy()
    leaf = int(np.nonzero(bad[:, 6] == -1)[0][0])
    bad[leaf, 8] = tris.shape[0]  # leaf range past the triangle array
    with pytest.raises(RuntimeError, match="leaf range"):
        raytrace_los_check_bvh(np.zeros(3), sat, bad, b._sorted_tris)
    bad = b._nodes_flat.copy()
    bad[0, 7] = 0.5
    with pytest.raises(RuntimeError, match="integers"):
        raytrace_los_check_bvh(np.zeros(3), sat, bad, b._sorted_tris)
    # a chain deeper than 62 levels
    chain = np.zeros((2 * 70 + 1, 10))
    for i in range(70):
        chain[2 * i, 6:] = [2 * i + 1, 2 * i + 2, 0, 0]
        chain[2 * i + 1, 6:] = [-1, -1, 0, 1]
    chain[140, 6:] = [-1, -1, 0, 1]
    with pytest.raises(RuntimeError, match="deeper"):
        raytrace_los_check_bvh(np.zeros(3), sat, chain, tris[:1])


def test_bvh_and_pf3d_argument_checks():
    from gnss_gpu_amd import ParticleFilter3D, ParticleFilter3DBVH
    from gnss_gpu_amd._gnss_gpu_pf3d_bvh import pf_weight_3d_bvh
    from gnss_gpu_amd.bvh import BVHAccelerator

    with pytest.raises(ValueError):
        BVHAccelerator(np.zeros((10, 2, 3)))
    b = BVHAccelerator.create_box([0, 0, 0], 10, 10, 10)
    assert b.n_triangles == 12 and b.n_nodes >= 1
    with pytest.raises(ValueError):
        b.check_los_batch(np.zeros((4, 3)), np.zeros((4, 3)))
    with pytest.raises(ValueError):
        b.check_los_batch(np.zeros((4, 2)), np.zeros((4, 2, 3)))
    with pytest.raises(ValueError):
        b.check_los_batch(np.zeros((3, 3)), np.zeros((4, 2, 3)))
    with pytest.raises(TypeError):
        ParticleFilter3DBVH(bvh="not_a_bvh", n_particles=100)
    with pytest.raises(TypeError):
        ParticleFilter3D(building_model="x", n_particles=100)
    pf = ParticleFilter3DBVH(bvh=b, n_particles=500, sigma_los=4.0, sigma_nlos=25.0,
                             nlos_bias=15.0)
    assert (pf.sigma_los, pf.sigma_nlos, pf.nlos_bias, pf.n_particles) == (4.0, 25.0, 15.0, 500)
    with pytest.raises(RuntimeError, match="not initialized"):
        pf.update(np.zeros((1, 3)), np.zeros(1))
    with pytest.raises(ValueError, match="log_weights"):
        pf_weight_3d_bvh(np.zeros(4), np.zeros(4), np.zeros(4), np.zeros(4), np.zeros(3),
                         np.zeros(1), np.ones(1), b._nodes_flat, b._sorted_tris,
                         np.zeros(4, dtype=np.float32), 4, 1, 3.0, 30.0, 20.0)


def test_ffbsi_rejects_forced_draws_outside_the_population():
    """gg_ffbsi_sample checks the forced draws before it touches the device."""
    import ctypes as C

    from gnss_gpu_amd._lib import call, dptr, iptr

    T, N = 3, 5
    lw = np.zeros((T, N))
    X = np.zeros((T, N, 4))
    vel, dt, sp = np.zeros((T, 3)), np.ones(T), np.ones(T)
    u = np.full((1, T), 0.5)
    idx = np.zeros((1, T), dtype=np.int64)
    st = np.zeros(1, dtype=np.int32)
    for bad in (N, -2):
        forced = np.array([[-1, bad, -1]], dtype=np.int32)
        with pytest.raises(RuntimeError, match="forced index"):
            call("gg_ffbsi_sample", dptr(lw), dptr(X), dptr(vel), dptr(dt), dptr(sp), 3.0, None,
                 T, N, 1, dptr(u), iptr(forced), C.c_void_p(idx.ctypes.data), iptr(st))


def test_systematic_count_fast_path_bound():
    """The resampling kernel counts the systematic targets (u0 + j) / n <= c
    as floor(c n - u0) + 1 whenever c n - u0 lies farther than
    1e-14 (n + |c n - u0|) from an integer, and compares exactly otherwise.
    Check that rule against the exact count on targets, their neighbours and
    random CDF values, for small and large n."""
    from oracle import detmath

    rng = np.random.default_rng(3)
    for n in (1, 7, 1000, 1_000_000, 5_000_000):
        u0 = float(rng.random()) / n
        j = np.unique(np.concatenate([rng.integers(0, n, 4000), [0, n - 1]]))
        t = (u0 + j.astype(np.float64)) / n
        c = np.concatenate([t, np.nextafter(t, np.inf), np.nextafter(t, -np.inf),
                            rng.random(4000), [0.0, 1.0, 1.0 + 1e-12, -1e-300]])
        x = c * n - u0
        est = np.floor(x)
        fr = x - est
        eps = 1e-14 * (n + np.abs(x))
        fast = (fr > eps) & (fr < 1.0 - eps)
        k_fast = np.clip(est + 1, 0, n).astype(np.int64)
        ref = np.searchsorted(detmath.systematic_targets(n, u0), c, side="right")
        assert fast[3 * t.size: 3 * t.size + 4000].sum() > 3900  # random CDF values: fast
        np.testing.assert_array_equal(k_fast[fast], ref[fast])
        # every exact target and its neighbours fall to the exact comparisons
        assert not np.any(fast[: 3 * t.size] & (np.abs(fr[: 3 * t.size] - 0.5) > 0.49))


def test_reference_module_aliases():
    """gnss_gpu.particle_filter_3d_bvh and gnss_gpu._gnss_gpu_pf3d resolve to the
    modules that hold their names here."""
    import importlib

    m = importlib.import_module("gnss_gpu_amd.particle_filter_3d_bvh")
    from gnss_gpu_amd.particle_filter_3d import ParticleFilter3DBVH

    assert m.ParticleFilter3DBVH is ParticleFilter3DBVH
    from gnss_gpu_amd._gnss_gpu_pf3d import pf_weight_3d
    from gnss_gpu_amd._gnss_gpu_pf3d_bvh import pf_weight_3d as direct

    assert pf_weight_3d is direct
