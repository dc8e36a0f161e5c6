"""GPU ray casting (BVH LOS, multipath, ray-traced NLOS weight) against the
brute-force binary64 oracle, on the scenarios of tests/test_bvh.py and
tests/test_pf3d_bvh.py."""

import numpy as np
import pytest

from oracle import raytrace as oray

pytestmark = pytest.mark.gpu


def _box():
    from gnss_gpu_amd.raytrace import BuildingModel

    return BuildingModel.create_box(center=[100.0, 0.0, 25.0], width=20.0, depth=20.0,
                                    height=50.0)


def _random_boxes(n_boxes, seed):
    from gnss_gpu_amd.raytrace import BuildingModel

    rng = np.random.RandomState(seed)
    out = []
    for _ in range(n_boxes):
        c = [rng.uniform(-500, 500), rng.uniform(-500, 500), rng.uniform(0, 100)]
        out.append(BuildingModel.create_box(c, rng.uniform(5, 30), rng.uniform(5, 30),
                                            rng.uniform(10, 60)).triangles)
    return np.concatenate(out, axis=0)


def _sky(rng, n, r_lo=100.0, r_hi=3.0e7, el_lo=-np.pi / 2):
    th = rng.uniform(0, 2 * np.pi, n)
    ph = rng.uniform(el_lo, np.pi / 2, n)
    r = rng.uniform(r_lo, r_hi, n)
    return np.stack([r * np.cos(ph) * np.cos(th), r * np.cos(ph) * np.sin(th), r * np.sin(ph)], 1)


def test_box_los_and_linear_equivalence(gpu):
    """test_bvh.py:26-47, 81-113: box cases; BVH and the triangle-list query
    agree with the oracle on 1200 triangles and 200 random directions."""
    from gnss_gpu_amd.bvh import BVHAccelerator
    from gnss_gpu_amd.raytrace import BuildingModel

    b = BVHAccelerator.from_building_model(_box())
    rx = np.zeros(3)
    assert b.check_los(rx, [[0.0, 0.0, 2e7]])[0]
    assert not b.check_los(rx, [[200.0, 0.0, 25.0]])[0]
    tris = _random_boxes(100, 42)
    bvh, lin = BVHAccelerator(tris), BuildingModel(tris)
    sat = _sky(np.random.RandomState(42), 200)
    ref = oray.los_check(rx, sat, tris)[0]
    np.testing.assert_array_equal(bvh.check_los(rx, sat), ref)
    np.testing.assert_array_equal(lin.check_los(rx, sat), ref)
    assert 0 < ref.sum() < 200


def test_los_batch_and_nan_padding(gpu):
    """test_bvh.py:147-191: batch equals per-epoch queries; NaN padding is not LOS."""
    from gnss_gpu_amd.bvh import BVHAccelerator

    b = BVHAccelerator.from_building_model(_box())
    rng = np.random.RandomState(7)
    rx = rng.uniform(-50, 50, size=(16, 3))
    sat = np.zeros((16, 8, 3))
    for s in range(8):
        sat[:, s] = [0.0, 0.0, 2e7 + s] if s % 2 == 0 else [200.0, float(s), 25.0]
    batch = b.check_los_batch(rx, sat)
    single = np.stack([b.check_los(rx[e], sat[e]) for e in range(16)])
    np.testing.assert_array_equal(batch, single)
    np.testing.assert_array_equal(batch, oray.los_check(rx, sat, b.triangles, batch=True))
    rx2 = np.array([[0.0, 0.0, 0.0], [10.0, 0.0, 0.0]])
    sat2 = np.array([[[0.0, 0.0, 2e7], [200.0, 0.0, 25.0]], [[0.0, 0.0, 2e7], [np.nan] * 3]])
    assert b.check_los_batch(rx2, sat2).tolist() == [[True, False], [True, False]]


def test_large_batch_los_matches_oracle(gpu):
    """20,000 rays through 3600 triangles (300 boxes)."""
    from gnss_gpu_amd.bvh import BVHAccelerator

    tris = _random_boxes(300, 5)
    b = BVHAccelerator(tris)
    rng = np.random.RandomState(3)
    rx = np.c_[rng.uniform(-400, 400, (500, 2)), rng.uniform(0, 40, 500)]
    sat = np.stack([_sky(rng, 40, 1e3, 3e7, el_lo=0.05) for _ in range(500)])
    got = b.check_los_batch(rx, sat)
    ref = oray.los_check(rx, sat, tris, batch=True)
    np.testing.assert_array_equal(got, ref)
    assert 0.05 < 1 - ref.mean() < 0.95


def test_multipath_matches_oracle(gpu):
    """test_bvh.py:193-222 scenario: batch, single and triangle-list multipath
    equal the brute-force oracle (delays 1e-9 m)."""
    from gnss_gpu_amd.bvh import BVHAccelerator

    b = BVHAccelerator.from_building_model(_box())
    rng = np.random.RandomState(11)
    rx = rng.uniform(-50, 50, size=(6, 3))
    rx[:, 2] = rng.uniform(0, 30, size=6)
    sat = np.zeros((6, 4, 3))
    for e in range(6):
        for s in range(4):
            sat[e, s] = [200.0 + 10.0 * s, 5.0 * (e - 3), 25.0 + s]
    delay, refl = b.compute_multipath_batch(rx, sat)
    rrefl, rdelay = oray.multipath(rx, sat, b._sorted_tris, batch=True)
    np.testing.assert_allclose(delay, rdelay, atol=1e-9)
    np.testing.assert_allclose(refl, rrefl, atol=1e-9)
    assert (rdelay > 0).any()
    d1, r1 = b.compute_multipath(rx[0], sat[0])
    np.testing.assert_allclose(d1, rdelay[0], atol=1e-9)
    from gnss_gpu_amd.raytrace import BuildingModel

    d2, r2 = BuildingModel(b.triangles).compute_multipath(rx[0], sat[0])
    np.testing.assert_allclose(d2, rdelay[0], atol=1e-9)
    wall = np.array([[[10.0, -100.0, -50.0], [10.0, 100.0, -50.0], [10.0, 100.0, 150.0]],
                     [[10.0, -100.0, -50.0], [10.0, 100.0, 150.0], [10.0, -100.0, 150.0]]])
    s = np.array([-2.0e7, 3.0e6, 1.5e7])
    d3, r3 = BuildingModel(wall).compute_multipath([0.0, 0.0, 1.5], s[None])
    img = np.array([20.0, 0.0, 1.5])
    np.testing.assert_allclose(d3[0], np.linalg.norm(img - s) - np.linalg.norm([0, 0, 1.5] - s),
                               rtol=1e-9)


def _particles(n, seed, lo=-10, hi=10):
    rng = np.random.RandomState(seed)
    return (rng.uniform(lo, hi, n), rng.uniform(lo, hi, n), rng.uniform(lo, hi, n),
            rng.uniform(50, 150, n))


def test_weight_closed_forms(gpu):
    """test_pf3d_bvh.py:94-214 through pf_weight_3d_bvh."""
    from gnss_gpu_amd._gnss_gpu_pf3d_bvh import pf_weight_3d_bvh
    from gnss_gpu_amd.bvh import BVHAccelerator

    b = BVHAccelerator.from_building_model(_box())
    z = np.zeros(100)
    cb = np.full(100, 100.0)

    def run(bv, sat, pr, **kw):
        lw = np.zeros(100)
        pf_weight_3d_bvh(z, z, z, cb, np.asarray(sat, float).ravel(), np.asarray(pr, float),
                         np.ones(len(pr)), bv._nodes_flat, bv._sorted_tris, lw, 100, len(pr),
                         3.0, 30.0, 20.0, **kw)
        return lw

    np.testing.assert_allclose(run(b, [0.0, 0.0, 20000.0], [20100.0]), 0.0, atol=1e-10)
    dist = float(np.hypot(200.0, 25.0))
    np.testing.assert_allclose(run(b, [200.0, 0.0, 25.0], [dist + 120.0]), 0.0, atol=1e-10)
    np.testing.assert_allclose(run(b, [200.0, 0.0, 25.0], [dist + 80.0]), -0.5 * 400 / 900,
                               atol=1e-10)
    assert np.all(run(b, [200.0, 0.0, 25.0], [dist + 150.0]) > -0.5 * 2500 / 9)
    empty = BVHAccelerator(np.zeros((0, 3, 3)))
    np.testing.assert_allclose(run(empty, [0.0, 0.0, 20000.0], [20100.0]), 0.0, atol=1e-10)
    lw = run(b, [200.0, 0.0, 25.0], [dist + 150.0], blocked_nlos_prob=0.3, clear_nlos_prob=0.05)
    ref = oray.weight_3d(z, z, z, cb, [[200.0, 0.0, 25.0]], [dist + 150.0], [1.0], b.triangles,
                         3.0, 30.0, 20.0, 0.3, 0.05)
    np.testing.assert_allclose(lw, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("case", ["box", "mesh1200", "varied"])
def test_weight_matches_oracle_and_linear(gpu, case):
    """test_pf3d_bvh.py:223-334: BVH weights equal the triangle-list weights and
    the brute-force oracle."""
    from gnss_gpu_amd._gnss_gpu_pf3d import pf_weight_3d
    from gnss_gpu_amd._gnss_gpu_pf3d_bvh import pf_weight_3d_bvh
    from gnss_gpu_amd.bvh import BVHAccelerator

    if case == "box":
        tris = _box().triangles
        px, py, pz, pcb = _particles(200, 99)
        sat = np.array([[0.0, 0.0, 20000.0], [200.0, 0.0, 25.0], [0.0, 200.0, 5000.0],
                        [-200.0, 0.0, 5000.0]])
    elif case == "mesh1200":
        tris = _random_boxes(105, 17)
        px, py, pz, pcb = _particles(500, 99)
        sat = _sky(np.random.RandomState(55), 8, 500, 3e7, el_lo=np.radians(15))
    else:
        tris = _box().triangles
        n = 300
        px = np.tile([0.0, 50.0, 110.0, -50.0, 0.0], 60)[:n].astype(float)
        py = np.tile([0.0, 0.0, 0.0, 0.0, 10.0], 60)[:n].astype(float)
        pz = np.linspace(0.0, 60.0, n)
        pcb = np.full(n, 100.0)
        sat = np.array([[0.0, 0.0, 20000.0], [200.0, 0.0, 25.0], [200.0, 0.0, 100.0],
                        [0.0, -200.0, 5000.0]])
    n, ns = len(px), len(sat)
    pr = np.sqrt(np.sum(sat ** 2, axis=1)) + 100.0
    ws = np.linspace(0.5, 1.5, ns)
    b = BVHAccelerator(tris)
    lw_b = np.zeros(n)
    pf_weight_3d_bvh(px, py, pz, pcb, sat.ravel(), pr, ws, b._nodes_flat, b._sorted_tris, lw_b,
                     n, ns, 3.0, 30.0, 20.0)
    lw_l = np.zeros(n)
    pf_weight_3d(px, py, pz, pcb, sat.ravel(), pr, ws, tris, lw_l, n, ns, 3.0, 30.0, 20.0)
    ref = oray.weight_3d(px, py, pz, pcb, sat, pr, ws, tris, 3.0, 30.0, 20.0)
    # ranges of up to 3e7 m carry ~4e-9 m of rounding (the device contracts
    # the squares into FMAs); through -0.5 w r^2 / sigma^2 that is ~1e-9 relative
    np.testing.assert_allclose(lw_b, ref, rtol=1e-9, atol=1e-7)
    np.testing.assert_allclose(lw_l, lw_b, rtol=0, atol=1e-10)


def test_particle_filter_3d_bvh_runs_and_converges(gpu):
    """test_pf3d_bvh.py:341-487: update / estimate / predict cycle, ESS range,
    and convergence within 30 m with one NLOS satellite."""
    from gnss_gpu_amd import ParticleFilter3D, ParticleFilter3DBVH
    from gnss_gpu_amd.bvh import BVHAccelerator

    b = BVHAccelerator.from_building_model(_box())
    sat = np.array([[0.0, 0.0, 20000.0], [0.0, 200.0, 5000.0], [-200.0, 0.0, 5000.0],
                    [0.0, -200.0, 5000.0], [200.0, 0.0, 25.0]])
    pr = np.sqrt(np.sum(sat ** 2, axis=1)) + 100.0
    pr[4] += 20.0
    pf = ParticleFilter3DBVH(bvh=b, n_particles=10000, sigma_pos=1.0, sigma_cb=300.0,
                             sigma_los=3.0, sigma_nlos=30.0, nlos_bias=20.0,
                             resampling="systematic", seed=42)
    pf.initialize(position_ecef=[0.0, 0.0, 0.0], clock_bias=100.0, spread_pos=20.0,
                  spread_cb=200.0)
    for _ in range(3):
        pf.predict(dt=0.0)
        pf.update(sat, pr, np.ones(5))
        assert 0 < pf.get_ess() <= pf.n_particles
    est = pf.estimate()
    assert est.shape == (4,) and np.all(np.isfinite(est))
    assert np.linalg.norm(est[:3]) < 30.0
    assert pf.get_particles().shape == (10000, 4)
    lin = ParticleFilter3D(building_model=_box(), n_particles=2000, seed=1)
    lin.initialize(position_ecef=[0.0, 0.0, 0.0], clock_bias=100.0, spread_pos=5.0,
                   spread_cb=50.0)
    lin.update(sat, pr)
    assert np.all(np.isfinite(lin.estimate()))


def test_device_weight_on_resident_particles(gpu):
    """gg_pf_weight_3d_bvh_device on a resident tree and device particles
    (1M particles, 12 satellites, 1200 triangles) against the oracle on a
    sample of particles."""
    import ctypes as C

    import torch

    from gnss_gpu_amd import _lib
    from gnss_gpu_amd.bvh import BVHAccelerator

    tris = _random_boxes(100, 9)
    b = BVHAccelerator(tris)
    h = C.c_void_p()
    _lib.call("gg_bvh_create", _lib.dptr(np.ascontiguousarray(b._nodes_flat)), b.n_nodes,
              _lib.dptr(np.ascontiguousarray(b._sorted_tris)), b.n_triangles, C.byref(h))
    try:
        n, ns = 1 << 20, 12
        rng = np.random.RandomState(4)
        p = [rng.uniform(-300, 300, n), rng.uniform(-300, 300, n), rng.uniform(0, 30, n),
             rng.uniform(90, 110, n)]
        sat = _sky(rng, ns, 2e7, 2.6e7, el_lo=np.radians(10))
        pr = np.sqrt(np.sum(sat ** 2, axis=1)) + 100.0 + rng.normal(0, 5, ns)
        ws = np.ones(ns)
        dev = torch.device("cuda:0")
        tp = [torch.from_numpy(a).to(dev) for a in p]
        ts, tr, tw = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (sat, pr, ws))
        lw = torch.empty(n, dtype=torch.float64, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        _lib.call("gg_pf_weight_3d_bvh_device", h, *[C.c_void_p(t.data_ptr()) for t in tp],
                  C.c_void_p(ts.data_ptr()), C.c_void_p(tr.data_ptr()), C.c_void_p(tw.data_ptr()),
                  C.c_void_p(lw.data_ptr()), n, ns, 3.0, 30.0, 20.0, 0.9, 0.1, C.c_void_p(stream))
        torch.cuda.synchronize()
        got = lw.cpu().numpy()
        idx = rng.choice(n, 4000, replace=False)
        ref = oray.weight_3d(*(a[idx] for a in p), sat, pr, ws, tris, 3.0, 30.0, 20.0, 0.9, 0.1)
        np.testing.assert_allclose(got[idx], ref, rtol=1e-9, atol=1e-7)  # range rounding
    finally:
        _lib.call("gg_bvh_destroy", h)


def _pf_states(pf):
    p = pf.get_particles()
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), p[:, 3].copy()


def test_device_filter_ray_traced_update_matches_oracle(gpu):
    """ParticleFilterDevice.update_3d_bvh on resident particles equals the
    oracle's ray-traced weight at the particles' ECEF states (log-weights are
    binary32 on the device)."""
    from gnss_gpu_amd import ParticleFilterDevice
    from gnss_gpu_amd.bvh import BVHAccelerator
    from gnss_gpu_amd.sim import enu_rotation

    anchor = np.array([-3957199.0, 3310205.0, 3737911.0])
    R = enu_rotation(anchor)
    enu_tris = _random_boxes(60, 12).reshape(-1, 3) * np.array([0.2, 0.2, 1.0])
    tris = (anchor + enu_tris @ R).reshape(-1, 3, 3)
    b = BVHAccelerator(tris)
    rng = np.random.RandomState(8)
    ns = 10
    u = _sky(rng, ns, 1.0, 1.0, el_lo=np.radians(10))
    sat = anchor + 2.2e7 * (u @ R)
    pr = np.linalg.norm(sat - anchor, axis=1) + 50.0 + rng.normal(0, 5, ns)
    ws = rng.uniform(0.5, 1.5, ns)
    pf = ParticleFilterDevice(n_particles=100_000, sigma_pos=1.0, sigma_cb=1.0, seed=3)
    pf.initialize(anchor + np.array([3.0, -4.0, 1.0]) @ R, clock_bias=50.0, spread_pos=40.0,
                  spread_cb=5.0)
    pf.update_3d_bvh(b, sat, pr, ws, sigma_los=4.0, sigma_nlos=25.0, nlos_bias=15.0,
                     blocked_nlos_prob=0.9, clear_nlos_prob=0.05, resample=False)
    lw = pf.get_log_weights()
    px, py, pz, pcb = _pf_states(pf)
    ref = oray.weight_3d(px, py, pz, pcb, sat, pr, ws, tris, 4.0, 25.0, 15.0, 0.9, 0.05)
    np.testing.assert_allclose(lw, ref.astype(np.float32), rtol=2e-6, atol=2e-5)
    blocked = ~oray.los_check(np.stack([px[:2000], py[:2000], pz[:2000]], 1),
                              np.broadcast_to(sat, (2000, ns, 3)), tris, batch=True)
    assert 0.02 < blocked.mean() < 0.98
    pf.update_3d_bvh(b, sat, pr, ws)
    est = pf.estimate()
    assert np.all(np.isfinite(est))


def test_device_filter_ray_traced_update_with_isb(gpu):
    """A second system's satellites carry the particle's inter-system bias."""
    from gnss_gpu_amd import ParticleFilterDevice
    from gnss_gpu_amd.bvh import BVHAccelerator

    b = BVHAccelerator.from_building_model(_box())
    sat = np.array([[0.0, 0.0, 20000.0], [200.0, 0.0, 25.0], [0.0, 200.0, 5000.0],
                    [-200.0, 0.0, 5000.0]])
    sys_ids = np.array([0, 0, 1, 1], dtype=np.int32)
    pr = np.linalg.norm(sat, axis=1) + 100.0 + np.array([0.0, 0.0, 7.0, 7.0])
    pf = ParticleFilterDevice(n_particles=4096, sigma_pos=1.0, sigma_cb=1.0, seed=5, n_isb=1)
    pf.initialize([0.0, 0.0, 0.0], clock_bias=100.0, spread_pos=5.0, spread_cb=2.0)
    pf.update_3d_bvh(b, sat, pr, None, resample=False, system_ids=sys_ids)
    from gnss_gpu_amd import _gnss_gpu_pf_device as dev

    lw = pf.get_log_weights()
    full = dev.pf_device_get_particles_full(pf._state)
    px, py, pz, pcb, isb = (full[:, k].copy() for k in range(5))
    ref = np.zeros_like(px)
    for i in range(0, 4096, 16):
        adj = pr - np.array([0.0, 0.0, isb[i], isb[i]])
        ref[i] = oray.weight_3d(px[i:i + 1], py[i:i + 1], pz[i:i + 1], pcb[i:i + 1], sat, adj,
                                np.ones(4), b.triangles, 3.0, 30.0, 20.0)[0]
    np.testing.assert_allclose(lw[::16], ref[::16].astype(np.float32), rtol=2e-6, atol=2e-5)
