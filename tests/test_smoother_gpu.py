"""Forward-backward smoothing on the device filter (ParticleFilterDevice.
enable_smoothing / store_epoch / smooth; reference
python/gnss_gpu/particle_filter_device.py:892-1251): the reference's smoke cases
(tests/test_cuda_streams.py:775-946), the backward pass equal to a filter driven
by hand through the same calls, and a smoothed track more accurate than the
forward one on a seeded constellation."""

from types import SimpleNamespace

import numpy as np
import pytest

from gnss_gpu_amd import sim
from gnss_gpu_amd.particle_filter_device import ParticleFilterDevice

pytestmark = pytest.mark.gpu
SEED = 42


def _dirs(seed, n):
    d = np.random.RandomState(seed).randn(n, 3)
    return d / np.linalg.norm(d, axis=1, keepdims=True) * 20_000_000.0


def test_smooth_dd_pseudorange_smoke(gpu):
    sat = _dirs(987, 5)
    rover, base = np.zeros(3), np.array([120.0, -35.0, 18.0])
    rr = np.linalg.norm(sat - rover, axis=1)
    br = np.linalg.norm(sat - base, axis=1)
    dd = SimpleNamespace(
        dd_pseudorange_m=(rr[1:] - rr[0]) - (br[1:] - br[0]), sat_ecef_k=sat[1:].copy(),
        sat_ecef_ref=np.repeat(sat[[0]], 4, axis=0), base_range_k=br[1:].copy(),
        base_range_ref=np.repeat(br[0], 4), dd_weights=np.ones(4),
        ref_sat_ids=("G01",) * 4, n_dd=4)
    pf = ParticleFilterDevice(n_particles=20_000, sigma_pos=1.0, sigma_cb=300.0, sigma_pr=5.0,
                              resampling="systematic", seed=SEED)
    pf.initialize(position_ecef=rover, clock_bias=250.0, spread_pos=20.0, spread_cb=150.0)
    pf.enable_smoothing()
    for _ in range(3):
        pf.predict(dt=1.0)
        pf.update_dd_pseudorange(dd, sigma_pr=0.5)
        pf.store_epoch(sat, rr + 250.0, np.ones(5), None, 1.0, spp_ref=None,
                       dd_pseudorange=dd, dd_pseudorange_sigma=0.5)
    smoothed, forward = pf.smooth(position_update_sigma=None)
    assert smoothed.shape == (3, 3) and forward.shape == (3, 3)
    assert np.all(np.isfinite(smoothed)) and np.all(np.isfinite(forward))


def test_smooth_carrier_afv_and_anchor_smoke(gpu):
    wl = 0.190293673
    sat = _dirs(4321, 6)
    rover = np.zeros(3)
    rr = np.linalg.norm(sat - rover, axis=1)
    afv = {"sat_ecef": sat, "carrier_phase_cycles": rr / wl, "weights": np.ones(6)}
    pf = ParticleFilterDevice(n_particles=20_000, sigma_pos=0.5, sigma_cb=300.0, sigma_pr=5.0,
                              resampling="systematic", seed=SEED)
    pf.initialize(position_ecef=rover, clock_bias=0.0, spread_pos=5.0, spread_cb=10.0)
    pf.enable_smoothing()
    for _ in range(3):
        pf.predict(dt=1.0)
        pf.update(sat, rr, np.ones(6))
        pf.resample_if_needed()
        pf.update_carrier_afv(sat, afv["carrier_phase_cycles"], weights=np.ones(6),
                              wavelength=wl, sigma_cycles=0.10)
        pf.store_epoch(sat, rr, np.ones(6), None, 1.0, carrier_afv=afv, carrier_afv_sigma=0.10,
                       carrier_afv_wavelength=wl)
    smoothed, forward = pf.smooth()
    assert smoothed.shape == (3, 3) and np.all(np.isfinite(smoothed))

    rover = np.array([5.0, -3.0, 2.0])
    sat = _dirs(2468, 6)
    rr = np.linalg.norm(sat - rover, axis=1)
    bias = np.linspace(1000.0, 1005.0, 6)
    cyc = (rr + 120.0) / wl + bias
    anchor = {"sat_ecef": sat, "pseudoranges": wl * (cyc - bias), "weights": np.ones(6)}
    pf = ParticleFilterDevice(n_particles=20_000, sigma_pos=1.0, sigma_cb=300.0, sigma_pr=5.0,
                              resampling="systematic", seed=SEED)
    pf.initialize(position_ecef=rover, clock_bias=120.0, spread_pos=5.0, spread_cb=20.0)
    pf.enable_smoothing()
    for _ in range(3):
        pf.predict(dt=1.0)
        pf.update(sat, anchor["pseudoranges"], weights=anchor["weights"], sigma_pr=0.25)
        pf.store_epoch(sat, anchor["pseudoranges"], anchor["weights"], None, 1.0,
                       carrier_anchor_pseudorange=anchor, carrier_anchor_sigma=0.25)
    smoothed, forward = pf.smooth()
    assert smoothed.shape == (3, 3) and np.all(np.isfinite(smoothed))


def _forward(sc, n, epochs, smoothing=True):
    pf = ParticleFilterDevice(n_particles=n, sigma_pos=1.0, sigma_cb=3.0, sigma_pr=5.0,
                              resampling="systematic", seed=SEED)
    pf.initialize(sc.truth[0] - sc.vel, clock_bias=sc.cb, spread_pos=10.0, spread_cb=10.0)
    if smoothing:
        pf.enable_smoothing()
    for t in range(epochs):
        pf.predict(velocity=sc.vel, dt=1.0)
        pf.update(sc.sat, sc.pr[t])
        if smoothing:
            pf.store_epoch(sc.sat, sc.pr[t], np.ones(sc.n_sat), sc.vel, 1.0)
    return pf


def test_backward_pass_equals_hand_driven_filter(gpu):
    """smooth() = (forward + backward) / 2 with the backward filter's estimates
    reproduced bit for bit by a second filter driven through the same calls."""
    sc = sim.scenario_plain(n_epochs=20, seed=3)
    pf = _forward(sc, 50_000, 20)
    fwd_expected = np.array([ep["estimate"] for ep in pf._smooth_epochs])
    smoothed, forward = pf.smooth()
    np.testing.assert_array_equal(forward, fwd_expected)

    start = forward[-1]
    cb0 = float(np.median(sc.pr[19] - np.linalg.norm(sc.sat - start, axis=1)))
    bwd = ParticleFilterDevice(n_particles=50_000, sigma_pos=1.0, sigma_cb=3.0, sigma_pr=5.0,
                               resampling="systematic", seed=SEED + 1)
    bwd.initialize(start, clock_bias=cb0, spread_pos=10.0, spread_cb=100.0)
    back = np.zeros_like(forward)
    for t in range(19, -1, -1):
        bwd.predict(velocity=-sc.vel, dt=1.0)
        bwd.correct_clock_bias(sc.sat, sc.pr[t])
        bwd.update(sc.sat, sc.pr[t], weights=np.ones(sc.n_sat))
        back[t] = bwd.estimate()[:3]
    np.testing.assert_array_equal(smoothed, (forward + back) / 2.0)
    # the records are consumed
    with pytest.raises(RuntimeError):
        pf.smooth()


def test_smoothed_track_beats_forward_track(gpu):
    sc = sim.scenario_plain(n_epochs=60, seed=11)
    pf = _forward(sc, 100_000, 60)
    smoothed, forward = pf.smooth()
    e_f = np.linalg.norm(forward - sc.truth[:60], axis=1)[5:-5]
    e_s = np.linalg.norm(smoothed - sc.truth[:60], axis=1)[5:-5]
    rms_f, rms_s = float(np.sqrt(np.mean(e_f ** 2))), float(np.sqrt(np.mean(e_s ** 2)))
    print(f"forward RMS {rms_f:.3f} m, smoothed RMS {rms_s:.3f} m")
    assert rms_s < rms_f
