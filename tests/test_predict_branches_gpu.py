"""Every predict branch of the device filter against the binary64 oracle
(pf_device.cu:106-187): the velocity-guide blend on per-particle velocities,
per-particle velocity noise, and the RBPF Cholesky draw from
sigma_pos^2 I + dt^2 Sigma_v with Sigma_v += q dt, on the uniform covariance
and on per-particle covariances.  The reference's own closed forms
(test_cuda_streams.py:239-283, 486-505) run through the same entry points.

Bars: positions and clock biases within 1 mm, velocities within 1e-5 m/s,
covariances within 1e-5 relative or 2e-6 absolute (binary32 storage of
entries up to a few m^2/s^2), of the oracle driven by
the same normals.
"""

import numpy as np
import pytest

from oracle import pf as opf

pytestmark = pytest.mark.gpu

SEED = 77
RX = np.array([-3957199.0, 3310205.0, 3737911.0])


def _pair(n, **kw):
    from gnss_gpu_amd import ParticleFilterDevice

    pf = ParticleFilterDevice(n_particles=n, sigma_pos=kw.pop("sigma_pos", 1.0),
                              sigma_cb=kw.pop("sigma_cb", 3.0), resampling="systematic",
                              seed=SEED, **kw)
    orc = opf.OraclePF(n, seed=SEED)
    return pf, orc


def _check(pf, orc, cov=False):
    s = pf.get_particle_states()
    np.testing.assert_allclose(s[:, :3], orc.state[:, :3], rtol=0, atol=1e-3)
    np.testing.assert_allclose(s[:, 3], orc.state[:, 3], rtol=0, atol=1e-3)
    np.testing.assert_allclose(s[:, 4:7], orc.vel, rtol=0, atol=1e-5)
    if cov:
        np.testing.assert_allclose(s[:, 7:16], orc.vcov.reshape(-1, 9), rtol=1e-5, atol=2e-6)
    return s


def test_guide_blend_on_per_particle_velocities(gpu):
    """alpha = 0.3 on velocities spread per particle (spread_vel > 0)."""
    n = 8192
    pf, orc = _pair(n, velocity_guide_alpha=0.3)
    v0 = np.array([1.0, 0.5, -0.2])
    pf.initialize(RX, clock_bias=100.0, spread_pos=20.0, spread_cb=10.0, velocity=v0,
                  spread_vel=0.7)
    orc.initialize(RX, 100.0, 20.0, 10.0, velocity=v0, spread_vel=0.7)
    _check(pf, orc)
    for step, vg in enumerate(([3.0, -1.0, 0.5], [2.5, -1.5, 0.0], [0.0, 0.0, 4.0]), start=1):
        pf.predict(velocity=np.array(vg), dt=0.5)
        orc.predict(np.array(vg), 0.5, 1.0, 3.0, step, alpha=0.3)
        s = _check(pf, orc)
    # the blend keeps the spread of the initial velocities: 0.7^3 of 0.7
    assert 0.8 * 0.7 * 0.343 < float(np.std(s[:, 4] - s[:, 4].mean())) < 1.2 * 0.7 * 0.343


def test_velocity_noise_from_uniform_and_per_particle_velocities(gpu):
    """sigma_vel > 0 (no velocity filter): the first predict leaves the uniform
    velocity mode, later ones add noise to per-particle velocities, with a
    guide blend alpha = 0.6."""
    n = 8192
    pf, orc = _pair(n, sigma_vel=0.5, velocity_guide_alpha=0.6)
    v0 = np.array([-1.0, 2.0, 0.0])
    pf.initialize(RX, clock_bias=50.0, spread_pos=10.0, spread_cb=5.0, velocity=v0)
    orc.initialize(RX, 50.0, 10.0, 5.0, velocity=v0)
    for step in range(1, 5):
        vg = np.array([0.5 * step, -1.0, 0.25])
        pf.predict(velocity=vg, dt=1.0)
        orc.predict(vg, 1.0, 1.0, 3.0, step, alpha=0.6, sigma_vel=0.5)
        s = _check(pf, orc)
    assert 0.4 < float(np.std(s[:, 4])) < 0.8


def test_rbpf_predict_reference_closed_forms(gpu):
    """test_cuda_streams.py:257-283 and 486-505 through this library's entry
    points: mu_v unchanged, position std in (1.6, 2.4), Sigma_v = 4.5 I; the
    wrapper case Sigma_v = 1.5 I."""
    from gnss_gpu_amd import ParticleFilterDevice
    from gnss_gpu_amd import _gnss_gpu_pf_device as dev

    n = 20_000
    st = dev.pf_device_create(n)
    dev.pf_device_initialize(st, 0.0, 0.0, 0.0, 3000.0, 0.0, 0.0, 42, 1.0, 0.0, 0.0, 0.0, 2.0)
    dev.pf_device_predict(st, 1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 42, 1, 100.0, 1.0, True, 0.5)
    s = dev.pf_device_get_particle_states(st)
    np.testing.assert_allclose(s[:, 4:7], np.tile([1.0, 0.0, 0.0], (n, 1)))
    assert 1.6 < float(np.std(s[:, 0] - 1.0)) < 2.4
    assert 1.6 < float(np.std(s[:, 1])) < 2.4
    assert 1.6 < float(np.std(s[:, 2])) < 2.4
    np.testing.assert_allclose(s[:, 7], 4.5)
    np.testing.assert_allclose(s[:, 11], 4.5)
    np.testing.assert_allclose(s[:, 15], 4.5)
    np.testing.assert_allclose(s[:, [8, 9, 10, 12, 13, 14]], 0.0, atol=1e-12)
    pf = ParticleFilterDevice(n_particles=4096, seed=42, rbpf_velocity_kf=True,
                              velocity_process_noise=0.25)
    pf.initialize(position_ecef=RX, clock_bias=3000.0, spread_pos=0.0, spread_cb=0.0,
                  velocity=np.array([0.5, -0.25, 0.0]), velocity_init_sigma=1.0)
    pf.predict(velocity=np.array([0.5, -0.25, 0.0]), dt=2.0, sigma_pos=0.0)
    s = pf.get_particle_states()
    np.testing.assert_allclose(s[:, 4:7], np.tile([0.5, -0.25, 0.0], (4096, 1)), atol=1e-6)
    np.testing.assert_allclose(s[:, [7, 11, 15]], 1.5, rtol=1e-6)


def test_rbpf_predict_uniform_covariance_matches_oracle(gpu):
    """Velocity filter on, one covariance for all particles (no Doppler update
    yet): the Cholesky factor of sigma_pos^2 I + dt^2 Sigma_v in the local frame
    and Sigma_v += q dt, over three predicts."""
    n = 8192
    pf, orc = _pair(n, sigma_pos=0.5, rbpf_velocity_kf=True, velocity_process_noise=0.3)
    v0 = np.array([1.0, -0.5, 0.2])
    pf.initialize(RX, clock_bias=100.0, spread_pos=5.0, spread_cb=2.0, velocity=v0,
                  velocity_init_sigma=2.0)
    orc.initialize(RX, 100.0, 5.0, 2.0, velocity=v0, velocity_init_sigma=2.0)
    for step in range(1, 4):
        pf.predict(velocity=v0, dt=1.5)
        orc.predict(v0, 1.5, 0.5, 3.0, step, velocity_kf=True, velocity_process_noise=0.3)
        _check(pf, orc, cov=True)


def test_rbpf_predict_per_particle_covariance_matches_oracle(gpu):
    """After a Doppler Kalman update every particle holds its own velocity mean
    and covariance; the next RBPF predicts factor each particle's covariance."""
    from tests.test_pf_gpu import _doppler_rows

    d, rx, sat, sv, tv, dop = _doppler_rows()
    n = 8192
    pf, orc = _pair(n, sigma_pos=0.7, rbpf_velocity_kf=True, velocity_process_noise=0.2)
    pf.initialize(rx, clock_bias=100.0, spread_pos=40.0, spread_cb=3.0,
                  velocity=np.array([1.0, 0.5, -0.2]), spread_vel=0.7, velocity_init_sigma=3.0)
    rng = np.random.default_rng(4)
    pf.update_doppler_kf(sat, sv, dop + rng.normal(0.0, 0.3, len(dop)), wavelength=d["wavelength_m"],
                         sigma_mps=d["sigma_mps"], resample=False)
    s0 = pf.get_particle_states()
    # the oracle takes the device's post-update states, then runs the predicts
    orc.initialize(rx, 100.0, 40.0, 3.0)
    orc.state = s0[:, :4].copy()
    orc.vel = s0[:, 4:7].copy()
    orc.vcov = s0[:, 7:16].reshape(n, 3, 3).copy()
    for step in range(1, 3):
        vg = np.array([0.3, 0.1, 0.0])
        pf.predict(velocity=vg, dt=1.0, velocity_guide_alpha=0.5)
        orc.predict(vg, 1.0, 0.7, 3.0, step, alpha=0.5, velocity_kf=True,
                    velocity_process_noise=0.2)
        _check(pf, orc, cov=True)
    # covariances that really differ per particle (random SPD, non-diagonal)
    rng = np.random.default_rng(8)
    M = rng.normal(0.0, 0.6, (n, 3, 3))
    C = M @ np.transpose(M, (0, 2, 1)) + 0.2 * np.eye(3)
    rows = pf.get_particle_states()
    rows[:, 4:7] = rng.normal(0.0, 1.0, (n, 3))
    rows[:, 7:16] = C.reshape(n, 9)
    pf.set_particle_states(rows)
    s1 = pf.get_particle_states()
    np.testing.assert_allclose(s1[:, 7:16], rows[:, 7:16], rtol=1e-6, atol=1e-6)
    orc.state = s1[:, :4].copy()
    orc.vel = s1[:, 4:7].copy()
    orc.vcov = s1[:, 7:16].reshape(n, 3, 3).copy()
    for step in range(3, 5):
        vg = np.array([-0.4, 0.2, 0.1])
        pf.predict(velocity=vg, dt=2.0, velocity_guide_alpha=0.25)
        orc.predict(vg, 2.0, 0.7, 3.0, step, alpha=0.25, velocity_kf=True,
                    velocity_process_noise=0.2)
        _check(pf, orc, cov=True)
