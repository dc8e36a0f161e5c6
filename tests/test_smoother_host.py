"""Forward-backward smoother orchestration (ParticleFilterDevice.smooth) on the
host, with a recording stand-in for the backward filter: which updates are
replayed, in which order and with which arguments
(reference: tests/test_particle_filter_smoother_metadata.py and
python/gnss_gpu/particle_filter_device.py:1070-1251).  No GPU work."""

from types import SimpleNamespace

import numpy as np
import pytest

import gnss_gpu_amd.particle_filter_device as pfd


class _Recorder:
    made = []

    def __init__(self, **kw):
        self.config = kw
        self.calls = []
        self.est = np.array([10.0, 20.0, 30.0, 0.0])
        type(self).made.append(self)

    def initialize(self, pos, **kw):
        self.calls.append(("initialize", np.asarray(pos, dtype=float).copy(), kw))

    def predict(self, **kw):
        v = kw.get("velocity")
        self.calls.append(("predict", None if v is None else np.asarray(v).copy(), kw["dt"]))

    def update_dd_pseudorange(self, dd, sigma_pr=0.75, resample=True):
        self.calls.append(("dd_pr", int(dd.n_dd), float(sigma_pr)))

    def update_dd_carrier_afv(self, dd, sigma_cycles=0.05):
        self.calls.append(("dd_cp", int(dd.n_dd), float(sigma_cycles)))

    def correct_clock_bias(self, sat, pr):
        self.calls.append(("clock", len(pr)))

    def update(self, sat, pr, weights=None, sigma_pr=None, system_ids=None):
        self.calls.append(("pr", len(pr), sigma_pr, system_ids))

    def resample_if_needed(self):
        self.calls.append(("resample",))

    def update_carrier_afv(self, sat, cycles, weights=None, wavelength=None, sigma_cycles=None):
        self.calls.append(("afv", len(cycles), wavelength, sigma_cycles))

    def position_update(self, ref, sigma_pos=None):
        self.calls.append(("pos", np.asarray(ref).copy(), sigma_pos))

    def estimate(self):
        return self.est


def _epoch(sat, est, **over):
    ep = {"estimate": np.asarray(est, dtype=float), "sat_ecef": sat,
          "pseudoranges": np.linalg.norm(sat - np.asarray(est), axis=1) + 50.0,
          "weights": np.ones(len(sat)), "velocity": np.array([1.0, -2.0, 0.5]), "dt": 0.5,
          "spp_ref": None, "dd_pseudorange": None, "dd_pseudorange_sigma": None,
          "dd_pseudorange_source": None, "dd_carrier": None, "dd_carrier_sigma": None,
          "carrier_anchor_pseudorange": None, "carrier_anchor_sigma": None,
          "carrier_afv": None, "carrier_afv_sigma": None, "carrier_afv_wavelength": None,
          "doppler_update": None, "doppler_sigma_mps": None,
          "doppler_velocity_update_gain": None, "doppler_max_velocity_update_mps": None}
    ep.update(over)
    return ep


def _shell(epochs):
    return SimpleNamespace(
        _smooth_enabled=True, _smooth_epochs=epochs, n_particles=16, sigma_pos=1.0,
        sigma_cb=300.0, sigma_pr=3.0, nu=0.0, resampling="systematic", ess_threshold=0.5,
        seed=7, per_particle_nlos_gate=False, per_particle_nlos_dd_pr_threshold_m=10.0,
        per_particle_nlos_dd_carrier_threshold_cycles=0.5,
        per_particle_nlos_undiff_pr_threshold_m=30.0, per_particle_huber=False,
        per_particle_huber_dd_pr_k=1.5, per_particle_huber_dd_carrier_k=1.5,
        per_particle_huber_undiff_pr_k=1.5, sigma_vel=0.0, velocity_guide_alpha=1.0,
        rbpf_velocity_kf=False, velocity_process_noise=0.0, _velocity_init_sigma=0.25)


SAT = np.array([[2.0e7, 0.0, 0.0], [0.0, 2.0e7, 0.0], [0.0, 0.0, 2.0e7], [2.0e7, 2.0e7, 0.0]])


@pytest.fixture()
def recorder(monkeypatch):
    _Recorder.made = []
    monkeypatch.setattr(pfd, "ParticleFilterDevice", _Recorder)
    return _Recorder


def test_backward_filter_config_start_and_average(recorder):
    eps = [_epoch(SAT, [1.0, 2.0, 3.0]), _epoch(SAT, [4.0, 5.0, 6.0])]
    sh = _shell(eps)
    smoothed, forward = _smooth(sh)
    bwd = recorder.made[-1]
    assert bwd.config["seed"] == 8 and bwd.config["n_particles"] == 16
    assert bwd.config["resampling"] == "systematic" and bwd.config["sigma_pr"] == 3.0
    name, pos, kw = bwd.calls[0]
    assert name == "initialize"
    np.testing.assert_array_equal(pos, [4.0, 5.0, 6.0])
    # clock bias: median residual of the last epoch at the last estimate (+50 m here)
    assert kw["clock_bias"] == pytest.approx(50.0, abs=1e-6)
    assert (kw["spread_pos"], kw["spread_cb"], kw["velocity_init_sigma"]) == (10.0, 100.0, 0.25)
    # epochs visited last to first, velocity negated, dt kept
    preds = [c for c in bwd.calls if c[0] == "predict"]
    assert len(preds) == 2
    np.testing.assert_array_equal(preds[0][1], [-1.0, 2.0, -0.5])
    assert preds[0][2] == 0.5
    np.testing.assert_array_equal(forward, [[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]])
    np.testing.assert_array_equal(smoothed, (forward + np.array([[10.0, 20.0, 30.0]] * 2)) / 2)
    assert sh._smooth_enabled is False and sh._smooth_epochs == []


_SMOOTH = pfd.ParticleFilterDevice.smooth  # taken before any test replaces the class


def _smooth(sh, **kw):
    return _SMOOTH(sh, **kw)


def test_widelane_dd_pseudorange_skip(recorder):
    dd = {"n_dd": 3}
    for skip, want_dd in ((False, True), (True, False)):
        recorder.made = []
        _smooth(_shell([_epoch(SAT, [1.0, 2.0, 3.0], dd_pseudorange=dd,
                               dd_pseudorange_sigma=0.1, dd_pseudorange_source="widelane")]),
                position_update_sigma=None, skip_widelane_dd_pseudorange=skip)
        names = [c[0] for c in recorder.made[-1].calls]
        assert ("dd_pr" in names) == want_dd
        assert ("clock" in names and "pr" in names) == (not want_dd)
    recorder.made = []
    _smooth(_shell([_epoch(SAT, [1.0, 2.0, 3.0], dd_pseudorange=dd, dd_pseudorange_sigma=0.1,
                           dd_pseudorange_source="dd_pseudorange")]),
            skip_widelane_dd_pseudorange=True)
    calls = recorder.made[-1].calls
    assert ("dd_pr", 3, 0.1) in calls


def test_replay_order_and_defaults(recorder):
    anchor = {"sat_ecef": SAT, "pseudoranges": np.ones(4), "weights": np.ones(4), "n_sat": 4}
    afv = {"sat_ecef": SAT, "carrier_phase_cycles": np.ones(4), "weights": np.ones(4),
           "n_sat": 4}
    ep = _epoch(SAT, [1.0, 2.0, 3.0], dd_carrier={"n_dd": 2}, carrier_anchor_pseudorange=anchor,
                carrier_afv=afv, spp_ref=np.array([7.0, 8.0, 9.0]))
    _smooth(_shell([ep]), position_update_sigma=4.0)
    names = [c[0] for c in recorder.made[-1].calls]
    assert names == ["initialize", "predict", "clock", "pr", "resample", "dd_cp", "pr",
                     "resample", "afv", "pos"]
    calls = recorder.made[-1].calls
    assert ("dd_cp", 2, 0.05) in calls                 # default DD carrier sigma
    assert ("pr", 4, 3.0, None) in calls               # anchor sigma defaults to sigma_pr
    assert ("afv", 4, 0.190293673, 0.05) in calls      # default wavelength and sigma
    pos = [c for c in calls if c[0] == "pos"][0]
    np.testing.assert_array_equal(pos[1], [7.0, 8.0, 9.0])
    assert pos[2] == 4.0
    # without a position sigma the SPP reference is not replayed
    recorder.made = []
    _smooth(_shell([_epoch(SAT, [1.0, 2.0, 3.0], spp_ref=np.array([7.0, 8.0, 9.0]))]))
    assert "pos" not in [c[0] for c in recorder.made[-1].calls]


def test_smooth_without_epochs_raises():
    sh = _shell([])
    with pytest.raises(RuntimeError):
        _smooth(sh)
    sh = _shell([_epoch(SAT, [1.0, 2.0, 3.0])])
    sh._smooth_enabled = False
    with pytest.raises(RuntimeError):
        _smooth(sh)
