"""The kernels behind the headline figure, against the host-driven drop-in path
and against the binary64 oracle.

The timed step is BatchedParticleFilter's device-driven step on one filter of
the multi-GNSS workload (GPS/Galileo/BeiDou/GLONASS, 3 satellites each, state
[e n u cb isb x3], Cauchy scale 5 m, 30% NLOS): the fused predict + likelihood
+ tile-partial launch, then the decision and the resampler (systematic, or
Megopolis B = 32) whose gather finalises.  Tests drive it through step_part, which
splits gg_pfb_step after its likelihood launch so that the binary32
log-weights the resampler consumes can be read.

Bars (tolerances are written in each assertion):
  * against ParticleFilterDevice (the drop-in path, separate predict, weight,
    decision and resampling kernels): log-weights, decisions, ancestors,
    particles and estimates equal bit for bit;
  * against OraclePF (binary64 absolute state, the library's random numbers),
    fed the GPU's binary32 log-weights for resampling, as the reference's
    semantics ask: ancestors equal; particles and estimates within 1 mm;
    log-weights within 1e-5 relative of the oracle's likelihood evaluated at
    the GPU's own particles; against the independently propagated oracle the
    log-weights differ by at most 1e-5 relative plus the Student-t slope bound
    times the state drift (stated in _drift_bound).
"""

import numpy as np
import pytest

from oracle import detmath
from oracle import pf as opf

pytestmark = pytest.mark.gpu

SEED = 12345
SIG_PR, NU = 5.0, 1.0
SIG_POS, SIG_CB, SIG_ISB = 1.0, 1.0, 0.1


def _scenario(T, seed=0):
    from gnss_gpu_amd import sim

    return sim.scenario_multi_gnss(n_epochs=T, seed=seed)


def _init(sc):
    return np.concatenate([sc.truth[0] - sc.vel, [sc.cb], sc.isb])


def _batched(sc, N, T):
    from gnss_gpu_amd._lib import LIK_STUDENT
    from gnss_gpu_amd.batched import BatchedParticleFilter

    S = sc.n_sat
    b = BatchedParticleFilter([N], n_isb=3, sat_cap=S, n_slots=T)
    b.load_epochs([[sc.sat]] * T, [[sc.pr[t]] for t in range(T)], [[S]] * T,
                  sys=[[sc.sys]] * T, vel=np.tile(sc.vel, (T, 1, 1)))
    b.initialize(_init(sc)[None], [SEED], spread_pos=10.0, spread_cb=10.0, spread_isb=5.0)
    st = b.step_args(sigma_pos=SIG_POS, sigma_cb=SIG_CB, sigma_isb=SIG_ISB, sigma_pr=SIG_PR,
                     nu=NU, lik=LIK_STUDENT, n_slots=T, ess_frac=0.5)
    return b, st


def _oracle(sc, N):
    orc = opf.OraclePF(N, seed=SEED, n_isb=3)
    x0 = _init(sc)
    orc.initialize(x0[:3], x0[3], 10.0, 10.0, isb=x0[4:], spread_isb=5.0)
    return orc


def _drift_bound(n_sat, delta_e):
    """|d lw| <= sum_s |d l_s / d e| |delta e| with the Student-t term
    l(e) = -(nu + 1)/2 log(1 + e^2 / (nu s^2)), whose slope is at most
    (nu + 1) / (2 s sqrt(nu)) (at e = s sqrt(nu))."""
    return n_sat * (NU + 1.0) / (2.0 * SIG_PR * np.sqrt(NU)) * delta_e


def _resample_oracle(orc, resampler, B, lw32):
    if resampler == "systematic":
        orc.resample_systematic(lw32=lw32)
    else:
        orc.resample_megopolis_chesser(B, lw32=lw32)


def _run_against_oracle(N, T, resampler, B=32):
    sc = _scenario(T)
    b, st = _batched(sc, N, T)
    orc = _oracle(sc, N)
    S = sc.n_sat
    w = np.ones(S)
    rep = dict(steps=T, resampled=0, anc_mismatch=0, max_state_m=0.0, max_est_m=0.0,
               max_lw_rel_given=0.0, max_lw_abs_indep=0.0, max_ess_rel=0.0)
    for t in range(T):
        b.step_part(st, 0, resampler, B)
        lw32 = b.get_log_weights_raw(0)
        orc.predict(sc.vel, 1.0, SIG_POS, SIG_CB, t + 1, sigma_isb=SIG_ISB)
        orc.weight(sc.sat, sc.pr[t], sys=sc.sys, sigma_pr=SIG_PR, nu=NU)
        pre_state = orc.state.copy()
        pre_lw = orc.lw.copy()
        b.step_part(st, 1, resampler, B)
        d = b.read_dyn()[0]
        res = bool(d.did_resample)
        ess_orc = orc.ess()
        if abs(ess_orc - 0.5 * N) > 1e-4 * N:
            assert res == (ess_orc < 0.5 * N), (t, ess_orc)
        if not res:
            rel = abs(d.ess - ess_orc) / ess_orc
            rep["max_ess_rel"] = max(rep["max_ess_rel"], rel)
            assert rel < 1e-4, (t, d.ess, ess_orc)
        p = b.get_particles(0)
        if res:
            rep["resampled"] += 1
            anc = b.get_ancestors(0)
            _resample_oracle(orc, resampler, B, lw32)
            rep["anc_mismatch"] += int(np.sum(anc != orc.ancestors))
            lw_at = lw32[anc].astype(np.float64)
            pre, lw_ind = pre_state[anc], pre_lw[anc]
        else:
            lw_at = lw32.astype(np.float64)
            pre, lw_ind = pre_state, pre_lw
        # particles (the resampled ones after a resampling) within 1 mm
        dpos = np.linalg.norm(p[:, :3] - orc.state[:, :3], axis=1)
        dstate = np.max(np.abs(p - orc.state))
        rep["max_state_m"] = max(rep["max_state_m"], float(dstate))
        assert dstate < 1e-3, (t, dstate)
        delta_e = dpos + np.abs(p[:, 3] - pre[:, 3]) + np.max(np.abs(p[:, 4:] - pre[:, 4:]), axis=1)
        # log-weights given the GPU's particles: 1e-5 relative (binary32 path)
        lw_given = opf.log_likelihood(p[:, :3], p[:, 3], p[:, 4:], sc.sat, sc.pr[t], w, sc.sys,
                                      SIG_PR, nu=NU)
        rel = np.max(np.abs(lw_at - lw_given) / np.maximum(np.abs(lw_given), 1.0))
        rep["max_lw_rel_given"] = max(rep["max_lw_rel_given"], float(rel))
        np.testing.assert_allclose(lw_at, lw_given, rtol=1e-5, atol=2e-5)
        # against the independently propagated oracle: 1e-5 relative + drift bound
        err = np.abs(lw_at - lw_ind)
        rep["max_lw_abs_indep"] = max(rep["max_lw_abs_indep"], float(err.max()))
        assert np.all(err <= 1e-5 * np.abs(lw_ind) + _drift_bound(S, delta_e) + 2e-5), t
        # estimate within 1 mm
        est = np.array([d.est[k] for k in range(7)])
        de = float(np.max(np.abs(est - orc.estimate())))
        rep["max_est_m"] = max(rep["max_est_m"], de)
        assert de < 1e-3, (t, est, orc.estimate())
    b.close()
    assert rep["anc_mismatch"] == 0, rep
    return rep


@pytest.mark.parametrize("resampler", ["systematic", "megopolis"])
def test_timed_step_equals_drop_in_path(gpu, resampler):
    """Batched graph-free steps vs ParticleFilterDevice, bit for bit, 25 steps."""
    from gnss_gpu_amd import ParticleFilterDevice
    from gnss_gpu_amd import _gnss_gpu_pf_device as dev

    N, T, B = 20_000, 25, 32
    sc = _scenario(T)
    b, st = _batched(sc, N, T)
    x0 = _init(sc)
    pf = ParticleFilterDevice(n_particles=N, sigma_pos=SIG_POS, sigma_cb=SIG_CB, sigma_pr=SIG_PR,
                              nu=NU, resampling=resampler, ess_threshold=0.5, seed=SEED, n_isb=3,
                              sigma_isb=SIG_ISB, megopolis_mode="chesser", megopolis_iterations=B)
    pf.initialize(x0[:3], clock_bias=x0[3], spread_pos=10.0, spread_cb=10.0, isb=x0[4:],
                  spread_isb=5.0)
    n_res = 0
    for t in range(T):
        b.step_part(st, 0, resampler, B)
        lw_b = b.get_log_weights_raw(0)
        pf.predict(velocity=sc.vel, dt=1.0)
        pf.update(sc.sat, sc.pr[t], system_ids=sc.sys, resample=False)
        np.testing.assert_array_equal(lw_b.astype(np.float64), pf.get_log_weights())
        b.step_part(st, 1, resampler, B)
        d = b.read_dyn()[0]
        res_d = pf.resample_if_needed()
        assert bool(d.did_resample) == res_d, t
        if res_d:
            n_res += 1
            a_d = np.empty(N, dtype=np.int32)
            dev.pf_device_get_resample_ancestors(pf._state, a_d)
            np.testing.assert_array_equal(b.get_ancestors(0), a_d)
        est_b = np.array([d.est[k] for k in range(7)])
        np.testing.assert_array_equal(est_b, pf.estimate_full())
        if t % 5 == 4:
            np.testing.assert_array_equal(b.get_particles(0),
                                          dev.pf_device_get_particles_full(pf._state))
    assert 3 <= n_res < T
    b.close()


@pytest.mark.parametrize("resampler", ["systematic", "megopolis"])
def test_timed_step_matches_oracle(gpu, resampler):
    """20,000 particles, 20 steps of the timed kernels against OraclePF."""
    rep = _run_against_oracle(20_000, 20, resampler)
    print(f"\n[timed-path oracle 20k {resampler}] {rep}", flush=True)
    assert rep["resampled"] >= 3


def test_timed_step_1M_error_against_fp64_oracle(gpu):
    """The full-size workload (1,000,000 particles x 12 satellites), 10 steps,
    systematic resampling, beside the binary64 oracle: no ancestor differs, and
    the estimate stays within 1 mm.  The measured maxima are printed."""
    rep = _run_against_oracle(1_000_000, 10, "systematic")
    print(f"\n[timed-path oracle 1M systematic] {rep}", flush=True)
    assert rep["resampled"] >= 2


def test_timed_step_1M_megopolis_against_fp64_oracle(gpu):
    rep = _run_against_oracle(1_000_000, 4, "megopolis")
    print(f"\n[timed-path oracle 1M megopolis B=32] {rep}", flush=True)
    assert rep["resampled"] >= 1


def test_batched_256_filter_step_matches_oracle(gpu):
    """Configuration 3: 256 filters (particles log-uniform in [1e3, 1e5],
    satellites U{6..20}, NLOS generator, Cauchy) in one step per stage; a
    sample of filters (the smallest, the largest and six others) is checked
    every step against its own OraclePF."""
    from gnss_gpu_amd import sim
    from gnss_gpu_amd._lib import LIK_STUDENT
    from gnss_gpu_amd.batched import BatchedParticleFilter

    F, T = 256, 6
    n, s = sim.batched_filters(F, seed=1)
    sc = sim.batched_scenarios(s, T, seed=2)
    b = BatchedParticleFilter(n, sat_cap=s, n_slots=T)
    b.load_epochs([[sc[f].sat for f in range(F)] for _ in range(T)],
                  [[sc[f].pr[t] for f in range(F)] for t in range(T)], [s] * T,
                  vel=np.tile(np.array([x.vel for x in sc])[None], (T, 1, 1)))
    init = np.array([np.concatenate([x.truth[0] - x.vel, [x.cb]]) for x in sc])
    seeds = np.arange(F) + 100
    b.initialize(init, seeds, spread_pos=10.0, spread_cb=10.0)
    st = b.step_args(sigma_pos=SIG_POS, sigma_cb=SIG_CB, sigma_pr=SIG_PR, nu=NU, lik=LIK_STUDENT,
                     n_slots=T, ess_frac=0.5)
    rng = np.random.default_rng(3)
    sample = sorted(set([int(np.argmin(n)), int(np.argmax(n))] +
                        rng.choice(F, 6, replace=False).tolist()))
    orcs = {}
    for f in sample:
        o = opf.OraclePF(int(n[f]), seed=int(seeds[f]))
        o.initialize(init[f, :3], init[f, 3], 10.0, 10.0)
        orcs[f] = o
    resampled = 0
    for t in range(T):
        b.step_part(st, 0)
        lw32 = {f: b.get_log_weights_raw(f) for f in sample}
        b.step_part(st, 1)
        dyn = b.read_dyn()
        for f in sample:
            o = orcs[f]
            o.predict(sc[f].vel, 1.0, SIG_POS, SIG_CB, t + 1)
            o.weight(sc[f].sat, sc[f].pr[t], sigma_pr=SIG_PR, nu=NU)
            res = bool(dyn[f].did_resample)
            ess_o = o.ess()
            if abs(ess_o - 0.5 * n[f]) > 1e-4 * n[f]:
                assert res == (ess_o < 0.5 * n[f]), (f, t)
            p = b.get_particles(f)
            if res:
                resampled += 1
                o.resample_systematic(lw32=lw32[f])
                np.testing.assert_array_equal(b.get_ancestors(f), o.ancestors)
                lw_at = lw32[f][o.ancestors].astype(np.float64)
            else:
                lw_at = lw32[f].astype(np.float64)
            assert np.max(np.abs(p - o.state)) < 1e-3, (f, t)
            lw_given = opf.log_likelihood(p[:, :3], p[:, 3], p[:, 4:], sc[f].sat, sc[f].pr[t],
                                          np.ones(int(s[f])), None, SIG_PR, nu=NU)
            np.testing.assert_allclose(lw_at, lw_given, rtol=1e-5, atol=2e-5)
            est = np.array([dyn[f].est[k] for k in range(4)])
            assert np.max(np.abs(est - o.estimate())) < 1e-3, (f, t)
    assert resampled >= 2
    b.close()


def _cfg5_features(anchor, R, sat, cn0, flag, ns, essf):
    """The learned head's per-satellite inputs as the weight kernel forms them
    (binary32 unit vector from the step frame's anchor in the filter's ENU
    frame; gg_mlp_features)."""
    d = (np.asarray(sat) - anchor) @ R.T
    u = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32).astype(np.float64)
    ch = np.hypot(u[:, 0], u[:, 1])
    S = len(sat)
    return np.stack([np.arcsin(np.clip(u[:, 2], -1, 1)) / (np.pi / 2), u[:, 0] / ch,
                     u[:, 1] / ch, (np.asarray(cn0, np.float64) - 35.0) * 0.1,
                     np.asarray(flag, np.float64), np.full(S, ns * 0.05),
                     np.full(S, essf)], axis=1).astype(np.float32)


def test_scale_config_step_reduced_matches_oracle(gpu):
    """Configuration 5 at reduced N: 65,536 particles x 32 satellites (8 each
    of 4 systems, ISB states), learned likelihood with FP8 E4M3 weights through
    its code tables, Megopolis B = 64, canyon drive.  Per step: the learned
    log-weights against the oracle's fp8-mode network at the GPU's particles
    (the bf16-operand tolerance of the learned tests: 1e-3 relative + 0.05),
    Megopolis ancestors equal to the oracle's twin on the GPU's binary32
    log-weights, particles and estimates within 1 mm of OraclePF."""
    from gnss_gpu_amd import urban
    from gnss_gpu_amd._lib import LIK_LEARNED
    from gnss_gpu_amd.batched import BatchedParticleFilter
    from gnss_gpu_amd.learned import LearnedLikelihood
    from oracle import learned as ol

    N, S, B, T = 65_536, 32, 64, 8
    drive = urban.drive_ecef(np.random.default_rng(92), T, n_sat=S, speed=1.5, n_sys=4,
                             canyon_kw=dict(height=(20.0, 150.0)))
    p = ol.init_params(21).astype(np.float64)
    p[ol.OFF_W3:ol.OFF_B3] *= 10.0
    p = p.astype(np.float32)
    m = LearnedLikelihood(p)
    m.set_fp8(True)
    b = BatchedParticleFilter([N], n_isb=3, sat_cap=S, n_slots=T)
    b.load_epochs([[drive.sat[t]] for t in range(T)], [[drive.pr[t]] for t in range(T)],
                  [[S]] * T, sys=[[drive.sys] for _ in range(T)], vel=drive.vel[:, None, :])
    b.set_learned(m)
    flags = np.array([drive.los_flags(drive.truth[t], t) for t in range(T)], np.float32)
    b.load_features([[drive.cn0[t]] for t in range(T)], [[flags[t]] for t in range(T)])
    x0 = np.concatenate([drive.truth[0], [drive.cb], drive.isb])
    b.initialize(x0[None], [4242], spread_pos=5.0, spread_cb=5.0, spread_isb=2.0)
    st = b.step_args(sigma_pos=1.0, sigma_cb=1.0, sigma_isb=0.05, sigma_pr=5.0, nu=1.0,
                     lik=LIK_LEARNED, n_slots=T, ess_frac=0.5, learned_mode=1)
    orc = opf.OraclePF(N, seed=4242, n_isb=3)
    orc.initialize(x0[:3], x0[3], 5.0, 5.0, isb=x0[4:], spread_isb=2.0)
    essf = 1.0
    resampled = 0
    for t in range(T):
        b.step_part(st, 0, "megopolis", B)
        lw32 = b.get_log_weights_raw(0)
        b.step_part(st, 1, "megopolis", B)
        d = b.read_dyn()[0]
        anchor = np.array(d.anchor[:3])
        R = np.array(d.R[:9]).reshape(3, 3)
        orc.predict(drive.vel[t], 1.0, 1.0, 1.0, t + 1, sigma_isb=0.05)
        feats = _cfg5_features(anchor, R, drive.sat[t], drive.cn0[t], flags[t], S, essf)
        res = bool(d.did_resample)
        if res:
            resampled += 1
            anc_ref = detmath.megopolis_ancestors(lw32, N, B, 4242, orc.resample_count)
            np.testing.assert_array_equal(b.get_ancestors(0), anc_ref)
            orc.resample_megopolis_chesser(B, lw32=lw32)
            lw_at = lw32[anc_ref].astype(np.float64)
        else:
            lw_at = lw32.astype(np.float64)
        pg = b.get_particles(0)
        assert np.max(np.abs(pg - orc.state)) < 1e-3, t
        # learned log-weights at the GPU's particles (the resampled ones map to
        # their ancestors' log-weights)
        sysk = np.asarray(drive.sys)
        isb_term = np.where(sysk[None, :] >= 1, pg[:, 4 + np.maximum(sysk - 1, 0)], 0.0)
        e = drive.pr[t][None, :] - (np.linalg.norm(drive.sat[t][None] - pg[:, None, :3], axis=2)
                                    + pg[:, 3:4] + isb_term)
        ref = np.zeros(N)
        for k in range(S):
            ref += ol.mlp_forward(p, e[:, k].astype(np.float32),
                                  np.broadcast_to(feats[k], (N, 7)), mode="fp8")["ll"]
        bad = np.abs(lw_at - ref) > 1e-3 * np.abs(ref) + 0.05
        # the network sees a residual only through the bf16 code of 0.1 e; a
        # particle may differ only where one of its residuals lies within the
        # binary32 rounding of the device's residual (1e-5 m here) of a code
        # boundary, so that the two sides take neighbouring codes
        xe_lo = ol.scaled_residual((e[bad] - 1e-5).astype(np.float32))
        xe_hi = ol.scaled_residual((e[bad] + 1e-5).astype(np.float32))
        on_edge = np.any(ol.bf16(xe_lo) != ol.bf16(xe_hi), axis=1)
        assert np.all(on_edge), (t, np.nonzero(bad)[0][~on_edge])
        assert bad.sum() <= max(2, N // 10_000), (t, int(bad.sum()))
        est = np.array([d.est[k] for k in range(7)])
        assert np.max(np.abs(est - orc.estimate())) < 1e-3, t
        essf = float(d.ess) / N
    assert resampled >= 2
    b.close()
