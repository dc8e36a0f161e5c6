"""A filter on the caller's stream (gg_pf_set_stream): its kernels and copies
queue behind the torch work already on that stream, results equal those of the
same filter on its own stream bit for bit, and it can go back to a stream of
its own.  The reference's device filter owns one stream per state
(src/particle_filter/pf_device.cu:1351) and never takes the caller's; this
library keeps that default and adds the caller-stream option.

The ordering check: a torch kernel that spins for ~0.2 s (torch.cuda._sleep)
is queued on stream s, then the filter's predict / update / estimate on s.
estimate() returns host values, so it synchronises the filter's stream; the
event recorded on s after the spin must then have completed, which holds only
if the filter's work queued behind it on s.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RX = np.array([-3957199.0, 3310205.0, 3737911.0])
SPIN_CYCLES = 400_000_000  # ~0.2 s of one spinning wave at gfx950 clocks


def _sats():
    from gnss_gpu_amd import sim

    rng = np.random.default_rng(3)
    sat = sim.place_satellites(RX, rng.uniform(15, 85, 8), rng.uniform(0, 360, 8))
    pr = np.linalg.norm(sat - RX, axis=1) + 100.0 + rng.normal(0, 3, 8)
    return sat, pr


def _device_pair():
    from gnss_gpu_amd import ParticleFilterDevice

    out = []
    for _ in range(2):
        pf = ParticleFilterDevice(n_particles=100_000, sigma_pos=1.0, sigma_cb=3.0, sigma_pr=5.0,
                                  resampling="systematic", seed=11)
        pf.initialize(RX, clock_bias=100.0, spread_pos=20.0, spread_cb=10.0)
        out.append(pf)
    return out


def _epoch(pf, sat, pr, k):
    pf.predict(velocity=np.array([1.0, -0.5, 0.25]) * k, dt=1.0)
    pf.update(sat, pr + 0.5 * k)
    return pf.estimate()


def test_device_filter_queues_behind_torch_work_on_its_stream(gpu):
    import torch

    sat, pr = _sats()
    pf, ref = _device_pair()
    s = torch.cuda.Stream()
    pf.set_stream(s)
    spun = torch.cuda.Event()
    with torch.cuda.stream(s):
        x = torch.arange(1 << 20, device="cuda", dtype=torch.float32)
        torch.cuda._sleep(SPIN_CYCLES)
        y = x * 2.0
        spun.record(s)
    assert not spun.query(), "the spin ended before the filter's work was queued"
    est = _epoch(pf, sat, pr, 1)
    assert spun.query(), "the filter's work did not queue behind the torch work on its stream"
    np.testing.assert_array_equal(est, _epoch(ref, sat, pr, 1))
    # torch work queued after the filter's on the same stream sees a consistent order too
    with torch.cuda.stream(s):
        z = y + 1.0
    s.synchronize()
    np.testing.assert_array_equal(z.cpu().numpy(), 2.0 * np.arange(1 << 20, dtype=np.float32) + 1)
    for k in (2, 3):
        np.testing.assert_array_equal(_epoch(pf, sat, pr, k), _epoch(ref, sat, pr, k))
    np.testing.assert_array_equal(pf.get_particles(), ref.get_particles())
    # back to a stream of its own
    pf.set_stream(None)
    np.testing.assert_array_equal(_epoch(pf, sat, pr, 4), _epoch(ref, sat, pr, 4))


def test_batched_graph_replays_on_a_torch_stream(gpu):
    """Capture and replay of the device-driven step on the caller's stream
    give the estimates of the same batch on its own stream, bit for bit."""
    import torch

    from gnss_gpu_amd import sim
    from gnss_gpu_amd._lib import LIK_STUDENT
    from gnss_gpu_amd.batched import BatchedParticleFilter

    T = 20
    sc = sim.scenario_multi_gnss(n_epochs=T, seed=2)
    S = sc.n_sat
    init = np.concatenate([sc.truth[0] - sc.vel, [sc.cb], sc.isb])[None, :]
    runs = []
    s = torch.cuda.Stream()
    for on_torch in (False, True):
        pf = BatchedParticleFilter([50_000], n_isb=3, sat_cap=S, n_slots=T)
        if on_torch:
            pf.set_stream(s)
        pf.load_epochs([[sc.sat]] * T, [[sc.pr[t]] for t in range(T)], [[S]] * T,
                       sys=[[sc.sys]] * T, vel=np.tile(sc.vel, (T, 1, 1)))
        pf.initialize(init, [99], 10.0, 10.0, 5.0)
        st = pf.step_args(sigma_pos=1.0, sigma_cb=1.0, sigma_isb=0.1, sigma_pr=5.0, nu=1.0,
                          lik=LIK_STUDENT, n_slots=T, ess_frac=0.5)
        pf.capture(st, "systematic", 32, steps=5)
        est = []
        for _ in range(T // 5):
            if on_torch:
                with torch.cuda.stream(s):
                    torch.cuda._sleep(SPIN_CYCLES // 8)
            pf.replay(1)
            est.append(pf.estimates().copy())
        runs.append((np.stack(est), pf.get_particles(0)))
        pf.close()
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
