#!/usr/bin/env python3
"""Benchmark: one particle-filter step at 1,000,000 particles x 12 satellites.

Workload (the 1M-particle configuration of the metric): GPS/Galileo/BeiDou/
GLONASS, 3 satellites each, state [e n u cb isb_E isb_C isb_R] as binary32
offsets around a binary64 anchor, 30% NLOS per epoch (bias Exp(20 m)), LOS
noise N(0, 3 m), Cauchy likelihood (Student-t nu = 1) with scale 5 m,
systematic resampling when ESS < N/2, weighted-mean estimate.  A step is
predict + weight + ESS/decision + [resample] + estimate, captured once in a
hipGraph and replayed; satellite data for the whole trajectory is resident in
HBM before the timed region starts.

Prints ONE JSON line.  Only one GPU is used: --gpus > 1 exits with a message.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "1M-particle PF step ms + %HBM roofline; learned vs plain at equal P50 err"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--resampler", default="systematic",
                    choices=["systematic", "megopolis", "megopolis_ref"])
    ap.add_argument("--mego-iters", type=int, default=32)
    ap.add_argument("--stage-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (counter calibration)")
    return ap.parse_args()


def copy_calibration(mib):
    """Plain copy kernel of known size: HBM read + write of `mib` MiB each."""
    import torch

    n = mib * 1024 * 1024 // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "launches": 5}), flush=True)


def cpu_info():
    model = platform.processor() or "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; the particle-filter "
                          "step is measured per GPU (--gpus 1)"}))
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    import torch

    from gnss_gpu_amd import sim
    from gnss_gpu_amd.batched import (BatchedParticleFilter, STAGE_CDF,
                                      STAGE_GATHER_ANCESTORS, STAGE_GATHER_SYSTEMATIC,
                                      STAGE_MEGOPOLIS, STAGE_WEIGHT_PREDICT)
    from gnss_gpu_amd._lib import LIK_STUDENT

    N = args.particles
    T = args.warmup + args.steps + args.stage_reps * 4 + 8
    sc = sim.scenario_multi_gnss(n_epochs=T, seed=0)
    S = sc.n_sat
    D = 7

    pf = BatchedParticleFilter([N], n_isb=3, sat_cap=S, n_slots=T)
    sats = [[sc.sat] for _ in range(T)]
    prs = [[sc.pr[t]] for t in range(T)]
    sys_ids = [[sc.sys] for _ in range(T)]
    pf.load_epochs(sats, prs, n_sat=[[S]] * T, sys=sys_ids, vel=np.tile(sc.vel, (T, 1, 1)))
    # every step predicts first, so the filter starts one epoch before the first
    # measurement epoch
    init = np.concatenate([sc.truth[0] - sc.vel, [sc.cb], sc.isb])[None, :]
    pf.initialize(init, [12345], spread_pos=10.0, spread_cb=10.0, spread_isb=5.0)
    st = pf.step_args(sigma_pos=1.0, sigma_cb=1.0, sigma_isb=0.1, sigma_pr=5.0, nu=1.0,
                      lik=LIK_STUDENT, n_slots=T, ess_frac=0.5)

    # ---- accuracy pass (untimed): first epochs, host reads after each step
    p50 = None
    if not args.no_accuracy:
        errs = []
        for t in range(min(60, T)):
            pf.step(st, args.resampler, args.mego_iters)
            est = pf.estimates()[0]
            errs.append(float(np.linalg.norm(est[:3] - sc.truth[t])))
        p50 = float(np.median(errs[10:])) if len(errs) > 10 else float(np.median(errs))
        pf.initialize(init, [12345], spread_pos=10.0, spread_cb=10.0, spread_isb=5.0)
        pf.set_epoch(0)

    # ---- timed region: graph replays of one full step
    pf.capture(st, args.resampler, args.mego_iters, steps=1)
    pf.replay(args.warmup)
    pf.sync()
    torch.cuda.synchronize()
    rc0 = pf.read_dyn()[0].resample_count
    ms_total = pf.time_replays(args.steps)
    pf.sync()
    torch.cuda.synchronize()
    rc1 = pf.read_dyn()[0].resample_count
    ms_step = ms_total / args.steps
    resample_fraction = (rc1 - rc0) / args.steps

    # ---- per-kernel timing (hipEvents on the handle's stream)
    reps = args.stage_reps
    stage_ms = {}
    stage_ms["predict_weight"] = pf.time_stage(st, STAGE_WEIGHT_PREDICT, reps) / reps
    if args.resampler == "systematic":
        stage_ms["cdf"] = pf.time_stage(st, STAGE_CDF, reps) / reps
        stage_ms["gather"] = pf.time_stage(st, STAGE_GATHER_SYSTEMATIC, reps) / reps
    elif args.resampler == "megopolis":
        stage_ms["megopolis"] = pf.time_stage(st, STAGE_MEGOPOLIS, reps, args.mego_iters) / reps
        stage_ms["gather"] = pf.time_stage(st, STAGE_GATHER_ANCESTORS, reps) / reps
    pf.sync()

    # algorithmic HBM bytes per particle and launch (binary32 state, D = 7)
    bytes_pp = {
        "predict_weight": 2 * D * 4 + 4,         # read + write state, write lw
        "cdf": 4 + 4,                            # read lw, write output ends
        "gather": 4 + 2 * D * 4 + 4,             # read ends, gather state, write ancestors
        "megopolis": 4 + 4 * args.mego_iters + 4,
    }
    dom = max(stage_ms, key=stage_ms.get)
    achieved = bytes_pp[dom] * N / (stage_ms[dom] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                         "hbm_traffic.json")
    if os.path.exists(tpath):
        try:
            rec = json.load(open(tpath)).get(dom)
            if rec and rec.get("particles") == N and rec.get("satellites") == S:
                traffic = rec["bytes_per_launch"]
        except (OSError, ValueError, KeyError):
            traffic = None
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "algorithmic_bytes": bytes_pp[dom] * N,
                "avg_launch_ms": round(stage_ms[dom], 5)}

    # ---- CPU baseline (bounded sample of the same workload, host cores)
    cpu = None
    if not args.no_cpu:
        from oracle import build_cpu

        lib = build_cpu.load()
        import ctypes as C

        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
        h = lib.cpu_pf_create(N, 3, 12345, threads)
        ini = np.ascontiguousarray(init[0])
        DP = C.POINTER(C.c_double)
        lib.cpu_pf_initialize(h, ini.ctypes.data_as(DP), 10.0, 10.0, 5.0)
        sat = np.ascontiguousarray(sc.sat.ravel())
        sysi = np.ascontiguousarray(sc.sys.astype(np.int32))
        vel = np.ascontiguousarray(sc.vel)
        est = np.zeros(8)
        times = []
        t_start = time.perf_counter()
        t = 0
        while time.perf_counter() - t_start < args.cpu_seconds and t < 200:
            pr = np.ascontiguousarray(sc.pr[t % T])
            t0 = time.perf_counter()
            lib.cpu_pf_step(h, vel.ctypes.data_as(DP), 1.0, 1.0, 1.0, 0.1, t + 1,
                            sat.ctypes.data_as(DP), pr.ctypes.data_as(DP),
                            sysi.ctypes.data_as(C.POINTER(C.c_int)), S, 5.0, 1.0, 0.5,
                            est.ctypes.data_as(DP))
            times.append((time.perf_counter() - t0) * 1e3)
            t += 1
        lib.cpu_pf_destroy(h)
        cpu = {"value": round(float(np.median(times)), 3), "unit": "ms",
               "cores": int(lib.cpu_pf_threads()), "kind": "port",
               "sample": f"{len(times)} steps of the same 1M x 12 workload (binary64, "
                         f"OpenMP C restatement), median ms/step; CPU {cpu_info()}"}

    out = {
        "metric": METRIC,
        "value": round(ms_step, 5),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "pf_step_1Mx12_mgnss_cauchy", "particles": N, "satellites": S,
                   "state_dim": D, "likelihood": "cauchy(scale=5m)",
                   "resampler": args.resampler, "ess_threshold": 0.5, "nlos_fraction": 0.3,
                   "graph_steps": 1},
        "particles_per_s": round(N / (ms_step * 1e-3), 1),
        "resample_fraction": round(resample_fraction, 3),
        "stage_ms": {k: round(v, 5) for k, v in stage_ms.items()},
        "p50_err_m": None if p50 is None else round(p50, 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
