"""Batched, device-driven particle filters: many independent filters advanced by
one hipGraph-captured step.

Each filter has its own particle count, satellite count, seed and anchor; all
of them share one structure-of-arrays arena in HBM and every stage of the step
is one launch over all filters (tile -> filter tables).  Satellite data for a
whole trajectory can be preloaded into epoch slots; the step reads slot
``epoch % n_slots`` from a device-side epoch counter, so graph replays advance
the filters with no host work.

There is no reference counterpart for the batch (the reference advances one
filter per ParticleFilterDevice); a batch of one reproduces
ParticleFilterDevice's step exactly.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, dptr, f64, i32, iptr

STAGE_WEIGHT_PREDICT = 0
STAGE_WEIGHT = 1
STAGE_SYSTEMATIC = 2  # decision, CDF ends, gather and finalize in one launch
STAGE_CDF = STAGE_SYSTEMATIC
STAGE_GATHER_SYSTEMATIC = 3  # the same launch
STAGE_MEGOPOLIS = 4
STAGE_GATHER_ANCESTORS = 5
STAGE_DECIDE = 6

RESAMPLERS = {"systematic": 0, "megopolis": 1, "megopolis_ref": 2}


class BatchedParticleFilter:
    def __init__(self, n_particles, n_isb=0, sat_cap=64, n_slots=1, with_velocity=False):
        n = np.atleast_1d(np.asarray(n_particles, dtype=np.int32))
        self.n_filters = int(n.size)
        self.n = n
        cap = np.broadcast_to(np.asarray(sat_cap, dtype=np.int32), n.shape).copy()
        self.sat_cap = cap
        self.sat_off = np.concatenate([[0], np.cumsum(cap)[:-1]]).astype(np.int64)
        self.sat_rows = int(cap.sum())
        self.n_isb = int(n_isb)
        self.D = 4 + self.n_isb
        self.n_slots = int(n_slots)
        h = C.c_void_p()
        nn = i32(n)
        cc = i32(cap)
        call("gg_pf_create", self.n_filters, iptr(nn), iptr(cc), self.n_isb, self.n_slots,
             1 if with_velocity else 0, C.byref(h))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None:
            _lib.lib.gg_pf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- setup -----------------------------------------------------------------
    def initialize(self, init, seeds, spread_pos=10.0, spread_cb=10.0, spread_isb=0.0):
        init = f64(init).reshape(self.n_filters, self.D)
        seeds = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64).reshape(self.n_filters))
        call("gg_pfb_initialize", self._h, dptr(init),
             seeds.ctypes.data_as(C.POINTER(C.c_uint64)), float(spread_pos), float(spread_cb),
             float(spread_isb))

    def load_epochs(self, sats, prs, n_sat, weights=None, sys=None, vel=None):
        """sats[t][f] -> (S_tf, 3), prs[t][f] -> (S_tf,), optional weights/sys
        the same shape, vel[t][f] -> (3,) ECEF.  Lists (epoch-major) of
        per-filter arrays, or dense arrays shaped [T, F, S, ...]."""
        T = len(prs)
        rows = self.sat_rows
        xyz = np.zeros((T, rows, 3))
        pr = np.zeros((T, rows))
        w = np.ones((T, rows))
        sy = np.zeros((T, rows), dtype=np.int32)
        ns = np.zeros((T, self.n_filters), dtype=np.int32)
        for t in range(T):
            for f in range(self.n_filters):
                k = int(n_sat[t][f])
                o = int(self.sat_off[f])
                if k > self.sat_cap[f]:
                    raise ValueError("more satellites than the filter's sat_cap")
                xyz[t, o:o + k] = np.asarray(sats[t][f], dtype=np.float64).reshape(-1, 3)[:k]
                pr[t, o:o + k] = np.asarray(prs[t][f], dtype=np.float64)[:k]
                if weights is not None:
                    w[t, o:o + k] = np.asarray(weights[t][f], dtype=np.float64)[:k]
                if sys is not None:
                    sy[t, o:o + k] = np.asarray(sys[t][f], dtype=np.int32)[:k]
                ns[t, f] = k
        v = np.zeros((T, self.n_filters, 3)) if vel is None else f64(vel).reshape(T, self.n_filters, 3)
        call("gg_pfb_load_epochs", self._h, T, dptr(xyz), dptr(pr), dptr(w), iptr(sy), iptr(ns),
             dptr(v))
        self.loaded_slots = T

    def set_learned(self, model):
        """Attach a LearnedLikelihood (None detaches); steps use it with
        step_args(lik=LIK_LEARNED).  The model must outlive its use."""
        self._mlp = model
        call("gg_pf_set_mlp", self._h, None if model is None else model.handle)

    def load_features(self, cn0, los_flag):
        """Per-satellite C/N0 [dB-Hz] and line-of-sight flags of the learned
        likelihood, epoch-major like load_epochs: cn0[t][f] -> (S_tf,)."""
        T = len(cn0)
        rows = self.sat_rows
        c = np.zeros((T, rows), dtype=np.float32)
        fl = np.zeros((T, rows), dtype=np.float32)
        for t in range(T):
            for f in range(self.n_filters):
                o = int(self.sat_off[f])
                v = np.asarray(cn0[t][f], dtype=np.float32).ravel()
                c[t, o:o + v.size] = v
                fl[t, o:o + v.size] = np.asarray(los_flag[t][f], dtype=np.float32).ravel()
        call("gg_pfb_load_features", self._h, T, c.ctypes.data_as(C.POINTER(C.c_float)),
             fl.ctypes.data_as(C.POINTER(C.c_float)))

    # -- stepping --------------------------------------------------------------
    def step_args(self, **kw):
        kw.setdefault("n_slots", getattr(self, "loaded_slots", self.n_slots))
        return _lib.default_step_args(**kw)

    def step(self, args, resampler="systematic", n_iter=32):
        call("gg_pfb_step", self._h, C.byref(args), RESAMPLERS[resampler], int(n_iter))

    def step_part(self, args, part, resampler="systematic", n_iter=32):
        """Half of step(): part 0 the likelihood launch (predict + weight + tile
        partials), part 1 the decision and resampling (with its finalize)."""
        call("gg_pfb_step_part", self._h, C.byref(args), RESAMPLERS[resampler], int(n_iter),
             int(part))

    def set_tuning(self, weight_tpw=None, mego_depth=None):
        """Launch shapes of the step (results do not depend on them); None
        keeps a shape as it is.  Capture after changing them."""
        cur = self.get_tuning()
        tpw = cur["weight_tpw"] if weight_tpw is None else int(weight_tpw)
        depth = cur["mego_depth"] if mego_depth is None else int(mego_depth)
        call("gg_pfb_set_tuning", self._h, tpw, depth)

    def get_tuning(self):
        w, d = C.c_int32(), C.c_int32()
        call("gg_pfb_get_tuning", self._h, C.byref(w), C.byref(d))
        return {"weight_tpw": int(w.value), "mego_depth": int(d.value)}

    def autotune(self, args, resampler="systematic", n_iter=32, steps=10, reps=3):
        """Time every candidate launch shape on graphs of `steps` steps from a
        snapshot of the filters (restored afterwards, so the trajectory is
        unchanged) and keep the fastest.  Returns {"chosen": shapes,
        "candidates": [(shapes, ms per step), ...]}.  Capture afterwards."""
        cap = 16
        cand = np.zeros(2 * cap, dtype=np.int32)
        ms = np.zeros(cap, dtype=np.float32)
        n = C.c_int32()
        call("gg_pfb_autotune", self._h, C.byref(args), RESAMPLERS[resampler], int(n_iter),
             int(steps), int(reps), cap, cand.ctypes.data_as(C.POINTER(C.c_int32)),
             ms.ctypes.data_as(C.POINTER(C.c_float)), C.byref(n))
        rows = [({"weight_tpw": int(cand[2 * k]), "mego_depth": int(cand[2 * k + 1])},
                 float(ms[k])) for k in range(n.value)]
        return {"chosen": self.get_tuning(), "candidates": rows}

    def capture(self, args, resampler="systematic", n_iter=32, steps=1):
        call("gg_pfb_capture", self._h, C.byref(args), RESAMPLERS[resampler], int(n_iter),
             int(steps))

    def replay(self, n=1):
        call("gg_pfb_replay", self._h, int(n))

    def time_replays(self, n):
        ms = C.c_float()
        call("gg_pfb_time_replays", self._h, int(n), C.byref(ms))
        return float(ms.value)

    def time_replays_each(self, n):
        """Device ms of each of n graph replays (hipEvents on the handle's stream)."""
        out = np.zeros(int(n), dtype=np.float32)
        call("gg_pfb_time_replays_each", self._h, int(n),
             out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def profile_steps(self, args, n_steps, resampler="systematic", n_iter=32):
        """Eager steps, every kernel timed at its own start and end (events
        recorded by the launch itself), a launch group from its first kernel's
        start to its last kernel's end.  Returns
        ({stage: ms} for steps that resampled, {stage: ms} for the others,
        (n_resampled, n_other)); stages are predict_weight and resample (the
        systematic decision, CDF ends, gather and finalize in one launch), or
        predict_weight, decide and megopolis_gather (weights and packed states,
        chains with the state copy and finalize)."""
        ms = np.zeros(6, dtype=np.float32)
        cnt = np.zeros(2, dtype=np.int32)
        call("gg_pfb_profile_steps", self._h, C.byref(args), RESAMPLERS[resampler], int(n_iter),
             int(n_steps), ms.ctypes.data_as(C.POINTER(C.c_float)),
             cnt.ctypes.data_as(C.POINTER(C.c_int32)))
        if resampler == "systematic":
            names = ("predict_weight", "resample")
        else:
            names = ("predict_weight", "decide", "megopolis_gather")
        k = len(names)
        return (dict(zip(names, map(float, ms[:k]))), dict(zip(names, map(float, ms[3:3 + k]))),
                (int(cnt[0]), int(cnt[1])))

    def time_stage(self, args, stage, reps, n_iter=32):
        ms = C.c_float()
        call("gg_pfb_time_stage", self._h, C.byref(args), int(stage), int(n_iter), int(reps),
             C.byref(ms))
        return float(ms.value)

    def set_epoch(self, epoch):
        call("gg_pfb_set_epoch", self._h, int(epoch))

    def sync(self):
        call("gg_pf_device_sync", self._h)

    def set_stream(self, stream=None):
        """Enqueue steps, captures and replays on ``stream`` (torch.cuda.Stream
        or raw hipStream_t) from now on; None: the handle's own stream."""
        from ._gnss_gpu_pf_device import stream_handle

        call("gg_pf_set_stream", self._h, stream_handle(stream))

    # -- results ---------------------------------------------------------------
    def read_dyn(self):
        arr = (_lib.Dyn * self.n_filters)()
        call("gg_pfb_read_dyn", self._h, arr)
        return arr

    def estimates(self):
        d = self.read_dyn()
        return np.array([[d[f].est[k] for k in range(self.D)] for f in range(self.n_filters)])

    def ess(self):
        d = self.read_dyn()
        return np.array([d[f].ess for f in range(self.n_filters)])

    def get_particles(self, f):
        out = np.empty((int(self.n[f]), self.D), dtype=np.float64)
        call("gg_pfb_get_particles", self._h, int(f), dptr(out))
        return out

    def get_log_weights_raw(self, f):
        """Filter f's binary32 log-weights as the last likelihood launch wrote
        them (also after a resampling, which only marks them zero)."""
        out = np.empty(int(self.n[f]), dtype=np.float32)
        call("gg_pfb_get_log_weights_raw", self._h, int(f),
             out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def get_ancestors(self, f):
        """Ancestor indices of filter f's last resampling: out[j] = source of slot j."""
        out = np.empty(int(self.n[f]), dtype=np.int32)
        call("gg_pfb_get_ancestors", self._h, int(f), iptr(out))
        return out

    def get_log_weights(self, f):
        out = np.empty(int(self.n[f]), dtype=np.float64)
        call("gg_pfb_get_log_weights", self._h, int(f), dptr(out))
        return out
