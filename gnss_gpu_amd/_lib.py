"""ctypes binding of libgnss_gpu_amd.so (the C ABI declared in include/gnss_gpu_amd.h).

There is no CPU fallback: importing this module fails with ImportError when the
HIP library has not been built (``python -m gnss_gpu_amd.build``), and every
call that returns a non-zero status raises RuntimeError with the library's
message.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG_DIR, "libgnss_gpu_amd.so")

GG_TILE = 1024
GG_MAX_ISB = 3
GG_MAX_D = 4 + GG_MAX_ISB

LIK_GAUSS = 0
LIK_STUDENT = 1
LIK_HUBER = 2
LIK_LEARNED = 3
LIK_GMM = 4

RESAMPLER_SYSTEMATIC = 0
RESAMPLER_MEGOPOLIS = 1
RESAMPLER_MEGOPOLIS_REF = 2


class StepArgs(C.Structure):
    """Mirror of gg_step_args (include/gnss_gpu_amd_layout.h)."""

    _fields_ = [
        ("dt", C.c_double),
        ("sigma_pos", C.c_float),
        ("sigma_cb", C.c_float),
        ("sigma_isb", C.c_float),
        ("sigma_vel", C.c_float),
        ("alpha", C.c_float),
        ("qv", C.c_float),
        ("velocity_kf", C.c_int32),
        ("step", C.c_int32),
        ("step_from_device", C.c_int32),
        ("sigma_pr", C.c_float),
        ("nu", C.c_float),
        ("gate_thr", C.c_float),
        ("huber_k", C.c_float),
        ("lik", C.c_int32),
        ("slot_from_device", C.c_int32),
        ("n_slots", C.c_int32),
        ("ess_frac", C.c_double),
        ("auto_flag", C.c_int32),
        ("u0_mode", C.c_int32),
        ("recenter", C.c_int32),
        ("learned_mode", C.c_int32),
        ("u0_value", C.c_double),
        ("gmm_w_los", C.c_float),
        ("gmm_mu_nlos", C.c_float),
        ("gmm_sigma_nlos", C.c_float),
        # derived by the library before every launch
        ("gmm_c_los", C.c_float),
        ("gmm_c_nlos", C.c_float),
        ("gmm_ratio", C.c_float),
    ]


class Dyn(C.Structure):
    """Mirror of gg_dyn (per-filter device state)."""

    _fields_ = [
        ("anchor", C.c_double * 3),
        ("cb_anchor", C.c_double),
        ("isb_anchor", C.c_double * GG_MAX_ISB),
        ("R", C.c_double * 9),
        ("vel_u", C.c_double * 3),
        ("vcov_u", C.c_double * 6),
        ("M", C.c_double),
        ("W", C.c_double),
        ("W2", C.c_double),
        ("ess", C.c_double),
        ("u0", C.c_double),
        ("mean", C.c_double * GG_MAX_D),
        ("est", C.c_double * GG_MAX_D),
        ("shift", C.c_float * GG_MAX_D),
        ("cur", C.c_int32),
        ("lw_zero", C.c_int32),
        ("flag", C.c_int32),
        ("vel_mode", C.c_int32),
        ("epoch", C.c_int32),
        ("resample_count", C.c_int32),
        ("did_resample", C.c_int32),
        ("est_valid", C.c_int32),
        ("cnt_a", C.c_uint32),
        ("cnt_b", C.c_uint32),
        ("cnt_c", C.c_uint32),
        ("aos_only", C.c_int32),
    ]


_P = C.c_void_p
_D = C.c_double
_I = C.c_int32
_U64 = C.c_uint64
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_FP = C.POINTER(C.c_float)

# name -> argtypes (all return int status)
_SIGS = {
    "gg_mlp_create": [C.POINTER(_P)],
    "gg_mlp_destroy": [_P],
    "gg_mlp_set_params": [_P, _FP],
    "gg_mlp_get_params": [_P, _FP],
    "gg_mlp_set_fp8": [_P, _I],
    "gg_mlp_eval": [_P, _I, _FP, _FP, _FP, _FP],
    "gg_mlp_load_data": [_P, _I, _I, _FP, _FP, _FP, _IP],
    "gg_mlp_train": [_P, _I, _I, C.c_float, _U64, C.c_float, C.c_float, C.c_float, _FP],
    "gg_mlp_loss_grad": [_P, _I, _FP, _FP, _FP],
    "gg_pf_set_mlp": [_P, _P],
    "gg_pfb_load_features": [_P, _I, _FP, _FP],
    "gg_pf_device_weight_learned": [_P, _DP, _DP, _DP, _IP, _I, _FP, _FP, _I],
    "gg_abi_version": [],
    "gg_struct_sizes": [_IP, _IP, _IP],
    "gg_device_count": [_IP],
    "gg_pf_create": [_I, _IP, _IP, _I, _I, _I, C.POINTER(_P)],
    "gg_pf_device_create": [_I, _I, C.POINTER(_P)],
    "gg_pf_destroy": [_P],
    "gg_pf_set_stream": [_P, _P],
    "gg_pf_device_sync": [_P],
    "gg_pf_info": [_P, _IP, _IP, C.POINTER(C.c_int64)],
    "gg_pf_device_initialize": [_P, _D, _D, _D, _D, _D, _D, _U64, _D, _D, _D, _D, _D, _DP, _D],
    "gg_pf_device_predict": [_P, _D, _D, _D, _D, _D, _D, _U64, _I, _D, _D, _I, _D, _D],
    "gg_pf_device_weight": [_P, _DP, _DP, _DP, _IP, _I, _D, _D, _D, _I, _D],
    "gg_pf_device_weight_gmm": [_P, _DP, _DP, _DP, _I, _D, _D, _D, _D],
    "gg_pf_device_weight_dd_pseudorange": [_P, _DP, _DP, _DP, _DP, _DP, _DP, _I, _D, _D, _I, _D],
    "gg_pf_device_weight_carrier_afv": [_P, _DP, _DP, _DP, _I, _D, _D],
    "gg_pf_device_weight_dd_carrier_afv": [_P, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _I, _D, _D, _I,
                                           _D],
    "gg_pf_device_weight_doppler": [_P, _DP, _DP, _DP, _DP, _I, _D, _D, _D, _D],
    "gg_pf_device_doppler_kf_update": [_P, _DP, _DP, _DP, _DP, _I, _D, _D],
    "gg_pf_device_ess": [_P, _DP],
    "gg_pf_device_position_spread": [_P, _D, _D, _D, _DP],
    "gg_pf_device_resample_systematic": [_P, _U64, _I],
    "gg_pf_device_resample_megopolis": [_P, _I, _U64],
    "gg_pf_device_resample_megopolis_chesser": [_P, _I],
    "gg_pf_device_estimate": [_P, _DP],
    "gg_pf_device_estimate_full": [_P, _DP],
    "gg_pf_device_get_particles": [_P, _DP],
    "gg_pf_device_get_particles_full": [_P, _DP],
    "gg_pf_device_get_particle_states": [_P, _DP],
    "gg_pf_device_set_particle_states": [_P, _DP],
    "gg_pf_device_get_log_weights": [_P, _DP],
    "gg_pf_device_get_positions": [_P, _DP, _DP, _DP, _DP],
    "gg_pf_device_set_positions": [_P, _DP, _DP, _DP, _DP],
    "gg_pf_device_set_log_weights": [_P, _DP],
    "gg_pf_device_get_resample_ancestors": [_P, _IP],
    "gg_pf_device_shift_clock_bias": [_P, _D],
    "gg_pf_device_shift_position": [_P, _D, _D, _D],
    "gg_pf_device_position_update": [_P, _D, _D, _D, _D],
    "gg_pfb_initialize": [_P, _DP, C.POINTER(_U64), _D, _D, _D],
    "gg_pfb_load_epochs": [_P, _I, _DP, _DP, _DP, _IP, _IP, _DP],
    "gg_pfb_step": [_P, C.POINTER(StepArgs), _I, _I],
    "gg_pfb_capture": [_P, C.POINTER(StepArgs), _I, _I, _I],
    "gg_pfb_replay": [_P, _I],
    "gg_pfb_time_replays": [_P, _I, C.POINTER(C.c_float)],
    "gg_pfb_time_replays_each": [_P, _I, C.POINTER(C.c_float)],
    "gg_pfb_profile_steps": [_P, C.POINTER(StepArgs), _I, _I, _I, C.POINTER(C.c_float),
                             C.POINTER(C.c_int32)],
    "gg_pfb_time_stage": [_P, C.POINTER(StepArgs), _I, _I, _I, C.POINTER(C.c_float)],
    "gg_pfb_read_dyn": [_P, C.POINTER(Dyn)],
    "gg_pfb_set_epoch": [_P, _I],
    "gg_pfb_get_particles": [_P, _I, _DP],
    "gg_pfb_get_log_weights": [_P, _I, _DP],
    "gg_pfb_step_part": [_P, C.POINTER(StepArgs), _I, _I, _I],
    "gg_pfb_get_log_weights_raw": [_P, _I, _FP],
    "gg_pfb_get_ancestors": [_P, _I, _IP],
    "gg_pfb_set_tuning": [_P, _I, _I],
    "gg_pfb_get_tuning": [_P, _IP, _IP],
    "gg_pfb_autotune": [_P, C.POINTER(StepArgs), _I, _I, _I, _I, _I, _IP, _FP, _IP],
    "gg_wls_batch": [_DP, _DP, _DP, _IP, _I, _I, _I, _D, _DP, _IP],
    "gg_wls_batch_device": [_P, _P, _P, _P, _I, _I, _I, _D, _P, _P, _I, _P],
    "gg_wls_batch_noalloc": [_DP, _DP, _DP, _IP, _I, _I, _I, _D, _DP, _IP],
    "gg_wls_position": [_DP, _DP, _DP, _I, _I, _D, _DP, _IP],
    "gg_wls_multi_gnss_batch": [_DP, _DP, _DP, _IP, _I, _I, _I, _I, _D, _DP, _IP],
    "gg_wls_multi_gnss": [_DP, _DP, _DP, _IP, _I, _I, _I, _D, _DP, _IP],
    "gg_wls_multi_gnss_batch_device": [_P, _P, _P, _P, _I, _I, _I, _I, _D, _P, _P, _I, _P],
    "gg_bvh_build": [_DP, _I, _DP, _IP, _IP],
    "gg_pf_device_weight_3d_bvh": [_P, _P, _DP, _DP, _DP, _IP, _I, _D, _D, _D, _D, _D],
    "gg_ffbsi_sample": [_DP, _DP, _DP, _DP, _DP, _D, _IP, _I, _I, _I, _DP, _IP, _P, _IP],
    "gg_transition_logpdf": [_DP, _DP, _I, _DP, _D, _D, _D, _DP],
    "gg_ffbsi_sample_device": [_P, _P, _P, _P, _P, _D, _P, _I, _I, _I, _P, _P, _P, _P, _P],
    "gg_pf_device_record": [_P, _P, _P],
    "gg_bvh_create": [_DP, _I, _DP, _I, _P],
    "gg_bvh_destroy": [_P],
    "gg_raytrace_los_check_bvh": [_DP, _DP, _DP, _I, _DP, _I, _IP, _I],
    "gg_raytrace_los_check_bvh_batch": [_DP, _DP, _DP, _I, _DP, _I, _IP, _I, _I],
    "gg_raytrace_multipath_bvh": [_DP, _DP, _DP, _I, _DP, _I, _DP, _DP, _I],
    "gg_raytrace_multipath_bvh_batch": [_DP, _DP, _DP, _I, _DP, _I, _DP, _DP, _I, _I],
    "gg_raytrace_los_check": [_DP, _DP, _DP, _I, _IP, _I],
    "gg_raytrace_multipath": [_DP, _DP, _DP, _I, _DP, _DP, _I],
    "gg_pf_weight_3d_bvh": [_DP, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _I, _DP, _I, _DP, _I, _I,
                            _D, _D, _D, _D, _D],
    "gg_pf_weight_3d": [_DP, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _I, _DP, _I, _I, _D, _D, _D, _D,
                        _D],
    "gg_pf_weight_3d_bvh_device": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _D, _D, _D, _D, _D,
                                   _P],
    "gg_debug_math": [_I, _DP, _DP, _I],
}

EXPORTED = tuple(["gg_last_error", *_SIGS.keys()])


def _load():
    # PyTorch ships its own HIP runtime; import it first so the process has a
    # single runtime and this library binds to it (torch tensors and our
    # kernels then share devices and streams).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library with `python -m gnss_gpu_amd.build`"
        )
    lib = C.CDLL(LIB_PATH)
    lib.gg_last_error.restype = C.c_char_p
    lib.gg_last_error.argtypes = []
    for name, argt in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = C.c_int
        fn.argtypes = argt
    return lib


lib = _load()


def _check_layout():
    sa, dy, fi = C.c_int32(), C.c_int32(), C.c_int32()
    lib.gg_struct_sizes(C.byref(sa), C.byref(dy), C.byref(fi))
    if sa.value != C.sizeof(StepArgs) or dy.value != C.sizeof(Dyn):
        raise ImportError(
            f"ABI layout mismatch: gg_step_args {sa.value} vs {C.sizeof(StepArgs)}, "
            f"gg_dyn {dy.value} vs {C.sizeof(Dyn)}"
        )


_check_layout()


def last_error() -> str:
    msg = lib.gg_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = last_error()
        raise RuntimeError(f"{what}: {msg}" if what else msg)


def call(name: str, *args) -> None:
    check(getattr(lib, name)(*args), name)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_DP)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_IP)


def fptr(a: np.ndarray):
    return a.ctypes.data_as(_FP)


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def device_count() -> int:
    n = C.c_int32()
    lib.gg_device_count(C.byref(n))
    return int(n.value)


def default_step_args(**kw) -> StepArgs:
    s = StepArgs()
    s.dt = 1.0
    s.sigma_pos = 1.0
    s.sigma_cb = 300.0
    s.sigma_isb = 0.0
    s.sigma_vel = 0.0
    s.alpha = 1.0
    s.qv = 0.0
    s.velocity_kf = 0
    s.step = 1
    s.step_from_device = 1
    s.sigma_pr = 5.0
    s.nu = 0.0
    s.gate_thr = 0.0
    s.huber_k = 1.5
    s.lik = LIK_GAUSS
    s.slot_from_device = 1
    s.n_slots = 1
    s.ess_frac = 0.5
    s.auto_flag = 1
    s.u0_mode = 0
    s.recenter = 1
    s.u0_value = 0.0
    s.gmm_w_los = 0.7
    s.gmm_mu_nlos = 15.0
    s.gmm_sigma_nlos = 30.0
    for k, v in kw.items():
        if not hasattr(s, k):
            raise AttributeError(f"StepArgs has no field {k!r}")
        setattr(s, k, v)
    return s
