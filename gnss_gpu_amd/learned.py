"""Learned pseudorange likelihood (no reference counterpart).

A small MLP per (particle, satellite) pair predicts the NLOS bias mean and
log-variance of the pseudorange residual from the residual itself and per
satellite features (elevation, azimuth, C/N0, line-of-sight flag at the
particle mean, satellite count, ESS fraction); the pair log-likelihood
-(e - mu)^2 exp(-l) / 2 - l / 2 replaces the Student-t term of
pfd_weight_kernel (src/particle_filter/pf_device.cu:262-266).  Evaluation and
training run on the GPU (bf16 operands, binary32 accumulation on the matrix
cores); see gnss_gpu_amd/csrc/gg_mlp.h for the layout.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, f32, fptr, i32, iptr

H = 64
NF = 8
N_PARAMS = H * NF + H + H * H + H + 2 * H + 2


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation; the output layer starts small with the
    log-variance bias at ln(25 m^2)."""
    r = np.random.default_rng(seed)
    W1 = r.normal(0, np.sqrt(2.0 / NF), (H, NF))
    W2 = r.normal(0, np.sqrt(2.0 / H), (H, H))
    W3 = r.normal(0, 0.1 * np.sqrt(1.0 / H), (2, H))
    return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), W3.ravel(),
                           [0.0, np.log(25.0)]]).astype(np.float32)


class LearnedLikelihood:
    def __init__(self, params=None, seed: int = 0):
        h = C.c_void_p()
        call("gg_mlp_create", C.byref(h))
        self._h = h
        self.fp8 = False
        self.set_params(init_params(seed) if params is None else params)

    def close(self):
        if getattr(self, "_h", None) is not None:
            _lib.lib.gg_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_params(self, p):
        p = f32(p).ravel()
        if p.size != N_PARAMS:
            raise ValueError(f"expected {N_PARAMS} parameters, got {p.size}")
        call("gg_mlp_set_params", self._h, fptr(p))

    def set_fp8(self, enable: bool = True):
        """FP8 E4M3 weights with per-tensor power-of-two scales (see
        gg_mlp_set_fp8); the master parameters stay binary32."""
        call("gg_mlp_set_fp8", self._h, 1 if enable else 0)
        self.fp8 = bool(enable)

    def get_params(self) -> np.ndarray:
        out = np.empty(N_PARAMS, dtype=np.float32)
        call("gg_mlp_get_params", self._h, fptr(out))
        return out

    def eval_groups(self, e, f):
        """e [G, 16] residuals [m], f [G, 7] satellite features ->
        (ll [G, 16], z3 [G, 16, 2])."""
        e = f32(e).reshape(-1, 16)
        f = f32(f).reshape(-1, NF - 1)
        if e.shape[0] != f.shape[0]:
            raise ValueError("e and f must have the same number of groups")
        g = e.shape[0]
        ll = np.empty((g, 16), dtype=np.float32)
        z3 = np.empty((g, 16, 2), dtype=np.float32)
        call("gg_mlp_eval", self._h, g, fptr(e), fptr(f), fptr(ll), fptr(z3))
        return ll, z3

    # -- training ------------------------------------------------------------
    def load_data(self, f, r0, u, n_sat):
        """f [E, S, 7], r0 [E, S], u [E, S, 3], n_sat [E] (S <= 16)."""
        f = f32(f)
        E, S = f.shape[0], f.shape[1]
        call("gg_mlp_load_data", self._h, E, S, fptr(f), fptr(f32(r0)), fptr(f32(u)),
             iptr(i32(n_sat)))
        self.n_train = E

    def train(self, n_steps, batch=256, lr=2e-3, seed=0, cand_sigma=(5.0, 3.0, 5.0)):
        """Adam on the contrastive loss; returns the per-step mean batch loss."""
        loss = np.empty(int(n_steps), dtype=np.float32)
        call("gg_mlp_train", self._h, int(n_steps), int(batch), float(lr), int(seed),
             float(cand_sigma[0]), float(cand_sigma[1]), float(cand_sigma[2]), fptr(loss))
        return loss

    def loss_grad(self, cand):
        """Loss per epoch and the gradient summed over the first len(cand)
        loaded epochs, for candidates cand [E, 64, 4] (dx dy dz dcb)."""
        cand = f32(cand)
        E = cand.shape[0]
        grad = np.empty(N_PARAMS, dtype=np.float32)
        loss = np.empty(E, dtype=np.float32)
        call("gg_mlp_loss_grad", self._h, E, fptr(cand), fptr(grad), fptr(loss))
        return loss, grad
