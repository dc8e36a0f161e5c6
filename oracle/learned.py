"""CPU oracle of the learned pseudorange likelihood (TEST INFRASTRUCTURE ONLY).

The learned likelihood has no counterpart in the reference: it replaces the
Cauchy/Student-t term of pfd_weight_kernel (src/particle_filter/pf_device.cu:
262-266) by a small MLP per (particle, satellite) pair that predicts the NLOS
bias mean mu and log-variance l of the pseudorange residual e:

    x  = [clamp(0.1 e, -20, 20), f_1..f_7]           (f: per-satellite features)
    z1 = W1 x + b1, h1 = relu(z1)                     (64)
    z2 = W2 h1 + b2, h2 = relu(z2)                    (64)
    z3 = W3 h2 + b3;  mu = 10 z3_0,  l = clamp(z3_1, -4, 12)
    ll = -(e - mu)^2 exp(-l) / 2 - l / 2

Training objective for one epoch with candidate offsets d_k of the receiver
(d_0 = 0) and L_k = sum_s ll(e_ks, f_s): the cross-entropy between the
normalised particle weights softmax(L) and the target softmax(-|d_k|^2 /
(2 tau^2)).  The residual at a candidate is e_ks = r0_s + u_s . dpos_k - dcb_k
(r0: residual at the truth, u: unit vector to the satellite).

Two modes: exact binary64 (pinned by central finite differences), and "bf16",
which rounds at the points where the device rounds (the first-layer operands,
h1, h2, the layer-2/3 weights and, in the backward pass, dz2) and accumulates
in binary64 where the device accumulates in binary32.
"""

from __future__ import annotations

import numpy as np

H = 64
NF = 8
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * NF
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + 2 * H
N_PARAMS = OFF_B3 + 2
LV_MIN, LV_MAX = -4.0, 12.0
MU_SCALE = 10.0
X_CLAMP = 20.0
TAU = 2.0


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(H, NF), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(2, H), p[OFF_B3:N_PARAMS])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, np.ravel(W3), b3]).astype(np.float32)


def init_params(seed=0):
    """He-style initialisation, z3 scaled down, b3 = [0, ln 25]."""
    r = np.random.default_rng(seed)
    W1 = r.normal(0, np.sqrt(2.0 / NF), (H, NF))
    W2 = r.normal(0, np.sqrt(2.0 / H), (H, H))
    W3 = r.normal(0, 0.1 * np.sqrt(1.0 / H), (2, H))
    return pack(W1, np.zeros(H), W2, np.zeros(H), W3, np.array([0.0, np.log(25.0)]))


def bf16(x):
    """Round binary32 values to the nearest bfloat16 (ties to even), as binary64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(a), np.nan, out)


def e4m3_scale(W):
    """Per-tensor scale of the FP8 weight mode (twin of gg_e4m3_scale): the
    smallest power of two 2^k with max|W| / 2^k <= 448 (binary32 quotient),
    1 for an all-zero tensor."""
    mm = np.float32(np.max(np.abs(np.asarray(W, dtype=np.float32)))) if np.size(W) else 0.0
    if not mm > 0:
        return np.float32(1.0)
    r = np.float32(mm / np.float32(448.0))
    _, e = np.frexp(r)
    sc = np.float32(np.ldexp(1.0, int(e)))
    if np.float32(np.ldexp(1.0, int(e) - 1)) >= r:
        sc = np.float32(np.ldexp(1.0, int(e) - 1))
    return sc


def q_e4m3(x, sc):
    """OCP FP8 E4M3FN value nearest to x / sc (ties to even, saturated at
    +-448), times sc (twin of gg_q_e4m3)."""
    x = np.asarray(x, dtype=np.float32)
    y = (x / np.float32(sc)).astype(np.float32)
    a = np.abs(y)
    _, e = np.frexp(a)
    lead = np.maximum(e.astype(np.int32) - 1, -6)
    step = np.ldexp(np.float32(1.0), lead - 3).astype(np.float32)
    with np.errstate(invalid="ignore"):
        q = np.minimum(np.rint(a / step) * step, np.float32(448.0)).astype(np.float32)
    q = np.where(a >= 448.0, np.float32(448.0), q)
    out = (np.copysign(q, y) * np.float32(sc)).astype(np.float32)
    out = np.where(a > 0, out, x * np.float32(0.0))
    return out.astype(np.float32)


def quantised(p, mode):
    """Parameters as the device uses them: in "fp8" mode W1, W2 and W3 are
    replaced by their per-tensor scaled E4M3 values (binary64 arrays)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    if mode == "fp8":
        W1 = q_e4m3(W1, e4m3_scale(W1)).astype(np.float64)
        W2 = q_e4m3(W2, e4m3_scale(W2)).astype(np.float64)
        W3 = q_e4m3(W3, e4m3_scale(W3)).astype(np.float64)
    return W1, b1, W2, b2, W3, b3


def first_layer_bias(W1, b1, f):
    """c = b1 + sum_j W1[:, 1+j] f_j in binary32, j left to right (device order).
    f [..., 7] -> [..., 64]."""
    f = np.asarray(f, dtype=np.float32)
    c = np.broadcast_to(np.asarray(b1, dtype=np.float32), f.shape[:-1] + (H,)).copy()
    for j in range(NF - 1):
        c = (c + (np.asarray(W1[:, 1 + j], dtype=np.float32) * f[..., j:j + 1])).astype(np.float32)
    return c


def scaled_residual(e):
    return np.clip(np.asarray(e, dtype=np.float32) * np.float32(0.1), -X_CLAMP, X_CLAMP)


def mlp_forward(p, e, f, mode="exact"):
    """e [...], f [..., 7] -> dict with ll [...] and the intermediates.
    mode: "exact" (binary64), "bf16" (device rounding), "fp8" (device
    rounding with the FP8 weight mode)."""
    W1, b1, W2, b2, W3, b3 = quantised(p, mode)
    e64 = np.asarray(e, dtype=np.float64)
    xe = scaled_residual(e).astype(np.float64)
    if mode in ("bf16", "fp8"):
        c = first_layer_bias(W1, b1, f).astype(np.float64)
        z1 = (bf16(W1[:, 0]) * bf16(xe)[..., None] + bf16(c)).astype(np.float32).astype(np.float64)
        h1 = bf16(np.maximum(z1, 0.0))
        z2 = (h1 @ bf16(W2).T + b2).astype(np.float32).astype(np.float64)
        h2 = bf16(np.maximum(z2, 0.0))
        z3 = (h2 @ bf16(W3).T + b3).astype(np.float32).astype(np.float64)
    else:
        x = np.concatenate([xe[..., None], np.asarray(f, dtype=np.float64)], axis=-1)
        z1 = x @ W1.T + b1
        h1 = np.maximum(z1, 0.0)
        z2 = h1 @ W2.T + b2
        h2 = np.maximum(z2, 0.0)
        z3 = h2 @ W3.T + b3
    mu = MU_SCALE * z3[..., 0]
    lv = np.clip(z3[..., 1], LV_MIN, LV_MAX)
    r = e64 - mu
    ll = -0.5 * r * r * np.exp(-lv) - 0.5 * lv
    return {"ll": ll, "z1": z1, "h1": h1, "z2": z2, "h2": h2, "z3": z3, "mu": mu, "lv": lv,
            "xe": xe, "e": e64}


def pair_loglik(e, mu, lv):
    """Log-density of e under N(mu, exp(lv)) without the 2 pi term."""
    r = np.asarray(e, dtype=np.float64) - mu
    return -0.5 * r * r * np.exp(-lv) - 0.5 * lv


def mlp_backward(p, f, fw, g_ll, mode="exact"):
    """Gradient of sum(g_ll * ll) with respect to the parameters, from the
    intermediates of mlp_forward.  Returns a flat binary64 vector."""
    W1, b1, W2, b2, W3, b3 = quantised(p, mode)
    z3 = fw["z3"]
    g = np.asarray(g_ll, dtype=np.float64)
    r = fw["e"] - fw["mu"]
    ex = np.exp(-fw["lv"])
    inside = (z3[..., 1] > LV_MIN) & (z3[..., 1] < LV_MAX)
    dz3 = np.stack([MU_SCALE * r * ex * g, np.where(inside, (0.5 * r * r * ex - 0.5) * g, 0.0)],
                   axis=-1)
    dev = mode in ("bf16", "fp8")
    W3u = bf16(W3) if dev else W3
    W2u = bf16(W2) if dev else W2
    dh2 = dz3 @ W3u
    dz2 = dh2 * (fw["z2"] > 0)
    dz2u = bf16(dz2) if dev else dz2
    dh1 = dz2u @ W2u
    dz1 = dh1 * (fw["z1"] > 0)
    lead = int(np.prod(g.shape))
    h1 = fw["h1"].reshape(lead, H)
    h2 = fw["h2"].reshape(lead, H)
    dW3 = dz3.reshape(lead, 2).T @ h2
    db3 = dz3.reshape(lead, 2).sum(0)
    dW2 = dz2u.reshape(lead, H).T @ h1
    db2 = dz2.reshape(lead, H).sum(0)
    xe = (bf16(fw["xe"]) if dev else fw["xe"]).reshape(lead)
    dz1f = dz1.reshape(lead, H)
    fl = np.broadcast_to(np.asarray(f, dtype=np.float64), g.shape + (NF - 1,)).reshape(lead, NF - 1)
    dW1 = np.concatenate([(dz1f * xe[:, None]).sum(0)[:, None], dz1f.T @ fl], axis=1)
    db1 = dz1f.sum(0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])


def candidate_residuals(r0, u, dpos, dcb):
    """e[k, s] = r0_s + u_s . dpos_k - dcb_k in binary32 (device order)."""
    r0 = np.asarray(r0, dtype=np.float32)
    u = np.asarray(u, dtype=np.float32)
    d = np.asarray(dpos, dtype=np.float32)
    ud = (u[None, :, 0] * d[:, None, 0] + u[None, :, 1] * d[:, None, 1]).astype(np.float32)
    ud = (ud + u[None, :, 2] * d[:, None, 2]).astype(np.float32)
    return ((r0[None, :] + ud).astype(np.float32) - np.asarray(dcb, np.float32)[:, None]).astype(
        np.float32)


def contrastive_loss_and_grad(p, r0, u, f, dpos, dcb, mode="exact"):
    """One epoch: r0 [S], u [S, 3], f [S, 7], candidates dpos [K, 3], dcb [K].
    Returns (loss, grad [N_PARAMS], dL [K])."""
    e = candidate_residuals(r0, u, dpos, dcb)
    fb = np.broadcast_to(np.asarray(f, dtype=np.float64)[None], e.shape + (NF - 1,))
    fw = mlp_forward(p, e, fb, mode)
    L = fw["ll"].sum(axis=1)
    Lm = L - L.max()
    pk = np.exp(Lm) / np.exp(Lm).sum()
    d2 = (np.asarray(dpos, np.float64) ** 2).sum(1) + np.asarray(dcb, np.float64) ** 2
    tt = np.exp(-(d2 - d2.min()) / (2 * TAU * TAU))
    tt /= tt.sum()
    loss = -np.sum(tt * (Lm - np.log(np.exp(Lm).sum())))
    dL = pk - tt
    g = mlp_backward(p, fb, fw, np.broadcast_to(dL[:, None], e.shape), mode)
    return float(loss), g, dL
