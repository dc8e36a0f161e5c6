"""NumPy twin of the Philox4x32-10 streams used by the GPU library.

Philox4x32-10 as published by Salmon et al., "Parallel Random Numbers: As Easy
as 1, 2, 3" (SC'11); pinned by the published known-answer vectors in
tests/test_oracle_pinning.py.  Stream layout (shared with gg_device.h):
key = (seed lo, seed hi), counter = (index, step, purpose, call).
"""

from __future__ import annotations

import numpy as np

M32 = np.uint64(0xFFFFFFFF)
MUL0 = np.uint64(0xD2511F53)
MUL1 = np.uint64(0xCD9E8D57)
WEYL0 = np.uint64(0x9E3779B9)
WEYL1 = np.uint64(0xBB67AE85)

PURPOSE_INIT = 1
PURPOSE_PREDICT = 2
PURPOSE_MEGO_REF = 3
PURPOSE_MEGO = 4
PURPOSE_U0 = 5
PURPOSE_MEGO_U = 6


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10; inputs broadcast, returns 4 uint64 arrays (< 2^32)."""
    c0, c1, c2, c3, k0, k1 = (np.asarray(v, dtype=np.uint64) & M32
                              for v in (c0, c1, c2, c3, k0, k1))
    c0, c1, c2, c3, k0, k1 = np.broadcast_arrays(c0, c1, c2, c3, k0, k1)
    c0, c1, c2, c3, k0, k1 = (v.copy() for v in (c0, c1, c2, c3, k0, k1))
    for r in range(10):
        if r:
            k0 = (k0 + WEYL0) & M32
            k1 = (k1 + WEYL1) & M32
        p0 = MUL0 * c0
        p1 = MUL1 * c2
        c0, c1, c2, c3 = ((p1 >> np.uint64(32)) ^ c1 ^ k0, p1 & M32,
                          (p0 >> np.uint64(32)) ^ c3 ^ k1, p0 & M32)
    return c0, c1, c2, c3


def draw(seed, idx, step, purpose, call):
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return philox4x32_10(idx, step, purpose, call, seed & 0xFFFFFFFF, seed >> 32)


def u01f(x):
    """uint32 -> binary32 uniform in (0, 1), exact."""
    x = np.asarray(x, dtype=np.uint64)
    return ((x >> np.uint64(9)).astype(np.float32) + np.float32(0.5)) * np.float32(2.0 ** -23)


def u01d(hi, lo):
    """two uint32 -> binary64 uniform in (0, 1], exact."""
    hi = np.asarray(hi, dtype=np.uint64)
    lo = np.asarray(lo, dtype=np.uint64)
    m = (hi << np.uint64(21)) | (lo >> np.uint64(11))
    return (m + np.uint64(1)).astype(np.float64) * (2.0 ** -53)


def normal4(seed, idx, step, purpose, call):
    """Four standard normals per index (binary64 Box-Muller of the same uniforms
    the GPU turns into binary32 normals).  Returns shape (len(idx), 4)."""
    r = draw(seed, idx, step, purpose, call)
    u = [u01f(v).astype(np.float64) for v in r]
    out = []
    for a, b in ((u[0], u[1]), (u[2], u[3])):
        rad = np.sqrt(-2.0 * np.log(a))
        out.append(rad * np.cos(2.0 * np.pi * b))
        out.append(rad * np.sin(2.0 * np.pi * b))
    return np.stack(out, axis=-1)


def xo_seed(seed, idx, step, purpose, call):
    """xoshiro128+ state (Blackman and Vigna, ACM TOMS 2021) seeded by the
    Philox block (seed, idx, step, purpose, call); the all-zero state, the
    generator's fixed point, becomes (1, 0, 0, 0) (twin of gg_xo_seed).
    Returns a list of four uint64 arrays holding 32-bit words."""
    s0, s1, s2, s3 = draw(seed, idx, step, purpose, call)
    zero = (s0 | s1 | s2 | s3) == 0
    return [s0 | zero.astype(np.uint64), s1, s2, s3]


def xo_next(st):
    """Next xoshiro128+ output of every stream in st (advanced in place):
    out = s0 + s3, then the xorshift-rotate state update (twin of gg_xo_next)."""
    s0, s1, s2, s3 = st
    out = (s0 + s3) & M32
    t = (s1 << np.uint64(9)) & M32
    s2 = s2 ^ s0
    s3 = s3 ^ s1
    s1 = s1 ^ s2
    s0 = s0 ^ s3
    s2 = s2 ^ t
    s3 = ((s3 << np.uint64(11)) | (s3 >> np.uint64(21))) & M32
    st[:] = [s0, s1, s2, s3]
    return out


def xo_normals(st, count):
    """count (even) standard normals per stream: consecutive Box-Muller pairs of
    the streams' next outputs (binary64 arithmetic on the exact binary32
    uniforms).  Returns [streams, count]."""
    out = np.empty((np.asarray(st[0]).size, count))
    for k in range(0, count, 2):
        a = u01f(xo_next(st)).astype(np.float64)
        b = u01f(xo_next(st)).astype(np.float64)
        rad = np.sqrt(-2.0 * np.log(a))
        out[:, k] = rad * np.cos(2.0 * np.pi * b)
        out[:, k + 1] = rad * np.sin(2.0 * np.pi * b)
    return out


def predict_normals(seed, n, step, n_isb):
    """The normals one predict call draws (gg_pred_normals): particles 4g..4g+3
    share the stream (seed, g, step, PURPOSE_PREDICT, call).  Call 0 gives
    particle 4g + q the normals 4 q D .. of its components (x, y, z in the
    local ENU frame, clock bias, ISBs; D = 4 + n_isb), call 1 its three
    velocity-noise normals.  Returns z [n, 4], zv [n, 3], zi [n, n_isb]."""
    n = int(n)
    D = 4 + int(n_isb)
    g = np.arange((n + 3) // 4)
    zp = xo_normals(xo_seed(seed, g, step, PURPOSE_PREDICT, 0), 4 * D).reshape(-1, D)[:n]
    zv = xo_normals(xo_seed(seed, g, step, PURPOSE_PREDICT, 1), 12).reshape(-1, 3)[:n]
    return zp[:, :4], zv, zp[:, 4:]

