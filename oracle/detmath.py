"""Bit-exact NumPy twins of the deterministic pieces of the GPU resampler.

Every function here uses only correctly rounded IEEE-754 operations in the
same order as the HIP code (gnss_gpu_amd/csrc/gg_device.h and
gg_pf_kernels*.h), so for identical inputs the outputs are identical bit for
bit.  The reference algorithm they implement is the systematic resampler of
src/particle_filter/pf_device.cu:1813-1878 (max, exp-shift, normalise,
inclusive scan, a_j = min{i : cdf_i >= (u0 + j)/N} with hi = N-1) and the
Megopolis kernel of pf_device.cu:1231-1298; the reference leaves the scan
order to thrust, so the order here is the library's own fixed one.
"""

from __future__ import annotations

import math

import numpy as np

from . import philox

TILE = 1024
BLOCK = 256
ITEMS = 4

_INV_LN2 = float.fromhex("0x1.71547652b82fep+0")
_LN2_HI = float.fromhex("0x1.62e42fee00000p-1")
_LN2_LO = float.fromhex("0x1.a39ef35793c76p-33")
_EXP_C = [float.fromhex(h) for h in (
    "0x1.6124613a86d09p-33", "0x1.1eed8eff8d898p-29", "0x1.ae64567f544e4p-26",
    "0x1.27e4fb7789f5cp-22", "0x1.71de3a556c734p-19", "0x1.a01a01a01a01ap-16",
    "0x1.a01a01a01a01ap-13", "0x1.6c16c16c16c17p-10", "0x1.1111111111111p-7",
    "0x1.5555555555555p-5", "0x1.5555555555555p-3", "0x1.0p-1", "0x1.0p+0", "0x1.0p+0")]

_F = np.float32
_INV_LN2_F = _F(float.fromhex("0x1.715476p+0"))
_LN2_HI_F = _F(float.fromhex("0x1.62e4p-1"))
_LN2_LO_F = _F(float.fromhex("0x1.7f7d1cp-20"))
_EXP_CF = [_F(float.fromhex(h)) for h in (
    "0x1.a01a02p-13", "0x1.6c16c2p-10", "0x1.111112p-7", "0x1.555556p-5",
    "0x1.555556p-3", "0x1.0p-1", "0x1.0p+0", "0x1.0p+0")]


def exp64(x):
    """Twin of gg_exp64: Cody-Waite reduction + degree-13 Taylor, binary64."""
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    nan = np.isnan(x)
    hi = x > 709.782712893384
    lo = x < -745.1332191019412
    mid = ~(nan | hi | lo)
    xm = x[mid]
    k = np.rint(xm * _INV_LN2)
    r = (xm - k * _LN2_HI) - k * _LN2_LO
    p = np.full_like(r, _EXP_C[0])
    for c in _EXP_C[1:]:
        p = p * r + c
    out[mid] = np.ldexp(p, k.astype(np.int32))
    out[nan] = x[nan]
    out[hi] = np.inf
    return out


def fmaf(a, b, c):
    """Correctly rounded binary32 fused multiply-add a*b + c (twin of
    __builtin_fmaf / v_fma_f32): the product of two binary32 values is exact in
    binary64, TwoSum gives the exact error of the binary64 sum, the sum is taken
    to binary64 with round-to-odd and then rounded to binary32 (exact because
    53 >= 24 + 2).  Finite inputs; arrays broadcast."""
    a = np.asarray(a, dtype=np.float32).astype(np.float64)
    b = np.asarray(b, dtype=np.float32).astype(np.float64)
    c = np.asarray(c, dtype=np.float32).astype(np.float64)
    p = a * b
    s = p + c
    bp = s - p
    ap = s - bp
    err = (p - ap) + (c - bp)
    # truncation toward zero of the exact p + c: one step toward zero when the
    # rounded sum lies beyond it
    away = ((s > 0) & (err < 0)) | ((s < 0) & (err > 0))
    t = np.where(away, np.nextafter(s, 0.0), s)
    bits = t.view(np.int64) | (err != 0).astype(np.int64)
    return bits.view(np.float64).astype(np.float32)


def exp32(x):
    """Twin of gg_exp32 (binary32, fused multiply-adds)."""
    x = np.asarray(x, dtype=np.float32)
    out = np.zeros_like(x)
    nan = np.isnan(x)
    hi = x > _F(88.72283)
    lo = x < _F(-103.97208)
    mid = ~(nan | hi | lo)
    xm = x[mid]
    k = np.rint(xm * _INV_LN2_F)
    r = fmaf(-k, _LN2_LO_F, fmaf(-k, _LN2_HI_F, xm))
    p = np.full_like(r, _EXP_CF[0])
    for c in _EXP_CF[1:]:
        p = fmaf(p, r, c)
    out[mid] = np.ldexp(p, k.astype(np.int32)).astype(np.float32)
    out[nan] = x[nan]
    out[hi] = np.inf
    return out


def _hs_wave(x):
    """Hillis-Steele inclusive scan over the last axis (64 lanes)."""
    x = np.array(x, dtype=np.float64, copy=True)
    d = 1
    while d < 64:
        y = x.copy()
        x[..., d:] = y[..., :-d] + y[..., d:]
        d <<= 1
    return x


def block_scan256(v):
    """Twin of gg_block_scan256 over the last axis (256 threads).
    Returns (inclusive, exclusive, total)."""
    v = np.asarray(v, dtype=np.float64)
    lead = v.shape[:-1]
    x = _hs_wave(v.reshape(*lead, 4, 64))
    t = x[..., :, 63]
    o1 = t[..., 0]
    o2 = o1 + t[..., 1]
    o3 = o2 + t[..., 2]
    tot = o3 + t[..., 3]
    incl = x.copy()
    incl[..., 1, :] = o1[..., None] + x[..., 1, :]
    incl[..., 2, :] = o2[..., None] + x[..., 2, :]
    incl[..., 3, :] = o3[..., None] + x[..., 3, :]
    incl = incl.reshape(*lead, 256)
    excl = np.concatenate([np.zeros((*lead, 1)), incl[..., :-1]], axis=-1)
    return incl, excl, tot


def block_total256(v):
    """Twin of gg_block_totals for one value per thread."""
    return block_scan256(v)[2]


def tile_view(lw, n):
    """Pad a filter's log-weights to whole tiles: returns (lw[nt,256,4], valid)."""
    nt = (n + TILE - 1) // TILE
    pad = np.zeros(nt * TILE, dtype=np.float32)
    pad[:n] = np.asarray(lw, dtype=np.float32)[:n]
    valid = np.zeros(nt * TILE, dtype=bool)
    valid[:n] = True
    return pad.reshape(nt, BLOCK, ITEMS), valid.reshape(nt, BLOCK, ITEMS)


def tile_partials(lw, n):
    """Per-tile m_b = max lw (binary32, exact), e_i = exp32(lw_i - m_b) and
    S_b = sum e_i accumulated in binary64 (pf_device.cu:924-965 pattern,
    library order: per thread left to right over its 4 particles, then
    gg_block_scan256's total).  Returns (m[nt] as binary64, S[nt],
    e[nt,256,4] as binary64)."""
    lwt, val = tile_view(lw, n)
    nt = lwt.shape[0]
    m32 = np.fmax.reduce(np.where(val, lwt, _F(-np.inf)).reshape(nt, -1), axis=1).astype(np.float32)
    e = np.zeros(lwt.shape, dtype=np.float32)
    ok = val & (m32[:, None, None] > -np.inf)
    diff = (lwt - m32[:, None, None]).astype(np.float32)
    e[ok] = exp32(diff[ok])
    ed = e.astype(np.float64)
    s = ed[..., 0].copy()
    for q in range(1, ITEMS):
        s = s + ed[..., q]
    S = block_total256(s)
    return m32.astype(np.float64), S, ed


def kcount(c, n, u0):
    """#{j in [0, N) : (u0 + j) / N <= c}, NaN -> 0 (twin of gg_kcount)."""
    c = np.asarray(c, dtype=np.float64)
    k = np.searchsorted(systematic_targets(n, u0), c, side="right").astype(np.int64)
    return np.where(np.isnan(c), 0, k)


def tile_scale(m):
    """Twin of gg_tile_scale: k_b = floor(m_b / ln 2) and c_b = exp64((m_b - k_b
    ln2_hi) - k_b ln2_lo) in [1, 2), so that exp(lw_i) is proportional to
    e_i c_b 2^k_b for e_i = exp32(lw_i - m_b); empty tiles (m_b = -inf) get
    k = -2^30, c = 0; a non-finite maximum k = 0, c = 1."""
    m = np.asarray(m, dtype=np.float64)
    k = np.zeros_like(m)
    c = np.ones_like(m)
    empty = ~(m > -np.inf)
    fin = ~empty & (m < np.inf)
    k[empty] = -1073741824.0
    c[empty] = 0.0
    kf = np.floor(m[fin] * _INV_LN2)
    k[fin] = kf
    c[fin] = exp64((m[fin] - kf * _LN2_HI) - kf * _LN2_LO)
    return k, c


def tile_shift(v, dk):
    """Twin of gg_tile_shift: 2^dk v, exactly, for dk >= -1000, else 0."""
    v = np.asarray(v, dtype=np.float64)
    dk = np.asarray(dk, dtype=np.float64)
    ok = dk >= -1000.0
    return np.where(ok, np.ldexp(v, np.where(ok, dk, 0.0).astype(np.int32)), 0.0)


def decide(m, S):
    """Twin of gg_combine: M = max m_b, K = max k_b, the exact power-of-two
    tile scales g_b = 2^(k_b - K) c_b, T_b = 2^(k_b - K) (c_b S_b), the tile
    offsets P_b and the total W in the fixed chunked order."""
    nt = m.shape[0]
    M = float(np.max(m))
    k, c = tile_scale(m)
    K = float(np.max(k))
    g = tile_shift(c, k - K)
    T = tile_shift(S * c, k - K)
    c = (nt + BLOCK - 1) // BLOCK
    TT = np.zeros((BLOCK, c))
    idx = np.arange(nt)
    TT[idx // c, idx % c] = T
    tot = np.zeros(BLOCK)
    for k in range(c):
        tot = tot + TT[:, k]
    _, excl, W = block_scan256(tot)
    P = np.zeros(nt)
    run = np.zeros(BLOCK)
    for k in range(c):
        sel = np.arange(BLOCK) * c + k
        okb = sel < nt
        P[sel[okb]] = excl[okb] + run[okb]
        run = run + TT[:, k]
    return M, g, P, float(W)


def systematic_targets(n, u0):
    """t_j = (u0 + j) / N in binary64 (pf_device.cu:1197)."""
    return (u0 + np.arange(n, dtype=np.float64)) / float(n)


def reference_u0(seed, n):
    """u0 = (seed % 1e6) / (1e6 * N) with the reference's int32 product
    (pf_device.cu:1852)."""
    den = (1000000 * int(n)) & 0xFFFFFFFF
    if den >= 2 ** 31:
        den -= 2 ** 32
    return float(int(seed) % 1000000) / float(den)


def fresh_u0(seed, resample_count):
    r = philox.draw(seed, 0, resample_count, philox.PURPOSE_U0, 0)
    return float(philox.u01d(r[0], r[1])) - 2.0 ** -53


def tile_ends(P, W, n, u0):
    """K_b = #{j : t_j <= P_b / W} for b = 0 .. nt (P_nt = W, so K_nt = N):
    the outputs that fall before tile b."""
    Pn = np.append(np.asarray(P, dtype=np.float64), W)
    return kcount(Pn / W, n, u0)


def systematic_ends(lw, n, u0):
    """Per-particle output ends E_i of the library's systematic resampler:
    cdf_i = (P_b + g_b L_i) / W with L_i the in-tile inclusive scan of e,
    k_i = #{j : t_j <= cdf_i}, clamped to the tile's end count K_{b+1}
    (tile_ends); the tile's last particle takes K_{b+1} itself and the
    filter's last particle N.  E is non-decreasing, so the ancestors
    a_j = #{i : E_i <= j} = min{i : E_i > j}, and every tile's outputs are
    fixed by the tile offsets alone."""
    m, S, e = tile_partials(lw, n)
    M, g, P, W = decide(m, S)
    nt = m.shape[0]
    l = e.copy()
    for q in range(1, ITEMS):
        l[..., q] = l[..., q - 1] + e[..., q]
    _, excl, _ = block_scan256(l[..., ITEMS - 1])
    L = excl[..., None] + l
    cdf = ((P[:, None, None] + g[:, None, None] * L) / W).reshape(nt * TILE)[:n]
    K = tile_ends(P, W, n, u0)
    tile = np.arange(n) // TILE
    k = np.minimum(kcount(cdf, n, u0), K[tile + 1])
    last = (np.arange(n) % TILE) == TILE - 1
    k[last] = K[tile[last] + 1]
    k[n - 1] = n
    return k, cdf


def systematic_ancestors(lw, n, u0):
    """Ancestor indices of the library's systematic resampler."""
    E, _ = systematic_ends(lw, n, u0)
    return np.searchsorted(E, np.arange(n), side="right").astype(np.int32)


def megopolis_ref_ancestors(lw, n, n_iter, seed):
    """Reference-rule Megopolis in ancestor form (pf_device.cu:1231-1298):
    per iteration, offset = floor(U1 (N-1)) + 1, j = (i + offset) mod N,
    accept if U2 < min(1, exp(lw_j - lw_i)) with the slot weights."""
    lw64 = np.asarray(lw, dtype=np.float32)[:n].astype(np.float64)
    i = np.arange(n)
    a = i.copy()
    for it in range(n_iter):
        r = philox.draw(seed, i, it, philox.PURPOSE_MEGO_REF, 0)
        u1 = philox.u01d(r[0], r[1])
        u2 = philox.u01d(r[2], r[3])
        off = (u1 * float(n - 1)).astype(np.int64) + 1
        j = (i + off) % n
        alpha = np.minimum(1.0, exp64(lw64[j] - lw64[i]))
        acc = u2 < alpha
        a = np.where(acc, a[j], a[i])
    return a.astype(np.int32)


def megopolis_weights(lw, n, lw_zero=False):
    """w_i = exp32(lw_i - M), M = max lw (binary32; NaN ignored), or 1 for
    zeroed weights (twin of gg_k_mego_w)."""
    lw32 = np.asarray(lw, dtype=np.float32)[:n]
    if lw_zero:
        return np.ones(n, dtype=np.float32)
    M = np.float32(np.nanmax(lw32)) if n else np.float32(0.0)
    return exp32((lw32 - M).astype(np.float32))


def megopolis_ancestors(lw, n, n_iter, seed, call, lw_zero=False):
    """Megopolis of Chesser et al. (2021) as implemented by gg_k_mego: Metropolis
    chains on the original weights, one offset per 256-particle segment; a
    chain at k moves to proposal j when u * w_k < w_j (binary32 product),
    w = megopolis_weights; iteration b's uniform is output b of the chain's
    xoshiro128+ stream."""
    w = megopolis_weights(lw, n, lw_zero)
    i = np.arange(n, dtype=np.int64)
    seg = i // 256
    k = i.copy()
    wk = w.copy()
    nseg = int(seg.max()) + 1 if n else 0
    segs = np.arange(nseg)
    xs = philox.xo_seed(seed, i, call, philox.PURPOSE_MEGO_U, 0)  # one stream per chain
    for b in range(n_iter):
        r = philox.draw(seed, segs, call, philox.PURPOSE_MEGO, b)
        O = ((r[0] * np.uint64(n)) >> np.uint64(32)).astype(np.int64)
        u = philox.u01f(philox.xo_next(xs))
        j = i + O[seg]
        j = np.where(j >= n, j - n, j)
        wj = w[j]
        with np.errstate(invalid="ignore"):
            acc = (u * wk).astype(np.float32) < wj
        k = np.where(acc, j, k)
        wk = np.where(acc, wj, wk)
    return k.astype(np.int32)


def sequential_systematic(w_norm, u0):
    """Textbook systematic resampling on normalised weights (sequential cumsum,
    first index with cdf >= target, clamped) — the reference semantics with a
    sequential scan; used as an independent check."""
    n = w_norm.shape[0]
    cdf = np.cumsum(np.asarray(w_norm, dtype=np.float64))
    t = systematic_targets(n, u0)
    a = np.searchsorted(cdf, t, side="left")
    return np.minimum(a, n - 1).astype(np.int32)


__all__ = [n for n in dir() if not n.startswith("_")]
_ = math
