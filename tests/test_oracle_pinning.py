"""Pins the CPU oracle before it is trusted as the GPU's checker.

Each oracle function is checked against something that does not restate its
own formula: published known-answer vectors, exact (rational) arithmetic,
textbook closed forms, brute force on tiny inputs, the reference's own test
constants (tests/golden/reference_cases.json) and invariants.
"""

import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import detmath, philox
from oracle import pf as opf
from oracle import wls as owls

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_cases.json")))


# ---------------------------------------------------------------- Philox ----
@pytest.mark.parametrize(
    "ctr,key,expect",
    [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ],
)
def test_philox_known_answers(ctr, key, expect):
    """Random123 known-answer vectors for philox4x32-10."""
    out = philox.philox4x32_10(*ctr, *key)
    assert tuple(int(v) for v in out) == expect


def _xo_transition_matrix():
    """The xoshiro128+ state update as a 128 x 128 matrix over GF(2), column k
    the update of the state with only bit k set (bit 32 w + b = bit b of s_w)."""
    T = np.zeros((128, 128), dtype=np.int64)
    for k in range(128):
        st = [np.zeros(1, dtype=np.uint64) for _ in range(4)]
        st[k // 32][0] = np.uint64(1) << np.uint64(k % 32)
        philox.xo_next(st)
        for w in range(4):
            v = int(st[w][0])
            for b in range(32):
                T[32 * w + b, k] = (v >> b) & 1
    return T


def test_xoshiro_transition_matches_published_jump_polynomial():
    """The published jump of xoshiro128+ (Blackman and Vigna's reference code,
    JUMP = {0x8764000b, 0xf542d2d3, 0x6fa035c3, 0x77f2db5b}) advances the state
    by 2^64 steps: sum_b c_b T^b must equal T^(2^64).  A wrong shift, rotation
    or xor order in xo_next changes T and fails."""
    T = _xo_transition_matrix()
    P = T.copy()
    for _ in range(64):
        P = (P @ P) % 2
    J = np.zeros_like(T)
    Tb = np.eye(128, dtype=np.int64)
    for i, word in enumerate((0x8764000B, 0xF542D2D3, 0x6FA035C3, 0x77F2DB5B)):
        for b in range(32):
            if (word >> b) & 1:
                J = (J + Tb) % 2
            Tb = (Tb @ T) % 2
    np.testing.assert_array_equal(J, P)


def test_xoshiro_outputs_and_seed_rules():
    """Output = s0 + s3 before the update (state (1, 2, 3, 4) gives 5, then
    7 + 12288 after one update by hand: s2 = 3^1, s3 = 4^2, s1 = 2^2, s0 = 1^6,
    s2 ^= 2 << 9, s3 = rotl(6, 11)); the all-zero seed becomes (1, 0, 0, 0);
    uniforms from the top bits are uniform (chi-square over 64 bins of 1e6
    draws) and consecutive outputs uncorrelated."""
    st = [np.array([v], dtype=np.uint64) for v in (1, 2, 3, 4)]
    assert int(philox.xo_next(st)[0]) == 5
    assert [int(v[0]) for v in st] == [7, 0, 1026, 12288]
    assert int(philox.xo_next(st)[0]) == 7 + 12288
    z = [np.zeros(1, dtype=np.uint64) for _ in range(4)]
    zero = (z[0] | z[1] | z[2] | z[3]) == 0
    assert int((z[0] | zero.astype(np.uint64))[0]) == 1
    xs = philox.xo_seed(7, np.arange(1000), 3, philox.PURPOSE_MEGO_U, 0)
    u = np.stack([philox.u01f(philox.xo_next(xs)) for _ in range(1000)], axis=1).astype(np.float64)
    counts = np.histogram(u.ravel(), bins=64, range=(0.0, 1.0))[0]
    chi2 = float(np.sum((counts - u.size / 64) ** 2 / (u.size / 64)))
    assert chi2 < 63 + 5 * np.sqrt(2 * 63)
    c = np.corrcoef(u[:, :-1].ravel(), u[:, 1:].ravel())[0, 1]
    assert abs(c) < 5.0 / np.sqrt(u.size)


def test_predict_normals_moments_and_independence():
    """Group streams: every component and particle slot is N(0, 1), and the
    normals of different components, of the four particles of one group and of
    neighbouring groups are uncorrelated."""
    n = 400_000
    z, zv, zi = philox.predict_normals(99, n, 5, 3)
    allz = np.concatenate([z, zi, zv], axis=1)
    np.testing.assert_allclose(allz.mean(axis=0), 0.0, atol=0.01)
    np.testing.assert_allclose(allz.std(axis=0), 1.0, atol=0.01)
    C = np.corrcoef(allz.T)
    assert np.max(np.abs(C - np.eye(C.shape[0]))) < 0.01
    for q in range(4):  # slot q of each group vs slot q + 1, and group g vs g + 1
        a = z[q::4, 0]
        b = z[(q + 1) % 4::4, 0][: a.size]
        assert abs(np.corrcoef(a, b)[0, 1]) < 0.01
    assert abs(np.corrcoef(z[:-4, 1], z[4:, 1])[0, 1]) < 0.01


def test_uniform_ranges_and_exactness():
    x = np.array([0, 1, 255, 256, 0xFFFFFFFF], dtype=np.uint64)
    u = philox.u01f(x)
    assert u.dtype == np.float32
    assert np.all(u > 0) and np.all(u < 1)
    assert u[0] == np.float32(2.0 ** -24) and u[-1] == np.float32(1 - 2.0 ** -24)
    d = philox.u01d(np.array([0, 0xFFFFFFFF], dtype=np.uint64),
                    np.array([0, 0xFFFFFFFF], dtype=np.uint64))
    assert d[0] == 2.0 ** -53 and d[1] == 1.0


def test_normals_are_standard():
    z = philox.normal4(7, np.arange(200_000), 3, philox.PURPOSE_PREDICT, 0).ravel()
    assert abs(z.mean()) < 0.01
    assert abs(z.std() - 1.0) < 0.01


# ------------------------------------------------------------- exp twins ----
def test_exp64_against_libm():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-745, 709, 20000), rng.uniform(-5, 5, 20000),
                        np.array([0.0, -0.0, 1.0, -1.0, 709.7, -708.0, -744.0])])
    got = detmath.exp64(x)
    ref = np.array([math.exp(v) for v in x])
    normal = ref >= np.finfo(np.float64).tiny
    ulp = np.abs(got[normal] - ref[normal]) / np.spacing(ref[normal])
    assert ulp.max() <= 2.0
    assert np.all(np.abs(got[~normal] - ref[~normal]) <= 2 * 5e-324)
    assert detmath.exp64(np.array([-800.0]))[0] == 0.0
    assert np.isinf(detmath.exp64(np.array([800.0]))[0])
    assert detmath.exp64(np.array([0.0]))[0] == 1.0


def _round_f32_exact(q):
    """Correctly rounded binary32 of a rational (ties to even), by comparing
    the neighbours of a candidate exactly."""
    from fractions import Fraction

    f = np.float32(float(q))
    best = None
    for cand in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        key = (abs(Fraction(float(cand)) - q), int(np.array(cand).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, cand)
    return best[1]


def test_fmaf_against_rational_arithmetic():
    """detmath.fmaf (the oracle's fused multiply-add) equals the correctly
    rounded binary32 of a*b + c computed exactly, on random triples over many
    magnitudes and on constructed ties and near-ties."""
    from fractions import Fraction

    rng = np.random.default_rng(11)
    n = 6000
    a = (rng.normal(size=n) * rng.choice([1e-3, 1.0, 1e3], n)).astype(np.float32)
    b = rng.normal(size=n).astype(np.float32)
    c = (rng.normal(size=n) * rng.choice([1e-8, 1.0, 1e8], n)).astype(np.float32)
    # ties: (1 + 2^-12)^2 = 1 + 2^-11 + 2^-24 sits halfway between two binary32
    # values; c = +-2^-60 pushes it just off the tie either way
    t = np.float32(1.0 + 2.0 ** -12)
    a = np.concatenate([a, [t, t, t, -t, np.float32(1.0 + 3 * 2.0 ** -12)]]).astype(np.float32)
    b = np.concatenate([b, [t, t, t, t, t]]).astype(np.float32)
    c = np.concatenate([c, [0.0, 2.0 ** -60, -(2.0 ** -60), 0.0, 0.0]]).astype(np.float32)
    got = detmath.fmaf(a, b, c)
    for i in range(a.size):
        want = _round_f32_exact(Fraction(float(a[i])) * Fraction(float(b[i])) +
                                Fraction(float(c[i])))
        assert np.array(got[i]).view(np.uint32) == np.array(want).view(np.uint32), i


def test_exp32_against_float64_exp():
    x = np.linspace(-87.0, 88.0, 50001, dtype=np.float32)
    got = detmath.exp32(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    rel = np.abs(got - ref) / ref
    assert rel.max() < 3.0 * 2.0 ** -24
    assert detmath.exp32(np.array([0.0], dtype=np.float32))[0] == 1.0


# ----------------------------------------------------------- block scans ----
def test_block_scan_exact_on_integers():
    """On small integers every association order is exact: the fixed-order
    scan must equal the sequential prefix sums."""
    v = np.random.default_rng(2).integers(0, 1000, (3, 256)).astype(np.float64)
    incl, excl, tot = detmath.block_scan256(v)
    np.testing.assert_array_equal(incl, np.cumsum(v, axis=1))
    np.testing.assert_array_equal(excl[:, 1:], np.cumsum(v, axis=1)[:, :-1])
    np.testing.assert_array_equal(tot, v.sum(axis=1))


def test_block_scan_close_to_exact_rational():
    v = np.random.default_rng(3).random(256)
    incl, _, tot = detmath.block_scan256(v)
    exact = [Fraction(0)]
    for x in v:
        exact.append(exact[-1] + Fraction(x))
    err = max(abs(float(Fraction(incl[i]) - exact[i + 1])) for i in range(256))
    assert err < 256 * 2.0 ** -52 * 256
    assert abs(float(Fraction(tot) - exact[-1])) < 1e-12


# ---------------------------------------------------- systematic resample ----
def test_tile_scale_and_shift_exact():
    """tile_scale: c 2^k = exp(m) to binary64 accuracy with c in [1, 2) (libm's
    exp as the reference); tile_shift: exactly v 2^dk (rational arithmetic) for
    dk >= -1000 and 0 below; decide: g_b / exp(m_b - M) is one constant over
    the tiles, i.e. the tiles keep their exact relative weights."""
    rng = np.random.default_rng(4)
    m = np.concatenate([rng.uniform(-700, 700, 2000), [0.0, -0.5, 1e-300, -1e-9]])
    k, c = detmath.tile_scale(m)
    assert np.all((c >= 1.0) & (c < 2.0 + 1e-15))
    np.testing.assert_array_equal(k, np.floor(k))
    got = c * np.exp2(np.minimum(k, 1000.0)) * np.exp2(np.maximum(k - 1000.0, 0.0))
    ref = np.array([math.exp(x) for x in m])
    np.testing.assert_allclose(got, ref, rtol=3e-16, atol=0)
    k2, c2 = detmath.tile_scale(np.array([-np.inf, np.inf]))
    assert list(k2) == [-1073741824.0, 0.0] and list(c2) == [0.0, 1.0]
    v = rng.uniform(1.0, 2048.0, 500)
    dk = rng.integers(-1000, 1, 500).astype(np.float64)
    out = detmath.tile_shift(v, dk)
    for a, b, o in zip(v[:100], dk[:100], out[:100]):
        assert Fraction(o) == Fraction(a) * Fraction(2) ** int(b)
    assert np.all(detmath.tile_shift(v, dk - 1001.0) == 0.0)
    mb = rng.uniform(-80.0, 0.0, 300).astype(np.float32).astype(np.float64)
    M, g, P, W = detmath.decide(mb, np.ones(300))
    ratio = g / np.exp(mb - M)
    np.testing.assert_allclose(ratio, ratio[0], rtol=4e-15)


def test_reference_u0_int32_product():
    """pf_device.cu:1852 multiplies 1000000 * N in int: N=10000 wraps."""
    assert detmath.reference_u0(42, 100) == 42 / 1e8
    assert detmath.reference_u0(42, 10000) == 42 / 1410065408.0
    assert detmath.reference_u0(42, 1_000_000) == 42 / -727379968.0


def _brute_systematic(w, u0):
    """Exact rational CDF, first index with cdf >= t, clamped to N-1."""
    n = len(w)
    tot = sum(Fraction(x) for x in w)
    cdf, run = [], Fraction(0)
    for x in w:
        run += Fraction(x)
        cdf.append(run / tot)
    import bisect

    out = []
    for j in range(n):
        t = (Fraction(u0) + j) / n
        out.append(min(bisect.bisect_left(cdf, t), n - 1))
    return np.array(out)


@pytest.mark.parametrize("n", [1, 2, 7, 64, 300, 1024, 1025, 1500, 4099])
def test_systematic_matches_brute_force(n):
    rng = np.random.default_rng(n)
    lw = rng.normal(0.0, 2.0, n).astype(np.float32)
    u0 = 0.37
    got = detmath.systematic_ancestors(lw, n, u0)
    e = np.exp(lw.astype(np.float64) - lw.max())
    ref = _brute_systematic(e, u0)
    # identical except where a target lands within rounding of a CDF value
    assert np.sum(got != ref) <= max(1, n // 1000)


def test_systematic_invariants_large():
    n = 20_000
    lw = np.random.default_rng(5).normal(0, 3, n).astype(np.float32)
    u0 = 0.5
    a = detmath.systematic_ancestors(lw, n, u0)
    assert a.dtype == np.int32
    assert np.all(np.diff(a) >= 0) and a.min() >= 0 and a.max() < n
    w = np.exp(lw.astype(np.float64) - lw.max())
    w /= w.sum()
    counts = np.bincount(a, minlength=n)
    assert np.all(counts >= np.floor(n * w) - 1) and np.all(counts <= np.ceil(n * w) + 1)
    assert np.mean((counts == np.floor(n * w)) | (counts == np.ceil(n * w))) > 0.999
    ref = detmath.sequential_systematic(w, u0)
    assert np.mean(ref == a) > 0.999


def test_systematic_flip_count_at_1M_against_binary64_reference():
    """How far the library's resampler (binary32 log-weights, fixed-order
    binary64 CDF) is from the reference's semantics on binary64 weights
    (sequential cumulative sum, first index with cdf >= (u0 + j) / N) at the
    full-size workload: 1,000,000 particles scattered 10 m around the truth of
    the 12-satellite multi-GNSS scenario, Cauchy log-weights.  Two counts are
    reported: outputs whose ancestor differs because of the binary32 rounding of
    the log-weights plus the summation order, and because of the summation
    order alone (the same binary32 log-weights in both).  Every flipped output
    has its target within 5e-9 of a binary64 CDF step of the reference (the
    binary32 rounding of the log-weights moves the CDF by a few 1e-10)."""
    from gnss_gpu_amd import sim

    n = 1_000_000
    sc = sim.scenario_multi_gnss(n_epochs=2, seed=0)
    rng = np.random.default_rng(11)
    pos = sc.truth[1][None, :] + rng.normal(0.0, 10.0, (n, 3))
    cb = sc.cb + rng.normal(0.0, 10.0, n)
    isb = sc.isb[None, :] + rng.normal(0.0, 1.0, (n, 3))
    lw64 = opf.log_likelihood(pos, cb, isb, sc.sat, sc.pr[1], np.ones(sc.n_sat), sc.sys, 5.0,
                              nu=1.0)
    lw32 = lw64.astype(np.float32)
    u0 = detmath.reference_u0(12345, n)
    got = detmath.systematic_ancestors(lw32, n, u0)
    w64 = np.exp(lw64 - lw64.max())
    ref64 = detmath.sequential_systematic(w64 / w64.sum(), u0)
    w32 = np.exp(lw32.astype(np.float64) - float(lw32.max()))
    ref32 = detmath.sequential_systematic(w32 / w32.sum(), u0)
    flips64 = int(np.sum(got != ref64))
    flips32 = int(np.sum(got != ref32))
    print(f"\n[systematic flips at N=1M] binary32 log-weights vs binary64 reference: {flips64}; "
          f"summation order alone: {flips32}", flush=True)
    cdf = np.cumsum(w64 / w64.sum())
    j = np.nonzero(got != ref64)[0]
    t = detmath.systematic_targets(n, u0)[j]
    r = ref64[j].astype(np.int64)
    near = np.minimum(np.abs(cdf[r] - t), np.abs(cdf[np.maximum(r - 1, 0)] - t))
    assert np.all(near < 5e-9), float(near.max())
    assert flips64 <= n // 1000 and flips32 <= flips64


def test_systematic_uniform_weights_identity():
    n = 5000
    a = detmath.systematic_ancestors(np.zeros(n, np.float32), n, detmath.reference_u0(42, n))
    np.testing.assert_array_equal(a, np.arange(n))


def test_systematic_degenerate_single_particle_weight():
    n = 3000
    lw = np.full(n, -1e4, np.float32)
    lw[1234] = 0.0
    a = detmath.systematic_ancestors(lw, n, 0.1)
    assert np.all(a == 1234)


# ------------------------------------------------------------- Megopolis ----
def test_megopolis_ref_zero_iterations_identity():
    lw = np.random.default_rng(6).normal(size=100).astype(np.float32)
    np.testing.assert_array_equal(detmath.megopolis_ref_ancestors(lw, 100, 0, 3),
                                  np.arange(100))


def test_megopolis_ref_uniform_weights_always_accept():
    """alpha = 1 for equal weights: each iteration moves to the proposal."""
    n = 97
    a = detmath.megopolis_ref_ancestors(np.zeros(n, np.float32), n, 1, 11)
    r = philox.draw(11, np.arange(n), 0, philox.PURPOSE_MEGO_REF, 0)
    off = (philox.u01d(r[0], r[1]) * (n - 1)).astype(np.int64) + 1
    np.testing.assert_array_equal(a, (np.arange(n) + off) % n)


def test_megopolis_concentrates_on_heavy_particles():
    """Metropolis chains on the original weights converge to the posterior."""
    n = 4096
    lw = np.full(n, -50.0, np.float32)
    heavy = np.random.default_rng(12).choice(n, n // 4, replace=False)
    lw[heavy] = 0.0
    a = detmath.megopolis_ancestors(lw, n, 60, 5, 0)
    assert np.mean(np.isin(a, heavy)) > 0.999


def test_megopolis_ref_slot_weight_rule():
    """The reference kernel compares the weights of the *slots* (pf_device.cu:
    1265-1270), not of their current contents: a heavy slot only takes content
    from heavy slots, so its content stays heavy; light slots accept anything."""
    n = 4096
    lw = np.full(n, -50.0, np.float32)
    heavy = np.arange(0, n, 64)
    lw[heavy] = 0.0
    a = detmath.megopolis_ref_ancestors(lw, n, 30, 5)
    assert np.all(np.isin(a[heavy], heavy))
    assert np.mean(np.isin(a, heavy)) > len(heavy) / n


def test_megopolis_acceptance_rule():
    """One Megopolis iteration: a chain on a heavy particle moves to a proposal
    four times lighter with probability 1/4, and a light chain always moves
    to a heavy proposal (Metropolis rule min(1, w_j / w_k))."""
    n = 1 << 16
    rng = np.random.default_rng(31)
    heavy = rng.random(n) < 0.5
    lw = np.where(heavy, 0.0, np.log(0.25)).astype(np.float32)
    w = detmath.megopolis_weights(lw, n)
    assert w.max() == 1.0
    np.testing.assert_allclose(w[~heavy], 0.25, rtol=1e-6)
    a = detmath.megopolis_ancestors(lw, n, 1, 17, 3)
    i = np.arange(n)
    r = philox.draw(17, i // 256, 3, philox.PURPOSE_MEGO, 0)
    prop = (i + ((r[0] * np.uint64(n)) >> np.uint64(32)).astype(np.int64)) % n
    assert np.all((a == i) | (a == prop))
    hl = heavy & ~heavy[prop]
    lh = ~heavy & heavy[prop]
    moved = a != i
    frac = moved[hl].mean()
    assert abs(frac - 0.25) < 4 * np.sqrt(0.25 * 0.75 / hl.sum())
    assert np.all(moved[lh])
    assert np.all(moved[heavy == heavy[prop]])  # equal weights: always accepted


def test_megopolis_chesser_matches_target_distribution():
    n = 8192
    rng = np.random.default_rng(8)
    lw = rng.normal(0, 1, n).astype(np.float32)
    a = detmath.megopolis_ancestors(lw, n, 64, 9, 0)
    w = np.exp(lw.astype(np.float64))
    w /= w.sum()
    # weighted mean of a smooth test function is preserved in expectation
    f = np.sin(np.arange(n) / 50.0)
    assert abs(f[a].mean() - (w * f).sum()) < 0.05


# -------------------------------------------------------- PF closed forms ----
def _stream_scenario(n_sat=5):
    rng = np.random.RandomState(123)
    d = rng.randn(n_sat, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    sat = d * 20_000_000.0
    pr = np.linalg.norm(sat, axis=1) + 100.0
    return sat, pr


def test_closed_form_gaussian_and_huber():
    g = GOLD["stream_closed_forms"]
    sat, pr = _stream_scenario()
    pr = pr.copy()
    pr[0] += 100.0
    pos, cb, isb = np.zeros((1, 3)), np.array([100.0]), np.zeros((1, 0))
    w = np.ones(5)
    lg = opf.log_likelihood(pos, cb, isb, sat, pr, w, None, 5.0)
    lh = opf.log_likelihood(pos, cb, isb, sat, pr, w, None, 5.0, huber=True, huber_k=1.5)
    assert lg[0] == pytest.approx(g["gaussian_log_weight"])
    assert lh[0] == pytest.approx(g["huber_log_weight"])


def test_gate_semantics():
    g = GOLD["stream_closed_forms"]
    sat, pr = _stream_scenario()
    pos, cb, isb = np.zeros((1, 3)), np.array([100.0]), np.zeros((1, 0))
    out = pr.copy()
    out[0] += 100.0
    l1 = opf.log_likelihood(pos, cb, isb, sat, out, np.ones(5), None, 5.0,
                            gate_thr=g["gate_threshold_m"])
    assert abs(l1[0]) < 1e-6
    l2 = opf.log_likelihood(pos, cb, isb, sat, pr + 100.0, np.ones(5), None, 5.0,
                            gate_thr=g["gate_threshold_m"])
    assert l2[0] < g["gate_reject_all_max_log_weight"]


def _isb_case():
    """Noise-free pseudoranges of 2 + 3 + 4 + 5 satellites of four systems
    (ids 0..3) with clock bias 100 m and inter-system biases 12 / -7.5 / 20 m."""
    from gnss_gpu_amd import sim

    rng = np.random.default_rng(17)
    counts = (2, 3, 4, 5)
    sys_ids = np.repeat(np.arange(4), counts).astype(np.int32)
    truth = sim.ANCHOR_CFG1.copy()
    sat = sim.place_satellites(truth, rng.uniform(15, 85, 14), rng.uniform(0, 360, 14))
    cb, isb = 100.0, np.array([12.0, -7.5, 20.0])
    pr = np.linalg.norm(sat - truth, axis=1) + cb + np.concatenate([[0.0], isb])[sys_ids]
    return truth, sat, pr, sys_ids, cb, isb, counts


def test_isb_closed_forms():
    """System k > 0 of a satellite uses ISB state k - 1, system 0 none: at the
    truth the Gaussian log-weight is 0; ISB j off by delta gives
    -n_{j+1} delta^2 / (2 sigma^2) with n_k the satellite count of system k
    (distinct counts, so a shifted index changes the value)."""
    truth, sat, pr, sys_ids, cb, isb, counts = _isb_case()
    w = np.ones(len(pr))
    lw0 = opf.log_likelihood(truth[None], np.array([cb]), isb[None], sat, pr, w, sys_ids, 5.0)
    assert abs(lw0[0]) < 1e-12
    for j, delta in enumerate((1.5, -2.0, 3.0)):
        off = isb.copy()
        off[j] += delta
        lw = opf.log_likelihood(truth[None], np.array([cb]), off[None], sat, pr, w, sys_ids, 5.0)
        assert lw[0] == pytest.approx(-counts[j + 1] * delta * delta / 50.0, rel=1e-9, abs=1e-12)
    # the clock bias moves every satellite
    lw = opf.log_likelihood(truth[None], np.array([cb + 2.0]), isb[None], sat, pr, w, sys_ids, 5.0)
    assert lw[0] == pytest.approx(-14 * 4.0 / 50.0, rel=1e-9)


def test_student_t_cauchy_brute_force():
    """nu = 1 (Cauchy): -log(1 + (r/s)^2) per satellite (pf_device.cu:264-266)."""
    sat, pr = _stream_scenario(4)
    pos, cb, isb = np.zeros((1, 3)), np.array([100.0]), np.zeros((1, 0))
    pr = pr + np.array([1.0, -2.0, 30.0, 0.5])
    got = opf.log_likelihood(pos, cb, isb, sat, pr, np.ones(4), None, 5.0, nu=1.0)
    r = np.array([1.0, -2.0, 30.0, 0.5])
    assert got[0] == pytest.approx(-np.sum(np.log1p((r / 5.0) ** 2)), rel=1e-12)


def test_ess_and_estimate_closed_forms():
    n = 1000
    assert opf.ess(np.zeros(n)) == pytest.approx(n)
    lw = np.full(n, -100.0)
    lw[3] = 0.0
    assert opf.ess(lw) < 2.0
    st = np.random.default_rng(9).normal(size=(n, 4))
    np.testing.assert_allclose(opf.weighted_mean(st, np.zeros(n)), st.mean(axis=0))
    w = np.random.default_rng(10).random(n)
    np.testing.assert_allclose(opf.weighted_mean(st, np.log(w)), (w[:, None] * st).sum(0) / w.sum())


# -------------------------------------------------------------------- WLS ----
def test_wls_position_no_sagnac_truth():
    g = GOLD["tokyo_constellation"]
    sat = np.array(g["sat_ecef"])
    tp = np.array(g["true_pos"])
    pr = np.linalg.norm(sat - tp, axis=1) + g["true_cb"]
    res, it = owls.wls_position_ref(sat, pr, np.ones(8))
    assert np.linalg.norm(res[:3] - tp) < GOLD["wls_bounds"]["single_epoch_error_m"]
    assert abs(res[3] - g["true_cb"]) < 0.01
    assert 1 <= it <= 10


def test_wls_batch_sagnac_truth_and_noise_bound():
    from gnss_gpu_amd.range_model import geometric_ranges_sagnac

    g = GOLD["tokyo_constellation"]
    sat = np.array(g["sat_ecef"])
    tp = np.array(g["true_pos"])
    pr = geometric_ranges_sagnac(tp, sat) + g["true_cb"]
    res, it = owls.wls_batch_ref(sat[None], pr[None], np.ones((1, 8)))
    assert np.linalg.norm(res[0, :3] - tp) < 1e-3
    rng = np.random.default_rng(42)
    E = 50
    prn = np.tile(pr, (E, 1)) + rng.normal(0, 3.0, (E, 8))
    res, it = owls.wls_batch_ref(np.tile(sat, (E, 1, 1)), prn, np.ones((E, 8)))
    assert np.all(np.linalg.norm(res[:, :3] - tp, axis=1) < GOLD["wls_bounds"]["batch_error_m_with_3m_noise"])


def test_wls_fewer_than_four_and_singular():
    sat = np.array(GOLD["tokyo_constellation"]["sat_ecef"])[:3]
    res, it = owls.wls_position_ref(sat, np.ones(3) * 2e7, np.ones(3))
    assert it == -1 and np.all(res == 0)
    # all satellites at one point: singular normal matrix -> early return
    s1 = np.tile(np.array([[1.5e7, 1.0e7, 2.0e7]]), (5, 1))
    res, it = owls.wls_position_ref(s1, np.full(5, 2.2e7), np.ones(5))
    assert it == 0 and np.all(np.isfinite(res))


def test_sagnac_rotation_closed_form():
    """Rotation by theta about z preserves |sat| and the z component, and theta
    equals omega_E * range / c (range_model.py:20-29)."""
    from gnss_gpu_amd.range_model import OMEGA_E, C_LIGHT, rotate_satellites_sagnac

    rx = np.array([-3957199.0, 3310205.0, 3737911.0])
    sat = np.array(GOLD["tokyo_constellation"]["sat_ecef"])
    rot = rotate_satellites_sagnac(rx, sat)
    np.testing.assert_allclose(np.linalg.norm(rot, axis=1), np.linalg.norm(sat, axis=1), rtol=1e-15)
    np.testing.assert_array_equal(rot[:, 2], sat[:, 2])
    th = OMEGA_E * np.linalg.norm(rx - sat, axis=1) / C_LIGHT
    ang = np.arctan2(sat[:, 1], sat[:, 0]) - np.arctan2(rot[:, 1], rot[:, 0])
    np.testing.assert_allclose(ang, th, rtol=1e-9)


def test_position_spread_two_point_cloud():
    """Two particles at +-d along x with weights w, 1-w around the origin: the
    weighted RMS distance is d whatever w; around (d, 0, 0) it is 2d*sqrt(1-w)."""
    d, w = 3.0, 0.25
    pos = np.array([[d, 0.0, 0.0], [-d, 0.0, 0.0]])
    lw = np.log([w, 1.0 - w])
    assert opf.position_spread(pos, lw, [0.0, 0.0, 0.0]) == pytest.approx(d, rel=1e-14)
    assert opf.position_spread(pos, lw, [d, 0.0, 0.0]) == pytest.approx(2 * d * math.sqrt(1 - w), rel=1e-14)


def test_solve4_against_numpy_and_pivot_rule():
    rng = np.random.default_rng(3)
    B = rng.normal(size=(64, 4, 4))
    A = np.einsum("eij,ekj->eik", B, B) + 0.1 * np.eye(4)
    b = rng.normal(size=(64, 4))
    x, ok = owls._solve4(np.concatenate([A, b[:, :, None]], axis=2), cpu_rule=False)
    np.testing.assert_allclose(x, np.linalg.solve(A, b[:, :, None])[:, :, 0], rtol=1e-9, atol=1e-12)
    assert ok.all()
    # a zero column: the CPU rule flags the system, the batch rule skips the pivot
    A0 = A[:1].copy()
    A0[0, :, 2] = 0.0
    _, ok0 = owls._solve4(np.concatenate([A0, b[:1, :, None]], axis=2), cpu_rule=True)
    assert not ok0[0]


# ------------------------------------------------------------ multi-GNSS ----
def _mg_scenario(isb_gal=0.0, isb_glo=0.0, sagnac=True):
    """test_multi_gnss.py:47-92: GPS / Galileo / GLONASS, four satellites each."""
    from gnss_gpu_amd.range_model import geometric_ranges_sagnac

    g = GOLD["multi_gnss_scenario"]
    sat = np.vstack([g["gps_sats"], g["galileo_sats"], g["glonass_sats"]])
    tp = np.array(g["true_pos"])
    sid = np.array([0] * 4 + [2] * 4 + [1] * 4)
    rng = geometric_ranges_sagnac(tp, sat) if sagnac else np.linalg.norm(sat - tp, axis=1)
    pr = rng + g["true_cb_gps"] + np.repeat([0.0, isb_gal, isb_glo], 4)
    return sat, pr, sid, tp


def test_multi_gnss_batch_recovers_reference_isbs():
    """test_multi_gnss.py:109-151: noise-free Sagnac ranges; the batch kernel's
    model recovers the position and every ISB (systems numbered in sorted order)."""
    g = GOLD["multi_gnss_scenario"]
    for case in g["isb_cases"]:
        sat, pr, sid, tp = _mg_scenario(case["isb_galileo"], case["isb_glonass"])
        n = case["n_sat"]
        systems = case["systems"]
        mapped = np.array([systems.index(s) for s in sid[:n]])
        res, it = owls.wls_multi_gnss_batch_ref(sat[None, :n], pr[None, :n], np.ones((1, n)),
                                                mapped[None], len(systems))
        assert np.linalg.norm(res[0, :3] - tp) < g["position_error_m"]
        cb = dict(zip(systems, res[0, 3:]))
        assert abs(cb[2] - cb[0] - case["isb_galileo"]) < g["isb_error_m"]
        if 1 in cb:
            assert abs(cb[1] - cb[0] - case["isb_glonass"]) < g["isb_error_m"]
        assert abs(cb[0] - g["true_cb_gps"]) < 1e-4
        assert 1 <= it[0] <= 10
    # GPS only, eight satellites (test_multi_gnss.py:23-44, 100-107)
    sat = np.array(GOLD["tokyo_constellation"]["sat_ecef"])
    from gnss_gpu_amd.range_model import geometric_ranges_sagnac
    pr = geometric_ranges_sagnac(np.array(g["true_pos"]), sat) + 3000.0
    res, it = owls.wls_multi_gnss_batch_ref(sat[None], pr[None], np.ones((1, 8)),
                                            np.zeros((1, 8), int), 1)
    assert np.linalg.norm(res[0, :3] - g["true_pos"]) < g["gps_only_error_m"]


def test_multi_gnss_single_epoch_has_no_sagnac_term():
    """multi_gnss.cu:101-112: the single-epoch solver takes plain ranges.  Plain
    ranges are recovered to the reference's 0.1 m bound; Sagnac ranges are not
    (the Earth-rotation term moves ranges by up to ~30 m here)."""
    g = GOLD["multi_gnss_scenario"]
    sat, pr, sid, tp = _mg_scenario(15.0, -8.0, sagnac=False)
    mapped = np.array([[0, 1, 2].index(s) for s in sid])
    res, it = owls.wls_multi_gnss_ref(sat, pr, np.ones(12), mapped, 3)
    assert np.linalg.norm(res[:3] - tp) < 1e-6
    np.testing.assert_allclose(res[3:], [3000.0, 2992.0, 3015.0], atol=1e-6)
    sat, pr, sid, tp = _mg_scenario(15.0, -8.0, sagnac=True)
    res, it = owls.wls_multi_gnss_ref(sat, pr, np.ones(12), mapped, 3)
    assert np.linalg.norm(res[:3] - tp) > 1.0


def test_multi_gnss_guards_and_one_system_equals_wls():
    g = GOLD["multi_gnss_scenario"]
    ins = g["insufficient"]
    res, it = owls.wls_multi_gnss_ref(np.array(ins["sat_ecef"]), ins["pr"], np.ones(2),
                                      [0, 1], 2)
    assert it == -1 and res.shape == (5,) and np.all(res == 0)
    res, it = owls.wls_multi_gnss_ref(np.zeros((6, 3)), np.full(6, 2e7), np.ones(6),
                                      np.zeros(6, int), 1)
    assert it == -1 and np.all(res == 0)
    # one system with every satellite in it is the pinned 4-state WLS
    from gnss_gpu_amd import sim

    sat, pr, w, ns, truth = sim.wls_epochs(n_epoch=40, s_min=8, s_max=8, seed=11)
    a, ia = owls.wls_multi_gnss_batch_ref(sat, pr, w, np.zeros((40, 8), int), 1)
    b, ib = owls.wls_batch_ref(sat, pr, w)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-7)
    np.testing.assert_array_equal(ia, ib)
    a, ia = owls.wls_multi_gnss_ref(sat[0], pr[0], w[0], np.zeros(8, int), 1)
    b, ib = owls.wls_position_ref(sat[0], pr[0], w[0])
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-7)


def test_multi_gnss_noisy_optimum_matches_least_squares():
    """With noise the Gauss-Newton fixed point is the weighted least-squares
    optimum: scipy's trust-region solver on the same Sagnac model and weights
    lands on it, and the gradient H^T W r vanishes there."""
    from scipy.optimize import least_squares

    from gnss_gpu_amd.range_model import geometric_ranges_sagnac

    sat, pr, sid, tp = _mg_scenario(15.0, -8.0)
    mapped = np.array([[0, 1, 2].index(s) for s in sid])
    rng = np.random.default_rng(5)
    E = 6
    prn = pr[None, :] + rng.normal(0, 2.0, (E, 12))
    w = rng.uniform(0.5, 2.0, (E, 12))
    res, it = owls.wls_multi_gnss_batch_ref(np.tile(sat, (E, 1, 1)), prn, w,
                                            np.tile(mapped, (E, 1)), 3, max_iter=20, tol=1e-9)
    for e in range(E):
        def f(s):
            return np.sqrt(w[e]) * (prn[e] - geometric_ranges_sagnac(s[:3], sat) - s[3:][mapped])
        x0 = np.concatenate([tp + 50.0, [2900.0, 2900.0, 2900.0]])
        ls = least_squares(f, x0, xtol=1e-15, ftol=1e-15, gtol=1e-15)
        np.testing.assert_allclose(res[e], ls.x, rtol=0, atol=1e-5)
        assert np.linalg.norm(res[e, :3] - tp) < GOLD["multi_gnss_scenario"]["batch_error_m"]


def test_multi_gnss_empty_system_and_foreign_id():
    """A system with no satellites leaves a zero column: its clock stays at 0
    and the others are unaffected; a system id outside [0, n_systems) uses
    clock 0's bias without a clock column (multi_gnss.cu:111, :120)."""
    sat, pr, sid, tp = _mg_scenario(15.0, 0.0)
    mapped = np.array([{0: 0, 2: 2, 1: 0}[s] for s in sid])  # GLONASS folded into GPS
    full, it = owls.wls_multi_gnss_batch_ref(sat[None], pr[None], np.ones((1, 12)),
                                             mapped[None], 3)
    assert full[0, 4] == 0.0
    assert np.linalg.norm(full[0, :3] - tp) < 1e-6
    np.testing.assert_allclose(full[0, [3, 5]], [3000.0, 3015.0], atol=1e-6)
    # the truth is still the fixed point, but the row's missing clock column
    # makes the Jacobian inexact: convergence turns linear (~48 iterations
    # instead of the 5 of the labelled case)
    foreign = mapped.copy()
    foreign[0] = 7
    res, it = owls.wls_multi_gnss_batch_ref(sat[None], pr[None], np.ones((1, 12)),
                                            foreign[None], 3, max_iter=100)
    assert np.linalg.norm(res[0, :3] - tp) < 1e-3
    assert 20 < it[0] < 100


def test_solve_n_against_numpy_and_zero_rows():
    rng = np.random.default_rng(8)
    for n in range(3, 9):
        B = rng.normal(size=(32, n, n))
        A = np.einsum("eij,ekj->eik", B, B) + 0.1 * np.eye(n)
        b = rng.normal(size=(32, n))
        x = owls._solve_n(np.concatenate([A, b[:, :, None]], axis=2))
        np.testing.assert_allclose(x, np.linalg.solve(A, b[:, :, None])[:, :, 0], rtol=1e-8,
                                   atol=1e-10)
    # zero row and column k: unknown k is 0, the rest solve the reduced system
    n, k = 6, 4
    B = rng.normal(size=(1, n, n))
    A = np.einsum("eij,ekj->eik", B, B) + 0.1 * np.eye(n)
    A[:, k, :] = 0.0
    A[:, :, k] = 0.0
    b = rng.normal(size=(1, n))
    b[:, k] = 0.0
    x = owls._solve_n(np.concatenate([A, b[:, :, None]], axis=2))
    keep = [i for i in range(n) if i != k]
    red = np.linalg.solve(A[0][np.ix_(keep, keep)], b[0, keep])
    assert x[0, k] == 0.0
    np.testing.assert_allclose(x[0, keep], red, rtol=1e-9)


# ------------------------------------------------------------ ray casting ----
def _box(center=(100.0, 0.0, 25.0), w=20.0, d=20.0, h=50.0):
    from gnss_gpu_amd.raytrace import BuildingModel

    return BuildingModel.create_box(list(center), w, d, h).triangles


def test_mt_hit_against_plane_and_barycentric_solve():
    """Moller-Trumbore against an independent form: the plane parameter
    t = n.(v0 - o) / n.d and barycentric coordinates from a 3x3 solve of
    o + t d = v0 + u e1 + v e2; agreement away from edges (margin 1e-9)."""
    from oracle import raytrace as oray

    rng = np.random.default_rng(21)
    tris = rng.normal(0, 10, (60, 3, 3))
    o = rng.normal(0, 30, (200, 3))
    d = rng.normal(0, 1, (200, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    hit, t = oray.mt_hit(o, d, tris)
    checked = 0
    for r in range(200):
        for k in range(60):
            v0, e1, e2 = tris[k, 0], tris[k, 1] - tris[k, 0], tris[k, 2] - tris[k, 0]
            M = np.stack([-d[r], e1, e2], axis=1)
            tt, u, v = np.linalg.solve(M, o[r] - v0)
            margin = min(u, v, 1 - u - v)
            if abs(margin) < 1e-9:
                continue
            assert hit[r, k] == (margin > 0), (r, k)
            if hit[r, k]:
                assert t[r, k] == pytest.approx(tt, rel=1e-9, abs=1e-9)
                checked += 1
    assert checked > 50


def test_los_box_cases_and_empty_mesh():
    """test_bvh.py:26-39: satellite overhead LOS, behind the box NLOS; an
    empty mesh blocks nothing (test_pf3d_bvh.py:196-214); NaN padding in the
    batch query is not LOS (test_bvh.py:178-191)."""
    from oracle import raytrace as oray

    tris = _box()
    los = oray.los_check([0.0, 0.0, 0.0], [[0.0, 0.0, 2e7], [200.0, 0.0, 25.0]], tris)
    assert los.tolist() == [[True, False]]
    assert oray.los_check([0.0, 0.0, 0.0], [[200.0, 0.0, 25.0]], np.zeros((0, 3, 3))).all()
    rx = np.array([[0.0, 0.0, 0.0], [10.0, 0.0, 0.0]])
    sat = np.array([[[0.0, 0.0, 2e7], [200.0, 0.0, 25.0]], [[0.0, 0.0, 2e7], [np.nan] * 3]])
    assert oray.los_check(rx, sat, tris, batch=True).tolist() == [[True, False], [True, False]]


def test_weight_3d_reference_closed_forms():
    """test_pf3d_bvh.py:94-194: zero LOS residual -> 0; NLOS residual equal to
    the bias -> 0; a negative NLOS residual keeps its value (-0.5 b^2 / s^2);
    the NLOS sigma penalises less than the LOS one; the 0 < p < 1 mixture is
    log((1-p) e^los + p e^nlos)."""
    from oracle import raytrace as oray

    tris = _box()
    z = np.zeros(8)
    cb = np.full(8, 100.0)
    up = np.array([[0.0, 0.0, 20000.0]])
    lw = oray.weight_3d(z, z, z, cb, up, [20000.0 + 100.0], [1.0], tris, 3.0, 30.0, 20.0)
    np.testing.assert_allclose(lw, 0.0, atol=1e-10)
    behind = np.array([[200.0, 0.0, 25.0]])
    dist = float(np.linalg.norm(behind))
    lw = oray.weight_3d(z, z, z, cb, behind, [dist + 120.0], [1.0], tris, 3.0, 30.0, 20.0)
    np.testing.assert_allclose(lw, 0.0, atol=1e-10)
    lw = oray.weight_3d(z, z, z, cb, behind, [dist + 80.0], [1.0], tris, 3.0, 30.0, 20.0)
    np.testing.assert_allclose(lw, -0.5 * 400.0 / 900.0, atol=1e-10)
    lw = oray.weight_3d(z, z, z, cb, behind, [dist + 150.0], [1.0], tris, 3.0, 30.0, 20.0)
    assert np.all(lw > -0.5 * 50.0 ** 2 / 9.0)
    p = 0.3
    lw = oray.weight_3d(z, z, z, cb, behind, [dist + 150.0], [2.0], tris, 3.0, 30.0, 20.0,
                        blocked_nlos_prob=p)
    los, nlos = -0.5 * 2.0 * 2500.0 / 9.0, -0.5 * 2.0 * 900.0 / 900.0
    np.testing.assert_allclose(lw, math.log((1 - p) * math.exp(los) + p * math.exp(nlos)),
                               rtol=1e-12)


def test_multipath_wall_closed_form():
    """A vertical wall x = 10 (two triangles): the receiver's mirror image is
    (20 - x, y, z); the reflection point is where the image-to-satellite line
    crosses the wall; the excess delay is |image - sat| - |rx - sat|."""
    from oracle import raytrace as oray

    wall = np.array([[[10.0, -100.0, -50.0], [10.0, 100.0, -50.0], [10.0, 100.0, 150.0]],
                     [[10.0, -100.0, -50.0], [10.0, 100.0, 150.0], [10.0, -100.0, 150.0]]])
    rx = np.array([0.0, 0.0, 1.5])
    sat = np.array([-2.0e7, 3.0e6, 1.5e7])
    refl, delay = oray.multipath(rx, sat[None], wall)
    img = np.array([20.0, 0.0, 1.5])
    t = (10.0 - img[0]) / (sat[0] - img[0])
    pt = img + t * (sat - img)
    np.testing.assert_allclose(refl[0, 0], pt, atol=1e-6)
    np.testing.assert_allclose(delay[0, 0], np.linalg.norm(img - sat) - np.linalg.norm(rx - sat),
                               rtol=1e-9)
    # a satellite on the receiver's side of the wall plane behind it has no reflection
    refl, delay = oray.multipath(rx, np.array([[2.0e7, 0.0, 1.0e7]]), wall)
    assert delay[0, 0] == 0.0 and np.all(refl == 0.0)


# ------------------------------------------------------------------ FFBSi ----
def test_choice_index_is_generator_choice():
    """numpy's Generator.choice(N, p) restated as one random() and a search of
    the normalised cumulative sum: same index for the same generator state."""
    from oracle import ffbsi as of

    for seed in range(40):
        w = np.random.default_rng(1000 + seed).random(257) ** 3
        a, b = np.random.default_rng(seed), np.random.default_rng(seed)
        assert of.choice_index(w, a.random()) == int(b.choice(257, p=w / w.sum()))


def test_transition_logpdf_is_gaussian_up_to_a_constant():
    from scipy.stats import norm

    from oracle import ffbsi as of

    rng = np.random.default_rng(2)
    xt = rng.normal(0, 5, (30, 4))
    xn = np.array([1.0, -2.0, 0.5, 40.0])
    v, dt, sp, sc = np.array([0.3, -0.1, 0.2]), 0.7, 1.3, 25.0
    lp = of.transition_logpdf(xn, xt, v, dt, sp, sc)
    ref = (norm.logpdf(xn[:3], xt[:, :3] + v * dt, sp).sum(axis=1) + norm.logpdf(xn[3], xt[:, 3], sc))
    np.testing.assert_allclose(lp - lp[0], ref - ref[0], rtol=1e-12, atol=1e-12)


def _history(T, N, seed):
    rng = np.random.default_rng(seed)
    X = np.cumsum(rng.normal(0, 1.0, (T, N, 4)), axis=0) * np.array([1.0, 1.0, 1.0, 3.0])
    lw = rng.normal(0, 2.0, (T, N))
    vel = rng.normal(0, 0.1, (T, 3))
    return lw, X, vel, np.full(T, 1.0), np.full(T, 1.5)


def test_ffbsi_backward_distribution_and_reference_shapes():
    """test_ffbsi.py:23-56 (bounds, genealogy trace); with T = 2 the backward
    draw of idx[0] follows w0_i p(x1[idx1] | x0_i) normalised (chi-square over
    20,000 trajectories); the generator is consumed one random() per step."""
    from scipy.stats import chisquare

    from oracle import ffbsi as of

    lw, X, vel, dt, sp = _history(5, 32, 0)
    I = of.ffbsi_indices(lw, X, vel, dt, sp, 10.0, np.random.default_rng(0))
    assert I.shape == (5,) and I.dtype == np.int64 and np.all((I >= 0) & (I < 32))
    rng = np.random.default_rng(1)
    anc = np.vstack([rng.integers(0, 16, (3, 16)), np.arange(16)[None]])
    lwg = rng.standard_normal((4, 16)) * 2.0
    J = of.genealogy_indices(lwg, anc, rng)
    for t in range(2, -1, -1):
        assert J[t] == anc[t, J[t + 1]]
    lw, X, vel, dt, sp = _history(2, 6, 3)
    X[1] = X[0] + rng.normal(0, 0.5, (6, 4))
    counts = np.zeros((6, 6))
    g = np.random.default_rng(7)
    for _ in range(20000):
        i = of.ffbsi_indices(lw, X, vel, dt, sp, 3.0, g)
        counts[i[1], i[0]] += 1
    for j in range(6):
        if counts[j].sum() < 500:
            continue
        logp = lw[0] + of.transition_logpdf(X[1, j], X[0], vel[1], dt[1], sp[1], 3.0)
        p = np.exp(logp - logp.max())
        p /= p.sum()
        keep = p * counts[j].sum() > 5
        exp_ = p[keep] * counts[j].sum()
        assert chisquare(counts[j][keep], exp_ * counts[j][keep].sum() / exp_.sum()).pvalue > 1e-4
    a, b = np.random.default_rng(11), np.random.default_rng(11)
    of.ffbsi_indices(lw, X, vel, dt, sp, 3.0, a)
    b.random(2)
    assert a.random() == b.random()


def test_ffbsi_degenerate_steps_consume_like_the_reference():
    """particle_ffbsi.py:73-75, 98-100: a degenerate step draws integers(0, N),
    a degenerate last step draws nothing and returns zeros."""
    from oracle import ffbsi as of

    lw, X, vel, dt, sp = _history(6, 20, 4)
    lw[2, 5] = np.nan
    a, b = np.random.default_rng(3), np.random.default_rng(3)
    I = of.ffbsi_indices(lw, X, vel, dt, sp, 10.0, a)
    b.random(3)
    assert I[2] == int(b.integers(0, 20))
    b.random(2)
    assert a.random() == b.random()
    lw[5] = -np.inf
    a, b = np.random.default_rng(3), np.random.default_rng(3)
    assert np.all(of.ffbsi_indices(lw, X, vel, dt, sp, 10.0, a) == 0)
    assert a.random() == b.random()


def test_oracle_predict_deterministic_closed_form():
    """sigma = 0: every particle moves by v*dt exactly (test_cuda_streams.py:239-255)."""
    pf = opf.OraclePF(100, seed=1)
    pf.initialize(np.array([1.0e6, 2.0e6, 3.0e6]), cb=5.0, spread_pos=4.0, spread_cb=2.0)
    before = pf.state.copy()
    pf.predict(np.array([1.5, -2.0, 0.25]), 2.0, 0.0, 0.0, step=1)
    np.testing.assert_allclose(pf.state[:, :3] - before[:, :3],
                               np.broadcast_to(np.array([3.0, -4.0, 0.5]), (100, 3)), atol=1e-9)
    np.testing.assert_array_equal(pf.state[:, 3], before[:, 3])


def test_oracle_predict_guide_blend_closed_form():
    """sigma = 0, alpha = 0.25: v <- 0.75 v + 0.25 v_guide exactly, then p += v dt
    (pf_device.cu:121-128)."""
    pf = opf.OraclePF(64, seed=3)
    v0 = np.array([2.0, -1.0, 0.5])
    pf.initialize(np.array([1.0e6, 2.0e6, 3.0e6]), cb=5.0, spread_pos=4.0, spread_cb=2.0,
                  velocity=v0)
    before = pf.state.copy()
    vg = np.array([-6.0, 3.0, 1.5])
    pf.predict(vg, 2.0, 0.0, 0.0, step=1, alpha=0.25)
    v = 0.75 * v0 + 0.25 * vg
    np.testing.assert_allclose(pf.vel, np.broadcast_to(v, (64, 3)), rtol=0, atol=1e-15)
    np.testing.assert_allclose(pf.state[:, :3] - before[:, :3], np.broadcast_to(2.0 * v, (64, 3)),
                               atol=1e-9)
    # alpha is clamped to [0, 1]: 1.7 acts as 1, -3 as 0
    pf.predict(vg, 1.0, 0.0, 0.0, step=2, alpha=1.7)
    np.testing.assert_allclose(pf.vel, np.broadcast_to(vg, (64, 3)), atol=1e-15)
    pf.predict(v0, 1.0, 0.0, 0.0, step=3, alpha=-3.0)
    np.testing.assert_allclose(pf.vel, np.broadcast_to(vg, (64, 3)), atol=1e-15)


def test_oracle_predict_velocity_noise_moments():
    """sigma_vel > 0 without the velocity filter: per-particle velocity noise
    N(0, sigma_vel^2 I) around the blend (pf_device.cu:129-133); positions move
    by the noisy velocity."""
    n = 200_000
    pf = opf.OraclePF(n, seed=9)
    pf.initialize(np.array([-3.96e6, 3.35e6, 3.70e6]), spread_pos=0.0, spread_cb=0.0)
    before = pf.state.copy()
    vg = np.array([1.0, -2.0, 0.5])
    pf.predict(vg, 0.5, 0.0, 0.0, step=1, sigma_vel=0.8)
    dv = pf.vel - vg
    np.testing.assert_allclose(dv.mean(axis=0), 0.0, atol=0.01)
    np.testing.assert_allclose(np.cov(dv.T), 0.64 * np.eye(3), atol=0.01)
    np.testing.assert_allclose(pf.state[:, :3] - before[:, :3], 0.5 * pf.vel, atol=1e-8)
    # with the velocity filter on, the velocity is not sampled (pf_device.cu:129)
    pf2 = opf.OraclePF(1000, seed=9)
    pf2.initialize(np.zeros(3), spread_pos=0.0, spread_cb=0.0, velocity=vg,
                   velocity_init_sigma=1.0)
    pf2.predict(vg, 1.0, 0.0, 0.0, step=1, sigma_vel=0.8, velocity_kf=True)
    np.testing.assert_array_equal(pf2.vel, np.broadcast_to(vg, (1000, 3)))


def test_oracle_rbpf_predict_reference_closed_forms():
    """The reference's RBPF predict pins (test_cuda_streams.py:257-283, 486-505):
    Sigma_v = 4 I, dt = 1, sigma_pos = 0, q = 0.5 -> Sigma_v = 4.5 I, mu_v
    unchanged, position std 2 (within (1.6, 2.4)); Sigma_v = I, dt = 2,
    q = 0.25 -> 1.5 I."""
    n = 20_000
    pf = opf.OraclePF(n, seed=42)
    pf.initialize(np.zeros(3), cb=0.0, spread_pos=0.0, spread_cb=0.0,
                  velocity=np.array([1.0, 0.0, 0.0]), velocity_init_sigma=2.0)
    pf.predict(np.array([1.0, 0.0, 0.0]), 1.0, 0.0, 0.0, step=1, sigma_vel=100.0,
               velocity_kf=True, velocity_process_noise=0.5)
    np.testing.assert_allclose(pf.vel, np.tile([1.0, 0.0, 0.0], (n, 1)))
    for k, c in enumerate((1.0, 0.0, 0.0)):
        assert 1.6 < float(np.std(pf.state[:, k] - c)) < 2.4
    np.testing.assert_allclose(pf.vcov, np.broadcast_to(4.5 * np.eye(3), (n, 3, 3)))
    pf = opf.OraclePF(4096, seed=42)
    pf.initialize(np.array([-3957199.0, 3310205.0, 3737911.0]), spread_pos=0.0, spread_cb=0.0,
                  velocity=np.array([0.5, -0.25, 0.0]), velocity_init_sigma=1.0)
    pf.predict(np.array([0.5, -0.25, 0.0]), 2.0, 0.0, 0.0, step=1, velocity_kf=True,
               velocity_process_noise=0.25)
    np.testing.assert_allclose(pf.vel, np.tile([0.5, -0.25, 0.0], (4096, 1)))
    np.testing.assert_allclose(pf.vcov[:, [0, 1, 2], [0, 1, 2]], 1.5)


def test_oracle_cholesky_against_numpy_and_noise_covariance():
    """_chol3 equals numpy's Cholesky factor on SPD matrices and keeps the
    reference's guards on singular ones; the RBPF position noise has covariance
    sigma_pos^2 I + dt^2 Sigma_v in ECEF for a non-diagonal Sigma_v (a
    transposed factor or a frame error fails this)."""
    rng = np.random.default_rng(0)
    M = rng.normal(size=(500, 3, 3))
    A = M @ np.transpose(M, (0, 2, 1)) + 0.1 * np.eye(3)
    np.testing.assert_allclose(opf.OraclePF._chol3(A), np.linalg.cholesky(A), rtol=0, atol=1e-12)
    S = np.zeros((1, 3, 3))
    S[0, 1, 1] = S[0, 2, 2] = 4.0   # first pivot zero: its column stays zero
    L = opf.OraclePF._chol3(S)
    np.testing.assert_allclose(L[0], np.diag([0.0, 2.0, 2.0]))
    n = 400_000
    pf = opf.OraclePF(n, seed=5)
    pf.initialize(np.array([-3.96e6, 3.35e6, 3.70e6]), spread_pos=0.0, spread_cb=0.0)
    C = np.array([[2.0, 0.8, -0.3], [0.8, 1.5, 0.4], [-0.3, 0.4, 1.0]])
    pf.vcov = np.broadcast_to(C, (n, 3, 3)).copy()
    before = pf.state.copy()
    pf.predict(np.zeros(3), 1.5, 0.7, 0.0, step=1, velocity_kf=True)
    d = pf.state[:, :3] - before[:, :3]
    np.testing.assert_allclose(np.cov(d.T), 0.49 * np.eye(3) + 2.25 * C, rtol=0, atol=0.03)


def test_every_oracle_function_has_a_pinning_entry():
    import inspect

    from oracle import pinning

    from oracle import learned as olearned

    from oracle import ffbsi as offbsi
    from oracle import raytrace as oray

    mods = {"philox": philox, "detmath": detmath, "pf": opf, "wls": owls, "learned": olearned,
            "raytrace": oray, "ffbsi": offbsi}
    missing = []
    for name, mod in mods.items():
        for fn, obj in vars(mod).items():
            if inspect.isfunction(obj) and obj.__module__ == mod.__name__:
                if f"{name}.{fn}" not in pinning.PINS:
                    missing.append(f"{name}.{fn}")
            if inspect.isclass(obj) and obj.__module__ == mod.__name__:
                for m, mo in vars(obj).items():
                    if inspect.isfunction(mo) and not m.startswith("_") and m not in ("ess", "estimate"):
                        if f"{name}.{fn}.{m}" not in pinning.PINS:
                            missing.append(f"{name}.{fn}.{m}")
    assert not missing, missing
    for key, (status, how) in pinning.PINS.items():
        assert status in ("pinned", "parity unpinned") and how


# -------------------------------------------------------- learned likelihood ----
def test_bf16_rounding_known_values():
    from oracle import learned as ol

    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties to even -> 1.0
    assert ol.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    # 1 + 3 * 2^-8 ties between 1 + 2^-7 and 1 + 2^-6: even mantissa is 1 + 2^-6
    assert ol.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6
    assert ol.bf16(np.float32(1.0 + 2.0 ** -8 + 2.0 ** -12)) == 1.0 + 2.0 ** -7
    assert ol.bf16(np.float32(-2.5)) == -2.5
    assert np.isnan(ol.bf16(np.float32(np.nan)))


def test_e4m3_quantiser_known_values():
    """OCP FP8 E4M3FN: values m/8 * 2^-6 (m < 8, subnormal) and (1 + m/8) 2^e
    for e in -6..8, the top code 1.875 * 2^8 being NaN, so 448 is the largest."""
    from oracle import learned as ol

    finite = [m / 8 * 2.0 ** -6 for m in range(8)]
    finite += [(1 + m / 8) * 2.0 ** e for e in range(-6, 9) for m in range(8)]
    finite = np.array(sorted(set(v for v in finite if v <= 448.0)), np.float32)
    assert finite[-1] == 448.0 and finite[1] == 2.0 ** -9 and finite.size == 127
    for sc in (1.0, 2.0 ** -5, 2.0 ** 3):
        np.testing.assert_array_equal(ol.q_e4m3(finite * np.float32(sc), sc), finite * sc)
        np.testing.assert_array_equal(ol.q_e4m3(-finite * np.float32(sc), sc), -finite * sc)
    # midpoints go to the neighbour with the even mantissa
    mid = ((finite[1:] + finite[:-1]) / 2).astype(np.float32)
    q = ol.q_e4m3(mid, 1.0)
    # 8 values per binade (and 8 subnormals), so a value's position in the
    # sorted list has the parity of its mantissa field
    idx = np.searchsorted(finite, q)
    np.testing.assert_array_equal(finite[idx], q)
    assert np.all(idx % 2 == 0)
    # anything just off a midpoint goes to the nearer value
    above = np.nextafter(mid, np.float32(np.inf))
    np.testing.assert_array_equal(ol.q_e4m3(above, 1.0), finite[1:])
    below = np.nextafter(mid, np.float32(-np.inf))
    np.testing.assert_array_equal(ol.q_e4m3(below, 1.0), finite[:-1])
    # saturation, subnormal quantum, signed zero
    assert ol.q_e4m3(np.float32(1e6), 1.0) == 448.0 and ol.q_e4m3(np.float32(-500.0), 1.0) == -448.0
    assert ol.q_e4m3(np.float32(2.0 ** -10), 1.0) == 0.0
    assert ol.q_e4m3(np.float32(3 * 2.0 ** -11), 1.0) == 2.0 ** -9
    assert np.signbit(ol.q_e4m3(np.float32(-0.0), 1.0))
    # per-tensor scale: the smallest power of two with max|W| / s <= 448
    for mx in (448.0, 449.0, 1.0, 0.3, 3e-4):
        W = np.array([mx, -0.5 * mx], np.float32)
        s = float(ol.e4m3_scale(W))
        assert s == 2.0 ** np.round(np.log2(s))
        assert mx / s <= 448.0 and mx / (s / 2) > 448.0
    assert ol.e4m3_scale(np.zeros(4)) == 1.0
    p = ol.init_params(3)
    assert all(np.array_equal(a, b) for a, b in zip(ol.quantised(p, "bf16"), ol.unpack(p)))
    W2q = ol.quantised(p, "fp8")[2]
    s2 = ol.e4m3_scale(ol.unpack(p)[2])
    assert np.all(np.isin(np.abs(W2q / s2).astype(np.float32), finite))


def test_first_layer_bias_close_to_binary64():
    from oracle import learned as ol

    p = ol.init_params(2)
    W1, b1 = ol.unpack(p)[:2]
    f = np.random.default_rng(1).normal(size=(5, 7))
    c = ol.first_layer_bias(W1, b1, f)
    np.testing.assert_allclose(c, b1 + f @ W1[:, 1:].T, rtol=1e-5, atol=1e-5)


def test_mlp_forward_affine_on_fixed_pattern():
    """With every unit active the exact network is affine: z3 = A x + c with
    A = W3 W2 W1, checked on inputs where all pre-activations are positive."""
    from oracle import learned as ol

    p = ol.init_params(0).astype(np.float64)
    W1, b1, W2, b2, W3, b3 = ol.unpack(p)
    W1, W2 = np.abs(W1) * 0.1, np.abs(W2) * 0.1
    b1, b2 = np.ones(ol.H), np.ones(ol.H)
    p = ol.pack(W1, b1, W2, b2, W3, b3).astype(np.float64)
    e = np.array([3.0, 17.0, 42.0])
    f = np.abs(np.random.default_rng(0).normal(size=(3, 7)))
    fw = ol.mlp_forward(p, e, f, mode="exact")
    x = np.concatenate([(e * np.float32(0.1))[:, None], f], axis=1)
    z3 = x @ (W3 @ W2 @ W1).T + (W3 @ (W2 @ b1 + b2) + b3)
    np.testing.assert_allclose(fw["z3"], z3, rtol=1e-6)
    np.testing.assert_allclose(fw["mu"], 10 * z3[:, 0], rtol=1e-6)
    # the pair log-likelihood is the Gaussian log-density of e under (mu, exp(l))
    from math import exp, log, pi, sqrt
    for i in range(3):
        sd = sqrt(exp(fw["lv"][i]))
        logpdf = -0.5 * ((e[i] - fw["mu"][i]) / sd) ** 2 - log(sd) - 0.5 * log(2 * pi)
        assert fw["ll"][i] == pytest.approx(logpdf + 0.5 * log(2 * pi), rel=1e-12, abs=1e-12)


def test_contrastive_gradient_finite_differences():
    from oracle import learned as ol

    p = ol.init_params(1).astype(np.float64)
    rng = np.random.default_rng(0)
    S, K = 5, 7
    r0 = rng.normal(0, 10, S)
    u = rng.normal(size=(S, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    f = rng.normal(size=(S, 7))
    dpos, dcb = rng.normal(0, 5, (K, 3)), rng.normal(0, 5, K)
    dpos[0], dcb[0] = 0, 0
    loss, g, _ = ol.contrastive_loss_and_grad(p, r0, u, f, dpos, dcb)
    assert loss > 0
    h = 1e-6
    for i in range(0, ol.N_PARAMS, 53):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        num = (ol.contrastive_loss_and_grad(pp, r0, u, f, dpos, dcb)[0]
               - ol.contrastive_loss_and_grad(pm, r0, u, f, dpos, dcb)[0]) / (2 * h)
        assert num == pytest.approx(g[i], abs=1e-7, rel=1e-5)


def test_cpu_wls_baseline_matches_oracle():
    """The C WLS restatement (timed CPU baseline) against the NumPy oracle."""
    import ctypes as C

    from gnss_gpu_amd import sim
    from oracle import build_cpu

    lib = build_cpu.load()
    sat, pr, w, ns, truth = sim.wls_epochs(200, seed=3)
    E, S = pr.shape
    out = np.zeros((E, 4))
    it = np.zeros(E, np.int32)
    DP, IP = C.POINTER(C.c_double), C.POINTER(C.c_int32)
    a = [np.ascontiguousarray(x) for x in (sat, pr, w)]
    nsi = np.ascontiguousarray(ns.astype(np.int32))
    lib.cpu_wls_batch(a[0].ctypes.data_as(DP), a[1].ctypes.data_as(DP), a[2].ctypes.data_as(DP),
                      nsi.ctypes.data_as(IP), E, S, 10, 1e-4, 1, 2, out.ctypes.data_as(DP),
                      it.ctypes.data_as(IP))
    ref, rit = owls.wls_batch_ref(sat, pr, w, n_sat=ns)
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-6)
    np.testing.assert_array_equal(it, rit)


def test_cpu_multi_gnss_baseline_matches_oracle():
    """The C multi-GNSS restatement (timed CPU baseline) against the NumPy
    oracle, batch rule and single-epoch rule."""
    import ctypes as C

    from gnss_gpu_amd import sim
    from oracle import build_cpu

    lib = build_cpu.load()
    sat, pr, w, ns, truth = sim.wls_epochs(120, s_min=14, s_max=14, seed=13)
    E, S = pr.shape
    ids = np.ascontiguousarray(np.tile(np.arange(S, dtype=np.int32) % 3, (E, 1)))
    ids[::4, 2] = -1
    pr = np.ascontiguousarray(pr + np.array([0.0, 9.0, -4.0])[np.clip(ids, 0, 2)])
    DP, IP = C.POINTER(C.c_double), C.POINTER(C.c_int32)
    for sagnac, guards, ref_fn in ((1, 0, owls.wls_multi_gnss_batch_ref), (0, 1, None)):
        out = np.zeros((E, 6))
        it = np.zeros(E, np.int32)
        lib.cpu_wls_mg_batch(sat.ctypes.data_as(DP), pr.ctypes.data_as(DP), w.ctypes.data_as(DP),
                             ids.ctypes.data_as(IP), E, S, 3, 40, 1e-4, sagnac, guards, 2,
                             out.ctypes.data_as(DP), it.ctypes.data_as(IP))
        if ref_fn is not None:
            ref, rit = ref_fn(sat, pr, w, ids, 3, max_iter=40)
        else:
            ref, rit = owls.wls_multi_gnss_solve(sat, pr, w, ids, 3, 40, 1e-4, sagnac=False,
                                                 cpu_rule=True)
        np.testing.assert_allclose(out, ref, rtol=0, atol=1e-6)
        np.testing.assert_array_equal(it, rit)


def _median_tree(tris, leaf=4):
    """Flat BVH rows [n, 10] (min xyz, max xyz, left, right, start, count; left -1
    at a leaf) by median splits of the centroids along the widest axis, and the
    triangles in the tree's order: a tree built apart from the library's."""
    cen = tris.mean(axis=1)
    rows, order = [], []

    def rec(idx):
        i = len(rows)
        rows.append(None)
        lo, hi = tris[idx].min(axis=(0, 1)), tris[idx].max(axis=(0, 1))
        if len(idx) <= leaf:
            rows[i] = [*lo, *hi, -1, -1, len(order), len(idx)]
            order.extend(idx.tolist())
            return i
        s = idx[np.argsort(cen[idx, int(np.argmax(hi - lo))], kind="stable")]
        left = rec(s[: len(s) // 2])
        right = rec(s[len(s) // 2:])
        rows[i] = [*lo, *hi, left, right, 0, 0]
        return i

    rec(np.arange(len(tris)))
    return np.array(rows, dtype=np.float64), np.ascontiguousarray(tris[order])


@pytest.mark.parametrize("probs", [(1.0, 0.0), (0.85, 0.05)])
def test_cpu_raytrace_baseline_matches_oracle(probs):
    """The C/OpenMP BVH walk (timed CPU baseline of the ray-traced weight) against
    the brute-force NumPy oracle on a 60-box city (720 triangles): the same
    log-weights, with rays both blocked and clear; an empty tree blocks nothing."""
    import ctypes as C

    from gnss_gpu_amd.raytrace import BuildingModel
    from oracle import build_cpu
    from oracle import raytrace as oray

    rng = np.random.RandomState(8)
    tris = np.concatenate([BuildingModel.create_box(
        [rng.uniform(-200, 200), rng.uniform(-200, 200), rng.uniform(0, 40)],
        rng.uniform(5, 30), rng.uniform(5, 30), rng.uniform(10, 60)).triangles
        for _ in range(60)])
    nodes, stris = _median_tree(tris)
    n, ns = 160, 10
    p = [rng.uniform(-200, 200, n), rng.uniform(-200, 200, n), rng.uniform(0, 30, n),
         rng.uniform(90, 110, n)]
    th, ph = rng.uniform(0, 2 * np.pi, ns), rng.uniform(np.radians(8), np.radians(60), ns)
    sat = 2.2e7 * np.stack([np.cos(ph) * np.cos(th), np.cos(ph) * np.sin(th), np.sin(ph)], 1)
    pr = np.sqrt((sat ** 2).sum(1)) + 100.0 + rng.normal(0, 20, ns)
    ws = rng.uniform(0.5, 1.5, ns)
    pos = np.stack(p[:3], 1)
    frac = np.mean([oray.los_blocked(pos, np.broadcast_to(sat[s], pos.shape), tris, tmin=1e-6,
                                     recip=True).mean() for s in range(ns)])
    assert 0.05 < frac < 0.95
    lib = build_cpu.load()
    DP = C.POINTER(C.c_double)

    def run(nd, tr):
        lw = np.zeros(n)
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (*p, sat, pr, ws, nd, tr)]
        rc = lib.cpu_weight_3d_bvh(*(x.ctypes.data_as(DP) for x in a[:4]), n,
                                   *(x.ctypes.data_as(DP) for x in a[4:7]), ns,
                                   a[7].ctypes.data_as(DP), len(nd), a[8].ctypes.data_as(DP),
                                   len(tr), 3.0, 30.0, 20.0, *probs, lw.ctypes.data_as(DP), 0)
        assert rc == 0
        return lw

    ref = oray.weight_3d(*p, sat, pr, ws, tris, 3.0, 30.0, 20.0, *probs)
    np.testing.assert_allclose(run(nodes, stris), ref, rtol=1e-12, atol=1e-12)
    clear = oray.weight_3d(*p, sat, pr, ws, np.zeros((0, 3, 3)), 3.0, 30.0, 20.0, *probs)
    np.testing.assert_allclose(run(np.zeros((0, 10)), np.zeros((0, 3, 3))), clear,
                               rtol=1e-12, atol=1e-12)
    assert np.abs(ref - clear).max() > 1.0

# ------------------------------------------- DD pseudorange / carrier AFV ----
def _dd_scenario():
    d = GOLD["dd_closed_forms"]
    rng = np.random.RandomState(123)
    dirs = rng.randn(5, 3)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    sat = dirs * 20_000_000.0
    rover = np.zeros(3)
    base = np.array(d["base_pos"])
    rr = np.linalg.norm(sat - rover, axis=1)
    br = np.linalg.norm(sat - base, axis=1)
    dd = (rr[1:] - rr[0]) - (br[1:] - br[0])
    return d, sat, rover, br, dd


def test_dd_closed_forms():
    """The reference's own DD constants (test_cuda_streams.py:604-771) from the
    oracle's binary64 DD pseudorange and DD carrier AFV terms."""
    d, sat, rover, br, dd = _dd_scenario()
    n = len(dd)
    sk, sr = sat[1:], np.repeat(sat[[0]], n, axis=0)
    brk, brr, w = br[1:], np.repeat(br[0], n), np.ones(n)
    pos = rover[None, :]
    k = d["huber_k"]
    ddo = dd.copy()
    ddo[0] += d["dd_pr_outlier_m"]
    s = d["dd_pr_sigma_m"]
    assert opf.dd_pseudorange_loglik(pos, sk, sr, ddo, brk, brr, w, s)[0] == pytest.approx(
        d["dd_pr_gaussian"])
    assert opf.dd_pseudorange_loglik(pos, sk, sr, ddo, brk, brr, w, s, huber=True,
                                     huber_k=k)[0] == pytest.approx(d["dd_pr_huber"])
    gated = opf.dd_pseudorange_loglik(pos, sk, sr, ddo, brk, brr, w, s, thr=d["dd_pr_gate_m"])
    assert abs(gated[0]) < d["gated_abs_max"]
    wl = np.full(n, d["wavelength_m"])
    ddc = dd / wl
    ddc[0] += d["dd_cp_outlier_cycles"]
    sc = d["dd_cp_sigma_cycles"]
    assert opf.dd_carrier_afv_loglik(pos, sk, sr, ddc, brk, brr, w, wl, sc)[0] == pytest.approx(
        d["dd_cp_gaussian"], abs=1e-9)
    assert opf.dd_carrier_afv_loglik(pos, sk, sr, ddc, brk, brr, w, wl, sc, huber=True,
                                     huber_k=k)[0] == pytest.approx(d["dd_cp_huber"], abs=1e-9)
    g2 = opf.dd_carrier_afv_loglik(pos, sk, sr, ddc, brk, brr, w, wl, sc,
                                   thr=d["dd_cp_gate_cycles"])
    assert abs(g2[0]) < d["gated_abs_max"]
    # every pair beyond the gate: fewer than min(n, 3) kept, so the gate is off
    ddo2 = dd + 20.0
    full = opf.dd_pseudorange_loglik(pos, sk, sr, ddo2, brk, brr, w, s, thr=10.0)
    assert full[0] == pytest.approx(-0.5 * 400.0 / (s * s) * n)
    # AFV: integer cycles vanish, x + k maps to x, |afv| <= 1/2
    x = np.array([0.0, 3.0, -7.0, 0.25, 2.25, -0.49, 0.51])
    np.testing.assert_allclose(opf.afv(x), [0, 0, 0, 0.25, 0.25, -0.49, -0.49], atol=1e-12)
    # undifferenced carrier AFV: zero at the truth, whole cycles invisible,
    # a 0.3-cycle offset costs -0.5 (0.3 / sigma)^2 per satellite
    cb = np.array([100.0])
    r = np.linalg.norm(sat - rover, axis=1)
    cp = (r + 100.0) / d["wavelength_m"]
    w5 = np.ones(5)
    assert abs(opf.carrier_afv_loglik(pos, cb, sat, cp, w5, d["wavelength_m"], 0.05)[0]) < 1e-9
    assert abs(opf.carrier_afv_loglik(pos, cb, sat, cp + 17.0, w5, d["wavelength_m"],
                                      0.05)[0]) < 1e-9
    cp3 = cp.copy()
    cp3[2] += 0.3
    assert opf.carrier_afv_loglik(pos, cb, sat, cp3, w5, d["wavelength_m"], 0.05)[0] == \
        pytest.approx(-0.5 * (0.3 / 0.05) ** 2, rel=1e-6)


# ------------------------------------------------------------------ Doppler ----
def _doppler_rows():
    d = GOLD["doppler_rows"]
    rx = np.array(d["rx_pos"])
    sat = np.array(d["sat_ecef"])
    sv = np.array(d["sat_vel"])
    tv = np.array(d["true_vel"])
    los = (sat - rx) / np.linalg.norm(sat - rx, axis=1)[:, None]
    rr = np.einsum("sk,sk->s", sv - tv[None, :], los) + d["true_cd"]
    dop = -rr / d["wavelength_m"]
    return d, rx, sat, sv, tv, dop


def test_doppler_oracle_known_answers():
    """Noise-free Doppler rows of the reference's tests (_make_doppler_data):
    the WLS nudge with gain 1 lands on the true velocity, the likelihood is
    zero at the truth for any clock drift, the step cap holds, and the KF
    update from a vague prior converges to the truth with a shrinking
    covariance; a single row matches the closed-form scalar Kalman update."""
    d, rx, sat, sv, tv, dop = _doppler_rows()
    wl, sig = d["wavelength_m"], d["sigma_mps"]
    w = np.ones(len(dop))
    pos = rx[None, :]
    add, v = opf.doppler_update(pos, np.zeros((1, 3)), sat, sv, dop, w, wl, sig, 1.0, 100.0)
    np.testing.assert_allclose(v[0], tv, atol=1e-9)
    add0, v0 = opf.doppler_update(pos, tv[None, :], sat, sv, dop, w, wl, sig, 0.0, 100.0)
    assert abs(add0[0]) < 1e-12 and np.array_equal(v0[0], tv)
    assert add[0] < 0.0
    _, vc = opf.doppler_update(pos, np.zeros((1, 3)), sat, sv, dop, w, wl, sig, 1.0, 0.25)
    assert np.linalg.norm(vc[0]) == pytest.approx(0.25, rel=1e-9)
    s0 = d["kf_init_sigma"] ** 2
    addk, mu, P = opf.doppler_kf_update(pos, np.zeros((1, 3)), s0 * np.eye(3)[None], sat, sv, dop,
                                        w, wl, sig)
    np.testing.assert_allclose(mu[0], tv, atol=d["kf_atol"])
    assert np.all(np.diag(P[0]) < s0) and np.all(np.diag(P[0]) > 0)
    # one satellite pair: two rows centred on their mean are +-(l1 - l0)/2
    sel = [0, 1]
    _, mu1, P1 = opf.doppler_kf_update(pos, np.zeros((1, 3)), s0 * np.eye(3)[None], sat[sel],
                                       sv[sel], dop[sel], w[sel], wl, sig)
    l = (sat[sel] - rx) / np.linalg.norm(sat[sel] - rx, axis=1)[:, None]
    y = -dop[sel] * wl - np.einsum("sk,sk->s", sv[sel], l)
    h = -(l[0] - l.mean(axis=0))
    yc = y[0] - y.mean()
    S = s0 * h @ h + sig * sig
    K = s0 * h / S
    mu_a = K * yc
    P_a = s0 * np.eye(3) - np.outer(K, h * s0)
    h2 = -(l[1] - l.mean(axis=0))
    yc2 = y[1] - y.mean()
    S2 = h2 @ P_a @ h2 + sig * sig
    K2 = P_a @ h2 / S2
    np.testing.assert_allclose(mu1[0], mu_a + K2 * (yc2 - h2 @ mu_a), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(P1[0], P_a - np.outer(K2, h2 @ P_a), rtol=1e-7, atol=1e-9)


def test_gmm_log_likelihood_pins():
    """pf.log_likelihood_gmm against what the mathematics fixes: with unit
    satellite weights each term is log of the normal-mixture density
    w N(e; 0, s_los) + (1 - w) N(e; mu, s_nlos) plus log sqrt(2 pi) (the
    reference drops that constant from both components, pf_device.cu:418-425);
    w_los -> 1 leaves the Gaussian term minus log s_los; two equal components
    make w_los irrelevant."""
    from scipy.stats import norm

    from oracle import pf as opf

    rng = np.random.default_rng(21)
    truth = np.array([-3957199.0, 3310205.0, 3737911.0])
    sat = truth + rng.normal(0, 1.5e7, (9, 3))
    pos = truth + rng.normal(0, 30.0, (64, 3))
    cb = rng.normal(100.0, 20.0, 64)
    pr = np.linalg.norm(sat - truth, axis=1) + 100.0 + rng.normal(0, 3, 9)
    pr[[1, 4]] += [25.0, 60.0]  # two NLOS-like biases
    e = pr[None, :] - (np.linalg.norm(pos[:, None, :] - sat[None, :, :], axis=2) + cb[:, None])
    ones = np.ones(9)
    got = opf.log_likelihood_gmm(pos, cb, sat, pr, ones, 5.0, 0.7, 15.0, 30.0)
    dens = 0.7 * norm.pdf(e, 0.0, 5.0) + 0.3 * norm.pdf(e, 15.0, 30.0)
    ref = np.sum(np.log(dens) + 0.5 * np.log(2 * np.pi), axis=1)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-9)
    # w_los = 1: the Gaussian log-density without its constant
    g1 = opf.log_likelihood_gmm(pos, cb, sat, pr, ones, 5.0, 1.0, 15.0, 30.0)
    np.testing.assert_allclose(g1, np.sum(-np.log(5.0) - 0.5 * e * e / 25.0, axis=1), rtol=1e-12)
    # identical components: any w_los gives the single Gaussian
    for wl in (0.1, 0.5, 0.9):
        gi = opf.log_likelihood_gmm(pos, cb, sat, pr, ones, 4.0, wl, 0.0, 4.0)
        np.testing.assert_allclose(gi, np.sum(-np.log(4.0) - 0.5 * e * e / 16.0, axis=1),
                                   rtol=1e-12)
    # at the truth the residuals are noise plus the positive NLOS biases: a
    # positive NLOS mean explains them better than a negative one
    t1 = opf.log_likelihood_gmm(truth[None, :], np.array([100.0]), sat, pr, ones, 5.0, 0.7,
                                15.0, 30.0)
    t2 = opf.log_likelihood_gmm(truth[None, :], np.array([100.0]), sat, pr, ones, 5.0, 0.7,
                                -15.0, 30.0)
    assert t1[0] > t2[0]


def test_tile_ends_pins():
    """detmath.tile_ends: the output count before each tile equals the count of
    systematic targets below that tile's share of an independent binary64
    cumulative sum of the weights (away from rounding ties); it is
    non-decreasing, ends at N, and starts at #{t_j <= 0}."""
    from oracle import detmath

    rng = np.random.default_rng(12)
    for n in (1, 1000, 1025, 5000, 70_000):
        lw = (rng.normal(0, 3.0, n) - 40.0).astype(np.float32)
        for u0 in (0.0, 0.37 / n, 0.999 / n):
            m, S, e = detmath.tile_partials(lw, n)
            _, _, P, W = detmath.decide(m, S)
            K = detmath.tile_ends(P, W, n, u0)
            assert K[-1] == n and np.all(np.diff(K) >= 0)
            assert K[0] == (1 if u0 == 0.0 else 0)
            w = np.exp(lw.astype(np.float64) - np.max(lw))
            C = np.concatenate([[0.0], np.cumsum(w)])[:: detmath.TILE][: len(P)] / np.sum(w)
            t = (u0 + np.arange(n)) / n
            Kref = np.searchsorted(t, C, side="right")
            near = np.array([np.min(np.abs(t - c)) < 1e-12 for c in C])
            assert np.all((K[:-1] == Kref) | near)
