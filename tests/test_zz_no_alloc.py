"""The allocation-free WLS batch (gg_wls_batch_noalloc, staged in a static device
area) returns exactly what gg_wls_batch returns, call after call; the second
call of each pair is smaller than the first, so stale contents of the shared
area would show."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _epochs(seed, n_epoch, s_max, ragged):
    from gnss_gpu_amd import sim

    rng = np.random.default_rng(seed)
    sat = np.zeros((n_epoch, s_max, 3))
    pr = np.zeros((n_epoch, s_max))
    w = np.zeros((n_epoch, s_max))
    ns = (rng.integers(6, s_max + 1, n_epoch) if ragged else np.full(n_epoch, s_max)).astype(np.int32)
    for e in range(n_epoch):
        truth = sim.ANCHOR_CFG1 + rng.normal(0, 500.0, 3)
        k = int(ns[e])
        el = rng.uniform(15, 85, k)
        az = rng.uniform(0, 360, k)
        sat[e, :k] = sim.place_satellites(truth, el, az)
        pr[e, :k] = np.linalg.norm(sat[e, :k] - truth, axis=1) + 30.0 + rng.normal(0, 3, k)
        w[e, :k] = rng.uniform(0.5, 2.0, k)
    return sat, pr, w, ns


def _run(name, sat, pr, w, ns, max_iter=10, tol=1e-4):
    from gnss_gpu_amd import _lib

    n_epoch, s_max = sat.shape[:2]
    res = np.full((n_epoch, 4), np.nan)
    it = np.full(n_epoch, -7, dtype=np.int32)
    sat = np.ascontiguousarray(sat)
    pr = np.ascontiguousarray(pr)
    wp = None if w is None else _lib.dptr(np.ascontiguousarray(w))
    nsp = None if ns is None else _lib.iptr(np.ascontiguousarray(ns))
    _lib.call(name, _lib.dptr(sat), _lib.dptr(pr), wp, nsp, n_epoch, s_max, max_iter, float(tol),
              _lib.dptr(res), _lib.iptr(it))
    return res, it


@pytest.mark.parametrize("ragged", [True, False])
def test_noalloc_matches_allocating_batch_twice(gpu, ragged):
    for n_epoch, s_max, seed in ((10_000, 20, 1), (3_000, 12, 2)):
        sat, pr, w, ns = _epochs(seed + (10 if ragged else 0), n_epoch, s_max, ragged)
        ns_arg = ns if ragged else None
        ref, rit = _run("gg_wls_batch", sat, pr, w, ns_arg)
        got, git = _run("gg_wls_batch_noalloc", sat, pr, w, ns_arg)
        np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))
        np.testing.assert_array_equal(git, rit)
        assert np.all(rit >= 1)


def test_noalloc_default_weights_after_weighted_call(gpu):
    """A call without weights (ones, written by a fill kernel) after a weighted
    call on the same area equals gg_wls_batch without weights."""
    sat, pr, w, ns = _epochs(5, 4000, 16, True)
    _run("gg_wls_batch_noalloc", sat, pr, w, ns)
    sat2, pr2, _, ns2 = _epochs(6, 1500, 10, True)
    ref, rit = _run("gg_wls_batch", sat2, pr2, None, ns2)
    got, git = _run("gg_wls_batch_noalloc", sat2, pr2, None, ns2)
    np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))
    np.testing.assert_array_equal(git, rit)


def test_noalloc_rejects_sizes_beyond_the_area(gpu):
    from gnss_gpu_amd import _lib

    n_epoch, s_max = 60_000, 8   # 60k x 8 x 40 B = 19.2 MB > 16 MiB
    sat = np.zeros((n_epoch, s_max, 3))
    pr = np.zeros((n_epoch, s_max))
    res = np.zeros((n_epoch, 4))
    it = np.zeros(n_epoch, dtype=np.int32)
    rc = _lib.lib.gg_wls_batch_noalloc(_lib.dptr(sat), _lib.dptr(pr), None, None, n_epoch, s_max,
                                       10, 1e-4, _lib.dptr(res), _lib.iptr(it))
    assert rc != 0
    assert "16 MiB" in _lib.last_error()
    _ = C
