"""Pins for the k-means oracle (oracle/kmeans_oracle.cpp), CPU only.

The reference's tests do not cover k-means, so the oracle is pinned by:
the C++ standard's mt19937 known answer; an independent Python restatement
of libstdc++'s shuffle; numpy brute force for the assignment and the
nearest-grid-point maps; exact-tie and sqrt-merge cases a d^2 argmin or a
last-index rule would get wrong; the weighted-mean fixed point and the
monotone Lloyd objective; and the loop-control behaviour of kmeans.cpp:233-257.
"""
import os

import numpy as np
import pytest

import libstdcxx_restated as R
from oracle import kmeans_oracle as KO

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_mt19937_standard_known_answer():
    want = int(open(os.path.join(GOLDEN, "mt19937_seed5489_10000.txt")).read())
    assert int(KO.mt19937(5489, 10000)[-1]) == want
    g = R.MT19937(5489)
    assert [g() for _ in range(10000)][-1] == want


@pytest.mark.parametrize("seed", [0, 1, 12345, 2**32 - 1])
def test_mt19937_restated(seed):
    g = R.MT19937(seed)
    np.testing.assert_array_equal(KO.mt19937(seed, 1500), np.array([g() for _ in range(1500)], dtype=np.uint32))


@pytest.mark.parametrize("n,k,seed", [(2, 1, 0), (1000, 50, 0), (1001, 60, 3), (32768, 128, 0),
                                      (65535, 40, 1), (65536, 40, 2), (70001, 100, 9)])
def test_initial_indices_restated(n, k, seed):
    got = KO.initial_indices(n, k, seed)
    assert list(got) == R.initial_indices(n, k, seed)
    assert np.all(np.diff(got) > 0)


def _brute_classify(grid, cen):
    # distances in the reference's operation order, first index on ties
    dx = cen[None, :, 0] - grid[:, None, 0]
    dy = cen[None, :, 1] - grid[:, None, 1]
    dz = cen[None, :, 2] - grid[:, None, 2]
    d = np.sqrt((dx * dx + dy * dy) + dz * dz)
    return np.argmin(d, axis=1).astype(np.int32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_classify_brute_force(seed):
    rng = np.random.default_rng(seed)
    grid = rng.uniform(-3, 5, size=(3000, 3))
    cen = rng.uniform(-3, 5, size=(77, 3))
    np.testing.assert_array_equal(KO.classify(grid, cen), _brute_classify(grid, cen))


def test_classify_exact_ties_first_index():
    cen = np.array([[0.0, 0, 0], [2.0, 0, 0], [1.0, 1.0, 0]])
    pts = np.array([[1.0, 0, 0], [1.0, 0.5, 0], [1.0, -5, 0]])
    # (1,0,0): d0 = d1 = 1 = d2 -> index 0; (1,.5,0): d2 = .5 smallest; (1,-5,0): d0 = d1 -> 0
    np.testing.assert_array_equal(KO.classify(pts, cen), [0, 2, 0])
    np.testing.assert_array_equal(KO.classify(pts, cen[::-1].copy()), [0, 0, 1])


def _find_sqrt_merge(rng):
    """Two offsets whose squared distances differ but round to the same sqrt:
    y = sqrt(T - x^2) puts d^2 = x^2 + y^2 within a few ulps of T, where sqrt
    maps neighbouring doubles onto one value."""
    for target in (2.0, 3.0, 5.0, 7.0):
        x = rng.uniform(0.3, 1.2, size=20000)
        y = np.sqrt(target - x * x)
        d2 = x * x + y * y
        s = np.sqrt(d2)
        for val in np.unique(s):
            sel = np.nonzero(s == val)[0]
            ds = d2[sel]
            if ds.min() != ds.max():
                lo, hi = sel[np.argmin(ds)], sel[np.argmax(ds)]
                return np.array([x[lo], y[lo]]), np.array([x[hi], y[hi]])
    raise AssertionError("no sqrt merge found")


def test_classify_sqrt_merge_first_index():
    """The reference compares rounded distances: a centroid listed first wins a
    tie in sqrt even when its d^2 is larger, which a d^2 argmin would miss."""
    small, large = _find_sqrt_merge(np.random.default_rng(0))
    p = np.array([[0.0, 0.0, 0.0]])
    cen = np.array([[large[0], large[1], 0.0], [small[0], small[1], 0.0]])
    assert KO.classify(p, cen)[0] == 0
    assert KO.classify(p, cen[::-1].copy())[0] == 0


def test_update_centroids_sequential_and_weighted_mean():
    rng = np.random.default_rng(4)
    n, k = 500, 9
    grid = rng.uniform(0, 3, size=(n, 3))
    rho = rng.uniform(0.01, 1, size=n)
    gmap = rng.integers(0, k, size=n).astype(np.int32)
    got, crit = KO.update_centroids(rho, grid, gmap, k)
    assert crit == 0
    # sequential sums in ascending index, product rounded then added
    w = [0.0] * k
    s = [[0.0] * 3 for _ in range(k)]
    for i in range(n):
        c = gmap[i]
        w[c] += rho[i]
        for d in range(3):
            s[c][d] += rho[i] * grid[i, d]
    seq = np.array([[s[c][d] / w[c] for d in range(3)] for c in range(k)])
    np.testing.assert_array_equal(got, seq)
    # and within a few ulps of the exactly weighted mean
    import math
    for c in range(k):
        m = gmap == c
        for d in range(3):
            exact = math.fsum(rho[m] * grid[m, d]) / math.fsum(rho[m])
            assert abs(got[c, d] - exact) <= 8 * np.spacing(abs(exact))


def test_update_centroids_empty_cluster_flag():
    grid = np.zeros((4, 3))
    _, crit = KO.update_centroids(np.ones(4), grid, np.array([0, 0, 1, 1], dtype=np.int32), 3)
    assert crit == 1


def test_normed_difference_closed_form():
    rng = np.random.default_rng(2)
    a = rng.normal(size=300)
    b = a + rng.normal(scale=1e-3, size=300)
    scale = max(abs(a.max()), abs(a.min())) + max(abs(b.max()), abs(b.min()))
    want = np.linalg.norm(a - b)
    assert abs(KO.normed_difference(a, b) - want) < 1e-14 * want
    assert KO.normed_difference(a, a) == 0.0
    # scale invariance: the scale factor cancels exactly for powers of two
    assert KO.normed_difference(4 * a, 4 * b) == 4 * KO.normed_difference(a, b)
    assert scale > 0


def test_map_to_grid_brute_force():
    rng = np.random.default_rng(8)
    grid = rng.uniform(0, 4, size=(2500, 3))
    cen = rng.uniform(0, 4, size=(60, 3))
    got, dups = KO.map_to_grid(grid, cen)
    d = np.linalg.norm(grid[None, :, :] - cen[:, None, :], axis=2)
    want = np.sort(np.argmin(d, axis=1))
    np.testing.assert_array_equal(got, want)
    assert dups == int(np.sum(want[1:] == want[:-1]))


def test_map_to_grid_reports_duplicates():
    grid = np.array([[0.0, 0, 0], [1.0, 0, 0], [5.0, 0, 0]])
    cen = np.array([[0.1, 0, 0], [-0.1, 0, 0], [4.0, 0, 0]])
    got, dups = KO.map_to_grid(grid, cen)
    np.testing.assert_array_equal(got, [0, 0, 2])
    assert dups == 1


def _lloyd_objective(grid, rho, cen, gmap):
    return float(np.sum(rho * np.sum((grid - cen[gmap]) ** 2, axis=1)))


def test_kmeans_lloyd_invariants():
    """Weighted Lloyd iterations never increase the weighted objective, every
    point sits in its nearest cluster, and the loop's centroids are the
    weighted means of the previous assignment."""
    from thcpp_amd import synthetic
    s = synthetic.make_system(reps=1, n=16, M=8, thc_cfac=4, max_it=12)
    r = KO.kmeans(s.grid, s.rho, s.K, s.max_it, -1.0, 0, keep_maps=True)
    assert r["status"] == 0 and r["iters"] == s.max_it
    objs = []
    cen = s.grid[KO.initial_indices(s.N, s.K, 0)]
    for it in range(s.max_it):
        gmap = r["maps"][it]
        np.testing.assert_array_equal(gmap, _brute_classify(s.grid, cen))
        objs.append(_lloyd_objective(s.grid, s.rho, cen, gmap))
        cen, _ = KO.update_centroids(s.rho, s.grid, gmap, s.K)
    assert all(b <= a * (1 + 1e-12) for a, b in zip(objs, objs[1:]))
    np.testing.assert_array_equal(cen, r["centroids"])
    np.testing.assert_array_equal(r["interp_idx"], KO.map_to_grid(s.grid, cen)[0])
    assert r["pcts"][0] == 0.0


def test_kmeans_threshold_semantics():
    from thcpp_amd import synthetic
    s = synthetic.make_system(reps=1, n=12, M=6, thc_cfac=4, max_it=10)
    full = KO.kmeans(s.grid, s.rho, s.K, 10, -1.0, 0)
    # break when diff < threshold: stops right after the first diff below it
    thr = float(full["diffs"][4]) * (1 + 1e-12)
    r = KO.kmeans(s.grid, s.rho, s.K, 10, thr, 0)
    first = int(np.argmax(full["diffs"] < thr))
    assert r["iters"] == first + 1
    np.testing.assert_array_equal(r["diffs"], full["diffs"][: first + 1])
    # diff == threshold on the last iteration: no mapping, indices stay 0
    eq = KO.kmeans(s.grid, s.rho, s.K, 6, float(full["diffs"][5]), 0)
    assert eq["iters"] == 6 and not eq["interp_idx"].any()
