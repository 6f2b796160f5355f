"""k-means on the GPU against the oracle's restatement of src/kmeans.cpp:
every assignment, centroid, normed difference and final index must be
bit-identical."""
import numpy as np
import pytest
import torch

from oracle import kmeans_oracle as KO
from thcpp_amd import _lib as L

pytestmark = pytest.mark.gpu


def _lattice(n, L_=5.0):
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return np.stack([i, j, k], -1).reshape(-1, 3) * (L_ / n)


@pytest.mark.parametrize("n,k,seed", [(5000, 97, 1), (20000, 1000, 2), (3, 1, 3), (777, 777, 4)])
def test_classify_random(ctx, n, k, seed):
    rng = np.random.default_rng(seed)
    grid = rng.uniform(0, 10, size=(n, 3))
    cen = rng.uniform(0, 10, size=(k, 3))
    got = ctx.classify_grid_points(grid, cen).cpu().numpy()
    np.testing.assert_array_equal(got, KO.classify(grid, cen))


def test_classify_lattice_ties(ctx):
    """Centroids on lattice points and midpoints: many exactly equal distances."""
    grid = _lattice(16)
    rng = np.random.default_rng(5)
    pick = rng.choice(grid.shape[0], 300, replace=False)
    cen = grid[pick].copy()
    cen[::3] += (5.0 / 16) / 2  # midpoints -> ties between neighbours
    got = ctx.classify_grid_points(grid, cen).cpu().numpy()
    np.testing.assert_array_equal(got, KO.classify(grid, cen))


def test_classify_duplicate_centroids(ctx):
    """Identical centroids: the first index must win."""
    rng = np.random.default_rng(6)
    grid = rng.uniform(0, 4, size=(3000, 3))
    cen = rng.uniform(0, 4, size=(50, 3))
    cen[30:40] = cen[5]
    got = ctx.classify_grid_points(grid, cen).cpu().numpy()
    ref = KO.classify(grid, cen)
    np.testing.assert_array_equal(got, ref)
    assert not np.isin(np.arange(30, 40), ref).any()


@pytest.mark.parametrize("n,k", [(10000, 64), (50000, 2000)])
def test_update_centroids(ctx, n, k):
    rng = np.random.default_rng(n)
    grid = rng.uniform(0, 7, size=(n, 3))
    rho = rng.uniform(0.1, 2.0, size=n)
    gmap = rng.integers(0, k, size=n).astype(np.int32)
    gmap[:k] = np.arange(k)  # no empty cluster
    got, st = ctx.update_centroids(rho, grid, gmap, k)
    ref, crit = KO.update_centroids(rho, grid, gmap, k)
    assert st == L.THC_OK and crit == 0
    np.testing.assert_array_equal(got.cpu().numpy(), ref)


def test_update_centroids_empty_cluster(ctx):
    rng = np.random.default_rng(9)
    grid = rng.uniform(0, 7, size=(1000, 3))
    rho = rng.uniform(0.1, 2.0, size=1000)
    gmap = rng.integers(0, 10, size=1000).astype(np.int32)
    gmap[gmap == 4] = 5
    _, st = ctx.update_centroids(rho, grid, gmap, 10)
    _, crit = KO.update_centroids(rho, grid, gmap, 10)
    assert crit == 1 and st == L.THC_ERR_EMPTY_CLUSTER


@pytest.mark.parametrize("n,k", [(4096, 40), (30000, 700)])
def test_map_to_grid(ctx, n, k):
    rng = np.random.default_rng(k)
    grid = rng.uniform(0, 9, size=(n, 3))
    cen = rng.uniform(0, 9, size=(k, 3))
    got, st = ctx.map_to_grid(grid, cen)
    ref, dups = KO.map_to_grid(grid, cen)
    np.testing.assert_array_equal(got.cpu().numpy(), ref)
    assert (st == L.THC_WARN_DUPLICATE_INDICES) == (dups > 0)


def test_map_to_grid_lattice(ctx):
    grid = _lattice(20, 6.0)
    rng = np.random.default_rng(3)
    cen = grid[rng.choice(grid.shape[0], 200, replace=False)] + (6.0 / 20) / 2  # equidistant to 8 points
    got, _ = ctx.map_to_grid(grid, cen)
    ref, _ = KO.map_to_grid(grid, cen)
    np.testing.assert_array_equal(got.cpu().numpy(), ref)


@pytest.mark.parametrize("n,k,seed", [(32768, 128, 0), (1000, 7, 3), (65536, 10, 11), (65535, 33, 5),
                                      (262144, 1024, 0), (884736, 3456, 0)])
def test_initial_indices(n, k, seed):
    from thcpp_amd.api import Context
    np.testing.assert_array_equal(Context.initial_indices(n, k, seed), KO.initial_indices(n, k, seed))


def test_kmeans_cfg0_every_iteration(ctx, cfg0, cfg0_kmeans):
    s, ref = cfg0, cfg0_kmeans
    out = ctx.kmeans(s.grid, s.rho, s.K, s.max_it, -1.0, 0, want_map=True, want_centroids=True)
    assert out["status"] == L.THC_OK
    assert out["iters"] == ref["iters"] == s.max_it
    np.testing.assert_array_equal(out["diffs"], ref["diffs"])
    np.testing.assert_array_equal(out["pcts"], ref["pcts"])
    np.testing.assert_array_equal(out["interp_idx"].cpu().numpy(), ref["interp_idx"])
    np.testing.assert_array_equal(out["grid_map"].cpu().numpy(), ref["maps"][-1])
    np.testing.assert_array_equal(out["centroids"].cpu().numpy(), ref["centroids"])


@pytest.mark.parametrize("iters", [1, 2, 7])
def test_kmeans_maps_per_iteration(ctx, cfg0, cfg0_kmeans, iters):
    s, ref = cfg0, cfg0_kmeans
    out = ctx.kmeans(s.grid, s.rho, s.K, iters, -1.0, 0, want_map=True)
    np.testing.assert_array_equal(out["grid_map"].cpu().numpy(), ref["maps"][iters - 1])
    np.testing.assert_array_equal(out["diffs"], ref["diffs"][:iters])


def test_kmeans_graph_vs_eager_and_replay(ctx, ctx_graph, cfg0, cfg0_kmeans):
    s = cfg0
    a = ctx_graph.kmeans(s.grid, s.rho, s.K, s.max_it, -1.0, 0)
    b = ctx_graph.kmeans(s.grid, s.rho, s.K, s.max_it, -1.0, 0)  # graph replay
    c = ctx.kmeans(s.grid, s.rho, s.K, s.max_it, -1.0, 0)
    for o in (a, b, c):
        np.testing.assert_array_equal(o["interp_idx"].cpu().numpy(), cfg0_kmeans["interp_idx"])
        np.testing.assert_array_equal(o["diffs"], cfg0_kmeans["diffs"])


def test_kmeans_threshold_converges(ctx, cfg0, cfg0_kmeans):
    s = cfg0
    thr = float(cfg0_kmeans["diffs"][8]) * 1.0000001
    ref = KO.kmeans(s.grid, s.rho, s.K, 50, thr, 0)
    out = ctx.kmeans(s.grid, s.rho, s.K, 50, thr, 0)
    assert out["iters"] == ref["iters"] and out["converged"]
    np.testing.assert_array_equal(out["interp_idx"].cpu().numpy(), ref["interp_idx"])
    np.testing.assert_array_equal(out["diffs"], ref["diffs"])


def test_kmeans_threshold_equal_leaves_zeros(ctx, cfg0, cfg0_kmeans):
    """diff == threshold on the last iteration: neither branch maps, indices stay 0
    (kmeans.cpp:239,250)."""
    s = cfg0
    it = 6
    thr = float(cfg0_kmeans["diffs"][it - 1])
    ref = KO.kmeans(s.grid, s.rho, s.K, it, thr, 0)
    out = ctx.kmeans(s.grid, s.rho, s.K, it, thr, 0)
    assert not ref["interp_idx"].any()
    assert out["interp_zeroed"] and not out["interp_idx"].cpu().numpy().any()


def test_kmeans_seeds(ctx, cfg0):
    s = cfg0
    for seed in (1, 12345):
        ref = KO.kmeans(s.grid, s.rho, 64, 8, -1.0, seed)
        out = ctx.kmeans(s.grid, s.rho, 64, 8, -1.0, seed)
        np.testing.assert_array_equal(out["interp_idx"].cpu().numpy(), ref["interp_idx"])
        np.testing.assert_array_equal(out["diffs"], ref["diffs"])


def test_kmeans_empty_cluster(ctx):
    """Zero density where a centroid starts -> weight 0 -> the reference exits."""
    g = _lattice(12, 4.0)
    rho = np.ones(g.shape[0])
    idx = KO.initial_indices(g.shape[0], 20, 0)
    first = KO.classify(g, g[idx])
    rho[first == 3] = 0.0
    ref = KO.kmeans(g, rho, 20, 5, -1.0, 0)
    out = ctx.kmeans(g, rho, 20, 5, -1.0, 0)
    assert ref["status"] == 1
    assert out["status"] == L.THC_ERR_EMPTY_CLUSTER


def test_kmeans_full_size_prefix(ctx):
    """The 64-atom, 64^3-grid system with N_mu = 1024: first 4 iterations and the
    points they give, against the oracle."""
    from thcpp_amd import synthetic
    s = synthetic.make_config("diamond64_n64_k1024")
    ref = KO.kmeans(s.grid, s.rho, s.K, 4, -1.0, 0)
    out = ctx.kmeans(s.grid, s.rho, s.K, 4, -1.0, 0)
    np.testing.assert_array_equal(out["diffs"], ref["diffs"])
    np.testing.assert_array_equal(out["pcts"], ref["pcts"])
    np.testing.assert_array_equal(out["interp_idx"].cpu().numpy(), ref["interp_idx"])


def test_device_sqrt_correctly_rounded(ctx):
    """The device FP64 square root of the distance kernels (thc_device_sqrt)
    against numpy's IEEE sqrt, bit for bit: log-uniform arguments over the
    whole normal range, subnormals, exact squares, the doubles next to exact
    squares and next to the midpoints between adjacent square roots (the
    hardest cases for a correctly rounded sqrt), 0, -0 and inf."""
    import ctypes
    rng = np.random.default_rng(11)
    parts = [np.exp(rng.uniform(np.log(1e-300), np.log(1e300), 1 << 20)),
             rng.uniform(0, 1, 1 << 14) * 2.0 ** -1022,  # subnormals
             np.array([0.0, -0.0, np.inf, 1.0, 4.0, 2.0, 0.5, 2.0 ** -1074, np.finfo(np.float64).max])]
    s = np.exp(rng.uniform(np.log(1e-150), np.log(1e150), 1 << 18))
    sq = s * s  # squares of doubles, rounded
    parts += [sq, np.nextafter(sq, 0), np.nextafter(sq, np.inf)]
    mid = (s + np.nextafter(s, np.inf)) / 2  # midpoints between adjacent roots
    m2 = mid * mid
    parts += [m2, np.nextafter(m2, 0), np.nextafter(m2, np.inf)]
    x = np.concatenate(parts)
    x = x[np.isfinite(x) | np.isinf(x)]
    xd = torch.from_numpy(x).to(ctx.device)
    yd = torch.empty_like(xd)
    print("device sqrt: %d arguments" % x.size, flush=True)
    st = L.lib().thc_device_sqrt(ctx.h, ctypes.c_void_p(xd.data_ptr()), x.size, ctypes.c_void_p(yd.data_ptr()))
    assert st == L.THC_OK
    got = yd.cpu().numpy()
    ref = np.sqrt(x)
    bad = np.flatnonzero(got.view(np.int64) != ref.view(np.int64))
    assert bad.size == 0, (bad.size, x[bad[:5]], got[bad[:5]], ref[bad[:5]])
