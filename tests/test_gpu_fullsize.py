"""The hot path at the sizes the benchmarks are quoted on, against the oracle.

- k-means on the 64-atom system (64^3 grid, N_mu = 1024 and 2048, 30
  iterations): every normed difference, every change percentage, the final
  assignment and the interpolation points bit-identical to the oracle's
  restatement of src/kmeans.cpp:187-264.  On the 216-atom system (96^3 grid,
  N_mu = 3456) the first two iterations and their points.
- ISDF: rows S (a seeded subset of 128 interpolation points) of xi and the
  block V[S, S], against the oracle's QR solve (pzgels, matrix_operations.h:
  184-202) applied to ZC^T built block by block of grid columns
  (oracle.isdf_oracle.xi_rows), then its FFT and Coulomb contraction
  (interpolating_vectors.cpp:331-404, 247-281).
- The sqrt-merge and sqrt-boundary cases of classify_grid_points and
  map_to_grid (src/kmeans.cpp:96-104, 146-157), which only a correctly rounded
  FP64 square root and a first-index rule on the rounded distances pass.
"""
import numpy as np
import pytest
import torch

from oracle import isdf_oracle as IO
from oracle import kmeans_oracle as KO
from thcpp_amd import _lib as L
from thcpp_amd import synthetic

pytestmark = pytest.mark.gpu

_SYSTEMS = {}


def system(name):
    if name not in _SYSTEMS:
        _SYSTEMS[name] = synthetic.make_config(name)
    return _SYSTEMS[name]


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ---------------------------------------------------------------- sqrt merges and boundaries

Q = 2.0 ** -46  # offsets on this grid keep (P + x) - P exact for lattice points P < 2^6


def _distance_pairs(rng, count):
    """Pairs of in-plane offsets (x1, y1), (x2, y2) on the Q grid whose squared
    distances, as the reference rounds them ((x*x + y*y) + 0*0), are adjacent
    doubles: 'merge' pairs share one correctly rounded sqrt, 'split' pairs
    straddle a rounding boundary of sqrt.  Offsets are aimed at a radius^2
    target T (y random, x = sqrt(T - y^2) on the grid), which lands d^2 within
    about a hundred ulps of T, so a few thousand draws give many adjacent pairs.
    Returns (merge, split), each a list of (offset_small_d2, offset_large_d2)."""
    merge, split = [], []
    for target in (1.7, 2.3, 3.1, 1.1, 2.9, 3.7, 1.3, 2.1, 3.3, 1.9, 2.7, 3.5):
        y = np.round(rng.uniform(0.3, 1.0, size=4000) / Q) * Q
        x = np.round(np.sqrt(target - y * y) / Q) * Q
        d2 = x * x + y * y
        d2u, first = np.unique(d2, return_index=True)
        for i in range(len(d2u) - 1):
            a, b = d2u[i], d2u[i + 1]
            if np.nextafter(a, np.inf) != b:
                continue
            pair = (np.array([x[first[i]], y[first[i]]]), np.array([x[first[i + 1]], y[first[i + 1]]]))
            (merge if np.sqrt(a) == np.sqrt(b) else split).append(pair)
    assert len(merge) >= count and len(split) >= count, (len(merge), len(split))
    order_m = rng.permutation(len(merge))[:count]
    order_s = rng.permutation(len(split))[:count]
    return [merge[i] for i in order_m], [split[i] for i in order_s]


def _lattice_points(count):
    side = int(np.ceil(count ** (1 / 3)))
    assert side <= 8
    i, j, k = np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij")
    return (np.stack([i, j, k], -1).reshape(-1, 3)[:count] * 8.0).astype(np.float64)


def _two_candidate_cases(pairs, big_first):
    """One lattice point per pair, and its two candidates at P + offset: the
    larger squared distance listed first when big_first."""
    pts = _lattice_points(len(pairs))
    cands = []
    for p, (small, large) in zip(pts, pairs):
        first, second = (large, small) if big_first else (small, large)
        cands.append(p + np.array([first[0], first[1], 0.0]))
        cands.append(p + np.array([second[0], second[1], 0.0]))
    cands = np.array(cands)
    # the offsets are exact: the kernel sees exactly the intended dx, dy
    for q, (small, large) in enumerate(pairs):
        d = cands[2 * q:2 * q + 2] - pts[q]
        want = np.array([large, small] if big_first else [small, large])
        assert np.all(d[:, 2] == 0) and np.array_equal(d[:, :2], want)
    return pts, cands


@pytest.mark.parametrize("big_first", [True, False])
def test_classify_sqrt_merge_and_boundary(ctx, big_first):
    """Merge pairs: the first-listed centroid wins even with the larger d^2.
    Boundary pairs: the smaller d^2 wins, which a sqrt rounded the wrong way
    at either argument would turn into a tie won by the first index."""
    merge, split = _distance_pairs(np.random.default_rng(1 + big_first), 500)
    for pairs, kind in ((merge, "merge"), (split, "split")):
        pts, cen = _two_candidate_cases(pairs, big_first)
        got = ctx.classify_grid_points(pts, cen).cpu().numpy()
        ref = KO.classify(pts, cen)
        np.testing.assert_array_equal(got, ref)
        first = 2 * np.arange(len(pairs))
        if kind == "merge":
            np.testing.assert_array_equal(ref, first)  # tie in sqrt -> first index
        else:
            np.testing.assert_array_equal(ref, first + (1 if big_first else 0))


@pytest.mark.parametrize("big_first", [True, False])
def test_map_to_grid_sqrt_merge_and_boundary(ctx, big_first):
    """map_to_grid swaps the roles (kmeans.cpp:149): each centroid picks among
    grid points, the grid points here being the two candidates per lattice
    point; a tie in sqrt goes to the lower grid index."""
    merge, split = _distance_pairs(np.random.default_rng(3 + big_first), 400)
    for pairs in (merge, split):
        cen, grid = _two_candidate_cases(pairs, big_first)
        got, _ = ctx.map_to_grid(grid, cen)
        ref, _ = KO.map_to_grid(grid, cen)
        np.testing.assert_array_equal(got.cpu().numpy(), ref)


# ---------------------------------------------------------------- k-means at full size

@pytest.mark.timeout(600)
@pytest.mark.parametrize("name", ["diamond64_n64_k1024", "diamond64_n64_k2048"])
def test_kmeans_full_size(ctx, name):
    s = system(name)
    ref = KO.kmeans(s.grid, s.rho, s.K, s.max_it, -1.0, 0, keep_maps=True)
    out = ctx.kmeans(s.grid, s.rho, s.K, s.max_it, -1.0, 0, want_map=True, want_centroids=True)
    assert out["status"] == L.THC_OK and ref["status"] == 0
    assert out["iters"] == ref["iters"] == s.max_it
    np.testing.assert_array_equal(out["diffs"], ref["diffs"])
    np.testing.assert_array_equal(out["pcts"], ref["pcts"])
    np.testing.assert_array_equal(out["centroids"].cpu().numpy(), ref["centroids"])
    np.testing.assert_array_equal(out["interp_idx"].cpu().numpy(), ref["interp_idx"])
    np.testing.assert_array_equal(out["grid_map"].cpu().numpy(), ref["maps"][-1])
    _SYSTEMS[name + "/idx"] = ref["interp_idx"]


@pytest.mark.timeout(600)
def test_kmeans_large_prefix(ctx):
    """216 atoms, 96^3 grid, N_mu = 3456: two iterations and their points."""
    s = system("diamond216_n96_k3456")
    ref = KO.kmeans(s.grid, s.rho, s.K, 2, -1.0, 0)
    out = ctx.kmeans(s.grid, s.rho, s.K, 2, -1.0, 0)
    assert out["status"] == L.THC_OK
    np.testing.assert_array_equal(out["diffs"], ref["diffs"])
    np.testing.assert_array_equal(out["pcts"], ref["pcts"])
    np.testing.assert_array_equal(out["interp_idx"].cpu().numpy(), ref["interp_idx"])
    _SYSTEMS["diamond216_n96_k3456/idx"] = ref["interp_idx"]


# ---------------------------------------------------------------- ISDF subset at full size

# (configuration, precision, xi bound, V[S, S] bound).  FP64: within the FP64
# conditioning floor of CC^T against the oracle's QR (two FP64 solvers of the
# same system differ there).  Mixed: the 1e-6 V target.
ISDF_CASES = [
    ("diamond64_n64_k1024", L.PREC_FP64, 1e-8, 1e-8),
    ("diamond64_n64_k2048", L.PREC_FP64, 1e-8, 1e-8),
    ("diamond64_n64_k2048", L.PREC_MIXED, None, 1e-6),
    ("diamond216_n96_k3456", L.PREC_MIXED, None, 1e-6),
]


def _subset_reference(s, idx, rows):
    xs = IO.xi_rows(s.aoR, idx, rows)
    vss = IO.construct_muv(IO.fft_vectors(xs, s.n), s.coulG).real
    return xs, vss


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name,prec,xi_tol,v_tol", ISDF_CASES)
def test_isdf_subset_vs_oracle(ctx, name, prec, xi_tol, v_tol):
    s = system(name)
    idx = _SYSTEMS.get(name + "/idx")
    if idx is None:
        it = 2 if name == "diamond216_n96_k3456" else s.max_it
        idx = KO.kmeans(s.grid, s.rho, s.K, it, -1.0, 0)["interp_idx"]
        _SYSTEMS[name + "/idx"] = idx
    rows = np.sort(np.random.default_rng(s.K).choice(s.K, 128, replace=False))
    key = "%s/ref%d" % (name, s.K)
    if key not in _SYSTEMS:
        _SYSTEMS[key] = _subset_reference(s, idx, rows)
    xs_ref, vss_ref = _SYSTEMS[key]
    out = ctx.isdf(s.aoR, idx, s.n, s.coulG, precision=prec, want_xi=xi_tol is not None)
    assert out["status"] == L.THC_OK, out["status"]
    r = torch.from_numpy(rows).to(ctx.device)
    vss = out["V"][r][:, r].cpu().numpy()
    ev = rel(vss, vss_ref)
    msg = "%s prec %d: rel V[S,S] vs oracle %.3e" % (name, prec, ev)
    if xi_tol is not None:
        ex = rel(out["xi"][r].cpu().numpy(), xs_ref)
        msg += ", rel xi[S] vs oracle %.3e" % ex
        assert ex < xi_tol, msg
    print(msg)
    assert ev < v_tol, msg
    v = out["V"].cpu().numpy()
    np.testing.assert_array_equal(v, v.T)
