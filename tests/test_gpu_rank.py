"""Numerical rank of CC^T on the GPU (thc_cct_rank, and thc_isdf with
check_rank) against the oracle's restatement of MatrixOperations::rank
(src/matrix_operations.h:713-735, numpy SVD with the same rule), as
IVecs::check_rank applies it (src/interpolating_vectors.cpp:192-208)."""
import os

import numpy as np
import pytest

from oracle import isdf_oracle as IO
from thcpp_amd import _lib as L

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _spectrum_matrix(k, eig, seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((k, k)))
    a = (q * eig) @ q.T
    return 0.5 * (a + a.T)


def test_rank_reference_fixture(ctx):
    """The reference-held etaPm (140 points, 10 orbitals): the Hadamard-squared
    Gram has rank <= 55 and 49 under the reference's rule."""
    eta = np.load(os.path.join(GOLDEN, "etaPm.npy"))
    c = IO.cct(IO.czt(eta, np.arange(eta.shape[0])), np.arange(eta.shape[0]))
    rank, smax, st = ctx.cct_rank(c)
    s = np.linalg.svd(c, compute_uv=False)
    assert rank == IO.rank(c) == 49
    assert st == L.THC_WARN_RANK_DEFICIENT
    assert abs(smax - s[0]) <= 1e-12 * s[0]


@pytest.mark.parametrize("k,r", [(1, 1), (40, 3), (64, 64), (300, 300), (301, 250), (600, 17), (1030, 1000)])
def test_rank_low_rank_products(ctx, k, r):
    """B B^T with B k x r: rank r, the null eigenvalues at rounding level.
    (k >= 40: for k of a few the threshold k s0 1e-16 is itself at rounding
    level, where no two algorithms agree.)"""
    b = np.random.default_rng(k + r).standard_normal((k, r))
    c = b @ b.T
    rank, smax, st = ctx.cct_rank(c)
    assert rank == IO.rank(c) == r
    assert (st == L.THC_WARN_RANK_DEFICIENT) == (r < k)
    s = np.linalg.svd(c, compute_uv=False)
    assert abs(smax - s[0]) <= 1e-12 * s[0]


def test_rank_threshold_sides(ctx):
    """Eigenvalues placed a factor 30 either side of k s0 1e-16, and negative
    ones (|lambda| is the singular value)."""
    k = 400
    thr = k * 1e-16
    eig = np.logspace(0, -8, k)
    eig[-10:] = 30 * thr          # counted
    eig[-20:-10] = thr / 30       # null
    eig[5] = -0.5                 # |eigenvalue| counted
    c = _spectrum_matrix(k, eig, 3)
    rank, _, _ = ctx.cct_rank(c)
    assert rank == IO.rank(c) == k - 10


def test_isdf_check_rank(ctx, cfg0, cfg0_kmeans):
    """thc_isdf reports CC^T's rank beside the solve: config 0 (K = 128 of a
    possible 136) is full rank under the rule; the V it returns is unchanged by
    the check."""
    s = cfg0
    idx = cfg0_kmeans["interp_idx"]
    ref = IO.isdf(s.aoR, idx, s.n, s.coulG, return_all=True)
    want = IO.rank(ref["C"])
    plain = ctx.isdf(s.aoR, idx, s.n, s.coulG)
    out = ctx.isdf(s.aoR, idx, s.n, s.coulG, check_rank=True)
    assert out["cct_rank"] == want
    assert out["status"] == (L.THC_OK if want == s.K else L.THC_WARN_RANK_DEFICIENT)
    assert plain["cct_rank"] == -1
    np.testing.assert_array_equal(out["V"].cpu().numpy(), plain["V"].cpu().numpy())


def test_isdf_check_rank_deficient(ctx):
    """N_mu above M (M + 1) / 2 for real orbitals: CC^T is singular in exact
    arithmetic.  The reference only warns (its QR solve goes on); here the
    FP64 Cholesky fails first, and the rank it reports matches the oracle."""
    from thcpp_amd import synthetic
    n = (16, 16, 16)
    N = int(np.prod(n))
    a = np.linalg.qr(np.random.default_rng(2).standard_normal((N, 6)))[0]
    idx = np.sort(np.random.default_rng(3).choice(N, 30, replace=False)).astype(np.int32)
    cg = synthetic.coulomb_kernel(6.0, n)
    c = IO.cct(IO.czt(a, idx), idx)
    out = ctx.isdf(a, idx, n, cg, check_rank=True)
    assert out["cct_rank"] == IO.rank(c) <= 21
    assert out["status"] in (L.THC_ERR_NOT_POSDEF, L.THC_WARN_RANK_DEFICIENT)
