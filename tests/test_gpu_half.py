"""Half-rotated THC path on the GPU (IVecs with rotate = true:
setup_CZt_half src/interpolating_vectors.cpp:143-161, the backward FFT
:384-391, Muv = IVMG^T IVG :258-273, HalfTransformedMuv :317-328) against the
oracle's restatement."""
import numpy as np
import pytest
import torch

from oracle import isdf_oracle as IO
from oracle import kmeans_oracle as KO
from thcpp_amd import _lib as L

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _rand_orbitals(n, m, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, m)))
    return q


@pytest.fixture(scope="module")
def half_case(cfg0):
    """Config 0 with the first M/2 orbitals occupied and N_mu = 4 M points
    (cond(CC^T) = 3.5e8)."""
    s = cfg0
    occ = np.ascontiguousarray(s.aoR[:, : s.M // 2])
    idx = KO.kmeans(s.grid, s.rho, 4 * s.M, s.max_it, -1.0, 0)["interp_idx"]
    return s, occ, idx


@pytest.mark.parametrize("n,m,mo,k", [(4096, 12, 5, 40), (3375, 7, 3, 21), (20000, 64, 32, 300)])
def test_czt_half(ctx, n, m, mo, k):
    a = _rand_orbitals(n, m, n)
    occ = np.ascontiguousarray(a[:, :mo])
    idx = np.sort(np.random.default_rng(k).choice(n, k, replace=False)).astype(np.int32)
    z = ctx.czt_half(a, occ, idx)
    assert rel(z.cpu(), IO.czt_half(a, occ, idx)) < 1e-14


@pytest.mark.parametrize("prec", [L.PREC_FP64, L.PREC_MIXED])
def test_isdf_half_all_occupied_equals_full(ctx, cfg0, cfg0_kmeans, prec):
    """With every orbital occupied, Pocc o P = P o P: the half-rotated Muv is
    V of the full path, bit for bit (same kernels, product order unchanged).
    Mixed: the full path writes the solve's ZC^T digits from the ZC^T GEMM's
    epilogue, the half-rotated path splits the stored P o Pocc afterwards
    (and its scale sqrt(rho rho) is rho exactly), so this also shows that the
    two give the same digits."""
    s = cfg0
    idx = cfg0_kmeans["interp_idx"]
    full = ctx.isdf(s.aoR, idx, s.n, s.coulG, precision=prec, want_luv=False)
    half = ctx.isdf_half(s.aoR, s.aoR, idx, s.n, s.coulG, precision=prec)
    assert full["status"] == L.THC_OK and half["status"] == L.THC_OK
    assert torch.equal(half["Muv"], full["V"])


def test_isdf_half_vs_oracle(ctx, half_case):
    s, occ, idx = half_case
    ref = IO.isdf_half(s.aoR, occ, idx, s.n, s.coulG, return_all=True)
    out = ctx.isdf_half(s.aoR, occ, idx, s.n, s.coulG, want_xi=True)
    assert out["status"] == L.THC_OK
    e = rel(out["Muv"].cpu(), ref["Muv"])
    ex = rel(out["xi"].cpu(), ref["X"])
    print("half-rotated cfg0 K=%d rel Muv %.3e rel xi %.3e" % (len(idx), e, ex))
    # QR (oracle) vs Cholesky-based inverse (GPU) at cond 3.5e8
    assert e < 1e-7 and ex < 1e-6
    # xi_mu(r_nu) = delta: CC^T (formed from Phi_mu beside the ZC^T GEMM) must be
    # the columns of ZC^T at the points
    xi = out["xi"].cpu().numpy()
    assert np.abs(xi[:, idx] - np.eye(len(idx))).max() < 1e-6
    m = out["Muv"].cpu().numpy()
    np.testing.assert_array_equal(m, m.T)


def test_isdf_half_mixed(ctx, half_case):
    """Mixed precision on the half-rotated pass (int8 digit solve with the
    Pocc o P scaling, FP32 FFT, int8 V) against the oracle: within 1e-6
    beyond the FP64 path's distance to it."""
    s, occ, idx = half_case
    ref = IO.isdf_half(s.aoR, occ, idx, s.n, s.coulG)
    fp64 = ctx.isdf_half(s.aoR, occ, idx, s.n, s.coulG)
    out = ctx.isdf_half(s.aoR, occ, idx, s.n, s.coulG, precision=L.PREC_MIXED)
    assert out["status"] == L.THC_OK
    e_mx, e_64 = rel(out["Muv"].cpu(), ref["Muv"]), rel(fp64["Muv"].cpu(), ref["Muv"])
    print("half-rotated rel Muv vs oracle: mixed %.3e, FP64 %.3e" % (e_mx, e_64))
    assert e_mx < e_64 + 1e-6


def test_isdf_half_rejects_empty_occupied_set(ctx, half_case):
    s, occ, idx = half_case
    out = ctx.isdf_half(s.aoR, np.zeros((s.N, 0)), idx, s.n, s.coulG)
    assert out["status"] == L.THC_ERR_ARG
