"""ISDF stages on the GPU against closed forms and the oracle's restatement
of src/interpolating_vectors.cpp."""
import numpy as np
import pytest
import scipy.linalg
import torch

from oracle import isdf_oracle as IO
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


@pytest.mark.parametrize("n,m,k", [(4096, 12, 40), (3375, 7, 130), (20000, 64, 300)])
def test_czt_cct(ctx, n, m, k):
    a = _rand_orbitals(n, m, n)
    idx = np.sort(np.random.default_rng(k).choice(n, k, replace=False)).astype(np.int32)
    z = ctx.czt(a, idx)
    zr = IO.czt(a, idx)
    assert rel(z.cpu(), zr) < 1e-14
    c = ctx.cct(z, idx).cpu().numpy()
    np.testing.assert_array_equal(c, z.cpu().numpy()[:, idx])


@pytest.mark.parametrize("k", [1, 63, 64, 65, 128, 200, 777, 1100, 2048, 2113])
def test_cholesky(ctx, k):
    rng = np.random.default_rng(k)
    b = rng.standard_normal((k, k + 10))
    a = b @ b.T / k + np.eye(k) * 0.1
    l, info = ctx.cholesky(a)
    assert info == 0
    lr = np.linalg.cholesky(a)
    assert rel(l.cpu(), lr) < 1e-12
    assert not np.triu(l.cpu().numpy(), 1).any()


def test_cholesky_ill_conditioned(ctx):
    """Eigenvalues spread over 1e-11 .. 1: the backward error L L^T - A stays
    at rounding level whatever the conditioning."""
    k = 700
    rng = np.random.default_rng(9)
    q, _ = np.linalg.qr(rng.standard_normal((k, k)))
    a = (q * np.logspace(-11, 0, k)) @ q.T
    a = 0.5 * (a + a.T)
    l, info = ctx.cholesky(a)
    assert info == 0
    l = l.cpu().numpy()
    assert np.linalg.norm(l @ l.T - a) / np.linalg.norm(a) < 1e-14


@pytest.mark.parametrize("k,bad", [(100, 70), (64, 63), (300, 0), (300, 5), (300, 200), (257, 256)])
def test_cholesky_not_posdef(ctx, k, bad):
    """info is the order of the first non-positive leading minor (LAPACK potrf)."""
    a = np.eye(k)
    a[bad, bad] = -1.0
    _, info = ctx.cholesky(a)
    assert info == bad + 1


@pytest.mark.parametrize("k,n", [(100, 5000), (513, 20000), (1100, 3000), (2048, 2500)])
def test_least_squares(ctx, k, n):
    rng = np.random.default_rng(k)
    b = rng.standard_normal((k, 3 * k))
    c = b @ b.T / k + np.eye(k)
    z = rng.standard_normal((k, n))
    x = ctx.least_squares(c, z).cpu().numpy()
    xr = scipy.linalg.solve(c, z, assume_a="pos")
    assert rel(x, xr) < 1e-12


def _half_spectrum_ref(x, n, coulg):
    k = x.shape[0]
    f = np.fft.rfftn(x.reshape(k, *n), axes=(1, 2, 3))
    kz = np.arange(n[2] // 2 + 1)
    w = np.where((kz == 0) | (2 * kz == n[2]), 1.0, 2.0)
    cg = coulg.reshape(n)[:, :, : n[2] // 2 + 1] * w
    f = f * np.sqrt(cg)[None]
    return np.stack([f.real, f.imag], axis=-1).reshape(k, -1)


@pytest.mark.parametrize("grid", [(8, 8, 8), (12, 10, 9), (16, 16, 16), (32, 32, 32), (24, 20, 30),
                                  (27, 25, 45), (64, 64, 64), (5, 3, 2), (1, 4, 6), (96, 96, 96), (48, 45, 40),
                                  (24, 24, 24), (25, 25, 25), (27, 27, 27), (30, 30, 30), (36, 36, 36),
                                  (40, 40, 40), (45, 45, 45), (48, 48, 48)])
def test_fft_vectors(ctx, grid):
    from thcpp_amd import synthetic
    rng = np.random.default_rng(sum(grid))
    k = 5
    n = int(np.prod(grid))
    x = rng.standard_normal((k, n))
    cg = synthetic.coulomb_kernel(7.0, grid)
    got = ctx.fft_vectors(x, grid, cg).cpu().numpy()
    ref = _half_spectrum_ref(x, grid, cg)
    assert rel(got, ref) < 1e-13


def test_fft_vectors_unaligned_output(ctx):
    """An output that is 8- but not 16-byte aligned takes scalar stores and gives
    the same bits as an aligned one."""
    import ctypes
    from thcpp_amd import synthetic
    grid = (12, 10, 9)
    rng = np.random.default_rng(3)
    x = ctx.f64(rng.standard_normal((4, int(np.prod(grid)))))
    cg = ctx.f64(synthetic.coulomb_kernel(7.0, grid)).reshape(-1)
    g2 = 2 * grid[0] * grid[1] * (grid[2] // 2 + 1)
    ref = ctx.fft_vectors(x, grid, cg)
    buf = torch.zeros(4 * g2 + 1, dtype=torch.float64, device=ctx.device)
    out = buf[1:]
    assert out.data_ptr() % 16 == 8
    n = (ctypes.c_int * 3)(*grid)
    st = L.lib().thc_fft_vectors(ctx.h, ctypes.c_void_p(x.data_ptr()), 4, n, ctypes.c_void_p(cg.data_ptr()),
                                 ctypes.c_void_p(out.data_ptr()))
    assert st == L.THC_OK
    assert torch.equal(out.reshape(4, g2).cpu(), ref.cpu())


def test_fft_vectors_matches_oracle_muv(ctx):
    """sum over the full grid of conj(A_mu) coulG A_nu (oracle, complex, reference
    order) equals the real Gram of the weighted half spectrum."""
    from thcpp_amd import synthetic
    grid = (12, 10, 8)
    rng = np.random.default_rng(0)
    x = rng.standard_normal((6, int(np.prod(grid))))
    cg = synthetic.coulomb_kernel(6.0, grid)
    b = ctx.fft_vectors(x, grid, cg).cpu().numpy()
    muv = IO.construct_muv(IO.fft_vectors(x, grid), cg)
    assert np.abs(muv.imag).max() < 1e-10 * np.abs(muv).max()
    assert rel(b @ b.T, muv.real) < 1e-13


@pytest.mark.parametrize("k,g2", [(50, 1000), (300, 70000), (1030, 5000), (520, 64000)])
def test_construct_muv(ctx, k, g2):
    """FP64 V = B B^T by split-K partial tiles; G2 a multiple of 16 (70000, 64000)
    takes the direct-to-LDS tile (520 rows: an 8-row last tile row, 134 splits of
    480 columns with a 160-column last one)."""
    rng = np.random.default_rng(k)
    b = rng.standard_normal((k, g2))
    v = ctx.construct_muv(b).cpu().numpy()
    assert rel(v, b @ b.T) < 1e-13
    np.testing.assert_array_equal(v, v.T)


def test_isdf_cfg0_vs_oracle(ctx, cfg0, cfg0_kmeans):
    s = cfg0
    idx = cfg0_kmeans["interp_idx"]
    ref = IO.isdf(s.aoR, idx, s.n, s.coulG, return_all=True)
    out = ctx.isdf(s.aoR, idx, s.n, s.coulG, want_xi=True)
    assert out["status"] == L.THC_OK
    ev = rel(out["V"].cpu(), ref["V"])
    ex = rel(out["xi"].cpu(), ref["X"])
    print("cfg0 FP64 rel V %.3e rel xi %.3e" % (ev, ex))
    # CC^T has condition number ~4.5e11 here; two FP64 solvers (QR vs Cholesky)
    # already differ by ~7e-8 in V.
    assert ev < 1e-6 and ex < 1e-5
    l = out["Luv"].cpu().numpy()
    assert rel(l @ l.T, out["V"].cpu()) < 1e-13


def test_isdf_fp32_cfg0(ctx, cfg0, cfg0_kmeans):
    s = cfg0
    idx = cfg0_kmeans["interp_idx"]
    ref = ctx.isdf(s.aoR, idx, s.n, s.coulG)
    out = ctx.isdf(s.aoR, idx, s.n, s.coulG, precision=L.PREC_FP32)
    e = rel(out["V"].cpu(), ref["V"].cpu())
    print("cfg0 FP32 FFT+contraction rel V vs FP64 %.3e" % e)
    assert e < 1e-5


def test_factorize_cfg0(ctx, cfg0, cfg0_kmeans):
    s = cfg0
    out = ctx.factorize(s.grid, s.rho, s.aoR, s.thc_cfac, s.max_it, s.n, s.coulG)
    assert out["status"] == L.THC_OK
    np.testing.assert_array_equal(out["interp_idx"].cpu().numpy(), cfg0_kmeans["interp_idx"])
    ref = IO.isdf(s.aoR, cfg0_kmeans["interp_idx"], s.n, s.coulG)
    assert rel(out["V"].cpu(), ref["V"]) < 1e-6


def test_isdf_refinement_cfg0(ctx, cfg0, cfg0_kmeans):
    """FP64 residual refinement (R = ZC^T - CC^T X, X += CC^T^-1 R) keeps V at the
    FP64 conditioning floor and moves xi towards the oracle's QR solution."""
    s = cfg0
    idx = cfg0_kmeans["interp_idx"]
    ref = IO.isdf(s.aoR, idx, s.n, s.coulG, return_all=True)
    r0 = ctx.isdf(s.aoR, idx, s.n, s.coulG, want_xi=True)
    r1 = ctx.isdf(s.aoR, idx, s.n, s.coulG, refine=1, want_xi=True)
    assert r1["status"] == L.THC_OK
    e0, e1 = rel(r0["V"].cpu(), ref["V"]), rel(r1["V"].cpu(), ref["V"])
    print("cfg0 rel V no refinement %.3e, one step %.3e" % (e0, e1))
    assert e1 < 1e-6
    c = IO.cct(ref["Z"], idx)
    x1 = r1["xi"].cpu().numpy()
    assert rel(c @ x1, ref["Z"]) <= 1.5 * rel(c @ r0["xi"].cpu().numpy(), ref["Z"]) + 1e-14


def assert_mixed_vs_oracle(v_mixed, v_fp64, v_oracle, label, target=1e-6):
    """The mixed path's V against the oracle's FP64 restatement: within the
    target beyond the FP64 path's own distance to the oracle (the conditioning
    floor of CC^T, two FP64 solvers of the same system)."""
    e_mx = rel(v_mixed, v_oracle)
    e_64 = rel(v_fp64, v_oracle)
    print("%s: rel V vs oracle mixed %.3e, FP64 %.3e" % (label, e_mx, e_64))
    assert e_mx < e_64 + target, (e_mx, e_64)


def test_isdf_mixed_int8_cfg0(ctx, cfg0, cfg0_kmeans):
    """Mixed path (int8 digit solve, FP32 FFT, int8 digit V) against the
    oracle: within 1e-6 beyond the FP64 floor."""
    s = cfg0
    idx = cfg0_kmeans["interp_idx"]
    ref = IO.isdf(s.aoR, idx, s.n, s.coulG)
    fp64 = ctx.isdf(s.aoR, idx, s.n, s.coulG)
    out = ctx.isdf(s.aoR, idx, s.n, s.coulG, precision=L.PREC_MIXED)
    assert out["status"] == L.THC_OK
    assert_mixed_vs_oracle(out["V"].cpu(), fp64["V"].cpu(), ref["V"], "cfg0 mixed")
    v = out["V"].cpu().numpy()
    np.testing.assert_array_equal(v, v.T)


@pytest.mark.parametrize("n,m,k", [((15, 12, 10), 9, 37), ((40, 36, 30), 24, 150), ((72, 6, 80), 24, 150)])
def test_isdf_mixed_uneven_sizes(ctx, n, m, k):
    """K not a multiple of the 128 tile (and spanning two tiles), grids with odd
    edges, grid lengths that are not a multiple of the 64-column solve tile, and
    a y-line of the weighted spectrum (2 * 72 * 41 = 5904 values, 5952 with its
    padding) longer than one chunk of the int8 V, so cut into three chunks that
    share the line's exponent; against the oracle."""
    from thcpp_amd import synthetic
    rng = np.random.default_rng(7)
    N = int(np.prod(n))
    a = np.linalg.qr(rng.standard_normal((N, m)))[0]
    idx = np.sort(rng.choice(N, k, replace=False)).astype(np.int32)
    cg = synthetic.coulomb_kernel(6.5, n)
    ref = IO.isdf(a, idx, n, cg)
    fp64 = ctx.isdf(a, idx, n, cg)
    out = ctx.isdf(a, idx, n, cg, precision=L.PREC_MIXED)
    assert out["status"] in (L.THC_OK, L.THC_WARN_LUV_NOT_POSDEF)
    assert_mixed_vs_oracle(out["V"].cpu(), fp64["V"].cpu(), ref["V"], "uneven %s K=%d mixed" % (n, k))


@pytest.mark.parametrize("prec", [L.PREC_MIXED, L.PREC_FP32])
@pytest.mark.parametrize("n,m,k", [((15, 12, 10), 9, 37), ((40, 36, 30), 24, 150), ((40, 36, 30), 64, 272)])
def test_isdf_f32_x_store_bitwise(ctx, prec, n, m, k):
    """Below FP64 the solve stores X as FP32 when no xi output is asked for (the
    FFT rounds X to FP32 as it reads it either way), so V with and without the
    FP64 xi output is bitwise the same."""
    from thcpp_amd import synthetic
    rng = np.random.default_rng(11)
    N = int(np.prod(n))
    a = np.linalg.qr(rng.standard_normal((N, m)))[0]
    idx = np.sort(rng.choice(N, k, replace=False)).astype(np.int32)
    cg = synthetic.coulomb_kernel(6.5, n)
    with_xi = ctx.isdf(a, idx, n, cg, precision=prec, want_xi=True)
    without = ctx.isdf(a, idx, n, cg, precision=prec)
    assert with_xi["status"] == without["status"]
    assert with_xi["status"] in (L.THC_OK, L.THC_WARN_LUV_NOT_POSDEF)
    assert torch.equal(with_xi["V"].cpu(), without["V"].cpu())


@pytest.mark.parametrize("n,k", [((40, 36, 30), 256), ((40, 36, 30), 272), ((50, 40, 36), 272)])
def test_isdf_solve_large_tiles(ctx, n, k):
    """The solve at sizes that take the direct-to-LDS 128 x 128 x 16 FP64 tile (A
    and B contiguous along k, k a multiple of 16, at least 512 tiles): k = 272
    leaves a 16-row last tile row, N = 43200 a 64-column last tile column, and
    N = 72000 (>= 65536) also takes the one-workgroup-per-CU Zt GEMM.  xi against
    scipy's solve of CC^T X = ZC^T, within the conditioning of CC^T."""
    from thcpp_amd import synthetic
    m = 64
    N = int(np.prod(n))
    rng = np.random.default_rng(k)
    a = np.linalg.qr(rng.standard_normal((N, m)))[0]
    idx = np.sort(rng.choice(N, k, replace=False)).astype(np.int32)
    cg = synthetic.coulomb_kernel(6.5, n)
    out = ctx.isdf(a, idx, n, cg, want_xi=True)
    assert out["status"] in (L.THC_OK, L.THC_WARN_LUV_NOT_POSDEF)
    z = IO.czt(a, idx)
    c = IO.cct(z, idx)
    xr = scipy.linalg.solve(c, z, assume_a="pos")
    cond = np.linalg.cond(c)
    e = rel(out["xi"].cpu().numpy(), xr)
    print("k %d cond(CC^T) %.2e rel xi %.3e" % (k, cond, e))
    assert e < max(1e-13, 1e-15 * cond) and e < 1e-6
