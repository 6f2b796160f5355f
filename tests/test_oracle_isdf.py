"""Pins for the ISDF oracle (oracle/isdf_oracle.py), CPU only.

Reference-held data: tests/golden/etaPm.npy and orb_prod.npy (from the
reference's tests/data/eri_thc.h5) fix the Hadamard-squared Gram CC^T =
(eta eta^T)^2 = orb_prod orb_prod^T and its numerical rank under the
reference's rank rule.  Mathematics fixes the rest: a brute-force DFT, the
Coulomb energy of a plane wave, invariance of V under a constant shift of a
vector (G = 0 is excluded), interpolation xi_mu(r_nu) = delta, and
Luv Luv^T = V.
"""
import os

import numpy as np
import pytest
import scipy.linalg

from oracle import isdf_oracle as IO
from thcpp_amd import synthetic

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_cct_fixture_hadamard_gram():
    eta = np.load(os.path.join(GOLDEN, "etaPm.npy"))
    prod = np.load(os.path.join(GOLDEN, "orb_prod.npy"))
    idx = np.arange(eta.shape[0])
    z = IO.czt(eta, idx)
    assert rel(z, prod @ prod.T) < 1e-14
    c = IO.cct(z, idx)
    np.testing.assert_array_equal(c, z)
    # M = 10 real orbitals: rank <= M (M + 1) / 2 = 55; the reference's rule gives 49
    assert IO.rank(c) == 49


def test_czt_is_square_of_pseudo_dm():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((300, 7))
    idx = np.array([3, 10, 77, 200])
    p = IO.pseudo_dm(a, idx)
    np.testing.assert_allclose(p, a[idx] @ a.T, rtol=0, atol=1e-14)
    np.testing.assert_array_equal(IO.czt(a, idx), p * p)


def _dft_bruteforce(x, n):
    n0, n1, n2 = n
    k0, k1, k2 = np.meshgrid(np.arange(n0), np.arange(n1), np.arange(n2), indexing="ij")
    r0, r1, r2 = k0.reshape(-1), k1.reshape(-1), k2.reshape(-1)
    ph = np.exp(-2j * np.pi * (np.outer(r0, r0) / n0 + np.outer(r1, r1) / n1 + np.outer(r2, r2) / n2))
    return x @ ph.T  # rows: vectors; columns: G (C order)


@pytest.mark.parametrize("n", [(4, 3, 5), (2, 6, 4), (3, 3, 3)])
def test_fft_vectors_bruteforce(n):
    rng = np.random.default_rng(sum(n))
    x = rng.standard_normal((3, int(np.prod(n))))
    ivg = IO.fft_vectors(x, n)
    assert rel(ivg.T, _dft_bruteforce(x, n)) < 1e-13
    # real input: X(-G) = conj X(G)
    a = ivg[:, 0].reshape(n)
    neg = a[(-np.arange(n[0])) % n[0]][:, (-np.arange(n[1])) % n[1]][:, :, (-np.arange(n[2])) % n[2]]
    assert np.abs(neg - a.conj()).max() < 1e-12


def test_plane_wave_coulomb_closed_form():
    """xi(r) = cos(G0 . r): its transform is N/2 at +-G0, so
    V = 2 (N/2)^2 coulG(G0) exactly."""
    L, n = 7.3, (8, 6, 10)
    N = int(np.prod(n))
    grid = synthetic.uniform_grid(L, n)
    cg = synthetic.coulomb_kernel(L, n)
    m = np.array([1, 2, 3])
    g0 = 2 * np.pi / L * m
    x = np.cos(grid @ g0)[None, :]
    v = IO.construct_muv(IO.fft_vectors(x, n), cg).real[0, 0]
    gidx = (m[0] * n[1] + m[1]) * n[2] + m[2]
    want = 2 * (N / 2) ** 2 * cg[gidx]
    assert abs(v - want) < 1e-12 * want
    assert abs(cg[gidx] - 4 * np.pi / g0.dot(g0) * L ** 3 / N ** 2) < 1e-15 * cg[gidx]


def test_g0_excluded_shift_invariance():
    rng = np.random.default_rng(3)
    n = (6, 5, 4)
    cg = synthetic.coulomb_kernel(5.0, n)
    assert cg[0] == 0.0
    x = rng.standard_normal((4, int(np.prod(n))))
    v1 = IO.construct_muv(IO.fft_vectors(x, n), cg)
    v2 = IO.construct_muv(IO.fft_vectors(x + rng.standard_normal((4, 1)), n), cg)
    assert rel(v2, v1) < 1e-13
    assert np.allclose(v1, v1.conj().T)
    assert np.linalg.eigvalsh(v1.real).min() > -1e-12 * np.abs(v1).max()


def test_least_squares_square_system():
    rng = np.random.default_rng(5)
    b = rng.standard_normal((40, 60))
    c = b @ b.T
    z = rng.standard_normal((40, 500))
    x = IO.least_squares(c, z)
    assert rel(x, scipy.linalg.solve(c, z)) < 1e-10
    assert rel(c @ x, z) < 1e-12


def test_isdf_interpolates_at_its_points():
    """X = CC^T^-1 ZC^T and CC^T = ZC^T[:, idx] give X[:, idx] = I."""
    s = synthetic.make_system(reps=1, n=12, M=8, thc_cfac=2, max_it=1)
    idx = np.sort(np.random.default_rng(1).choice(s.N, s.K, replace=False))
    out = IO.isdf(s.aoR, idx, s.n, s.coulG, return_all=True)
    assert np.abs(out["X"][:, idx] - np.eye(s.K)).max() < 1e-8
    assert out["imag_max"] < 1e-10 * np.abs(out["V"]).max()


def test_cholesky_luv():
    rng = np.random.default_rng(6)
    b = rng.standard_normal((30, 50))
    v = b @ b.T
    l, info = IO.cholesky_luv(v)
    assert info == 0 and rel(l @ l.T, v) < 1e-14 and not np.triu(l, 1).any()
    bad = np.eye(10)
    bad[6, 6] = -2
    _, info = IO.cholesky_luv(bad)
    assert info == 7


# ---------------------------------------------------------------- half-rotated path

def test_czt_half_khatri_rao_fixture():
    """Half-rotated Gram on the reference-held eta: (Pocc o P) restricted to the
    points is (A A^T) o (B B^T) = (A (.) B)(A (.) B)^T, with A the occupied
    columns, B all columns and (.) the row-wise Kronecker product (Schur product
    theorem); with every orbital occupied it is the squared Gram."""
    eta = np.load(os.path.join(GOLDEN, "etaPm.npy"))
    occ = eta[:, : eta.shape[1] // 2]
    idx = np.arange(eta.shape[0])
    z = IO.czt_half(eta, occ, idx)
    kr = np.einsum("ma,mi->mai", occ, eta).reshape(eta.shape[0], -1)
    assert rel(z, kr @ kr.T) < 1e-14
    np.testing.assert_array_equal(IO.czt_half(eta, eta, idx), IO.czt(eta, idx))
    # a transposed or swapped factor breaks the identity
    kr_wrong = np.einsum("ma,mi->mai", occ, eta[::-1]).reshape(eta.shape[0], -1)
    assert rel(z, kr_wrong @ kr_wrong.T) > 1e-3


@pytest.mark.parametrize("n", [(4, 3, 5), (3, 3, 3)])
def test_fft_vectors_backward_bruteforce(n):
    rng = np.random.default_rng(7 + sum(n))
    x = rng.standard_normal((3, int(np.prod(n))))
    ivmg = IO.fft_vectors_backward(x, n)
    # e^{+2 pi i k.r / n}: the conjugate of the forward kernel
    assert rel(ivmg.T, np.conj(_dft_bruteforce(np.conj(x), n))) < 1e-13
    # for real X the backward transform is the conjugate of the forward one
    assert rel(ivmg, np.conj(IO.fft_vectors(x, n))) < 1e-13


def test_isdf_half_reduces_to_full_and_interpolates():
    s = synthetic.make_system(reps=1, n=12, M=8, thc_cfac=2, max_it=1)
    idx = np.sort(np.random.default_rng(1).choice(s.N, s.K, replace=False))
    full = IO.isdf(s.aoR, idx, s.n, s.coulG)
    same = IO.isdf_half(s.aoR, s.aoR, idx, s.n, s.coulG)
    assert rel(same["Muv"], full["V"]) < 1e-12
    occ = s.aoR[:, : s.M // 2]
    half = IO.isdf_half(s.aoR, occ, idx, s.n, s.coulG, return_all=True)
    assert np.abs(half["X"][:, idx] - np.eye(s.K)).max() < 1e-8
    assert half["imag_max"] < 1e-10 * np.abs(half["Muv"]).max()
    m = half["Muv"]
    assert rel(m, m.T) < 1e-13
    assert np.linalg.eigvalsh(m).min() > -1e-10 * np.abs(m).max()


# ---------------------------------------------------------------- complex orbitals

def test_complex_orbitals_phase_invariance():
    """ZC^T = |Phi_mu^H Phi|^2 is unchanged by a phase per grid point
    (phi_i(r) -> e^{i theta(r)} phi_i(r)) and by a phase per orbital, so the
    oracle's complex path must reproduce the real path's V; a conjugation slip
    (Phi_mu^T Phi, which a per-point phase cannot expose) breaks the
    per-orbital case."""
    s = synthetic.make_system(reps=1, n=12, M=8, thc_cfac=2, max_it=1)
    idx = np.sort(np.random.default_rng(1).choice(s.N, s.K, replace=False))
    rng = np.random.default_rng(2)
    theta = rng.uniform(0, 2 * np.pi, s.N)
    alpha = rng.uniform(0, 2 * np.pi, s.M)
    a_pt = s.aoR * np.exp(1j * theta)[:, None]
    a_orb = s.aoR * np.exp(1j * alpha)[None, :]
    z_real = IO.czt(s.aoR, idx)
    for a in (a_pt, a_orb):
        z = IO.czt(a, idx)
        assert np.abs(z.imag).max() < 1e-15 * np.abs(z).max()
        assert rel(z.real, z_real) < 1e-13
    wrong = a_orb[idx] @ a_orb.T
    assert rel(np.abs(wrong) ** 2, z_real) > 1e-3
    v_real = IO.isdf(s.aoR, idx, s.n, s.coulG)["V"]
    v_c = IO.isdf(a_pt, idx, s.n, s.coulG)["V"]
    assert rel(v_c.real, v_real) < 1e-9


# ---------------------------------------------------------------- subset solve at full size

@pytest.mark.parametrize("n,m,k", [(5000, 20, 60), (4096, 12, 70)])
def test_xi_rows_matches_full_qr_solve(n, m, k):
    """Rows of X from the QR solve applied to ZC^T built block by block equal
    the rows of the full pzgels-style solve to rounding; a transposed Q, a
    missing conjugate on the CC^T Gram or a block offset slip fails by orders
    of magnitude."""
    rng = np.random.default_rng(k)
    a = np.linalg.qr(rng.standard_normal((n, m)))[0]
    idx = np.sort(rng.choice(n, k, replace=False))
    z = IO.czt(a, idx)
    c = IO.cct(z, idx)
    x = IO.least_squares(c, z)
    rows = np.sort(rng.choice(k, 15, replace=False))
    xr = IO.xi_rows(a, idx, rows, block=1000)
    assert rel(xr, x[rows]) < 1e-15 * np.linalg.cond(c) + 1e-14


def test_xi_rows_cfg0_within_conditioning_floor():
    """At the 8-atom configuration (cond(CC^T) ~ 4.5e11) the blockwise solve
    reproduces the full solve's rows of xi and the block V[S, S] to rounding."""
    from oracle import kmeans_oracle as KO
    s = synthetic.make_config("diamond8_n32")
    idx = KO.kmeans(s.grid, s.rho, s.K, s.max_it, -1.0, 0)["interp_idx"]
    ref = IO.isdf(s.aoR, idx, s.n, s.coulG, return_all=True)
    rows = np.sort(np.random.default_rng(0).choice(s.K, 32, replace=False))
    xr = IO.xi_rows(s.aoR, idx, rows, block=5000)
    assert rel(xr, ref["X"][rows]) < 1e-12
    v = IO.construct_muv(IO.fft_vectors(xr, s.n), s.coulG).real
    assert rel(v, ref["V"][np.ix_(rows, rows)]) < 1e-12


def test_fft_and_muv_fast_options_agree():
    """The threaded transform and the real-part Muv (the CPU baseline's options)
    give the same numbers as the defaults."""
    n = (8, 6, 10)
    x = np.random.default_rng(4).standard_normal((5, int(np.prod(n))))
    cg = synthetic.coulomb_kernel(5.0, n)
    a = IO.fft_vectors(x, n)
    b = IO.fft_vectors(x, n, workers=2)
    assert rel(b, a) < 1e-15
    m = IO.construct_muv(a, cg)
    assert rel(IO.construct_muv(a, cg, real_part=True), m.real) < 1e-14



# ---------------------------------------------------------------- complex half-rotated path

def _complex_half_case(seed=3, n=(3, 4, 5), m=3, m_occ=2, k=None):
    rng = np.random.default_rng(seed)
    N = int(np.prod(n))
    phi = rng.standard_normal((N, m)) + 1j * rng.standard_normal((N, m))
    occ = rng.standard_normal((N, m_occ)) + 1j * rng.standard_normal((N, m_occ))
    k = k or m * m_occ
    idx = np.sort(rng.choice(N, k, replace=False))
    f = [np.fft.fftfreq(q) * q for q in n]
    g2 = f[0][:, None, None] ** 2 + f[1][None, :, None] ** 2 + f[2][None, None, :] ** 2
    c = np.where(g2 > 0, 1.0 / np.where(g2 > 0, g2, 1.0), 0.0).reshape(-1)
    # W[r, r'] = sum_G c(G) e^{i G.(r - r')} by brute force over the grid
    pts = np.stack(np.meshgrid(*[np.arange(q) for q in n], indexing="ij"), -1).reshape(-1, 3)
    e = np.exp(2j * np.pi * sum(np.outer(pts[:, d], pts[:, d]) / n[d] for d in range(3)))
    w = (e * c[None, :]) @ e.conj().T
    # pair functions rho_ik(r) = phi_occ,i(r) conj(phi_k(r)), one row per (i, k)
    rho = np.einsum("ri,rk->rik", occ, phi.conj()).reshape(N, -1).T
    return phi, occ, idx, n, c, w, rho


def test_complex_half_solves_the_pair_least_squares():
    """Pocc o conj(P) = C^H Z and CC^T = C^H C for the pair functions
    rho_ik(r) = phi_occ,i(r) conj(phi_k(r)) (C = rho at the points, Z = rho on
    the grid), so X must be the least-squares fit rho ~ C X computed from the
    explicit pair matrix; a conjugation slip in P or in the Hadamard product
    fits other functions and fails."""
    phi, occ, idx, n, c, w, rho = _complex_half_case(k=5)
    out = IO.isdf_half(phi, occ, idx, n, c, return_all=True)
    x_ls = np.linalg.lstsq(rho[:, idx], rho, rcond=None)[0]
    assert rel(out["X"], x_ls) < 1e-11
    z_wrong = IO.pseudo_dm(occ, idx) * IO.pseudo_dm(phi, idx)
    assert rel(IO.least_squares(z_wrong[:, idx], z_wrong), x_ls) > 1e-3


def test_complex_half_muv_is_the_coulomb_matrix_of_the_vectors():
    """Muv = IVMG^T IVG with the forward / backward transforms of conj(X) is
    sum_{r, r'} conj(X_mu(r)) W(r - r') conj(X_nu(r')) with the brute-force
    kernel W(d) = sum_G coulG(G) e^{i G.d}; complex symmetric for an even
    coulG.  With as many points as pair functions the interpolation is exact
    and conj(C) Muv conj(C)^T reproduces the brute-force Coulomb integrals of
    the conjugated pairs; Muv of X (no conjugation) does not."""
    phi, occ, idx, n, c, w, rho = _complex_half_case()
    out = IO.isdf_half(phi, occ, idx, n, c, return_all=True)
    muv = out["Muv"]
    v = np.conj(out["X"])
    assert rel(muv, v @ w @ v.T) < 1e-13
    assert rel(muv, muv.T) < 1e-13
    cr = np.conj(rho)
    c_pts = rho[:, idx]
    assert rel(np.conj(c_pts) @ muv @ np.conj(c_pts).T, cr @ w @ cr.T) < 1e-11
    assert rel(out["X"] @ w @ out["X"].T, muv) > 1e-2


def test_complex_half_real_orbitals_reduce_to_real_half():
    s = synthetic.make_system(reps=1, n=8, M=6, thc_cfac=2, max_it=1)
    idx = np.sort(np.random.default_rng(4).choice(s.N, s.K, replace=False))
    occ = s.aoR[:, :3]
    real = IO.isdf_half(s.aoR, occ, idx, s.n, s.coulG)["Muv"]
    cplx = IO.isdf_half(s.aoR.astype(np.complex128), occ.astype(np.complex128), idx, s.n, s.coulG)["Muv"]
    assert rel(cplx.real, real) < 1e-12
    assert np.abs(cplx.imag).max() < 1e-12 * np.abs(real).max()
