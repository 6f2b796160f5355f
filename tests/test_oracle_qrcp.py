"""Pins for the QRCP oracle (oracle/qrcp_oracle.py), CPU only: the pivots of
LAPACK's column-pivoted QR of ZT equal a largest-residual Gram-Schmidt written
from the definition; ZT's Gram matrix is the Hadamard square of the pseudo
density matrix (the identity the GPU path is built on), and pivoted Cholesky of
that Gram matrix reproduces the QRCP pivots and |R_jj|."""
import numpy as np
import pytest
import scipy.linalg

from oracle import qrcp_oracle as QO


def _orbitals(n, m, seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((n, m)))
    return q


def _greedy_gram_schmidt(z, k):
    r = np.array(z, dtype=np.float64, copy=True)
    piv, diag = [], []
    for _ in range(k):
        norms = np.einsum("ij,ij->j", r, r)
        norms[piv] = -1.0
        p = int(np.argmax(norms))  # first index of the maximum
        piv.append(p)
        nrm = np.sqrt(norms[p])
        diag.append(nrm)
        q = r[:, p] / nrm
        r -= np.outer(q, q @ r)
    return np.array(piv), np.array(diag)


def _pivoted_cholesky_gram(aoR, k, occ=None):
    """Pivoted Cholesky of G = P o Pocc (the HIP path's algorithm) in numpy."""
    p_full = aoR @ aoR.T
    g = p_full * (occ @ occ.T if occ is not None else p_full)
    d = np.diag(g).copy()
    l = np.zeros((k, g.shape[0]))
    piv = []
    for j in range(k):
        dd = d.copy()
        dd[piv] = -np.inf
        p = int(np.argmax(dd))
        piv.append(p)
        l[j] = (g[:, p] - l[:j].T @ l[:j, p]) / np.sqrt(d[p])
        d -= l[j] ** 2
    return np.array(piv)


@pytest.mark.parametrize("n,m,k,seed", [(300, 5, 10, 0), (500, 7, 20, 1), (240, 6, 15, 2)])
def test_qrcp_matches_greedy_gram_schmidt(n, m, k, seed):
    a = _orbitals(n, m, seed)
    idx, rd, piv = QO.qrcp_points(a, k)
    gpiv, gdiag = _greedy_gram_schmidt(QO.z_matrix(a), k)
    np.testing.assert_array_equal(piv, gpiv)
    np.testing.assert_allclose(rd, gdiag, rtol=1e-10)
    np.testing.assert_array_equal(idx, np.sort(gpiv))
    assert np.all(np.diff(rd) <= 1e-12 * rd[0])  # |R_jj| non-increasing


def test_z_gram_is_hadamard_square_and_cholesky_pivots_agree():
    a = _orbitals(400, 6, 3)
    z = QO.z_matrix(a)
    assert z.shape == (36, 400)
    g = z.T @ z
    np.testing.assert_allclose(g, (a @ a.T) ** 2, rtol=0, atol=1e-15 * np.abs(g).max())
    _, _, piv = QO.qrcp_points(a, 18)
    np.testing.assert_array_equal(_pivoted_cholesky_gram(a, 18), piv)


def test_half_rotated_z_and_pivots():
    a = _orbitals(350, 6, 4)
    occ = a[:, :3]
    z = QO.z_half_matrix(a, occ)
    assert z.shape == (18, 350)
    np.testing.assert_allclose(z.T @ z, (occ @ occ.T) * (a @ a.T), rtol=0, atol=1e-15)
    _, _, piv = QO.qrcp_points(a, 12, occ)
    gpiv, _ = _greedy_gram_schmidt(z, 12)
    np.testing.assert_array_equal(piv, gpiv)
    np.testing.assert_array_equal(_pivoted_cholesky_gram(a, 12, occ), piv)


def _reference_literal_half_z(aoR, aoR_half):
    """ZT as src/qrcp.cpp:49-66 literally fills it: row i*nelec + k (stride
    nelec, not nbasis) for i < nelec, k < nbasis in a matrix of nelec*nbasis
    rows, later writes overwriting earlier ones."""
    n, m = aoR.shape
    mo = aoR_half.shape[1]
    z = np.zeros((mo * m, n))
    for i in range(mo):
        for k in range(m):
            z[i * mo + k] = aoR_half[:, i].conj() * aoR[:, k]
    return z


def test_half_rotated_z_departs_from_reference_indexing():
    """The reference writes the half-rotated ZT with row stride nelec where the
    (i, k) pair index needs nbasis (src/qrcp.cpp:62).  With nbasis > nelec the
    rows overlap: every product with k >= nelec is overwritten except for the
    last occupied orbital, and the last nelec*nbasis - (nelec-1)*nelec - nbasis
    rows stay zero.  The oracle (and the GPU path) use the Khatri-Rao rows the
    comment there states, [ZT]_{(ik)a} = phi_i^occ*(r_a) phi_k(r_a): both are
    shown here, and the pivots they give differ."""
    a = _orbitals(600, 8, 21)
    occ = np.ascontiguousarray(a[:, :3])
    lit = _reference_literal_half_z(a, occ)
    kr = QO.z_half_matrix(a, occ)
    assert kr.shape == lit.shape == (24, 600)
    # rows of the literal matrix: products (i, k < nelec) for i < nelec - 1, then
    # all of (nelec - 1, k), then zeros
    mo, m = 3, 8
    last = (mo - 1) * mo + m
    assert not lit[last:].any()
    np.testing.assert_array_equal(lit[:last], np.concatenate(
        [kr[i * m:i * m + mo] for i in range(mo - 1)] + [kr[(mo - 1) * m:mo * m]]))
    k = 12
    piv_lit = np.sort(scipy.linalg.qr(lit, mode="r", pivoting=True)[1][:k])
    piv_kr, _, _ = QO.qrcp_points(a, k, aoR_half=occ)
    assert not np.array_equal(piv_lit, piv_kr)
    # the literal matrix has rank (nelec - 1) nelec + nbasis = 14 at most
    assert np.linalg.matrix_rank(lit) <= (mo - 1) * mo + m


# ---------------------------------------------------------------- complex orbitals

def _complex_orbitals(n, m, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m)))
    return q


def _greedy_gram_schmidt_complex(z, k):
    r = np.array(z, dtype=np.complex128, copy=True)
    piv, diag = [], []
    for _ in range(k):
        norms = np.einsum("ij,ij->j", r.conj(), r).real
        norms[piv] = -1.0
        p = int(np.argmax(norms))
        piv.append(p)
        nrm = np.sqrt(norms[p])
        diag.append(nrm)
        q = r[:, p] / nrm
        r -= np.outer(q, q.conj() @ r)
    return np.array(piv), np.array(diag)


def _pivoted_cholesky_hermitian(g, k):
    """Pivoted Cholesky G = L L^H with complex columns (the HIP path's
    half-rotated algorithm) in numpy."""
    d = np.diag(g).real.copy()
    l = np.zeros((k, g.shape[0]), dtype=np.complex128)
    piv = []
    for j in range(k):
        dd = d.copy()
        dd[piv] = -np.inf
        p = int(np.argmax(dd))
        piv.append(p)
        l[j] = (g[:, p] - l[:j].T @ l[:j, p].conj()) / np.sqrt(d[p])
        d -= np.abs(l[j]) ** 2
    return np.array(piv)


@pytest.mark.parametrize("half", [False, True])
def test_complex_qrcp_matches_greedy_gram_schmidt_and_gram_identity(half):
    """Complex orbitals: LAPACK's pivots (zgeqp3) equal a largest-residual
    complex Gram-Schmidt; ZT^H ZT = conj(Pocc) o P with P = Phi^* Phi^T
    (|P|^2, real, without half rotation), and pivoted Cholesky of that Gram
    matrix picks the same columns.  Transposing a conjugation (P o P) breaks
    the identity."""
    a = _complex_orbitals(420, 6, 9)
    occ = np.ascontiguousarray(a[:, :3]) if half else None
    k = 14
    z = QO.z_half_matrix(a, occ) if half else QO.z_matrix(a)
    idx, rd, piv = QO.qrcp_points(a, k, occ)
    gpiv, gdiag = _greedy_gram_schmidt_complex(z, k)
    np.testing.assert_array_equal(piv, gpiv)
    np.testing.assert_allclose(rd, gdiag, rtol=1e-10)
    p = a.conj() @ a.T
    po = (occ.conj() @ occ.T) if half else p
    g = z.conj().T @ z
    np.testing.assert_allclose(g, np.conj(po) * p, rtol=0, atol=1e-14 * np.abs(g).max())
    if not half:
        assert np.abs(g.imag).max() < 1e-15 * np.abs(g).max()
    assert np.abs(g - p * p).max() > 1e-3 * np.abs(g).max()
    np.testing.assert_array_equal(_pivoted_cholesky_hermitian(g, k), piv)
