"""Pins for the QRCP oracle (oracle/qrcp_oracle.py), CPU only: the pivots of
LAPACK's column-pivoted QR of ZT equal a largest-residual Gram-Schmidt written
from the definition; ZT's Gram matrix is the Hadamard square of the pseudo
density matrix (the identity the GPU path is built on), and pivoted Cholesky of
that Gram matrix reproduces the QRCP pivots and |R_jj|."""
import numpy as np
import pytest

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
