"""ORACLE - TEST INFRASTRUCTURE ONLY.

FP64 CPU restatement of the reference's ISDF interpolating-vector and Coulomb
stages, used as the checker for the HIP path and as the CPU baseline.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import it;
the product package never does.

It follows the reference's algorithm and its complex arithmetic (the orbitals
are real here, so every imaginary part is zero, but the steps are the same):

  pseudo_dm          src/interpolating_vectors.cpp:59-125  P = aoR_mu^H aoR (pzgemm 'C','N')
  czt                src/interpolating_vectors.cpp:127-141 ZC^T = P * conj(P)
  cct                src/interpolating_vectors.cpp:163-190 CC^T = ZC^T[:, idx]
  rank               src/matrix_operations.h:713-735       SVD rank, rcond = max(m,n) s0 1e-16
  least_squares      src/matrix_operations.h:184-202       pzgels('N'): QR of CC^T, X = R^-1 Q^H ZC^T
  fft_vectors        src/interpolating_vectors.cpp:331-404 X^H, unnormalised forward 3-D DFT per mu
  construct_muv      src/interpolating_vectors.cpp:247-281 A = IVG sqrt(coulG); Muv = A^H A (lower)
  cholesky_luv       src/interpolating_vectors.cpp:283-315 pzpotrf('L'), upper triangle zeroed
  czt_half           src/interpolating_vectors.cpp:143-161 ZC^T = Pocc * conj(P) (half-rotated)
  fft_vectors_backward src/interpolating_vectors.cpp:384-391 unnormalised FFTW_BACKWARD of X^H
  construct_muv_half src/interpolating_vectors.cpp:258-273 Muv = IVMG^T IVG, both times sqrt(coulG)
  isdf_half          src/main.cpp:63-72 + IVecs(rotate=true): HalfTransformedMuv, no Cholesky
  xi_rows            rows of least_squares' X at full size without forming ZC^T: the
                     QR of CC^T applied as pzgels applies it (Q^H, then R^-1) to ZC^T
                     built block by block of grid columns

Parity pins (tests/test_oracle_isdf.py): numpy/scipy closed forms on small
inputs, brute-force DFT (both signs), plane-wave Coulomb closed form, the G=0
shift invariance, the reference-held etaPm fixture for the Hadamard-squared Gram
and, for the half-rotated Gram, the Khatri-Rao identity
(A A^T) o (B B^T) = (A (.) B)(A (.) B)^T on the same fixture.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg


def pseudo_dm(aoR: np.ndarray, idx: np.ndarray) -> np.ndarray:
    """P[mu, r] = sum_i conj(phi_i(r_mu)) phi_i(r)  (interpolating_vectors.cpp:96-117).

    aoR is (N, M) C-order as stored in the orbital file (rows = grid points).
    """
    aoR = np.asarray(aoR)
    mu = aoR[np.asarray(idx)]
    return mu.conj() @ aoR.T


def czt(aoR: np.ndarray, idx: np.ndarray) -> np.ndarray:
    """ZC^T = P * conj(P), elementwise (interpolating_vectors.cpp:133-135)."""
    p = pseudo_dm(aoR, idx)
    z = p * p.conj()
    return z.real if np.isrealobj(aoR) else z


def cct(z: np.ndarray, idx: np.ndarray) -> np.ndarray:
    """CC^T = columns idx of ZC^T (interpolating_vectors.cpp:172-176)."""
    return np.ascontiguousarray(z[:, np.asarray(idx)])


def rank(a: np.ndarray) -> int:
    """Numerical rank as MatrixOperations::rank computes it (matrix_operations.h:713-735)."""
    s = np.linalg.svd(a, compute_uv=False)
    rcond = max(a.shape) * s[0] * 1e-16
    return int(a.shape[0] - np.count_nonzero(s < rcond))


def least_squares(c: np.ndarray, z: np.ndarray) -> np.ndarray:
    """X = argmin ||C X - Z|| by Householder QR, as pzgels('N') on a square C
    (matrix_operations.h:184-202, called at interpolating_vectors.cpp:416)."""
    q, r = scipy.linalg.qr(c, mode="economic")
    return scipy.linalg.solve_triangular(r, q.conj().T @ z, lower=False)


def xi_rows(aoR: np.ndarray, idx: np.ndarray, rows: np.ndarray, block: int = 32768) -> np.ndarray:
    """Rows 'rows' of the least-squares X = CC^T^-1 ZC^T (K x N) for systems
    whose ZC^T does not fit in memory: CC^T = (Phi_mu Phi_mu^H) * conj(.)
    (setup_CCt, interpolating_vectors.cpp:163-190), its Householder QR, then
    for one block of grid columns at a time ZC^T (czt on that block), Q^H
    applied to it and the triangular solve with R, as pzgels('N') treats every
    right-hand side (matrix_operations.h:184-202).  The columns of ZC^T are
    independent right-hand sides, so the rows kept equal those of
    least_squares(cct(Z), Z)."""
    aoR = np.asarray(aoR)
    idx = np.asarray(idx)
    mu = aoR[idx]
    pc = mu.conj() @ mu.T
    c = (pc * pc.conj())
    c = c.real if np.isrealobj(aoR) else c
    q, r = scipy.linalg.qr(c, mode="economic")
    qh = np.ascontiguousarray(q.conj().T)
    rows = np.asarray(rows)
    n = aoR.shape[0]
    out = np.empty((rows.shape[0], n), dtype=c.dtype)
    for b0 in range(0, n, block):
        p = mu.conj() @ aoR[b0:b0 + block].T
        z = p * p.conj()
        z = z.real if np.isrealobj(aoR) else z
        out[:, b0:b0 + block] = scipy.linalg.solve_triangular(r, qh @ z, lower=False)[rows]
    return out


def fft_vectors(x: np.ndarray, fft_grid, workers=None) -> np.ndarray:
    """IVG[G, mu] = DFT_forward(conj(X[mu, :]))  (interpolating_vectors.cpp:342-383).

    X (K, N) is the least-squares solution; the reference Hermitian-transposes it
    (pzgeadd 'C') and runs an unnormalised FFTW_FORWARD 3-D transform over the
    row-major fft_grid on each column.  Returns (N, K) complex, k-space C order.
    workers: threads for the transforms (scipy.fft, the same pocketfft
    transform numpy runs on one thread).
    """
    n0, n1, n2 = (int(v) for v in fft_grid)
    k, n = x.shape
    assert n0 * n1 * n2 == n
    cols = np.conj(x).reshape(k, n0, n1, n2)
    if workers:
        import scipy.fft
        return scipy.fft.fftn(cols, axes=(1, 2, 3), workers=workers).reshape(k, n).T
    return np.fft.fftn(cols, axes=(1, 2, 3)).reshape(k, n).T


def construct_muv(ivg: np.ndarray, coulg: np.ndarray, real_part: bool = False) -> np.ndarray:
    """Muv = A^H A with A[G, mu] = IVG[G, mu] sqrt(coulG[G])
    (interpolating_vectors.cpp:258-278, pzherk 'L','C').  Returns the full
    Hermitian matrix (the reference keeps only the lower triangle); with
    real_part only Re Muv = Re(A)^T Re(A) + Im(A)^T Im(A), all that real
    orbitals keep (their Muv is real), in two real products."""
    a = ivg * np.sqrt(np.asarray(coulg, dtype=np.float64).reshape(-1))[:, None]
    if real_part:
        re = np.ascontiguousarray(a.real)
        im = np.ascontiguousarray(a.imag)
        del a
        return re.T @ re + im.T @ im
    return a.conj().T @ a


def cholesky_luv(muv: np.ndarray) -> tuple[np.ndarray, int]:
    """Luv = chol(Muv) lower, strict upper triangle zeroed
    (interpolating_vectors.cpp:289,307-312).  Returns (Luv, info) where info > 0
    names the first non-positive leading minor, as pzpotrf reports it."""
    m = np.array(muv, copy=True)
    try:
        l = np.linalg.cholesky(m)
        return l, 0
    except np.linalg.LinAlgError:
        # find the failing leading minor (1-based), like xPOTRF's info
        for j in range(1, m.shape[0] + 1):
            try:
                np.linalg.cholesky(m[:j, :j])
            except np.linalg.LinAlgError:
                return np.tril(m), j
        return np.tril(m), m.shape[0]


def isdf(aoR, idx, fft_grid, coulg, return_all=False):
    """The reference's single-rank IVecs pipeline: ctor then kernel()
    (interpolating_vectors.cpp:17-57, 406-455).  Returns V (real part of Muv for
    real orbitals) and Luv; with return_all also Z, C, X, IVG."""
    idx = np.asarray(idx)
    z = czt(aoR, idx)
    c = cct(z, idx)
    x = least_squares(c, z)
    ivg = fft_vectors(x, fft_grid)
    muv = construct_muv(ivg, coulg)
    v = muv.real if np.isrealobj(aoR) else muv
    luv, info = cholesky_luv(v)
    out = {"V": v, "Luv": luv, "chol_info": info, "imag_max": float(np.abs(muv.imag).max())}
    if return_all:
        out.update({"Z": z, "C": c, "X": x, "IVG": ivg})
    return out


def czt_half(aoR: np.ndarray, aoR_half: np.ndarray, idx: np.ndarray) -> np.ndarray:
    """Half-rotated ZC^T = Pocc * conj(P), elementwise (setup_CZt_half,
    interpolating_vectors.cpp:143-161): P from the full orbital set (aoR),
    Pocc from the occupied set (aoR_half, (N, M_occ))."""
    p = pseudo_dm(aoR, idx)
    pocc = pseudo_dm(aoR_half, idx)
    z = pocc * p.conj()
    return z.real if (np.isrealobj(aoR) and np.isrealobj(aoR_half)) else z


def fft_vectors_backward(x: np.ndarray, fft_grid) -> np.ndarray:
    """IVMG[G, mu] = unnormalised FFTW_BACKWARD (e^{+2 pi i k.r/n}) of conj(X[mu, :])
    (interpolating_vectors.cpp:384-391, the half-rotated pass).  (N, K) complex."""
    n0, n1, n2 = (int(v) for v in fft_grid)
    k, n = x.shape
    assert n0 * n1 * n2 == n
    cols = np.conj(x).reshape(k, n0, n1, n2)
    return (np.fft.ifftn(cols, axes=(1, 2, 3)) * n).reshape(k, n).T


def construct_muv_half(ivg: np.ndarray, ivmg: np.ndarray, coulg: np.ndarray) -> np.ndarray:
    """Muv = IVMG^T IVG, both scaled by sqrt(coulG) (interpolating_vectors.cpp:258-273,
    product 'T','N'); the full K x K matrix."""
    w = np.sqrt(np.asarray(coulg, dtype=np.float64).reshape(-1))[:, None]
    return (ivmg * w).T @ (ivg * w)


def isdf_half(aoR, aoR_half, idx, fft_grid, coulg, return_all=False):
    """IVecs with rotate = true (interpolating_vectors.cpp:17-57, 143-161, 406-455):
    the half-rotated ZC^T and CC^T, QR least squares, forward and backward FFTs,
    Muv = IVMG^T IVG (HalfTransformedMuv, no Cholesky, :317-328)."""
    idx = np.asarray(idx)
    z = czt_half(aoR, aoR_half, idx)
    c = cct(z, idx)
    x = least_squares(c, z)
    ivg = fft_vectors(x, fft_grid)
    ivmg = fft_vectors_backward(x, fft_grid)
    muv = construct_muv_half(ivg, ivmg, coulg)
    real = np.isrealobj(aoR) and np.isrealobj(aoR_half)
    out = {"Muv": muv.real if real else muv, "imag_max": float(np.abs(muv.imag).max())}
    if return_all:
        out.update({"Z": z, "C": c, "X": x, "IVG": ivg, "IVMG": ivmg})
    return out
