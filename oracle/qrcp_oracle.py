"""ORACLE - TEST INFRASTRUCTURE ONLY.

CPU restatement of the reference's QRCP interpolation-point selection, the
checker for the HIP path; only tests/ may import it.

  z_matrix       src/qrcp.cpp:33-48   ZT[(i k), a] = conj(phi_i(r_a)) phi_k(r_a)   (M^2 x N)
  z_half_matrix  src/qrcp.cpp:49-66   ZT[(i k), a] = conj(phi_i^occ(r_a)) phi_k(r_a) (M_occ M x N)
                 -- the matrix that function's comment defines, with row i*M + k.
                 The reference's loop writes row i*nelec + k instead (qrcp.cpp:62),
                 so with M > nelec its rows overlap and the tail stays zero; that
                 indexing is not reproduced (tests/test_oracle_qrcp.py shows both
                 matrices and that their pivots differ), so the half-rotated
                 pivots are pinned to the stated definition, not to the
                 reference's output.
  qrcp_points    src/qrcp.cpp:100-173 QR with column pivoting of ZT (pzgeqpf,
                 src/matrix_operations.h:756-788, every column free), the first
                 K pivots sorted ascending; |R_jj| in pivot order.

The factorisation is LAPACK's pivoted QR through scipy (dgeqp3; the
reference's pzgeqpf is the ScaLAPACK counterpart).  Pinned in
tests/test_oracle_qrcp.py against a plain Gram-Schmidt with largest-residual
pivoting written from the definition.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg


def z_matrix(aoR: np.ndarray) -> np.ndarray:
    """(M^2, N): row i*M + k holds conj(phi_i) phi_k over the grid (src/qrcp.cpp:33-48)."""
    a = np.asarray(aoR)          # (N, M), the orbital file layout
    n, m = a.shape
    return (a.conj()[:, :, None] * a[:, None, :]).reshape(n, m * m).T


def z_half_matrix(aoR: np.ndarray, aoR_half: np.ndarray) -> np.ndarray:
    """(M_occ M, N): row i*M + k holds conj(phi_i^occ) phi_k, the Khatri-Rao
    rows src/qrcp.cpp:53-56 states (its loop at :58-64 indexes the rows with
    stride nelec instead of M; see the module docstring)."""
    a, h = np.asarray(aoR), np.asarray(aoR_half)
    n = a.shape[0]
    return (h.conj()[:, :, None] * a[:, None, :]).reshape(n, -1).T


def qrcp_points(aoR, k, aoR_half=None):
    """First k pivot columns of QRCP of ZT, ascending, and |R_jj| (j < k)."""
    z = z_matrix(aoR) if aoR_half is None else z_half_matrix(aoR, aoR_half)
    r, piv = scipy.linalg.qr(z, mode="r", pivoting=True)
    r = r[0] if isinstance(r, tuple) else r
    return np.sort(piv[:k]).astype(np.int32), np.abs(np.diag(r)[:k]), piv[:k]
