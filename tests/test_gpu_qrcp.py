"""QRCP interpolation points on the GPU (thc_qrcp_points) against the oracle's
column-pivoted QR of the explicit orbital-product matrix
(src/qrcp.cpp:100-173)."""
import numpy as np
import pytest

from oracle import qrcp_oracle as QO
from thcpp_amd import _lib as L

pytestmark = pytest.mark.gpu


def _orbitals(n, m, seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((n, m)))
    return q


@pytest.mark.parametrize("n,m,k,seed", [(4096, 8, 24, 0), (3375, 10, 40, 1), (20000, 16, 100, 2)])
def test_qrcp_vs_oracle(ctx, n, m, k, seed):
    print("qrcp n=%d m=%d k=%d" % (n, m, k), flush=True)
    a = _orbitals(n, m, seed)
    st, idx, rd = ctx.qrcp_points(a, k)
    assert st == L.THC_OK
    ref_idx, ref_rd, _ = QO.qrcp_points(a, k)
    np.testing.assert_array_equal(idx.cpu().numpy(), ref_idx)
    # the Gram downdate carries the late residuals to about half the digits
    np.testing.assert_allclose(rd, ref_rd, rtol=1e-7)


def test_qrcp_half_rotated(ctx):
    a = _orbitals(5000, 12, 5)
    occ = np.ascontiguousarray(a[:, :5])
    st, idx, rd = ctx.qrcp_points(a, 30, occ)
    assert st == L.THC_OK
    ref_idx, ref_rd, _ = QO.qrcp_points(a, 30, occ)
    np.testing.assert_array_equal(idx.cpu().numpy(), ref_idx)
    np.testing.assert_allclose(rd, ref_rd, rtol=1e-7)


def test_qrcp_cfg0(ctx, cfg0):
    s = cfg0
    k = 4 * s.M
    st, idx, rd = ctx.qrcp_points(s.aoR, k)
    assert st == L.THC_OK
    ref_idx, ref_rd, _ = QO.qrcp_points(s.aoR, k)
    print("cfg0 qrcp |R_kk|/|R_00| = %.3e" % (rd[-1] / rd[0]))
    np.testing.assert_array_equal(idx.cpu().numpy(), ref_idx)


def test_qrcp_rank_exhausted(ctx):
    """M = 3 real orbitals span at most M (M + 1) / 2 = 6 product columns."""
    a = _orbitals(2000, 3, 7)
    st, _, _ = ctx.qrcp_points(a, 8)
    assert st == L.THC_ERR_NOT_POSDEF


def _complex_orbitals(n, m, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m)))
    return q


@pytest.mark.parametrize("n,m,mo,k", [(4096, 8, 0, 30), (3375, 10, 4, 36), (20000, 12, 5, 60)])
def test_qrcp_complex_vs_oracle(ctx, n, m, mo, k):
    """Complex orbitals: |P|^2 pivots (mo = 0) and the complex Hermitian
    half-rotated Gram (mo occupied orbitals) against zgeqp3 of the explicit ZT."""
    print("complex qrcp n=%d m=%d m_occ=%d k=%d" % (n, m, mo, k), flush=True)
    a = _complex_orbitals(n, m, n + m)
    occ = np.ascontiguousarray(a[:, :mo]) if mo else None
    st, idx, rd = ctx.qrcp_points_complex(a, k, occ)
    assert st == L.THC_OK
    ref_idx, ref_rd, _ = QO.qrcp_points(a, k, occ)
    np.testing.assert_array_equal(idx.cpu().numpy(), ref_idx)
    np.testing.assert_allclose(rd, ref_rd, rtol=1e-7)


def test_qrcp_complex_real_orbitals_match_real_path(ctx):
    a = _orbitals(5000, 9, 6)
    occ = np.ascontiguousarray(a[:, :4])
    for o in (None, occ):
        st_r, idx_r, rd_r = ctx.qrcp_points(a, 25, o)
        st_c, idx_c, rd_c = ctx.qrcp_points_complex(a.astype(np.complex128), 25,
                                                    o.astype(np.complex128) if o is not None else None)
        assert st_r == L.THC_OK and st_c == L.THC_OK
        np.testing.assert_array_equal(idx_c.cpu().numpy(), idx_r.cpu().numpy())
        np.testing.assert_allclose(rd_c, rd_r, rtol=1e-12)
