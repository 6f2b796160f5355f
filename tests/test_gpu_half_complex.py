"""Half-rotated THC path for complex orbitals on the GPU (setup_CZt_half with
complex aoR / aoR_half, src/interpolating_vectors.cpp:143-161; both FFT
directions :384-391; Muv = IVMG^T IVG :258-273) against the oracle's complex
restatement, whose conjugation conventions the CPU tests pin (explicit pair
least squares, brute-force Coulomb kernel, exact interpolation)."""
import numpy as np
import pytest
import torch

from oracle import isdf_oracle as IO
from oracle import kmeans_oracle as KO
from thcpp_amd import _lib as L
from thcpp_amd import synthetic

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _complex_orbitals(n, m, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m)))
    return q


@pytest.mark.parametrize("n,m,mo,k", [(4096, 6, 3, 17), (3375, 9, 4, 33), (20000, 16, 7, 111)])
def test_czt_half_complex(ctx, n, m, mo, k):
    a = _complex_orbitals(n, m, n)
    occ = np.ascontiguousarray(a[:, :mo])
    idx = np.sort(np.random.default_rng(k).choice(n, k, replace=False)).astype(np.int32)
    z = ctx.czt_half_complex(a, occ, idx).cpu().numpy()
    assert rel(z, IO.czt_half(a, occ, idx)) < 1e-14


@pytest.fixture(scope="module")
def complex_half_case():
    """16^3 grid, 8 complex orbitals of which 4 occupied (32 pair functions),
    24 points."""
    n = (16, 16, 16)
    N = int(np.prod(n))
    a = _complex_orbitals(N, 8, 11)
    occ = np.ascontiguousarray(a[:, :4])
    idx = np.sort(np.random.default_rng(5).choice(N, 24, replace=False)).astype(np.int32)
    cg = synthetic.coulomb_kernel(7.0, n)
    ref = IO.isdf_half(a, occ, idx, n, cg, return_all=True)
    return a, occ, idx, n, cg, ref


def test_isdf_half_complex_vs_oracle(ctx, complex_half_case):
    a, occ, idx, n, cg, ref = complex_half_case
    out = ctx.isdf_half_complex(a, occ, idx, n, cg, want_xi=True, check_rank=True)
    assert out["status"] == L.THC_OK, out["status"]
    muv = out["Muv"].cpu().numpy()
    xi = out["xi"].cpu().numpy()
    e, ex = rel(muv, ref["Muv"]), rel(xi, ref["X"])
    print("complex half-rotated vs oracle: rel Muv %.3e rel xi %.3e" % (e, ex))
    assert e < 1e-9 and ex < 1e-9
    assert np.abs(xi[:, idx] - np.eye(len(idx))).max() < 1e-9
    # complex symmetric (coulG is even in G), not Hermitian
    assert rel(muv, muv.T) < 1e-12
    assert rel(muv, muv.conj().T) > 1e-3
    assert out["cct_rank"] == len(idx)


@pytest.mark.parametrize("prec,tol", [(L.PREC_MIXED, 1e-6), (L.PREC_FP32, 1e-4)])
def test_isdf_half_complex_reduced_precision(ctx, complex_half_case, prec, tol):
    """The int8 digit solve on the 2k-row embedding (point scales of idx, twice)
    and the FP32 path, against the oracle beyond the FP64 path's distance."""
    a, occ, idx, n, cg, ref = complex_half_case
    fp64 = ctx.isdf_half_complex(a, occ, idx, n, cg)
    out = ctx.isdf_half_complex(a, occ, idx, n, cg, precision=prec)
    assert out["status"] == L.THC_OK
    e, e64 = rel(out["Muv"].cpu(), ref["Muv"]), rel(fp64["Muv"].cpu(), ref["Muv"])
    print("complex half-rotated precision %d: rel Muv %.3e (FP64 %.3e)" % (prec, e, e64))
    assert e < e64 + tol


def test_isdf_half_complex_refinement(ctx, complex_half_case):
    """One FP64 refinement step of the 2k-row solve brings X to the FP64
    path's accuracy; Muv keeps the FP32 FFT and int8 V of the mixed path."""
    a, occ, idx, n, cg, ref = complex_half_case
    out = ctx.isdf_half_complex(a, occ, idx, n, cg, precision=L.PREC_MIXED, refine=1, want_xi=True)
    assert out["status"] == L.THC_OK
    e, ex = rel(out["Muv"].cpu(), ref["Muv"]), rel(out["xi"].cpu(), ref["X"])
    print("complex half-rotated mixed + 1 refinement: rel Muv %.3e rel xi %.3e" % (e, ex))
    assert ex < 1e-12 and e < 1e-6


def test_isdf_half_complex_real_orbitals_match_real_path(ctx, cfg0):
    """Real orbitals passed as complex: Im X = 0 and Muv is the real
    half-rotated path's (checked to rounding; the solve, FFT and G contraction
    run on twice the rows).  Config 0 with M/2 orbitals occupied and 4 M
    points (M^2/2 pair functions)."""
    s = cfg0
    idx = KO.kmeans(s.grid, s.rho, 4 * s.M, s.max_it, -1.0, 0)["interp_idx"]
    occ = np.ascontiguousarray(s.aoR[:, : s.M // 2])
    real = ctx.isdf_half(s.aoR, occ, idx, s.n, s.coulG)
    cplx = ctx.isdf_half_complex(s.aoR.astype(np.complex128), occ.astype(np.complex128), idx, s.n, s.coulG)
    assert real["status"] == L.THC_OK and cplx["status"] == L.THC_OK
    m = cplx["Muv"].cpu().numpy()
    r = real["Muv"].cpu().numpy()
    print("complex half-rotated on real orbitals vs real path: rel %.3e, max |Im| %.3e" %
          (rel(m.real, r), np.abs(m.imag).max()))
    assert rel(m.real, r) < 1e-9
    assert np.abs(m.imag).max() < 1e-9 * np.abs(r).max()


def test_isdf_half_complex_rejects_bad_input(ctx, complex_half_case):
    a, occ, idx, n, cg, ref = complex_half_case
    out = ctx.isdf_half_complex(a, np.zeros((a.shape[0], 0), dtype=np.complex128), idx, n, cg)
    assert out["status"] == L.THC_ERR_ARG
    bad = idx.copy()
    bad[3] = a.shape[0]
    out = ctx.isdf_half_complex(a, occ, bad, n, cg)
    assert out["status"] == L.THC_ERR_ARG
    assert torch.cuda.is_available()
