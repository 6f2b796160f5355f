"""Complex orbitals on the GPU (thc_czt_complex, thc_isdf_complex): ZC^T =
|Phi_mu^H Phi|^2 (src/interpolating_vectors.cpp:59-141) against the oracle's
complex restatement, and V against the real path for phase-rotated real
orbitals, whose ZC^T is the same."""
import numpy as np
import pytest

from oracle import isdf_oracle as IO
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


@pytest.mark.parametrize("n,m,k", [(4096, 6, 30), (3375, 9, 70), (20000, 16, 200)])
def test_czt_complex(ctx, n, m, k):
    a = _complex_orbitals(n, m, n)
    idx = np.sort(np.random.default_rng(k).choice(n, k, replace=False)).astype(np.int32)
    z = ctx.czt_complex(a, idx).cpu().numpy()
    zr = IO.czt(a, idx)
    assert np.abs(zr.imag).max() < 1e-14 * np.abs(zr).max()
    assert rel(z, zr.real) < 1e-14


@pytest.fixture(scope="module")
def small_system():
    s = synthetic.make_system(reps=1, n=16, M=8, thc_cfac=2, max_it=1)
    idx = np.sort(np.random.default_rng(3).choice(s.N, s.K, replace=False)).astype(np.int32)
    return s, idx


def test_isdf_complex_phase_rotated_equals_real(ctx, small_system):
    """Per-point phases leave |P|^2, hence ZC^T, xi and V, unchanged."""
    s, idx = small_system
    theta = np.random.default_rng(4).uniform(0, 2 * np.pi, s.N)
    a = s.aoR * np.exp(1j * theta)[:, None]
    ref = ctx.isdf(s.aoR, idx, s.n, s.coulG)
    out = ctx.isdf_complex(a, idx, s.n, s.coulG)
    assert ref["status"] == L.THC_OK and out["status"] == L.THC_OK
    e = rel(out["V"].cpu(), ref["V"].cpu())
    print("complex (phase-rotated) vs real rel V %.3e" % e)
    assert e < 1e-9


def test_isdf_complex_vs_oracle(ctx):
    """Genuinely complex orbitals: M = 8 give |P|^2 of rank <= M^2 = 64."""
    n = (16, 16, 16)
    N = int(np.prod(n))
    a = _complex_orbitals(N, 8, 11)
    idx = np.sort(np.random.default_rng(5).choice(N, 40, replace=False)).astype(np.int32)
    cg = synthetic.coulomb_kernel(7.0, n)
    ref = IO.isdf(a, idx, n, cg, return_all=True)
    out = ctx.isdf_complex(a, idx, n, cg, want_xi=True)
    assert out["status"] == L.THC_OK
    ev = rel(out["V"].cpu(), ref["V"].real)
    ex = rel(out["xi"].cpu(), ref["X"].real)
    print("complex orbitals vs oracle rel V %.3e rel xi %.3e" % (ev, ex))
    assert ref["imag_max"] < 1e-9 * np.abs(ref["V"]).max()
    assert ev < 1e-8 and ex < 1e-7
    l = out["Luv"].cpu().numpy()
    assert rel(l @ l.T, out["V"].cpu()) < 1e-13
    mixed = ctx.isdf_complex(a, idx, n, cg, precision=L.PREC_MIXED)
    e_mx = rel(mixed["V"].cpu(), ref["V"].real)
    print("complex orbitals mixed rel V vs oracle %.3e" % e_mx)
    assert e_mx < ev + 1e-6
