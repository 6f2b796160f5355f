"""The drop-in driver thcpp_amd.x: orbital file + input.json in the
reference's formats, QMCPACK THC file out (src/main.cpp:18-77,
src/interpolating_vectors.cpp:72-90, 210-245, 283-315)."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import isdf_oracle as IO
from oracle import kmeans_oracle as KO
from thcpp_amd import synthetic

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "thcpp_amd")
vp = ctypes.c_void_p


def host():
    h = ctypes.CDLL(os.path.join(PKG, "libthcpp_amd_host.so"))
    h.thc_host_write_orbital_file.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.c_int, vp, vp, vp, vp, vp, vp,
                                              vp, ctypes.c_double, vp, ctypes.c_int, vp, vp]
    h.thc_host_read_f64.argtypes = [ctypes.c_char_p, ctypes.c_char_p, vp, ctypes.c_long]
    h.thc_host_read_f64.restype = ctypes.c_long
    h.thc_host_read_int.argtypes = [ctypes.c_char_p, ctypes.c_char_p, vp, ctypes.c_long]
    h.thc_host_read_int.restype = ctypes.c_long
    return h


def read(h, path, name, dtype=np.float64, cap=1 << 24):
    out = np.empty(cap, dtype=dtype)
    f = h.thc_host_read_f64 if dtype == np.float64 else h.thc_host_read_int
    n = f(path.encode(), name.encode(), out.ctypes.data_as(vp), cap)
    assert n >= 0, name
    return out[:n]


@pytest.mark.parametrize("precision,tol", [(None, 1e-9), ("mixed", 1e-6)])
def test_dropin_driver(tmp_path, precision, tol):
    """The driver's IVecs::kernel is one thc_isdf call; "thc_precision" picks
    the path (absent: FP64).  CC^T's rank is checked as check_rank does."""
    h = host()
    s = synthetic.make_system(reps=1, n=16, M=8, thc_cfac=3, max_it=7)
    orb = str(tmp_path / "orbitals.h5")
    thc = str(tmp_path / "thc.h5")
    P = lambda a: a.ctypes.data_as(vp)
    fg = np.array(s.n, dtype=np.int32)
    hc = np.arange(s.M * s.M, dtype=np.float64).reshape(s.M, s.M)
    ne = np.array([4, 4], dtype=np.int32)
    assert h.thc_host_write_orbital_file(orb.encode(), s.N, s.M, P(s.grid), P(s.rho), P(s.aoR), P(fg), P(s.coulG),
                                         P(hc), P(ne), -2.5, None, 0, None, None) == 0
    inp = {"orbital_file": orb, "output_file": thc, "thc_cfac": s.thc_cfac,
           "interpolating_points": {"kmeans": {"threshold": -1.0, "max_it": s.max_it, "rng_seed": 0}}}
    if precision:
        inp["thc_precision"] = precision
    (tmp_path / "input.json").write_text(json.dumps(inp))
    r = subprocess.run([os.path.join(PKG, "thcpp_amd.x"), str(tmp_path / "input.json")], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    ref = KO.kmeans(s.grid, s.rho, s.K, s.max_it, -1.0, 0)
    idx = read(h, thc, "/Hamiltonian/THC/InterpIndx", np.int32)
    np.testing.assert_array_equal(idx, ref["interp_idx"])
    orbs = read(h, thc, "/Hamiltonian/THC/Orbitals").reshape(s.M, s.K, 2)
    np.testing.assert_array_equal(orbs[:, :, 0], s.aoR[idx].T)
    assert not orbs[:, :, 1].any()
    luv = read(h, thc, "/Hamiltonian/THC/Luv").reshape(s.K, s.K, 2)[:, :, 0]
    assert not np.triu(luv, 1).any()
    v_ref = IO.isdf(s.aoR, idx, s.n, s.coulG)["V"]
    err = np.linalg.norm(luv @ luv.T - v_ref) / np.linalg.norm(v_ref)
    assert err < tol, err
    c = IO.cct(IO.czt(s.aoR, idx), idx)
    rank = IO.rank(c)
    assert (" * CCt matrix has full rank : %d" % rank if rank == s.K else
            " * WARNING: CCt matrix is rank deficient. Rank = %d" % rank) in r.stdout
    assert ("(mixed)" if precision == "mixed" else "(fp64)") in r.stdout
    np.testing.assert_array_equal(read(h, thc, "/Hamiltonian/dims", np.int32), [-1, 0, 0, s.M, 4, 4, 0, 0])
    np.testing.assert_array_equal(read(h, thc, "/Hamiltonian/THC/dims", np.int32), [s.M, s.K, 0])
    np.testing.assert_array_equal(read(h, thc, "/Hamiltonian/Energies"), [-2.5])
    np.testing.assert_array_equal(read(h, thc, "/Hamiltonian/occups", np.int32), np.zeros(8))
    np.testing.assert_array_equal(read(h, thc, "/Hamiltonian/hcore").reshape(s.M, s.M, 2)[:, :, 0], hc)
    assert "Step" in r.stdout and "Dumping Luv" in r.stdout


def test_dropin_driver_half_rotated(tmp_path):
    """half_rotated = true (src/main.cpp:57-72): the full pass writes Luv, a second
    k-means pass with thc_half_cfac picks new points, the half-rotated pass
    writes the HalfTransformed* datasets (Pocc o P, Muv with no Cholesky) and
    THC/dims = [M, K, K_half]."""
    h = host()
    s = synthetic.make_system(reps=1, n=16, M=8, thc_cfac=3, max_it=7)
    occ = np.ascontiguousarray(s.aoR[:, :4])
    half_cfac = 2
    kh = half_cfac * s.M
    orb = str(tmp_path / "orbitals.h5")
    thc = str(tmp_path / "thc.h5")
    P = lambda a: a.ctypes.data_as(vp)
    fg = np.array(s.n, dtype=np.int32)
    hc = np.eye(s.M)
    ne = np.array([4, 4], dtype=np.int32)
    assert h.thc_host_write_orbital_file(orb.encode(), s.N, s.M, P(s.grid), P(s.rho), P(s.aoR), P(fg), P(s.coulG),
                                         P(hc), P(ne), 0.0, P(occ), occ.shape[1], None, None) == 0
    inp = {"orbital_file": orb, "output_file": thc, "thc_cfac": s.thc_cfac, "half_rotated": True,
           "thc_half_cfac": half_cfac,
           "interpolating_points": {"kmeans": {"threshold": -1.0, "max_it": s.max_it, "rng_seed": 0}}}
    (tmp_path / "input.json").write_text(json.dumps(inp))
    r = subprocess.run([os.path.join(PKG, "thcpp_amd.x"), str(tmp_path / "input.json")], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    # full pass
    idx = read(h, thc, "/Hamiltonian/THC/InterpIndx", np.int32)
    np.testing.assert_array_equal(idx, KO.kmeans(s.grid, s.rho, s.K, s.max_it, -1.0, 0)["interp_idx"])
    luv = read(h, thc, "/Hamiltonian/THC/Luv").reshape(s.K, s.K, 2)[:, :, 0]
    v_ref = IO.isdf(s.aoR, idx, s.n, s.coulG)["V"]
    assert np.linalg.norm(luv @ luv.T - v_ref) / np.linalg.norm(v_ref) < 1e-9
    # half-rotated pass
    idx_h = read(h, thc, "/Hamiltonian/THC/HalfTransformedFullInterpIndx", np.int32)
    np.testing.assert_array_equal(idx_h, KO.kmeans(s.grid, s.rho, kh, s.max_it, -1.0, 0)["interp_idx"])
    np.testing.assert_array_equal(read(h, thc, "/Hamiltonian/THC/HalfTransformedOccInterpIndx", np.int32), idx_h)
    full = read(h, thc, "/Hamiltonian/THC/HalfTransformedFullOrbitals").reshape(s.M, kh, 2)
    np.testing.assert_array_equal(full[:, :, 0], s.aoR[idx_h].T)
    occo = read(h, thc, "/Hamiltonian/THC/HalfTransformedOccOrbitals").reshape(4, kh, 2)
    np.testing.assert_array_equal(occo[:, :, 0], occ[idx_h].T)
    muv = read(h, thc, "/Hamiltonian/THC/HalfTransformedMuv").reshape(kh, kh, 2)
    assert not muv[:, :, 1].any()
    m_ref = IO.isdf_half(s.aoR, occ, idx_h, s.n, s.coulG)["Muv"]
    err = np.linalg.norm(muv[:, :, 0] - m_ref) / np.linalg.norm(m_ref)
    assert err < 1e-9, err
    np.testing.assert_array_equal(read(h, thc, "/Hamiltonian/THC/dims", np.int32), [s.M, s.K, kh])
    assert "Dumping half transformed Muv" in r.stdout


def test_dropin_driver_qrcp(tmp_path):
    """interpolating_points.qrcp (src/interpolating_points.cpp:40-49,
    src/qrcp.cpp:100-173): the points are the first thc_cfac * M QRCP pivots,
    ascending, and the rest of the file follows from them."""
    from oracle import qrcp_oracle as QO
    h = host()
    s = synthetic.make_system(reps=1, n=16, M=8, thc_cfac=3, max_it=1)
    orb = str(tmp_path / "orbitals.h5")
    thc = str(tmp_path / "thc.h5")
    P = lambda a: a.ctypes.data_as(vp)
    fg = np.array(s.n, dtype=np.int32)
    hc = np.eye(s.M)
    ne = np.array([4, 4], dtype=np.int32)
    assert h.thc_host_write_orbital_file(orb.encode(), s.N, s.M, P(s.grid), P(s.rho), P(s.aoR), P(fg), P(s.coulG),
                                         P(hc), P(ne), 0.0, None, 0, None, None) == 0
    inp = {"orbital_file": orb, "output_file": thc, "thc_cfac": s.thc_cfac, "interpolating_points": {"qrcp": {}}}
    (tmp_path / "input.json").write_text(json.dumps(inp))
    r = subprocess.run([os.path.join(PKG, "thcpp_amd.x"), str(tmp_path / "input.json")], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    idx = read(h, thc, "/Hamiltonian/THC/InterpIndx", np.int32)
    ref_idx, ref_rd, _ = QO.qrcp_points(s.aoR, s.K)
    np.testing.assert_array_equal(idx, ref_idx)
    luv = read(h, thc, "/Hamiltonian/THC/Luv").reshape(s.K, s.K, 2)[:, :, 0]
    v_ref = IO.isdf(s.aoR, idx, s.n, s.coulG)["V"]
    assert np.linalg.norm(luv @ luv.T - v_ref) / np.linalg.norm(v_ref) < 1e-9
    diag = [l for l in r.stdout.splitlines() if l.startswith("full_ix_diag")]
    assert len(diag) == s.K
    np.testing.assert_allclose([float(l.split()[2]) for l in diag], ref_rd, rtol=1e-7)


def test_dropin_driver_complex_orbitals(tmp_path):
    """aoR with imaginary parts (k-point orbitals, the reference's complex128
    input): Orbitals are written complex, ZC^T = |P|^2 and V follow the oracle."""
    h = host()
    s = synthetic.make_system(reps=1, n=16, M=8, thc_cfac=3, max_it=5)
    theta = np.random.default_rng(9).uniform(0, 2 * np.pi, s.M)
    a = s.aoR * np.exp(1j * theta)[None, :]
    re, im = np.ascontiguousarray(a.real), np.ascontiguousarray(a.imag)
    orb = str(tmp_path / "orbitals.h5")
    thc = str(tmp_path / "thc.h5")
    P = lambda x: x.ctypes.data_as(vp)
    fg = np.array(s.n, dtype=np.int32)
    hc = np.eye(s.M)
    ne = np.array([4, 4], dtype=np.int32)
    assert h.thc_host_write_orbital_file(orb.encode(), s.N, s.M, P(s.grid), P(s.rho), P(re), P(fg), P(s.coulG),
                                         P(hc), P(ne), 0.0, None, 0, P(im), None) == 0
    inp = {"orbital_file": orb, "output_file": thc, "thc_cfac": s.thc_cfac,
           "interpolating_points": {"kmeans": {"threshold": -1.0, "max_it": s.max_it, "rng_seed": 0}}}
    (tmp_path / "input.json").write_text(json.dumps(inp))
    r = subprocess.run([os.path.join(PKG, "thcpp_amd.x"), str(tmp_path / "input.json")], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    idx = read(h, thc, "/Hamiltonian/THC/InterpIndx", np.int32)
    np.testing.assert_array_equal(idx, KO.kmeans(s.grid, s.rho, s.K, s.max_it, -1.0, 0)["interp_idx"])
    orbs = read(h, thc, "/Hamiltonian/THC/Orbitals").reshape(s.M, s.K, 2)
    np.testing.assert_array_equal(orbs[:, :, 0], re[idx].T)
    np.testing.assert_array_equal(orbs[:, :, 1], im[idx].T)
    luv = read(h, thc, "/Hamiltonian/THC/Luv").reshape(s.K, s.K, 2)[:, :, 0]
    v_ref = IO.isdf(a, idx, s.n, s.coulG)["V"].real
    assert np.linalg.norm(luv @ luv.T - v_ref) / np.linalg.norm(v_ref) < 1e-9
    assert "complex orbitals" in r.stdout


def test_dropin_missing_key_fails(tmp_path):
    (tmp_path / "input.json").write_text(json.dumps({"orbital_file": "x.h5", "output_file": "y.h5"}))
    r = subprocess.run([os.path.join(PKG, "thcpp_amd.x"), str(tmp_path / "input.json")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0 and "thc_cfac" in r.stderr


@pytest.mark.parametrize("points", ["kmeans", "qrcp"])
def test_dropin_driver_complex_half_rotated(tmp_path, points):
    """Complex orbitals with half_rotated = true: the half-rotated pass runs on
    Pocc o conj(P) (complex ZC^T and Hermitian CC^T) and writes a complex
    HalfTransformedMuv; with qrcp the points are the pivots of the complex ZT.
    Orbitals: complex, non-unitary combinations of the real ones (phases alone
    cancel in Pocc o conj(P))."""
    from oracle import qrcp_oracle as QO
    h = host()
    s = synthetic.make_system(reps=1, n=16, M=8, thc_cfac=3, max_it=7)
    rng = np.random.default_rng(13)
    c = np.eye(s.M) + 0.3 * (rng.standard_normal((s.M, s.M)) + 1j * rng.standard_normal((s.M, s.M)))
    a = s.aoR @ c
    occ = np.ascontiguousarray(a[:, :4])
    half_cfac = 2
    kh = half_cfac * s.M
    orb = str(tmp_path / "orbitals.h5")
    thc = str(tmp_path / "thc.h5")
    P = lambda x: np.ascontiguousarray(x).ctypes.data_as(vp)
    re, im = np.ascontiguousarray(a.real), np.ascontiguousarray(a.imag)
    ore, oim = np.ascontiguousarray(occ.real), np.ascontiguousarray(occ.imag)
    fg = np.array(s.n, dtype=np.int32)
    hc = np.eye(s.M)
    ne = np.array([4, 4], dtype=np.int32)
    assert h.thc_host_write_orbital_file(orb.encode(), s.N, s.M, P(s.grid), P(s.rho), P(re), P(fg), P(s.coulG),
                                         P(hc), P(ne), 0.0, P(ore), 4, P(im), P(oim)) == 0
    ip = {"kmeans": {"threshold": -1.0, "max_it": s.max_it, "rng_seed": 0}} if points == "kmeans" else {"qrcp": {}}
    inp = {"orbital_file": orb, "output_file": thc, "thc_cfac": s.thc_cfac, "half_rotated": True,
           "thc_half_cfac": half_cfac, "interpolating_points": ip}
    (tmp_path / "input.json").write_text(json.dumps(inp))
    r = subprocess.run([os.path.join(PKG, "thcpp_amd.x"), str(tmp_path / "input.json")], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    idx_h = read(h, thc, "/Hamiltonian/THC/HalfTransformedFullInterpIndx", np.int32)
    if points == "kmeans":
        ref_idx = KO.kmeans(s.grid, s.rho, kh, s.max_it, -1.0, 0)["interp_idx"]
    else:
        ref_idx = QO.qrcp_points(a, kh, occ)[0]
    np.testing.assert_array_equal(idx_h, ref_idx)
    occo = read(h, thc, "/Hamiltonian/THC/HalfTransformedOccOrbitals").reshape(4, kh, 2)
    np.testing.assert_array_equal(occo[:, :, 0], ore[idx_h].T)
    np.testing.assert_array_equal(occo[:, :, 1], oim[idx_h].T)
    muv = read(h, thc, "/Hamiltonian/THC/HalfTransformedMuv").reshape(kh, kh, 2)
    m = muv[:, :, 0] + 1j * muv[:, :, 1]
    m_ref = IO.isdf_half(a, occ, idx_h, s.n, s.coulG)["Muv"]
    assert np.abs(m_ref.imag).max() > 1e-3 * np.abs(m_ref).max()
    err = np.linalg.norm(m - m_ref) / np.linalg.norm(m_ref)
    print("complex half-rotated drop-in (%s points): rel Muv %.3e" % (points, err))
    assert err < 1e-8, err  # cond(CC^T) 4e6 with the k-means points
    assert "complex orbitals" in r.stdout
