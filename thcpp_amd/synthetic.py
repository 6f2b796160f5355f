"""Synthetic ISDF inputs in the layout the reference's orbital file holds.

The reference reads these datasets from its orbital file (written by
tools/thcpy/thcpy/transform_basis.py:51-148):

  real_space_grid  (N, 3) f64, C-order uniform grid r = L (i, j, k) / n
  density          (N, 1) f64, rho(r) = sum_m |phi_m(r)|^2 (unnormalised)
  aoR              (N, M) orbitals on the grid (real here)
  fft_grid         (3, 1) int
  fft_coulomb      (N, 1) f64, 4 pi / |G|^2 * vol / N^2, 0 at G = 0

pyscf, which writes the real file, is not available, so the orbitals are
built here: every orbital is a sum of 4 isotropic Gaussians per atom with
exponents ~ U[0.5, 3.0] bohr^-2 and coefficients ~ N(0, 1) (numpy
default_rng(seed)), each Gaussian made periodic by summing its images in the
neighbouring cells, then the N x M matrix is orthonormalised by QR.
Atoms sit on conventional diamond sites (a = 6.74 bohr), replicated per axis.
"""
from __future__ import annotations

import dataclasses

import numpy as np

DIAMOND_A = 6.74  # bohr, conventional cubic cell of diamond carbon
_DIAMOND_FRAC = np.array(
    [[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0],
     [0.25, 0.25, 0.25], [0.25, 0.75, 0.75], [0.75, 0.25, 0.75], [0.75, 0.75, 0.25]])


@dataclasses.dataclass
class System:
    """One ISDF problem: arrays named after the orbital-file datasets."""
    L: float
    n: tuple
    atoms: np.ndarray          # (natoms, 3)
    grid: np.ndarray           # (N, 3) f64
    aoR: np.ndarray            # (N, M) f64, orthonormal columns
    rho: np.ndarray            # (N,) f64
    coulG: np.ndarray          # (N,) f64
    thc_cfac: int
    max_it: int
    seed: int = 0
    threshold: float = -1.0

    @property
    def N(self):
        return self.grid.shape[0]

    @property
    def M(self):
        return self.aoR.shape[1]

    @property
    def K(self):
        return self.thc_cfac * self.M


def diamond_sites(reps: int, a: float = DIAMOND_A) -> np.ndarray:
    cells = np.array([[i, j, k] for i in range(reps) for j in range(reps) for k in range(reps)], dtype=float)
    return ((cells[:, None, :] + _DIAMOND_FRAC[None, :, :]) * a).reshape(-1, 3)


def uniform_grid(L: float, n) -> np.ndarray:
    """C-order grid, first index slowest (pyscf gen_uniform_grids order for a cubic cell)."""
    n0, n1, n2 = n
    i, j, k = np.meshgrid(np.arange(n0), np.arange(n1), np.arange(n2), indexing="ij")
    g = np.stack([i * (L / n0), j * (L / n1), k * (L / n2)], axis=-1)
    return np.ascontiguousarray(g.reshape(-1, 3))


def coulomb_kernel(L: float, n) -> np.ndarray:
    """4 pi / |G|^2 * vol / N^2 with G = 2 pi / L * fftfreq index, G = 0 -> 0, C order."""
    n0, n1, n2 = n
    N = n0 * n1 * n2
    g = [2 * np.pi / L * np.fft.fftfreq(m, d=1.0 / m) for m in n]
    g2 = (g[0][:, None, None] ** 2 + g[1][None, :, None] ** 2 + g[2][None, None, :] ** 2).reshape(-1)
    out = np.zeros(N)
    nz = g2 > 0
    out[nz] = 4 * np.pi / g2[nz] * (L ** 3) / (N ** 2)
    return out


def _periodic_gauss_1d(x: np.ndarray, centre: np.ndarray, alpha: np.ndarray, L: float, images: int = 2) -> np.ndarray:
    """(T, n) values of sum_{m=-images..images} exp(-alpha (x - c + m L)^2)."""
    d = x[None, :] - centre[:, None]
    out = np.zeros((centre.shape[0], x.shape[0]))
    for m in range(-images, images + 1):
        out += np.exp(-alpha[:, None] * (d + m * L) ** 2)
    return out


def build_orbitals(atoms: np.ndarray, L: float, n, M: int, seed: int = 0, per_atom: int = 4) -> np.ndarray:
    """(N, M) orthonormal orbitals.  Each raw orbital is
    sum_t c_t g_t(x) g_t(y) g_t(z) over T = per_atom * natoms periodic Gaussians,
    evaluated separably as (n0 n1, T) @ (T, n2) per orbital."""
    rng = np.random.default_rng(seed)
    nat = atoms.shape[0]
    alpha = rng.uniform(0.5, 3.0, size=(M, nat, per_atom))
    coef = rng.standard_normal(size=(M, nat, per_atom))
    n0, n1, n2 = n
    xs = [np.arange(m) * (L / m) for m in n]
    N = n0 * n1 * n2
    raw = np.empty((M, N))
    cen = np.repeat(atoms, per_atom, axis=0)  # (T, 3)
    for m in range(M):
        a = alpha[m].reshape(-1)
        c = coef[m].reshape(-1)
        gx = _periodic_gauss_1d(xs[0], cen[:, 0], a, L)
        gy = _periodic_gauss_1d(xs[1], cen[:, 1], a, L)
        gz = _periodic_gauss_1d(xs[2], cen[:, 2], a, L)
        gxy = (gx[:, :, None] * gy[:, None, :]).reshape(a.shape[0], n0 * n1)  # (T, n0 n1)
        raw[m] = (gxy.T @ (c[:, None] * gz)).reshape(-1)
    # CholeskyQR2: two passes of R = chol(A^T A), A <- A R^-1 give columns
    # orthonormal to working precision for well-conditioned A.
    q = raw
    for _ in range(2):
        r = np.linalg.cholesky(q @ q.T)          # lower, = R^T
        q = np.linalg.inv(r) @ q                 # R^-T applied on the row side
    return np.ascontiguousarray(q.T)


def make_system(reps: int, n: int, M: int, thc_cfac: int, max_it: int, seed: int = 0) -> System:
    L = DIAMOND_A * reps
    atoms = diamond_sites(reps)
    nn = (n, n, n)
    aoR = build_orbitals(atoms, L, nn, M, seed=seed)
    return System(L=L, n=nn, atoms=atoms, grid=uniform_grid(L, nn), aoR=aoR,
                  rho=np.einsum("ij,ij->i", aoR, aoR), coulG=coulomb_kernel(L, nn),
                  thc_cfac=thc_cfac, max_it=max_it, seed=0)


# The configurations the benchmarks are quoted on.
CONFIGS = {
    # 8-atom conventional diamond cell, 32^3 grid, 16 orbitals, N_mu = 128, 20 iterations
    "diamond8_n32": dict(reps=1, n=32, M=16, thc_cfac=8, max_it=20),
    # 2x2x2 supercell (64 atoms), 64^3 grid, 128 orbitals, N_mu = 1024, 30 iterations
    "diamond64_n64_k1024": dict(reps=2, n=64, M=128, thc_cfac=8, max_it=30),
    # same system, N_mu = 2048
    "diamond64_n64_k2048": dict(reps=2, n=64, M=128, thc_cfac=16, max_it=30),
    # 3x3x3 supercell (216 atoms), 96^3 grid, 432 orbitals, N_mu = 3456, 40 iterations
    "diamond216_n96_k3456": dict(reps=3, n=96, M=432, thc_cfac=8, max_it=40),
}


def make_config(name: str, seed: int = 0) -> System:
    return make_system(seed=seed, **CONFIGS[name])


_SMOOTH = (24, 25, 27, 30, 32, 36, 40, 45, 48)


def make_batch(count: int = 32, seed: int = 0, max_it: int = 20):
    """Uneven batch of small cells: side L ~ U[6, 14] bohr holding 8-64 carbon
    atoms on jittered diamond sites (the count follows the cell volume at
    diamond density, 8 atoms per (6.74 bohr)^3, rounded to a multiple of 8 and
    clamped to [8, 64], so every cell has ~90 grid points per interpolation
    point), grid edge the 2,3,5-smooth size in [24, 48] nearest to 3.5 points
    per bohr, M = 2 per atom, N_mu = cfac M with cfac ~ {8, 9, 10} capped so
    that N_mu <= M (M + 1) / 2 (above that CC^T is singular for real
    orbitals)."""
    rng = np.random.default_rng(seed)
    systems = []
    base = diamond_sites(2, 1.0) / 2.0  # 64 fractional sites in the unit cube
    for s in range(count):
        L = float(rng.uniform(6.0, 14.0))
        nat = int(min(8, max(1, round((L / DIAMOND_A) ** 3)))) * 8
        frac = base[rng.permutation(64)[:nat]]
        atoms = (frac + rng.normal(scale=0.02, size=frac.shape)) * L
        target = 3.5 * L
        n = min(_SMOOTH, key=lambda v: abs(v - target))
        M = 2 * nat
        cfac = int(rng.integers(8, 11))
        while cfac * M > M * (M + 1) // 2:
            cfac -= 1
        nn = (n, n, n)
        aoR = build_orbitals(atoms, L, nn, M, seed=seed * 1000 + s)
        systems.append(System(L=L, n=nn, atoms=atoms, grid=uniform_grid(L, nn), aoR=aoR,
                              rho=np.einsum("ij,ij->i", aoR, aoR), coulG=coulomb_kernel(L, nn),
                              thc_cfac=cfac, max_it=max_it, seed=0))
    return systems
