// Writer for orbital files in the layout the reference reads
// (tools/thcpy/thcpy/transform_basis.py:73-95): real_space_grid (N,3),
// density (N,1), aoR (N,M) as h5py complex {r,i}, fft_grid (3,1),
// fft_coulomb (N,1), hcore (M,M) complex, num_electrons (1,2),
// constant_energy_factors (1,1).  Used to feed the drop-in driver synthetic
// systems (pyscf, which writes the real files, is not available here).
#include <vector>

#include "h5io.h"

// aoR_half (n x m_occ, nullable) adds the occupied orbitals the half-rotated path reads;
// aoR_im (n x m, nullable) the imaginary parts of complex orbitals.
extern "C" int thc_host_write_orbital_file(const char* path, long n, int m, const double* grid, const double* rho,
                                           const double* aoR, const int* fft_grid, const double* coulG,
                                           const double* hcore, const int* nelec, double e0,
                                           const double* aoR_half, int m_occ, const double* aoR_im,
                                           const double* aoR_half_im) {
  try {
    thc::h5::File f(path, thc::h5::File::Truncate);
    const uint64_t N = n, M = m;
    f.write_f64("real_space_grid", grid, {N, 3});
    f.write_f64("density", rho, {N, 1});
    std::vector<double> z(2 * N * M, 0.0);
    for (uint64_t i = 0; i < N * M; ++i) {
      z[2 * i] = aoR[i];
      if (aoR_im) z[2 * i + 1] = aoR_im[i];
    }
    f.write_compound_complex("aoR", z.data(), {N, M});
    if (aoR_half && m_occ > 0) {
      const uint64_t MO = m_occ;
      std::vector<double> zh(2 * N * MO, 0.0);
      for (uint64_t i = 0; i < N * MO; ++i) {
        zh[2 * i] = aoR_half[i];
        if (aoR_half_im) zh[2 * i + 1] = aoR_half_im[i];
      }
      f.write_compound_complex("aoR_half", zh.data(), {N, MO});
    }
    f.write_int("fft_grid", fft_grid, {3, 1});
    f.write_f64("fft_coulomb", coulG, {N, 1});
    std::vector<double> h(2 * M * M, 0.0);
    for (uint64_t i = 0; i < M * M; ++i) h[2 * i] = hcore[i];
    f.write_compound_complex("hcore", h.data(), {M, M});
    f.write_int("num_electrons", nelec, {1, 2});
    f.write_f64("constant_energy_factors", &e0, {1, 1});
    return 0;
  } catch (...) {
    return 1;
  }
}

// Reads a real f64 dataset (or the real part of QMCPACK's trailing-2 complex layout is left to the
// caller) into out; returns the element count, or -1.  For tests of the written THC file.
extern "C" long thc_host_read_f64(const char* path, const char* name, double* out, long cap) {
  try {
    thc::h5::File f(path, thc::h5::File::ReadOnly);
    std::vector<double> v = f.read_f64(name);
    if ((long)v.size() > cap) return -1;
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
    return (long)v.size();
  } catch (...) {
    return -1;
  }
}

extern "C" long thc_host_read_int(const char* path, const char* name, int* out, long cap) {
  try {
    thc::h5::File f(path, thc::h5::File::ReadOnly);
    std::vector<int> v = f.read_int(name);
    if ((long)v.size() > cap) return -1;
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
    return (long)v.size();
  } catch (...) {
    return -1;
  }
}

extern "C" int thc_host_hdf5_available() { return thc::h5::available() ? 1 : 0; }
