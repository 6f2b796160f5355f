// Drop-in host: the reference's single-rank driver sequence over the C ABI.
#include "thcpp.h"

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "h5io.h"

namespace {

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
T* dev_alloc(size_t count) {
  void* p = nullptr;
  hip_ok(hipMalloc(&p, count * sizeof(T) + 16), "hipMalloc");
  return reinterpret_cast<T*>(p);
}

template <class T>
T* to_dev(const std::vector<T>& h) {
  T* d = dev_alloc<T>(h.size());
  hip_ok(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
  return d;
}

template <class T>
std::vector<T> to_host(const T* d, size_t count) {
  std::vector<T> h(count);
  hip_ok(hipMemcpy(h.data(), d, count * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
  return h;
}

struct DevFree {
  void* p;
  explicit DevFree(void* q) : p(q) {}
  ~DevFree() {
    if (p) (void)hipFree(p);
  }
};

void status_ok(int st, const char* where) {
  if (st == THC_OK || st >= 100) return;
  throw std::runtime_error(std::string(where) + ": " + thc_status_string(st) + " " + thc_last_error());
}

// Real orbitals from the complex aoR the reference reads (src/h5helper.cpp:19-26).
bool has_imag(const std::vector<double>& z) {
  double im = 0.0, re = 0.0;
  for (size_t i = 0; i + 1 < z.size(); i += 2) {
    re = std::max(re, std::fabs(z[i]));
    im = std::max(im, std::fabs(z[i + 1]));
  }
  return im > 1e-12 * std::max(1.0, re);
}

std::vector<double> real_part_checked(const std::vector<double>& z, const std::string& name) {
  std::vector<double> r(z.size() / 2);
  double im = 0.0, re = 0.0;
  for (size_t i = 0; i < r.size(); ++i) {
    r[i] = z[2 * i];
    re = std::max(re, std::fabs(z[2 * i]));
    im = std::max(im, std::fabs(z[2 * i + 1]));
  }
  if (im > 1e-12 * std::max(1.0, re))
    throw std::runtime_error(name + " has imaginary parts where real orbitals are required");
  return r;
}

// rows idx of a row-major (rows x width) host matrix
std::vector<double> gather_rows(const std::vector<double>& a, int width, const std::vector<int>& idx) {
  std::vector<double> out(idx.size() * (size_t)width);
  for (size_t mu = 0; mu < idx.size(); ++mu)
    std::copy(a.begin() + (size_t)idx[mu] * width, a.begin() + (size_t)(idx[mu] + 1) * width,
              out.begin() + mu * width);
  return out;
}

}  // namespace

namespace ContextHandler {
BlacsHandler::BlacsHandler() {
  const char* d = std::getenv("THC_DEVICE");
  device = d ? std::atoi(d) : 0;
  status_ok(thc_ctx_create(device, nullptr, &ctx), "thc_ctx_create");
}
BlacsHandler::~BlacsHandler() {
  if (ctx) thc_ctx_destroy(ctx);
}
}  // namespace ContextHandler

namespace UTILS {
void parse_simple_opts(thc::Json& input, int rank, int& thc_cfac, int& thc_half_cfac, bool& half_rotated,
                       bool& find_interp_vec) {
  (void)rank;
  thc_cfac = input.at("thc_cfac").as_int();
  find_interp_vec = input.contains("find_interp_vec") ? input.at("find_interp_vec").as_bool() : true;
  half_rotated = false;
  thc_half_cfac = 0;
  if (find_interp_vec) {
    if (input.contains("half_rotated")) {
      half_rotated = input.at("half_rotated").as_bool();
    } else {
      std::cout << " * Performing THC on full orbital basis." << std::endl;
    }
    if (half_rotated) {
      if (input.contains("thc_half_cfac")) {
        thc_half_cfac = input.at("thc_half_cfac").as_int();
      } else {
        std::cout << " * thc_half_cfac not specified." << std::endl;
        std::cout << " * Setting to be the same as thc_cfac." << std::endl;
        thc_half_cfac = thc_cfac;
      }
    }
  }
}
}  // namespace UTILS

namespace KMeans {

KMeansSolver::KMeansSolver(thc::Json& input, ContextHandler::BlacsHandler& BH) {
  (void)BH;
  filename = input.at("orbital_file").as_string();
  const thc::Json& km = input.at("interpolating_points").at("kmeans");
  max_it = km.at("max_it").as_int();
  threshold = km.at("threshold").as_double();
  if (km.contains("rng_seed")) {
    rng_seed = km.at("rng_seed").as_uint();
  } else {
    std::cout << " * RNG seed not set in input file." << std::endl;
    rng_seed = (unsigned int)std::chrono::system_clock::now().time_since_epoch().count();
    std::cout << " * Setting RNG seed to : " << rng_seed << std::endl;
    std::cout << std::endl;
  }
  if (km.contains("density")) {
    density_label = km.at("density").as_string();
  } else {
    density_label = "density";
    std::cout << " * Density not specified." << std::endl;
    std::cout << " * Using full basis set for density." << std::endl;
  }
}

void KMeansSolver::kernel(ContextHandler::BlacsHandler& BH, std::vector<int>& interp_indxs, int thc_cfac) {
  thc::h5::File f(filename, thc::h5::File::ReadOnly);
  std::vector<uint64_t> gd, dd;
  std::vector<double> grid = f.read_f64("real_space_grid", &gd);
  std::vector<double> rho = f.read_f64(density_label, &dd);
  const std::vector<uint64_t> ad = f.dims("aoR");
  const long n = (long)gd[0];
  const int k = thc_cfac * (int)ad[1];
  if (gd.size() != 2 || gd[1] != 3 || (long)rho.size() != n) throw std::runtime_error("bad grid/density shapes");
  std::cout << "#################################################" << std::endl;
  std::cout << "## Finding interpolation points using K-Means. ##" << std::endl;
  std::cout << "#################################################" << std::endl;
  std::cout << std::endl;
  std::cout << " * Number of interpolating points : " << k << std::endl;
  std::cout << " * Number of grid points : " << n << std::endl;
  double* d_grid = to_dev(grid);
  DevFree f1(d_grid);
  double* d_rho = to_dev(rho);
  DevFree f2(d_rho);
  int* d_idx = dev_alloc<int>(k);
  DevFree f3(d_idx);
  std::vector<double> diffs(std::max(1, max_it)), pcts(std::max(1, max_it));
  thc_kmeans_stats st = {};
  st.diffs = diffs.data();
  st.pcts = pcts.data();
  const auto t0 = std::chrono::steady_clock::now();
  const int s = thc_kmeans(BH.ctx, d_grid, d_rho, n, k, max_it, threshold, rng_seed, nullptr, d_idx, nullptr,
                           nullptr, &st);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::cout << "  *   Step        Error     % change" << std::endl;
  for (int i = 0; i < st.iterations; ++i)
    if (i % 10 == 0)
      std::cout << std::setw(10) << i << "  " << std::setprecision(5) << std::scientific << diffs[i] << "  "
                << pcts[i] << std::endl;
  std::cout.unsetf(std::ios::floatfield);
  if (s == THC_ERR_EMPTY_CLUSTER) {
    std::cout << " CRITICAL ERROR: Found centroid with weight ~ 0" << std::endl;
    std::exit(1);
  }
  status_ok(s, "thc_kmeans");
  if (!st.converged && st.last_diff > threshold)
    std::cout << " * Threshold not breached. Final Error: " << st.last_diff << std::endl;
  interp_indxs = to_host(d_idx, k);
  if (st.duplicates)
    for (int i = 0; i + 1 < k; ++i)
      if (interp_indxs[i] == interp_indxs[i + 1]) std::cout << "ERROR: Found repeated indices: " << i << std::endl;
  std::cout << " * Time taken for K-Means solution: " << secs << " seconds" << std::endl;
  std::cout << std::endl;
}

}  // namespace KMeans

namespace QRCP {

QRCPSolver::QRCPSolver(thc::Json& input, ContextHandler::BlacsHandler& BH) {
  (void)BH;
  input_file = input.at("orbital_file").as_string();
  const thc::Json& ip = input.at("interpolating_points").at("qrcp");
  if (ip.contains("sub_sample")) sub_sample = ip.at("sub_sample").as_bool();
  else std::cout << " * Performing full QRCP decomposition." << std::endl;
}

// First thc_cfac * M pivots, ascending (src/qrcp.cpp:100-173); the pivot
// magnitudes are printed as the reference prints the diagonal of R, for the
// pivots that are computed.
void QRCPSolver::kernel(ContextHandler::BlacsHandler& BH, std::vector<int>& interp_indxs, int thc_cfac,
                        bool half_rotate) {
  std::cout << " * Finding interpolating points via QRCP." << std::endl;
  std::vector<double> aoR, aoR_half;
  long n = 0;
  int m = 0, mocc = 0;
  {
    thc::h5::File fin(input_file, thc::h5::File::ReadOnly);
    const auto d = fin.dims("aoR");
    n = (long)d.at(0);
    m = (int)d.at(1);
    aoR = fin.read_complex("aoR");
    if (half_rotate) {
      const auto dh = fin.dims("aoR_half");
      mocc = (int)dh.at(1);
      aoR_half = fin.read_complex("aoR_half");
    }
  }
  // complex orbitals (any non-zero imaginary part) stay interleaved
  const bool cplx = has_imag(aoR) || (half_rotate && has_imag(aoR_half));
  if (!cplx) {
    aoR = real_part_checked(aoR, "aoR");
    if (half_rotate) aoR_half = real_part_checked(aoR_half, "aoR_half");
  }
  const int k = thc_cfac * m;
  std::cout << " * Constructing Z matrix." << std::endl;
  double* d_ao = to_dev(aoR);
  DevFree fa(d_ao);
  double* d_half = half_rotate ? to_dev(aoR_half) : nullptr;
  DevFree fh(d_half);
  int* d_idx = dev_alloc<int>((size_t)k);
  DevFree fi(d_idx);
  std::vector<double> rdiag(k);
  if (cplx)
    status_ok(thc_qrcp_points_complex(BH.ctx, d_ao, d_half, n, m, mocc, k, d_idx, rdiag.data()),
              "thc_qrcp_points_complex");
  else
    status_ok(thc_qrcp_points(BH.ctx, d_ao, d_half, n, m, mocc, k, d_idx, rdiag.data()), "thc_qrcp_points");
  interp_indxs = to_host(d_idx, (size_t)k);
  const std::string prefix = half_rotate ? "half_ix" : "full_ix";
  for (int i = 0; i < k; ++i)
    std::cout << prefix + "_diag" << " " << i << " " << std::setprecision(16) << rdiag[i] << std::endl;
}

}  // namespace QRCP

namespace InterpolatingPoints {

IPoints::IPoints(thc::Json& input, ContextHandler::BlacsHandler& BH) {
  (void)BH;
  std::cout << "#################################################" << std::endl;
  std::cout << "##   Setting up interpolating point solver.    ##" << std::endl;
  std::cout << "#################################################" << std::endl;
  std::cout << std::endl;
  const thc::Json& base = input.at("interpolating_points");
  if (base.contains("kmeans")) {
    solver = "kmeans";
  } else if (base.contains("qrcp")) {
    solver = "qrcp";
  } else {
    std::cout << "No solver specified." << std::endl;
    solver = "kmeans";
  }
}

std::vector<int> IPoints::kernel(thc::Json& input, ContextHandler::BlacsHandler& BH, int thc_cfac, bool half_rotate) {
  std::vector<int> interp_indxs;
  if (solver == "qrcp") {
    std::cout << " * Using QRCP solver." << std::endl;
    std::cout << std::endl;
    QRCP::QRCPSolver QR(input, BH);
    QR.kernel(BH, interp_indxs, thc_cfac, half_rotate);
    return interp_indxs;
  }
  // the k-means solver does not depend on half_rotate (src/interpolating_points.cpp:50-58)
  (void)half_rotate;
  std::cout << " * Using K-Means solver." << std::endl;
  std::cout << std::endl;
  KMeans::KMeansSolver KM(input, BH);
  KM.kernel(BH, interp_indxs, thc_cfac);
  return interp_indxs;
}

}  // namespace InterpolatingPoints

namespace InterpolatingVectors {

IVecs::IVecs(thc::Json& input, ContextHandler::BlacsHandler& BH, std::vector<int>& interp_indxs, bool half_rotated,
             bool append) {
  input_file = input.at("orbital_file").as_string();
  output_file = input.at("output_file").as_string();
  if (input.contains("thc_precision")) {
    const std::string p = input.at("thc_precision").as_string();
    if (p == "fp64") precision = THC_PREC_FP64;
    else if (p == "fp32") precision = THC_PREC_FP32;
    else if (p == "mixed") precision = THC_PREC_MIXED;
    else throw std::runtime_error("thc_precision must be \"fp64\", \"fp32\" or \"mixed\", not \"" + p + "\"");
  }
  if (input.contains("thc_refine_steps")) refine_steps = input.at("thc_refine_steps").as_int();
  std::cout << "#################################################" << std::endl;
  std::cout << "##   Setting up interpolative vector solver.   ##" << std::endl;
  std::cout << "#################################################" << std::endl;
  std::cout << std::endl;
  {
    thc::h5::File fin(input_file, thc::h5::File::ReadOnly);
    const auto d = fin.dims("aoR");
    ngrid = (long)d[0];
    nbasis = (int)d[1];
    fft_grid = fin.read_int("fft_grid");
  }
  if (append) thc::h5::File(output_file, thc::h5::File::Truncate);
  nmu = (int)interp_indxs.size();
  d_idx = to_dev(interp_indxs);
  if (half_rotated) {
    // to distinguish Luv between the half rotated and rotated case
    prefix = "HalfTransformed";
    half_rotate = true;
    setup_CZt_half(interp_indxs, BH);
  } else {
    prefix = "";
    half_rotate = false;
    setup_CZt(interp_indxs, BH);
  }
  setup_CCt(interp_indxs, BH);
  check_rank(BH);
}

// <prfx>Orbitals (M, K) and <prfx>InterpIndx (src/interpolating_vectors.cpp:72-90)
void IVecs::write_orbitals_at_points(const std::vector<double>& ao, int m, const std::vector<int>& interp_indxs,
                                     const std::string& prfx, bool interleaved) {
  const int k = (int)interp_indxs.size();
  std::vector<double> orb(2 * (size_t)m * k, 0.0);
  for (int i = 0; i < m; ++i)
    for (int mu = 0; mu < k; ++mu) {
      const size_t o = 2 * ((size_t)i * k + mu), src = (size_t)interp_indxs[mu] * m + i;
      if (interleaved) {
        orb[o] = ao[2 * src];
        orb[o + 1] = ao[2 * src + 1];
      } else {
        orb[o] = ao[src];
      }
    }
  thc::h5::File fo(output_file, thc::h5::File::ReadWrite);
  fo.ensure_group("/Hamiltonian/THC");
  std::cout << " * Writing aoR_mu to file" << std::endl;
  fo.write_complex_as_f64("/Hamiltonian/THC/" + prfx + "Orbitals", orb.data(), {(uint64_t)m, (uint64_t)k}, true);
  std::cout << " * Writing indices of interpolating points to file" << std::endl;
  fo.write_int("/Hamiltonian/THC/" + prfx + "InterpIndx", interp_indxs.data(), {(uint64_t)k});
}

// Half-rotated set-up (src/interpolating_vectors.cpp:143-161): the full-set and
// occupied-set orbitals at the points are written under the
// HalfTransformedFull / HalfTransformedOcc prefixes, and both orbital sets go
// to the device (interleaved when either has a non-zero imaginary part); ZC^T
// = Pocc o conj(P) itself is formed inside kernel().
void IVecs::setup_CZt_half(std::vector<int>& interp_indxs, ContextHandler::BlacsHandler& BH) {
  (void)BH;
  std::vector<double> aoR, aoR_half;
  {
    thc::h5::File fin(input_file, thc::h5::File::ReadOnly);
    aoR = fin.read_complex("aoR");
    const auto d = fin.dims("aoR_half");
    if (d.size() < 2 || (long)d[0] != ngrid) throw std::runtime_error("aoR_half must be (N, M_occ)");
    nocc = (int)d[1];
    aoR_half = fin.read_complex("aoR_half");
  }
  cplx = has_imag(aoR) || has_imag(aoR_half);
  if (!cplx) {
    aoR = real_part_checked(aoR, "aoR");
    aoR_half = real_part_checked(aoR_half, "aoR_half");
  }
  const int w = cplx ? 2 : 1;
  write_orbitals_at_points(aoR, nbasis, interp_indxs, "HalfTransformedFull", cplx);
  write_orbitals_at_points(aoR_half, nocc, interp_indxs, "HalfTransformedOcc", cplx);
  ao_mu = gather_rows(aoR, w * nbasis, interp_indxs);
  aoh_mu = gather_rows(aoR_half, w * nocc, interp_indxs);
  std::cout << " * Constructing CZt matrix" << (cplx ? " (complex orbitals)" : "") << std::endl;
  std::cout << " * Matrix Shape: (" << nmu << ", " << ngrid << ")" << std::endl;
  d_ao = to_dev(aoR);
  d_aoh = to_dev(aoR_half);
}

IVecs::~IVecs() {
  if (d_ao) (void)hipFree(d_ao);
  if (d_aoh) (void)hipFree(d_aoh);
  if (d_cct) (void)hipFree(d_cct);
  if (d_idx) (void)hipFree(d_idx);
}

// ZC^T = P o conj(P) (src/interpolating_vectors.cpp:127-141): the orbitals go
// to the device (complex ones, any non-zero imaginary part, interleaved);
// ZC^T itself is formed inside kernel().
void IVecs::setup_CZt(std::vector<int>& interp_indxs, ContextHandler::BlacsHandler& BH) {
  (void)BH;
  std::vector<double> z;
  {
    thc::h5::File fin(input_file, thc::h5::File::ReadOnly);
    z = fin.read_complex("aoR");
  }
  cplx = has_imag(z);
  std::vector<double> aoR;
  if (!cplx) aoR = real_part_checked(z, "aoR");
  write_orbitals_at_points(cplx ? z : aoR, nbasis, interp_indxs, prefix, cplx);
  ao_mu = gather_rows(cplx ? z : aoR, cplx ? 2 * nbasis : nbasis, interp_indxs);
  std::cout << " * Constructing CZt matrix" << (cplx ? " (complex orbitals)" : "") << std::endl;
  std::cout << " * Matrix Shape: (" << nmu << ", " << ngrid << ")" << std::endl;
  d_ao = to_dev(cplx ? z : aoR);
}

// CC^T = ZC^T[:, idx] (src/interpolating_vectors.cpp:163-190), formed directly
// from the orbitals at the points: the ZC^T product of the K gathered rows.
void IVecs::setup_CCt(std::vector<int>& interp_indxs, ContextHandler::BlacsHandler& BH) {
  (void)interp_indxs;
  const int k = nmu;
  std::vector<int> iota(k);
  for (int mu = 0; mu < k; ++mu) iota[mu] = mu;
  double* d_rows = to_dev(ao_mu);
  DevFree fr(d_rows);
  int* d_iota = to_dev(iota);
  DevFree fi(d_iota);
  cct_dim = k;
  if (half_rotate && cplx) {
    // complex Hermitian CC^T: kept as its real embedding [[Re, -Im], [Im, Re]],
    // which holds every singular value twice (check_rank halves the count)
    double* d_rh = to_dev(aoh_mu);
    DevFree fh(d_rh);
    double* d_c = dev_alloc<double>(2 * (size_t)k * k);
    DevFree fc(d_c);
    status_ok(thc_czt_half_complex(BH.ctx, d_rows, d_rh, k, nbasis, nocc, d_iota, k, d_c), "thc_czt_half_complex");
    const std::vector<double> c = to_host(d_c, 2 * (size_t)k * k);
    cct_dim = 2 * k;
    std::vector<double> e((size_t)cct_dim * cct_dim);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        const double re = c[2 * ((size_t)i * k + j)], im = c[2 * ((size_t)i * k + j) + 1];
        e[(size_t)i * cct_dim + j] = re;
        e[(size_t)(k + i) * cct_dim + k + j] = re;
        e[(size_t)i * cct_dim + k + j] = -im;
        e[(size_t)(k + i) * cct_dim + j] = im;
      }
    d_cct = to_dev(e);
    return;
  }
  d_cct = dev_alloc<double>((size_t)k * k);
  if (half_rotate) {
    double* d_rh = to_dev(aoh_mu);
    DevFree fh(d_rh);
    status_ok(thc_czt_half(BH.ctx, d_rows, d_rh, k, nbasis, nocc, d_iota, k, d_cct), "thc_czt_half");
  } else if (cplx) {
    status_ok(thc_czt_complex(BH.ctx, d_rows, k, nbasis, d_iota, k, d_cct), "thc_czt_complex");
  } else {
    status_ok(thc_czt(BH.ctx, d_rows, k, nbasis, d_iota, k, d_cct), "thc_czt");
  }
}

// IVecs::check_rank (src/interpolating_vectors.cpp:192-208): singular values of
// CC^T below K s0 1e-16 count as zero (MatrixOperations::rank).
void IVecs::check_rank(ContextHandler::BlacsHandler& BH) {
  std::cout << " * Checking rank of CCt matrix." << std::endl;
  int rank = 0;
  status_ok(thc_cct_rank(BH.ctx, d_cct, cct_dim, &rank, nullptr), "thc_cct_rank");
  rank /= cct_dim / nmu;
  if (rank < nmu) {
    std::cout << " * WARNING: CCt matrix is rank deficient. Rank = " << rank << std::endl;
    std::cout << std::endl;
  } else {
    std::cout << " * CCt matrix has full rank : " << rank << std::endl;
    std::cout << std::endl;
  }
}

void IVecs::kernel(ContextHandler::BlacsHandler& BH) {
  const int k = nmu;
  const long n = ngrid;
  std::cout << "#################################################" << std::endl;
  std::cout << "##        Finding interpolating vectors.       ##" << std::endl;
  std::cout << "#################################################" << std::endl;
  std::cout << std::endl;
  if ((long)fft_grid[0] * fft_grid[1] * fft_grid[2] != n)
    std::cout << " * WARNING: FFT grid not consitent with number of real space grid points." << std::endl;
  std::vector<double> coulG;
  {
    thc::h5::File fin(input_file, thc::h5::File::ReadOnly);
    coulG = fin.read_f64("fft_coulomb");
  }
  double* d_cg = to_dev(coulG);
  DevFree fc(d_cg);
  const bool half_cplx = half_rotate && cplx;
  double* d_v = dev_alloc<double>((size_t)k * k * (half_cplx ? 2 : 1));  // complex Muv interleaved
  DevFree fv(d_v);
  double* d_l = half_rotate ? nullptr : dev_alloc<double>((size_t)k * k);
  DevFree fl(d_l);
  thc_isdf_opts opts = {precision, refine_steps, half_rotate ? 0 : 1, 0};
  thc_isdf_stats st = {};
  const auto t0 = std::chrono::steady_clock::now();
  int s;
  if (half_cplx)
    s = thc_isdf_half_complex(BH.ctx, d_ao, d_aoh, n, nbasis, nocc, d_idx, k, fft_grid.data(), d_cg, &opts, d_v,
                              nullptr, &st);
  else if (half_rotate)
    s = thc_isdf_half(BH.ctx, d_ao, d_aoh, n, nbasis, nocc, d_idx, k, fft_grid.data(), d_cg, &opts, d_v, nullptr,
                      &st);
  else if (cplx)
    s = thc_isdf_complex(BH.ctx, d_ao, n, nbasis, d_idx, k, fft_grid.data(), d_cg, &opts, d_v, d_l, nullptr, &st);
  else
    s = thc_isdf(BH.ctx, d_ao, n, nbasis, d_idx, k, fft_grid.data(), d_cg, &opts, d_v, d_l, nullptr, &st);
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (s == THC_ERR_NOT_POSDEF) {
    std::cout << " * Parallel least squares failed." << std::endl;
    std::cout << " * Error code: " << st.cct_chol_info << std::endl;
  }
  status_ok(s, half_rotate ? "thc_isdf_half" : "thc_isdf");
  // the reference's stage lines, from the library's device times
  std::cout << " * Performing least squares solve." << std::endl;
  std::cout << " * Time for least squares solve : " << (st.stage_ms[0] + st.stage_ms[1] + st.stage_ms[2]) * 1e-3
            << " seconds" << std::endl;
  std::cout << std::endl;
  std::cout << "#################################################" << std::endl;
  std::cout << "##        Constructing Muv matrix.             ##" << std::endl;
  std::cout << "#################################################" << std::endl;
  std::cout << std::endl;
  std::cout << " * Performing FFT on interpolating vectors." << std::endl;
  std::cout << std::endl;
  std::cout << " * Time to FFT interpolating vectors : " << st.stage_ms[3] * 1e-3 << " seconds" << std::endl;
  std::cout << std::endl;
  std::cout << " * Constructing Muv." << std::endl;
  std::cout << "  * Time to construct Muv: " << st.stage_ms[4] * 1e-3 << " seconds." << std::endl;
  std::cout << " * Factorisation (" << (precision == THC_PREC_MIXED ? "mixed" : precision == THC_PREC_FP32 ? "fp32"
                                                                                                       : "fp64")
            << ") in " << wall << " seconds." << std::endl;
  // dump_thc_data (src/interpolating_vectors.cpp:283-328)
  if (half_rotate) {
    // Muv = IVMG^T IVG, written full with no Cholesky: real for real
    // interpolating vectors (IVMG = conj(IVG)), complex symmetric otherwise
    std::vector<double> mh = to_host(d_v, (size_t)k * k * (half_cplx ? 2 : 1));
    std::cout << " * Dumping half transformed Muv data to: " << output_file << "." << std::endl;
    std::cout << std::endl;
    std::vector<double> muv(2 * (size_t)k * k, 0.0);
    if (half_cplx) muv = mh;
    else
      for (size_t i = 0; i < (size_t)k * k; ++i) muv[2 * i] = mh[i];
    thc::h5::File fo(output_file, thc::h5::File::ReadWrite);
    fo.ensure_group("/Hamiltonian/THC");
    fo.write_complex_as_f64("/Hamiltonian/THC/" + prefix + "Muv", muv.data(), {(uint64_t)k, (uint64_t)k}, true);
    return;
  }
  std::cout << " * Performing Cholesky decomposition on Muv." << std::endl;
  std::cout << " * Time to perform cholesky decomposition on Muv: " << st.stage_ms[5] * 1e-3 << " seconds."
            << std::endl;
  std::vector<double> l = to_host(d_l, (size_t)k * k);
  std::cout << " * Dumping Luv to: " << output_file << "." << std::endl;
  if (s == THC_WARN_LUV_NOT_POSDEF) {
    std::cout << "  * WARNING: Cholesky decomposition failed!" << std::endl;
    std::cout << "  * SCALAPACK Error code: " << st.luv_chol_info << std::endl;
    std::cout << "   * The leading minor of order " << st.luv_chol_info << " is not positive definite."
              << std::endl;
  }
  std::cout << std::endl;
  std::vector<double> luv(2 * (size_t)k * k, 0.0);
  for (size_t i = 0; i < (size_t)k * k; ++i) luv[2 * i] = l[i];
  thc::h5::File fo(output_file, thc::h5::File::ReadWrite);
  fo.ensure_group("/Hamiltonian/THC");
  fo.write_complex_as_f64("/Hamiltonian/THC/" + prefix + "Luv", luv.data(), {(uint64_t)k, (uint64_t)k}, true);
}

void IVecs::dump_qmcpack_data(int thc_cfac, int thc_half_cfac, ContextHandler::BlacsHandler& BH) {
  (void)BH;
  thc::h5::File fin(input_file, thc::h5::File::ReadOnly);
  thc::h5::File fo(output_file, thc::h5::File::ReadWrite);
  fo.ensure_group("/Hamiltonian");
  std::vector<uint64_t> hd;
  std::vector<double> hcore = fin.read_complex("hcore", &hd);
  fo.write_complex_as_f64("/Hamiltonian/hcore", hcore.data(), hd, false);
  std::vector<int> ne = fin.read_int("num_electrons");
  const int nup = ne.at(0), ndown = ne.at(1);
  std::vector<double> e0 = fin.read_f64("constant_energy_factors");
  fo.write_f64("/Hamiltonian/Energies", e0.data(), {(uint64_t)e0.size()});
  std::vector<int> occups(nup + ndown, 0);
  fo.write_int("/Hamiltonian/occups", occups.data(), {(uint64_t)occups.size()});
  std::vector<int> dims = {-1, 0, 0, nbasis, nup, ndown, 0, 0};
  fo.write_int("/Hamiltonian/dims", dims.data(), {8});
  // thc_half_cfac is only defined on the half-rotated path (src/main.cpp:38); 0 otherwise
  std::vector<int> thc_dims = {nbasis, thc_cfac * nbasis, thc_half_cfac * nbasis};
  fo.ensure_group("/Hamiltonian/THC");
  fo.write_int("/Hamiltonian/THC/dims", thc_dims.data(), {3});
}

}  // namespace InterpolatingVectors

int thcpp_main(const std::string& input_path) {
  ContextHandler::BlacsHandler BH;
  const auto t0 = std::chrono::steady_clock::now();
  std::ifstream in(input_path);
  if (!in) throw std::runtime_error("cannot read " + input_path);
  std::stringstream ss;
  ss << in.rdbuf();
  thc::Json input = thc::Json::parse(ss.str());
  std::cout << " * Running on 1 GPU (" << thc_version() << ")." << std::endl;
  int thc_cfac = 0, thc_half_cfac = 0;
  bool half_rotate = false, find_interp_vec = true;
  UTILS::parse_simple_opts(input, BH.rank, thc_cfac, thc_half_cfac, half_rotate, find_interp_vec);
  InterpolatingPoints::IPoints IPSolver(input, BH);
  std::vector<int> interp_indxs = IPSolver.kernel(input, BH, thc_cfac, false);
  if (find_interp_vec) {
    InterpolatingVectors::IVecs IVSolver(input, BH, interp_indxs, false, true);
    IVSolver.kernel(BH);
    if (!half_rotate) IVSolver.dump_qmcpack_data(thc_cfac, thc_half_cfac, BH);
  }
  if (find_interp_vec && half_rotate) {
    // second pass over the half-rotated products (src/main.cpp:63-72)
    if (thc_half_cfac != thc_cfac) {
      InterpolatingPoints::IPoints IP(input, BH);
      interp_indxs = IP.kernel(input, BH, thc_half_cfac, half_rotate);
    }
    InterpolatingVectors::IVecs IVSolver(input, BH, interp_indxs, half_rotate, false);
    IVSolver.kernel(BH);
    IVSolver.dump_qmcpack_data(thc_cfac, thc_half_cfac, BH);
  }
  std::cout << " * Total simulation time : "
            << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() << " seconds."
            << std::endl;
  return 0;
}
