// Drop-in host classes for the single-rank ISDF path.  Names, arguments and
// error behaviour follow the reference's classes; the work underneath is the
// C ABI of libthcpp_amd (include/thcpp_amd.h) on one GPU.
//
//   ContextHandler::BlacsHandler          src/context_handler.h:28-39 (one rank)
//   UTILS::parse_simple_opts              src/utils.h:73-113
//   InterpolatingPoints::IPoints          src/interpolating_points.h:12-20
//   KMeans::KMeansSolver                  src/kmeans.h:13-40
//   InterpolatingVectors::IVecs           src/interpolating_vectors.h:12-41
#pragma once
#include <string>
#include <vector>

#include "../../include/thcpp_amd.h"
#include "json.h"

namespace ContextHandler {
// One process, one GPU: every rank-0 branch of the reference is taken.
struct BlacsHandler {
  int rank = 0;
  int nprocs = 1;
  int device = 0;
  BlacsHandler();
  ~BlacsHandler();
  thc_ctx* ctx = nullptr;
};
}  // namespace ContextHandler

namespace UTILS {
void parse_simple_opts(thc::Json& input, int rank, int& thc_cfac, int& thc_half_cfac, bool& half_rotated,
                       bool& find_interp_vec);
}

namespace KMeans {
class KMeansSolver {
 public:
  KMeansSolver(thc::Json& input, ContextHandler::BlacsHandler& BH);
  void kernel(ContextHandler::BlacsHandler& BH, std::vector<int>& interp_indxs, int thc_cfac);

 private:
  std::string filename;
  int max_it = 0;
  double threshold = 0;
  unsigned int rng_seed = 0;
  std::string density_label = "density";
};
}  // namespace KMeans

namespace QRCP {
// QR with column pivoting of the orbital-product matrix (src/qrcp.h, src/qrcp.cpp).
class QRCPSolver {
 public:
  QRCPSolver(thc::Json& input, ContextHandler::BlacsHandler& BH);
  void kernel(ContextHandler::BlacsHandler& BH, std::vector<int>& interp_indxs, int thc_cfac,
              bool half_rotate = false);

 private:
  std::string input_file;
  bool sub_sample = false;  // read for compatibility; the reference does not use it either
};
}  // namespace QRCP

namespace InterpolatingPoints {
class IPoints {
 public:
  IPoints(thc::Json& input, ContextHandler::BlacsHandler& BH);
  std::vector<int> kernel(thc::Json& input, ContextHandler::BlacsHandler& BH, int thc_cfac,
                          bool half_rotate = false);

 private:
  std::string solver = "kmeans";
};
}  // namespace InterpolatingPoints

namespace InterpolatingVectors {
// The constructor reads the orbitals, writes Orbitals / InterpIndx, forms CC^T
// and checks its rank as the reference's does; kernel() runs the whole
// factorisation (ZC^T, solve, FFT, Muv, Cholesky) as one thc_isdf call.
// Input keys beyond the reference's: "thc_precision" ("fp64" default, "fp32",
// "mixed") and "thc_refine_steps" (FP64 refinement steps, default 0).
class IVecs {
 public:
  IVecs(thc::Json& input, ContextHandler::BlacsHandler& BH, std::vector<int>& interp_indxs, bool half_rotated,
        bool append);
  ~IVecs();
  void kernel(ContextHandler::BlacsHandler& BH);
  void dump_qmcpack_data(int thc_cfac, int thc_half_cfac, ContextHandler::BlacsHandler& BH);

 private:
  void setup_CZt(std::vector<int>& interp_indxs, ContextHandler::BlacsHandler& BH);
  void setup_CZt_half(std::vector<int>& interp_indxs, ContextHandler::BlacsHandler& BH);
  void write_orbitals_at_points(const std::vector<double>& ao, int m, const std::vector<int>& interp_indxs,
                                const std::string& prfx, bool interleaved = false);
  void setup_CCt(std::vector<int>& interp_indxs, ContextHandler::BlacsHandler& BH);
  void check_rank(ContextHandler::BlacsHandler& BH);
  std::string input_file, output_file;
  int nbasis = 0;
  long ngrid = 0;
  int nmu = 0;
  int nocc = 0;
  bool cplx = false;
  std::vector<int> fft_grid;
  std::string prefix;
  bool half_rotate = false;
  int precision = THC_PREC_FP64;
  int refine_steps = 0;
  std::vector<double> ao_mu, aoh_mu;  // orbital rows at the points (host), for CC^T
  // device buffers
  double* d_ao = nullptr;   // aoR (n x m real, or n x 2m interleaved complex)
  double* d_aoh = nullptr;  // aoR_half (n x nocc, or n x 2 nocc interleaved), half-rotated only
  double* d_cct = nullptr;  // CC^T (nmu x nmu), or the real 2 nmu x 2 nmu embedding of a complex CC^T
  int cct_dim = 0;
  int* d_idx = nullptr;
};
}  // namespace InterpolatingVectors

// Run the reference driver's sequence (src/main.cpp:18-77) on one GPU.
int thcpp_main(const std::string& input_path);
