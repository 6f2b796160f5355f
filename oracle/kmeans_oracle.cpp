// ORACLE - TEST INFRASTRUCTURE ONLY.
//
// CPU FP64 restatement of the reference's weighted k-means (centroidal Voronoi)
// interpolation-point selector, used as the checker for the HIP path and as the
// CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it; the product library never links or calls it.
//
// Every function follows the reference line by line in *arithmetic order*
// (that order is what makes the result bit-exact), written from scratch:
//   oracle_initial_indices    src/kmeans.cpp:159-175 (guess_initial_centroids)
//   oracle_classify           src/kmeans.cpp:88-106  (classify_grid_points)
//   oracle_update_centroids   src/kmeans.cpp:108-144 (update_centroids)
//   oracle_normed_difference  src/matrix_operations.h:42-57
//   oracle_percentage_change  src/kmeans.cpp:177-185
//   oracle_map_to_grid        src/kmeans.cpp:146-157
//   oracle_kmeans             src/kmeans.cpp:187-264 (KMeansSolver::kernel, 1 rank)
//
// Build: g++ -O2 -ffp-contract=off -fopenmp (no -march: baseline x86-64 has no
// FMA, so the distance sums round after every multiply and add, exactly as the
// reference compiled without -march does).  Parity pins: tests/test_oracle_kmeans.py.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

namespace {

// Nearest centroid per point; ties resolved to the first index, as
// std::min_element over the distance vector does (kmeans.cpp:102-104).  Points
// are independent, so they are split over OpenMP threads; each point's scan
// runs in the reference's order in one thread, so the map does not depend on
// the thread count.
void classify(const double* pts, long npts, const double* cen, long ncen, int* out) {
#pragma omp parallel
  {
    std::vector<double> deltas(ncen);
#pragma omp for schedule(static)
    for (long i = 0; i < npts; ++i) {
      for (long j = 0; j < ncen; ++j) {
        // centroid minus point, C order xyz (kmeans.cpp:96-99)
        double dx = cen[j * 3 + 0] - pts[i * 3 + 0];
        double dy = cen[j * 3 + 1] - pts[i * 3 + 1];
        double dz = cen[j * 3 + 2] - pts[i * 3 + 2];
        deltas[j] = std::sqrt(dx * dx + dy * dy + dz * dz);
      }
      out[i] = (int)(std::min_element(deltas.begin(), deltas.end()) - deltas.begin());
    }
  }
}

}  // namespace

extern "C" {

// Fisher-Yates via libstdc++ std::shuffle driven by std::mt19937(seed); keep
// the first K indices, sorted ascending (kmeans.cpp:161-169).
void oracle_initial_indices(long num_grid, long num_interp, unsigned int seed,
                            int* out_idx) {
  std::vector<int> perm(num_grid);
  std::iota(perm.begin(), perm.end(), 0);
  std::mt19937 mt(seed);
  std::shuffle(perm.begin(), perm.end(), mt);
  std::vector<int> head(perm.begin(), perm.begin() + num_interp);
  std::sort(head.begin(), head.end());
  std::memcpy(out_idx, head.data(), sizeof(int) * num_interp);
}

// First 'count' outputs of std::mt19937(seed) (for the standard's known answer).
void oracle_mt19937(unsigned int seed, long count, unsigned int* out) {
  std::mt19937 mt(seed);
  for (long i = 0; i < count; ++i) out[i] = (unsigned int)mt();
}

void oracle_classify(const double* grid, long n, const double* centroids, long k,
                     int* grid_map) {
  classify(grid, n, centroids, k, grid_map);
}

// Weighted centroid update.  Sums run in ascending grid index with the product
// rounded before the add (kmeans.cpp:121-126), then one IEEE divide per
// coordinate (kmeans.cpp:131-136).  Returns 1 when some weight < 1e-8 (the
// reference prints "CRITICAL ERROR" and exits, kmeans.cpp:132,137-143).
int oracle_update_centroids(const double* rho, const double* grid, long n,
                            const int* grid_map, long k, double* centroids) {
  std::vector<double> w(k, 0.0);
  for (long j = 0; j < 3 * k; ++j) centroids[j] = 0.0;
  for (long i = 0; i < n; ++i) {
    const long c = grid_map[i];
    w[c] += rho[i];
    centroids[c * 3 + 0] += rho[i] * grid[i * 3 + 0];
    centroids[c * 3 + 1] += rho[i] * grid[i * 3 + 1];
    centroids[c * 3 + 2] += rho[i] * grid[i * 3 + 2];
  }
  int critical = 0;
  for (long c = 0; c < k; ++c) {
    if (w[c] < 1e-8) critical = 1;
    centroids[c * 3 + 0] /= w[c];
    centroids[c * 3 + 1] /= w[c];
    centroids[c * 3 + 2] /= w[c];
  }
  return critical;
}

// scale = max|a| + max|b|; diff = scale * sqrt(sum_i ((a_i-b_i)/scale)^2) with a
// sequential sum (matrix_operations.h:42-57).
double oracle_normed_difference(const double* a, const double* b, long len) {
  double maxa = *std::max_element(a, a + len), mina = *std::min_element(a, a + len);
  double maxb = *std::max_element(b, b + len), minb = *std::min_element(b, b + len);
  double sa = std::max(std::abs(maxa), std::abs(mina));
  double sb = std::max(std::abs(maxb), std::abs(minb));
  double scale = sa + sb;
  double acc = 0.0;
  for (long i = 0; i < len; ++i) {
    double t = (a[i] - b[i]) / scale;
    acc += t * t;
  }
  return scale * std::sqrt(acc);
}

double oracle_percentage_change(const int* a, const int* b, long n) {
  long changed = 0;
  for (long i = 0; i < n; ++i) changed += (a[i] != b[i]);
  return 100 * ((double)changed) / n;
}

// Nearest grid point per centroid (roles of points and centroids swapped, as
// kmeans.cpp:149 does), then sorted ascending.  Returns the number of adjacent
// duplicates (the reference only prints them, kmeans.cpp:151-155).
int oracle_map_to_grid(const double* grid, long n, const double* centroids, long k,
                       int* out_idx) {
  classify(centroids, k, grid, n, out_idx);
  std::sort(out_idx, out_idx + k);
  int dups = 0;
  for (long i = 0; i + 1 < k; ++i) dups += (out_idx[i] == out_idx[i + 1]);
  return dups;
}

// KMeansSolver::kernel at one rank (kmeans.cpp:187-264).
//   init_idx        optional K initial indices; NULL = derive them from seed
//   interp_idx      out K (left all zero when the loop ends with diff == threshold,
//                   the reference's own behaviour at kmeans.cpp:239,250)
//   diffs, pcts     out max_it entries (iterations that ran)
//   maps            optional out max_it x n grid_map per iteration
//   final_centroids optional out K x 3 (the centroids passed to map_to_grid)
//   iters_run       out number of iterations executed
// Returns 0, or 1 on the empty-cluster critical error, or 2 if duplicate
// interpolation indices were found (warning only).
int oracle_kmeans(const double* grid, const double* rho, long n, long k, int max_it,
                  double threshold, unsigned int seed, const int* init_idx,
                  int* interp_idx, double* diffs, double* pcts, int* maps,
                  double* final_centroids, int* iters_run) {
  std::vector<int> idx0(k);
  if (init_idx) {
    std::memcpy(idx0.data(), init_idx, sizeof(int) * k);
  } else {
    oracle_initial_indices(n, k, seed, idx0.data());
  }
  std::vector<double> cur(3 * k), nxt(3 * k);
  for (long j = 0; j < k; ++j)
    for (int d = 0; d < 3; ++d) cur[j * 3 + d] = grid[(long)idx0[j] * 3 + d];
  for (long j = 0; j < k; ++j) interp_idx[j] = 0;
  std::vector<int> map(n), old_map(n);
  classify(grid, n, cur.data(), k, map.data());
  old_map = map;
  double diff = 0.0;
  int status = 0;
  int it = 0;
  bool broke = false;
  for (int i = 0; i < max_it; ++i) {
    classify(grid, n, cur.data(), k, map.data());
    if (maps) std::memcpy(maps + (long)i * n, map.data(), sizeof(int) * n);
    if (oracle_update_centroids(rho, grid, n, map.data(), k, nxt.data())) {
      *iters_run = i + 1;
      return 1;
    }
    diff = oracle_normed_difference(nxt.data(), cur.data(), 3 * k);
    double pct = oracle_percentage_change(map.data(), old_map.data(), n);
    diffs[i] = diff;
    pcts[i] = pct;
    it = i + 1;
    if (diff < threshold) {
      if (final_centroids) std::memcpy(final_centroids, nxt.data(), sizeof(double) * 3 * k);
      if (oracle_map_to_grid(grid, n, nxt.data(), k, interp_idx)) status = 2;
      broke = true;
      break;
    }
    nxt.swap(cur);
    old_map.swap(map);
  }
  if (!broke && diff > threshold) {
    // the loop's last swap left the newest centroids in 'cur' (kmeans.cpp:250-256)
    if (final_centroids) std::memcpy(final_centroids, cur.data(), sizeof(double) * 3 * k);
    if (oracle_map_to_grid(grid, n, cur.data(), k, interp_idx)) status = 2;
  }
  *iters_run = it;
  return status;
}

}  // extern "C"
