// k-means constants, device state and host entry points (kmeans.hip).
#pragma once
#include "common.h"

namespace thc {
namespace km {

constexpr int TP = 256;        // grid points per assignment tile
constexpr int CS_TILE = 1024;  // points per counting-sort tile (index order)
constexpr int CS_SCAN_W = 16;  // waves of a column-scan workgroup
constexpr int MAXB = 4096;     // max bins: clusters, and Morton cells of the spatial sort
constexpr int CAND_CAP = 512;  // candidate centroids cached in LDS per tile

enum { MODE_LATEST = 0, MODE_FINAL = 1, MODE_ZEROS = 2 };

// Device-side loop state (one per run).
struct State {
  int done;        // loop ended: diff < threshold, or an error
  int final_buf;   // centroid buffer with the converged centroids (MODE_FINAL)
  int error;       // 1: a cluster weight < 1e-8 in an iteration the loop ran; 2: bad cluster key
  int changed[2];  // points whose cluster changed, by iteration parity
  int serr[2];     // a cluster weight < 1e-8 in the centroid sums, by iteration parity
  int iters;       // iterations executed
  int mode;        // which centroids map_to_grid uses (MODE_*)
  int dups;        // repeated indices after map_to_grid
  double last_diff;
};

}  // namespace km
}  // namespace thc

// Instantiated HIP graph of one k-means run, reused while the arguments match.
struct KmGraphCache {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* key_ptr[5] = {};
  long n = 0;
  int k = 0, max_it = 0;
  double threshold = 0;
  void* base = nullptr;
  bool overlap = false;
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

namespace thc {

int kmeans_run(thc_ctx* ctx, const double* grid, const double* rho, long n, int k, int max_it,
               double threshold, const int* host_init_idx, int* interp_idx, int* grid_map_out,
               double* centroids_out, thc_kmeans_stats* stats);
int kmeans_classify(thc_ctx* ctx, const double* grid, long n, const double* centroids, int k, int* grid_map);
int kmeans_update(thc_ctx* ctx, const double* rho, const double* grid, long n, const int* grid_map, int k,
                  double* centroids);
int kmeans_map_to_grid(thc_ctx* ctx, const double* grid, long n, const double* centroids, int k,
                       int* interp_idx);

// One system of a batched k-means (kmeans_run_many).
struct KmSpec {
  const double* grid;
  const double* rho;
  long n;
  int k, max_it;
  double threshold;
  unsigned int seed;
  int* interp_idx;  // out, k ints
};
// The k-means of nj systems in one set of launches on the context's stream
// (each launch covers one step of every system).  status[j] gets what
// thc_kmeans would return for system j; stats (optional, nj entries) its
// statistics, elapsed_ms being the device time of the whole batch.
int kmeans_run_many(thc_ctx* ctx, int nj, const KmSpec* spec, int* status, thc_kmeans_stats* stats);

}  // namespace thc
