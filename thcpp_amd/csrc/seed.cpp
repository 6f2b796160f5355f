// Initial k-means indices.  Compiled with g++ so that std::shuffle and
// std::mt19937 are libstdc++'s, the library the reference is built against:
// iota(N), std::shuffle driven by std::mt19937(seed), first K, sorted
// (KMeansSolver::guess_initial_centroids, src/kmeans.cpp:159-175).
#include <algorithm>
#include <numeric>
#include <random>
#include <vector>

extern "C" int thc_seed_indices_impl(long n, int k, unsigned int seed, int* out) {
  if (n < 1 || k < 1 || k > n || !out) return 1;
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::mt19937 mt(seed);
  std::shuffle(perm.begin(), perm.end(), mt);
  std::sort(perm.begin(), perm.begin() + k);
  std::copy(perm.begin(), perm.begin() + k, out);
  return 0;
}
