// Initial k-means indices: the first K entries of iota(N) shuffled by
// libstdc++'s std::shuffle driven by std::mt19937(seed), sorted
// (KMeansSolver::guess_initial_centroids, src/kmeans.cpp:159-175).  Compiled
// with g++ so that the small-N path is libstdc++'s own std::shuffle, the
// library the reference is built against.
//
// For N > 65535 libstdc++ 11 draws one swap position per element: step i
// (i = 1 .. N-1) swaps a[i] with a[j], j uniform in [0, i] by Lemire's nearly
// divisionless method on one 32-bit engine output (bits/stl_algo.h,
// bits/uniform_int_dist.h).  Only the first K positions are kept, and before
// step i position i still holds i (a step never touches a later position), so
// for i >= K the step only matters when j < K, and then writes i into a[j].
// The K-entry window plus the N draws reproduce the full shuffle's first K
// entries exactly; the engine runs in whole 624-word blocks (twist, then
// tempering) instead of one call per draw.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <random>
#include <vector>

namespace {

// MT19937 (Matsumoto & Nishimura 1998) with the C++ standard's parameters and
// seeding, emitting whole blocks of 624 tempered outputs.
struct Mt19937Blocks {
  static constexpr int N = 624, M = 397;
  uint32_t x[N], out[N];
  int pos = N;
  explicit Mt19937Blocks(uint32_t seed) {
    x[0] = seed;
    for (int i = 1; i < N; ++i) x[i] = 1812433253u * (x[i - 1] ^ (x[i - 1] >> 30)) + (uint32_t)i;
  }
  static uint32_t mix(uint32_t a, uint32_t b) {
    const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
    return (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
  }
  // The in-place twist x[i] = x[i+M mod N] ^ mix(x[i], x[i+1 mod N]) in three
  // runs, each reading exactly the generation the sequential loop would.
  void refill() {
    for (int i = 0; i < N - M; ++i) x[i] = x[i + M] ^ mix(x[i], x[i + 1]);
    for (int i = N - M; i < N - 1; ++i) x[i] = x[i + M - N] ^ mix(x[i], x[i + 1]);
    x[N - 1] = x[M - 1] ^ mix(x[N - 1], x[0]);
    for (int i = 0; i < N; ++i) {
      uint32_t y = x[i];
      y ^= y >> 11;
      y ^= (y << 7) & 0x9d2c5680u;
      y ^= (y << 15) & 0xefc60000u;
      y ^= y >> 18;
      out[i] = y;
    }
    pos = 0;
  }
  uint32_t operator()() {
    if (pos == N) refill();
    return out[pos++];
  }
};

// uniform in [0, range) from 32-bit outputs, Lemire's method (rejection of
// the low word below 2^32 mod range)
inline uint32_t lemire(Mt19937Blocks& g, uint32_t range) {
  uint64_t p = (uint64_t)g() * range;
  uint32_t low = (uint32_t)p;
  if (low < range) {
    const uint32_t thr = (0u - range) % range;
    while (low < thr) {
      p = (uint64_t)g() * range;
      low = (uint32_t)p;
    }
  }
  return (uint32_t)(p >> 32);
}

}  // namespace

extern "C" int thc_seed_indices_impl(long n, int k, unsigned int seed, int* out) {
  if (n < 1 || k < 1 || k > n || !out) return 1;
  if (n > 0x7fffffffL) return 1;
  if (0xffffffffUL / (unsigned long)n >= (unsigned long)n) {
    // paired draws (two swap positions per engine call): libstdc++ itself
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937 mt(seed);
    std::shuffle(perm.begin(), perm.end(), mt);
    std::sort(perm.begin(), perm.begin() + k);
    std::copy(perm.begin(), perm.begin() + k, out);
    return 0;
  }
  Mt19937Blocks g(seed);
  std::vector<int> a(k);
  std::iota(a.begin(), a.end(), 0);
  for (long i = 1; i < k; ++i) std::swap(a[i], a[lemire(g, (uint32_t)(i + 1))]);
  for (long i = k; i < n; ++i) {
    const uint32_t j = lemire(g, (uint32_t)(i + 1));
    if (j < (uint32_t)k) a[j] = (int)i;
  }
  std::sort(a.begin(), a.end());
  std::copy(a.begin(), a.end(), out);
  return 0;
}
