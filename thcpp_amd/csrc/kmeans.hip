// Weighted k-means (centroidal Voronoi) interpolation points on gfx950,
// bit-identical to the reference's sequential FP64 loop (src/kmeans.cpp).
//
// Exactness rules kept by every kernel here:
//  * distances are sqrt((dx*dx + dy*dy) + dz*dz) with dx = centroid - point,
//    every multiply and add rounded separately (no FMA contraction, below), and
//    sqrt correctly rounded (the device sqrt is; tests check it);
//  * the argmin takes the first index among equal rounded distances
//    (std::min_element), evaluated as a lexicographic (distance, index) minimum;
//    sqrt is monotone, so a sqrt is only needed when d^2 improves;
//  * centroid sums run per cluster in ascending grid index with the product
//    rounded before the add (a stable counting sort supplies that order), and
//    are divided once (IEEE double division);
//  * pruning of centroids (or grid tiles) is conservative: a candidate is
//    dropped only if its distance bound exceeds the best bound by a relative
//    1e-8 on squares, far above the few-ulp rounding of any distance.
#pragma clang fp contract(off)
#include <atomic>
#include <cfloat>
#include <climits>
#include <vector>

#include "common.h"
#include "kmeans.h"

namespace thc {
namespace km {

__device__ inline double dist2(double cx, double cy, double cz, double gx, double gy, double gz) {
  const double dx = cx - gx;
  const double dy = cy - gy;
  const double dz = cz - gz;
  return (dx * dx + dy * dy) + dz * dz;
}

// Lexicographic (rounded distance, index) update for candidates visited in
// any order.  bd2 is the exact d^2 of the current best.
__device__ inline void consider_any(double d2, int idx, double& bd2, double& bs, int& bi) {
  if (d2 < bd2) {
    const double s = sqrt(d2);
    if (s < bs || (s == bs && idx < bi)) { bs = s; bd2 = d2; bi = idx; }
  } else if (d2 == bd2) {
    if (idx < bi) bi = idx;
  } else if (idx < bi && d2 <= bd2 * (1.0 + 0x1p-50)) {
    // d^2 slightly above the best may still round to the same distance
    const double s = sqrt(d2);
    if (s == bs) bi = idx;
  }
}

__device__ inline double lb2_box(double cx, double cy, double cz, const double* b) {
  const double gx = fmax(0.0, fmax(b[0] - cx, cx - b[3]));
  const double gy = fmax(0.0, fmax(b[1] - cy, cy - b[4]));
  const double gz = fmax(0.0, fmax(b[2] - cz, cz - b[5]));
  return gx * gx + gy * gy + gz * gz;
}
__device__ inline double ub2_box(double cx, double cy, double cz, const double* b) {
  const double ex = fmax(fabs(cx - b[0]), fabs(cx - b[3]));
  const double ey = fmax(fabs(cy - b[1]), fabs(cy - b[4]));
  const double ez = fmax(fabs(cz - b[2]), fabs(cz - b[5]));
  return ex * ex + ey * ey + ez * ez;
}
constexpr double PRUNE = 1.0 + 1e-8;

__device__ inline double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}
__device__ inline int wave_sum(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// ------------------------------------------------------------------ setup

__device__ __forceinline__ void bbox_partial_b(const double* grid, long n, double* part, int bid, int nblk) {
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (long i = bid * 256L + threadIdx.x; i < n; i += 256L * nblk)
    for (int d = 0; d < 3; ++d) {
      const double v = grid[i * 3 + d];
      lo[d] = fmin(lo[d], v);
      hi[d] = fmax(hi[d], v);
    }
  __shared__ double s[6][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int d = 0; d < 3; ++d) {
    double a = wave_min(lo[d]), b = -wave_min(-hi[d]);
    if (lane == 0) { s[d][w] = a; s[3 + d][w] = b; }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int d = threadIdx.x;
    double r = s[d][0];
    for (int q = 1; q < 4; ++q) r = d < 3 ? fmin(r, s[d][q]) : fmax(r, s[d][q]);
    part[bid * 6 + d] = r;
  }
}
__global__ __launch_bounds__(256) void bbox_partial(const double* grid, long n, double* part) {
  bbox_partial_b(grid, n, part, (int)blockIdx.x, (int)gridDim.x);
}

__device__ __forceinline__ void bbox_final_b(const double* part, int nb, double* box, int bid, int nblk) {
  const int d = threadIdx.x;
  if (d >= 6) return;
  double r = part[d];
  for (int b = 1; b < nb; ++b) r = d < 3 ? fmin(r, part[b * 6 + d]) : fmax(r, part[b * 6 + d]);
  box[d] = r;
}
__global__ void bbox_final(const double* part, int nb, double* box) {
  bbox_final_b(part, nb, box, (int)blockIdx.x, (int)gridDim.x);
}

__device__ inline unsigned spread3(unsigned v) {
  unsigned r = 0;
  for (int b = 0; b < 10; ++b) r |= ((v >> b) & 1u) << (3 * b);
  return r;
}

__device__ __forceinline__ void morton_keys_b(const double* grid, long n, const double* box, int cells,
                                                   int* keys, int bid, int nblk) {
  const long i = bid * 256L + threadIdx.x;
  if (i >= n) return;
  unsigned c[3];
  for (int d = 0; d < 3; ++d) {
    const double lo = box[d], ext = box[3 + d] - box[d];
    int v = 0;
    if (ext > 0) v = (int)((grid[i * 3 + d] - lo) / ext * cells);
    c[d] = (unsigned)min(max(v, 0), cells - 1);
  }
  keys[i] = (int)((spread3(c[0]) << 2) | (spread3(c[1]) << 1) | spread3(c[2]));
}
__global__ __launch_bounds__(256) void morton_keys(const double* grid, long n, const double* box, int cells,
                                                   int* keys) {
  morton_keys_b(grid, n, box, cells, keys, (int)blockIdx.x, (int)gridDim.x);
}

__device__ __forceinline__ void gather_sorted_b(const double* grid, const int* perm, long n, double* sx,
                                                     double* sy, double* sz, int bid, int nblk) {
  const long p = bid * 256L + threadIdx.x;
  if (p >= n) return;
  const long i = perm[p];
  sx[p] = grid[i * 3 + 0];
  sy[p] = grid[i * 3 + 1];
  sz[p] = grid[i * 3 + 2];
}
__global__ __launch_bounds__(256) void gather_sorted(const double* grid, const int* perm, long n, double* sx,
                                                     double* sy, double* sz) {
  gather_sorted_b(grid, perm, n, sx, sy, sz, (int)blockIdx.x, (int)gridDim.x);
}

__device__ __forceinline__ void tile_boxes_b(const double* sx, const double* sy, const double* sz, long n,
                                                  double* boxes, int bid, int nblk) {
  const long p = bid * (long)TP + threadIdx.x;
  double v[3];
  const bool ok = p < n;
  v[0] = ok ? sx[p] : 0;
  v[1] = ok ? sy[p] : 0;
  v[2] = ok ? sz[p] : 0;
  __shared__ double s[6][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int d = 0; d < 3; ++d) {
    double a = wave_min(ok ? v[d] : DBL_MAX), b = -wave_min(ok ? -v[d] : DBL_MAX);
    if (lane == 0) { s[d][w] = a; s[3 + d][w] = b; }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int d = threadIdx.x;
    double r = s[d][0];
    for (int q = 1; q < 4; ++q) r = d < 3 ? fmin(r, s[d][q]) : fmax(r, s[d][q]);
    boxes[bid * 6 + d] = r;
  }
}
__global__ __launch_bounds__(256) void tile_boxes(const double* sx, const double* sy, const double* sz, long n,
                                                  double* boxes) {
  tile_boxes_b(sx, sy, sz, n, boxes, (int)blockIdx.x, (int)gridDim.x);
}

__device__ __forceinline__ void init_centroids_b(const double* grid, const int* idx, int k, double* cen, int bid, int nblk) {
  const int j = bid * 256 + threadIdx.x;
  if (j >= k) return;
  const long i = idx[j];
  cen[j * 3 + 0] = grid[i * 3 + 0];
  cen[j * 3 + 1] = grid[i * 3 + 1];
  cen[j * 3 + 2] = grid[i * 3 + 2];
}
__global__ __launch_bounds__(256) void init_centroids(const double* grid, const int* idx, int k, double* cen) {
  init_centroids_b(grid, idx, k, cen, (int)blockIdx.x, (int)gridDim.x);
}

__device__ __forceinline__ void init_state_b(State* st, int mode, int bid, int nblk) {
  if (threadIdx.x == 0) {
    st->done = 0;
    st->final_buf = 0;
    st->error = 0;
    st->changed[0] = st->changed[1] = 0;
    st->serr[0] = st->serr[1] = 0;
    st->iters = 0;
    st->mode = mode;
    st->dups = 0;
    st->last_diff = 0.0;
  }
}
__global__ void init_state(State* st, int mode) {
  init_state_b(st, mode, (int)blockIdx.x, (int)gridDim.x);
}

// ------------------------------------------------------------------ stable counting sort

// Per index-order tile of CS_TILE points: histogram of keys.
__device__ __forceinline__ void cs_hist_b(const int* keys, long n, int nbins, int* tile_count,
                                               const State* st, int bid, int nblk) {
  if (st && st->done) return;
  __shared__ int h[MAXB];
  for (int b = threadIdx.x; b < nbins; b += 256) h[b] = 0;
  __syncthreads();
  const long base = (long)bid * CS_TILE;
  for (int e = threadIdx.x; e < CS_TILE; e += 256) {
    const long i = base + e;
    if (i < n) {
      const int key = keys[i];
      if ((unsigned)key < (unsigned)nbins) atomicAdd(&h[key], 1);
    }
  }
  __syncthreads();
  int* out = tile_count + (long)bid * nbins;
  for (int b = threadIdx.x; b < nbins; b += 256) out[b] = h[b];
}
__global__ __launch_bounds__(256) void cs_hist(const int* keys, long n, int nbins, int* tile_count,
                                               const State* st) {
  cs_hist_b(keys, n, nbins, tile_count, st, (int)blockIdx.x, (int)gridDim.x);
}

// Column scan over tiles, in place: tile_count[t][b] -> exclusive offset of
// tile t within bin b; total[b] = count of bin b.  A workgroup takes 64 bins;
// its CS_SCAN_W waves scan consecutive runs of the tiles, then add the earlier
// runs' totals (16 waves: two dependent load rounds per wave at 256 tiles,
// where 4 waves needed eight).
__device__ __forceinline__ void cs_colscan_b(int* tile_count, int ntiles, int nbins, int* total,
                                                  const State* st, int bid, int nblk) {
  if (st && st->done) return;
  __shared__ int part[CS_SCAN_W][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int b = bid * 64 + lane;
  const int per = (ntiles + CS_SCAN_W - 1) / CS_SCAN_W;
  const int t0 = wave * per, t1 = min(ntiles, t0 + per);
  // chunks of 16 tiles: every load of a chunk is issued before its first use
  // and before the chunk's stores (on this target stores and loads share one
  // completion counter, so a load after a store would wait for the store)
  constexpr int CH = 16;
  int run = 0;
  if (b < nbins)
    for (int t = t0; t < t1; t += CH) {
      int c[CH];
#pragma unroll
      for (int q = 0; q < CH; ++q) c[q] = (t + q < t1) ? tile_count[(long)(t + q) * nbins + b] : 0;
#pragma unroll
      for (int q = 0; q < CH; ++q) run += c[q];
    }
  part[wave][lane] = run;
  __syncthreads();
  int off = 0;
  for (int w = 0; w < wave; ++w) off += part[w][lane];
  if (b < nbins) {
    for (int t = t0; t < t1; t += CH) {
      int c[CH];
#pragma unroll
      for (int q = 0; q < CH; ++q) c[q] = (t + q < t1) ? tile_count[(long)(t + q) * nbins + b] : 0;
#pragma unroll
      for (int q = 0; q < CH; ++q)
        if (t + q < t1) {
          tile_count[(long)(t + q) * nbins + b] = off;
          off += c[q];
        }
    }
    if (wave == CS_SCAN_W - 1) total[b] = off;
  }
}
__global__ __launch_bounds__(64 * CS_SCAN_W) void cs_colscan(int* tile_count, int ntiles, int nbins, int* total,
                                                  const State* st) {
  cs_colscan_b(tile_count, ntiles, nbins, total, st, (int)blockIdx.x, (int)gridDim.x);
}

// Exclusive scan of total[0..nbins) into start (one workgroup of 1024).
__device__ __forceinline__ void cs_totscan_b(const int* total, int nbins, int* start, const State* st, int bid, int nblk) {
  if (st && st->done) return;
  __shared__ int s[1024];
  const int t = threadIdx.x;
  const int per = (nbins + 1023) / 1024;
  int loc = 0;
  for (int q = 0; q < per; ++q) {
    const int b = t * per + q;
    if (b < nbins) loc += total[b];
  }
  s[t] = loc;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = t >= off ? s[t - off] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  int run = s[t] - loc;  // exclusive
  for (int q = 0; q < per; ++q) {
    const int b = t * per + q;
    if (b < nbins) {
      start[b] = run;
      run += total[b];
    }
  }
}
__global__ __launch_bounds__(1024) void cs_totscan(const int* total, int nbins, int* start, const State* st) {
  cs_totscan_b(total, nbins, start, st, (int)blockIdx.x, (int)gridDim.x);
}

// Stable scatter: out[start[key] + rank] = index, rank counting earlier
// indices with the same key.  Chunks of 256 go in order; inside a chunk the
// four waves go in order; inside a wave, lanes with equal keys are ranked by
// lane through a 'nbits'-ballot match.
__device__ __forceinline__ void cs_scatter_b(const int* keys, long n, int nbins, int nbits,
                                                  const int* tile_off, const int* start, int* out,
                                                  const State* st, int bid, int nblk) {
  if (st && st->done) return;
  __shared__ int cnt[MAXB];
  __shared__ int slot[256];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int* toff = tile_off + (long)bid * nbins;
  for (int b = tid; b < nbins; b += 256) cnt[b] = start[b] + toff[b];
  __syncthreads();
  const long base = (long)bid * CS_TILE;
  for (int ch = 0; ch < CS_TILE / 256; ++ch) {
    const long i = base + ch * 256 + tid;
    int key = i < n ? keys[i] : 0;
    const bool valid = i < n && (unsigned)key < (unsigned)nbins;
    if (!valid) key = 0;
    unsigned long long peers = __ballot(valid);
    for (int bit = 0; bit < nbits; ++bit) {
      const bool on = (key >> bit) & 1;
      const unsigned long long bal = __ballot(valid && on);
      peers &= on ? bal : ~bal;
    }
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int rank = __popcll(peers & below);
    const int leader = valid ? (__ffsll((long long)peers) - 1) : 0;
    for (int w = 0; w < 4; ++w) {
      if (wave == w && valid && rank == 0) {
        const int b0 = cnt[key];
        cnt[key] = b0 + __popcll(peers);
        slot[tid] = b0;
      }
      __syncthreads();
    }
    if (valid) out[slot[wave * 64 + leader] + rank] = (int)i;
    __syncthreads();
  }
}
__global__ __launch_bounds__(256) void cs_scatter(const int* keys, long n, int nbins, int nbits,
                                                  const int* tile_off, const int* start, int* out,
                                                  const State* st) {
  cs_scatter_b(keys, n, nbins, nbits, tile_off, start, out, st, (int)blockIdx.x, (int)gridDim.x);
}

// ------------------------------------------------------------------ Lloyd iteration

__device__ __forceinline__ void assign_b(const double* sx, const double* sy, const double* sz,
                                              const int* perm, long n, const double* boxes,
                                              const double* cen, int k, int* map_new, const int* map_old,
                                              State* st, int par, int bid, int nblk) {
  if (st && st->done) return;
  __shared__ int cidx[MAXB];
  __shared__ double ccx[CAND_CAP], ccy[CAND_CAP], ccz[CAND_CAP];
  __shared__ double redd[4];
  __shared__ int redi[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double* b = boxes + (long)bid * 6;
  double box[6];
  for (int d = 0; d < 6; ++d) box[d] = b[d];

  // this thread's centroids j = q * 256 + tid (k <= MAXB), loaded once
  constexpr int CPT = MAXB / 256;
  double rx[CPT], ry[CPT], rz[CPT];
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    const int j = q * 256 + tid;
    const bool ok = j < k;
    rx[q] = ok ? cen[j * 3] : 0.0;
    ry[q] = ok ? cen[j * 3 + 1] : 0.0;
    rz[q] = ok ? cen[j * 3 + 2] : 0.0;
  }
  double u = DBL_MAX;
#pragma unroll
  for (int q = 0; q < CPT; ++q)
    if (q * 256 + tid < k) u = fmin(u, ub2_box(rx[q], ry[q], rz[q], box));
  u = wave_min(u);
  if (lane == 0) redd[wave] = u;
  __syncthreads();
  u = fmin(fmin(redd[0], redd[1]), fmin(redd[2], redd[3]));
  const double thr = u * PRUNE;

  int ncand = 0;
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    if (q * 256 >= k) break;  // uniform over the workgroup
    const int j = q * 256 + tid;
    const double cx = rx[q], cy = ry[q], cz = rz[q];
    const bool flag = j < k && lb2_box(cx, cy, cz, box) <= thr;
    const unsigned long long bal = __ballot(flag);
    __syncthreads();
    if (lane == 0) redi[wave] = __popcll(bal);
    __syncthreads();
    int off = ncand;
    for (int w = 0; w < wave; ++w) off += redi[w];
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    off += __popcll(bal & below);
    if (flag) {
      cidx[off] = j;
      if (off < CAND_CAP) { ccx[off] = cx; ccy[off] = cy; ccz[off] = cz; }
    }
    ncand += redi[0] + redi[1] + redi[2] + redi[3];
  }
  __syncthreads();

  const long p = (long)bid * TP + tid;
  int changed = 0;
  if (p < n) {
    const double gx = sx[p], gy = sy[p], gz = sz[p];
    // ascending candidate order: only a strictly smaller rounded distance wins
    int bj = cidx[0];
    double bd2 = dist2(ccx[0], ccy[0], ccz[0], gx, gy, gz);
    double bs = sqrt(bd2);
    // eight candidates' distances at a time (their LDS reads in flight
    // together), then the updates in ascending order as one by one
    constexpr int U = 8;
    const int cached = min(ncand, CAND_CAP);
    int c = 1;
    for (; c + U <= cached; c += U) {
      double d2v[U];
      int jv[U];
#pragma unroll
      for (int q = 0; q < U; ++q) {
        jv[q] = cidx[c + q];
        d2v[q] = dist2(ccx[c + q], ccy[c + q], ccz[c + q], gx, gy, gz);
      }
#pragma unroll
      for (int q = 0; q < U; ++q) {
        if (d2v[q] < bd2) {
          const double sq = sqrt(d2v[q]);
          if (sq < bs) { bs = sq; bd2 = d2v[q]; bj = jv[q]; }
        }
      }
    }
    for (; c < ncand; ++c) {
      double cx, cy, cz;
      const int j = cidx[c];
      if (c < CAND_CAP) { cx = ccx[c]; cy = ccy[c]; cz = ccz[c]; }
      else { cx = cen[j * 3]; cy = cen[j * 3 + 1]; cz = cen[j * 3 + 2]; }
      const double d2 = dist2(cx, cy, cz, gx, gy, gz);
      if (d2 < bd2) {
        const double s = sqrt(d2);
        if (s < bs) { bs = s; bd2 = d2; bj = j; }
      }
    }
    const long orig = perm[p];
    if (map_old) changed = (map_old[orig] != bj);
    map_new[orig] = bj;
  }
  if (st && map_old) {
    changed = wave_sum(changed);
    if (lane == 0 && changed) atomicAdd(&st->changed[par], changed);
  }
}
__global__ __launch_bounds__(256) void assign(const double* sx, const double* sy, const double* sz,
                                              const int* perm, long n, const double* boxes,
                                              const double* cen, int k, int* map_new, const int* map_old,
                                              State* st, int par) {
  assign_b(sx, sy, sz, perm, n, boxes, cen, k, map_new, map_old, st, par, (int)blockIdx.x, (int)gridDim.x);
}

// One wave per cluster.  Lanes gather the members' rho and rho * x (each
// product rounded on its own, as the reference's multiply-then-add does) into
// LDS in ascending index order, then lanes 0..3 each run one of the four
// sequential sums (w, x, y, z) over them.
constexpr int CS_CHUNK = 256;

__device__ __forceinline__ void centroid_sums_b(const int* order, const int* start, const int* total,
                                                     const double* rho, const double* grid, int k,
                                                     double* cen_out, State* st, int par, int bid, int nblk) {
  if (st && st->done) return;
  __shared__ double buf[4][4][CS_CHUNK];  // [wave][w, x, y, z][member]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = bid * 4 + wave;
  if (c >= k) return;
  const int s = start[c], cnt = total[c];
  double acc = 0.0;  // lane q < 4 owns accumulator q
  for (int base = 0; base < cnt; base += CS_CHUNK) {
    const int len = min(CS_CHUNK, cnt - base);
    for (int t = lane; t < len; t += 64) {
      const long i = order[s + base + t];
      const double r = rho[i];
      buf[wave][0][t] = r;
      buf[wave][1][t] = r * grid[i * 3 + 0];
      buf[wave][2][t] = r * grid[i * 3 + 1];
      buf[wave][3][t] = r * grid[i * 3 + 2];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (lane < 4) {
      const double* v = buf[wave][lane];
      for (int t = 0; t < len; ++t) acc += v[t];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  const double w = __shfl(acc, 0);
  if (lane == 0 && w < 1e-8) atomicOr(&st->serr[par], 1);
  if (lane >= 1 && lane < 4) cen_out[c * 3 + lane - 1] = acc / w;
}
__global__ __launch_bounds__(256) void centroid_sums(const int* order, const int* start, const int* total,
                                                     const double* rho, const double* grid, int k,
                                                     double* cen_out, State* st, int par) {
  centroid_sums_b(order, start, total, rho, grid, k, cen_out, st, par, (int)blockIdx.x, (int)gridDim.x);
}

// normed_difference(new, cur) with the reference's sequential sum, % changed,
// and the loop control of KMeansSolver::kernel (kmeans.cpp:233-257).
__device__ __forceinline__ void diff_control_b(const double* nxt, const double* cur, int k, long n,
                                                     double threshold, int it, int max_it, int nxt_buf,
                                                     State* st, double* diffs, double* pcts, int bid, int nblk) {
  if (st->done) return;
  extern __shared__ double terms[];
  __shared__ double red[4][16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int len = 3 * k, par = it & 1;
  if (st->serr[par]) {
    if (tid == 0) { st->error = 1; st->done = 1; st->iters = it + 1; }
    return;
  }
  double mxa = -DBL_MAX, mna = DBL_MAX, mxb = -DBL_MAX, mnb = DBL_MAX;
  for (int i = tid; i < len; i += 1024) {
    mxa = fmax(mxa, nxt[i]); mna = fmin(mna, nxt[i]);
    mxb = fmax(mxb, cur[i]); mnb = fmin(mnb, cur[i]);
  }
  mxa = -wave_min(-mxa); mna = wave_min(mna); mxb = -wave_min(-mxb); mnb = wave_min(mnb);
  if (lane == 0) { red[0][wave] = mxa; red[1][wave] = mna; red[2][wave] = mxb; red[3][wave] = mnb; }
  __syncthreads();
  mxa = red[0][0]; mna = red[1][0]; mxb = red[2][0]; mnb = red[3][0];
  for (int w = 1; w < 16; ++w) {
    mxa = fmax(mxa, red[0][w]); mna = fmin(mna, red[1][w]);
    mxb = fmax(mxb, red[2][w]); mnb = fmin(mnb, red[3][w]);
  }
  const double scale = fmax(fabs(mxa), fabs(mna)) + fmax(fabs(mxb), fabs(mnb));
  for (int i = tid; i < len; i += 1024) {
    const double t = (nxt[i] - cur[i]) / scale;
    terms[i] = t * t;
  }
  __syncthreads();
  if (tid == 0) {
    // strictly left-to-right sum (the reference's order), with the LDS reads
    // of the next 16 terms issued before the adds of the current 16
    double acc = 0.0;
    const int nb = len / 16;
    double a[16], c[16];
    if (nb > 0) {
#pragma unroll
      for (int q = 0; q < 16; ++q) a[q] = terms[q];
    }
    for (int bi = 0; bi < nb; bi += 2) {
      const bool more = bi + 1 < nb;
      if (more) {
#pragma unroll
        for (int q = 0; q < 16; ++q) c[q] = terms[(bi + 1) * 16 + q];
      }
#pragma unroll
      for (int q = 0; q < 16; ++q) acc += a[q];
      if (!more) break;
      if (bi + 2 < nb) {
#pragma unroll
        for (int q = 0; q < 16; ++q) a[q] = terms[(bi + 2) * 16 + q];
      }
#pragma unroll
      for (int q = 0; q < 16; ++q) acc += c[q];
    }
    for (int i = nb * 16; i < len; ++i) acc += terms[i];
    const double diff = scale * sqrt(acc);
    const double pct = 100 * ((double)st->changed[par]) / (double)n;
    diffs[it] = diff;
    pcts[it] = pct;
    st->iters = it + 1;
    st->last_diff = diff;
    st->changed[par] = 0;  // the next iteration of this parity waits for this kernel
    if (diff < threshold) {
      st->done = 1;
      st->mode = MODE_FINAL;
      st->final_buf = nxt_buf;
    } else if (it == max_it - 1) {
      st->mode = (diff > threshold) ? MODE_LATEST : MODE_ZEROS;
    }
  }
}
__global__ __launch_bounds__(1024) void diff_control(const double* nxt, const double* cur, int k, long n,
                                                     double threshold, int it, int max_it, int nxt_buf,
                                                     State* st, double* diffs, double* pcts) {
  diff_control_b(nxt, cur, k, n, threshold, it, max_it, nxt_buf, st, diffs, pcts, (int)blockIdx.x, (int)gridDim.x);
}

// ------------------------------------------------------------------ map_to_grid

// Nearest grid point per centroid (lexicographic distance/index minimum over
// the grid tiles the bounds cannot exclude).
__device__ __forceinline__ void map_to_grid_b(const double* sx, const double* sy, const double* sz,
                                                   const int* perm, long n, const double* boxes, int ntiles,
                                                   const double* cen3, long cstride, int latest_buf,
                                                   const State* st, int* out, int bid, int nblk) {
  const int c = bid;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int mode = st ? st->mode : MODE_LATEST;
  if (mode == MODE_ZEROS || (st && st->error)) {
    if (tid == 0) out[c] = 0;
    return;
  }
  const int buf = (mode == MODE_FINAL) ? st->final_buf : latest_buf;
  const double* cen = cen3 + buf * cstride;
  const double cx = cen[c * 3], cy = cen[c * 3 + 1], cz = cen[c * 3 + 2];
  __shared__ double redd[4];
  __shared__ int redi[4];
  __shared__ int tl[MAXB];
  double u = DBL_MAX;
  for (int t = tid; t < ntiles; t += 256) u = fmin(u, ub2_box(cx, cy, cz, boxes + (long)t * 6));
  u = wave_min(u);
  if (lane == 0) redd[wave] = u;
  __syncthreads();
  u = fmin(fmin(redd[0], redd[1]), fmin(redd[2], redd[3]));
  const double thr = u * PRUNE;
  double bd2 = DBL_MAX, bs = DBL_MAX;
  int bi = INT_MAX;
  __shared__ int cntw;
  // candidate tiles, collected in windows of MAXB
  for (int t0 = 0; t0 < ntiles; t0 += MAXB) {
    const int tend = min(ntiles, t0 + MAXB);
    if (tid == 0) cntw = 0;
    __syncthreads();
    for (int t = t0 + tid; t < tend; t += 256)
      if (lb2_box(cx, cy, cz, boxes + (long)t * 6) <= thr) tl[atomicAdd(&cntw, 1)] = t;
    __syncthreads();
    const int nt = cntw;
    for (int q = 0; q < nt; ++q) {
      const long p = (long)tl[q] * TP + tid;
      if (p < n) consider_any(dist2(sx[p], sy[p], sz[p], cx, cy, cz), perm[p], bd2, bs, bi);
    }
    __syncthreads();
  }
  // lexicographic reduction of (bs, bi)
  for (int o = 32; o > 0; o >>= 1) {
    const double os = __shfl_xor(bs, o);
    const int oi = __shfl_xor(bi, o);
    if (os < bs || (os == bs && oi < bi)) { bs = os; bi = oi; }
  }
  if (lane == 0) { redd[wave] = bs; redi[wave] = bi; }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < 4; ++w)
      if (redd[w] < bs || (redd[w] == bs && redi[w] < bi)) { bs = redd[w]; bi = redi[w]; }
    // redd[0]/redi[0] already folded through bs/bi of wave 0
    out[c] = bi;
  }
}
__global__ __launch_bounds__(256) void map_to_grid(const double* sx, const double* sy, const double* sz,
                                                   const int* perm, long n, const double* boxes, int ntiles,
                                                   const double* cen3, long cstride, int latest_buf,
                                                   const State* st, int* out) {
  map_to_grid_b(sx, sy, sz, perm, n, boxes, ntiles, cen3, cstride, latest_buf, st, out, (int)blockIdx.x, (int)gridDim.x);
}

// Sort the k indices ascending (bitonic in LDS) and count adjacent repeats.
__device__ __forceinline__ void sort_indices_b(int* idx, int k, State* st, int bid, int nblk) {
  __shared__ int s[MAXB];
  int p2 = 1;
  while (p2 < k) p2 <<= 1;
  for (int i = threadIdx.x; i < p2; i += 1024) s[i] = i < k ? idx[i] : INT_MAX;
  __syncthreads();
  for (int size = 2; size <= p2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < p2; i += 1024) {
        const int j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          const int a = s[i], b = s[j];
          if ((a > b) == up) { s[i] = b; s[j] = a; }
        }
      }
      __syncthreads();
    }
  }
  int d = 0;
  for (int i = threadIdx.x; i < k; i += 1024) {
    idx[i] = s[i];
    if (i + 1 < k && s[i] == s[i + 1]) ++d;
  }
  d = wave_sum(d);
  if ((threadIdx.x & 63) == 0 && d && st) atomicAdd(&st->dups, d);
}
__global__ __launch_bounds__(1024) void sort_indices(int* idx, int k, State* st) {
  sort_indices_b(idx, k, st, (int)blockIdx.x, (int)gridDim.x);
}

}  // namespace km
}  // namespace thc

// ====================================================================== host

namespace thc {
namespace km {

__global__ void copy_final(const int* map0, const int* map1, long n, const double* cen3, int k, int latest_buf,
                           const State* st, int* map_out, double* cen_out) {
  const int iters = st->iters;
  const int mbuf = iters > 0 ? (iters - 1) & 1 : 0;
  const int cbuf = st->mode == MODE_FINAL ? st->final_buf : latest_buf;
  const int* m = mbuf ? map1 : map0;
  const double* c = cen3 + cbuf * 3L * k;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += 256L * gridDim.x) {
    if (map_out) map_out[i] = m[i];
    if (cen_out && i < 3L * k) cen_out[i] = c[i];
  }
  if (cen_out && n < 3L * k)
    for (long i = n + blockIdx.x * 256L + threadIdx.x; i < 3L * k; i += 256L * gridDim.x) cen_out[i] = c[i];
}

__global__ void check_keys(const int* keys, long n, int nbins, State* st) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += 256L * gridDim.x)
    if ((unsigned)keys[i] >= (unsigned)nbins) atomicOr(&st->error, 2);
}

struct Layout {
  long n;
  int k, max_it;
  int cells, nbins_sp, nbits_sp, nbits_k;
  int ntiles, ncs, nbb;
};

static int bits_for(int nb) {
  int b = 1;
  while ((1 << b) < nb) ++b;
  return b;
}

static Layout plan(long n, int k, int max_it) {
  Layout L;
  L.n = n; L.k = k; L.max_it = max_it;
  int cells = 1;
  while (cells < 16 && (long)(2 * cells) * (2 * cells) * (2 * cells) * 64 <= n) cells *= 2;
  L.cells = cells;
  L.nbins_sp = cells * cells * cells;
  L.nbits_sp = bits_for(L.nbins_sp);
  L.nbits_k = bits_for(k);
  L.ntiles = ceil_div(n, TP);
  L.ncs = ceil_div(n, CS_TILE);
  L.nbb = (int)std::min<long>(256, (n + 255) / 256);
  return L;
}

struct Bufs {
  double *part, *box, *sx, *sy, *sz, *boxes, *cen3, *diffs, *pcts;  // cen3: three centroid buffers
  int *keys, *perm, *map[2], *order, *tc, *total, *start, *idx0;
  State* st;
};

static size_t bytes_for(const Layout& L, bool full) {
  const long n = L.n;
  const int nb = std::max(L.nbins_sp, L.k);
  size_t b = 0;
  b += carve_bytes<double>(L.nbb * 6) + carve_bytes<double>(6);
  b += 3 * carve_bytes<double>(n) + carve_bytes<double>((size_t)L.ntiles * 6);
  b += 2 * carve_bytes<int>(n);                       // keys, perm
  b += carve_bytes<int>((size_t)L.ncs * nb) + 2 * carve_bytes<int>(nb);
  b += carve_bytes<State>(1);
  if (full) {
    b += carve_bytes<double>(9L * L.k) + 2 * carve_bytes<double>(std::max(1, L.max_it));
    b += 3 * carve_bytes<int>(n) + carve_bytes<int>(L.k);
  }
  return b;
}

static bool carve(Arena& a, const Layout& L, Bufs& b, bool full) {
  const long n = L.n;
  const int nb = std::max(L.nbins_sp, L.k);
  b = Bufs{};
  b.part = a.take<double>(L.nbb * 6);
  b.box = a.take<double>(6);
  b.sx = a.take<double>(n);
  b.sy = a.take<double>(n);
  b.sz = a.take<double>(n);
  b.boxes = a.take<double>((size_t)L.ntiles * 6);
  b.keys = a.take<int>(n);
  b.perm = a.take<int>(n);
  b.tc = a.take<int>((size_t)L.ncs * nb);
  b.total = a.take<int>(nb);
  b.start = a.take<int>(nb);
  b.st = a.take<State>(1);
  bool ok = b.part && b.box && b.sx && b.sy && b.sz && b.boxes && b.keys && b.perm && b.tc && b.total &&
            b.start && b.st;
  if (full) {
    b.cen3 = a.take<double>(9L * L.k);
    b.diffs = a.take<double>(std::max(1, L.max_it));
    b.pcts = a.take<double>(std::max(1, L.max_it));
    b.map[0] = a.take<int>(n);
    b.map[1] = a.take<int>(n);
    b.order = a.take<int>(n);
    b.idx0 = a.take<int>(L.k);
    ok = ok && b.cen3 && b.diffs && b.pcts && b.map[0] && b.map[1] && b.order && b.idx0;
  }
  return ok;
}

// Stable counting sort of keys[0..n) (values < nbins) into out (indices).
static void enqueue_sort(const Layout& L, const Bufs& b, const int* keys, int nbins, int nbits, int* out,
                         const State* st, hipStream_t s) {
  hipLaunchKernelGGL(cs_hist, dim3(L.ncs), dim3(256), 0, s, keys, L.n, nbins, b.tc, st);
  hipLaunchKernelGGL(cs_colscan, dim3(ceil_div(nbins, 64)), dim3(64 * CS_SCAN_W), 0, s, b.tc, L.ncs, nbins, b.total,
                     st);
  hipLaunchKernelGGL(cs_totscan, dim3(1), dim3(1024), 0, s, b.total, nbins, b.start, st);
  hipLaunchKernelGGL(cs_scatter, dim3(L.ncs), dim3(256), 0, s, keys, L.n, nbins, nbits, b.tc, b.start, out, st);
}

// Spatially sorted copy of the grid (Morton cells, stable) and tile boxes.
static void enqueue_setup(const Layout& L, const Bufs& b, const double* grid, hipStream_t s) {
  hipLaunchKernelGGL(bbox_partial, dim3(L.nbb), dim3(256), 0, s, grid, L.n, b.part);
  hipLaunchKernelGGL(bbox_final, dim3(1), dim3(64), 0, s, b.part, L.nbb, b.box);
  hipLaunchKernelGGL(morton_keys, dim3(ceil_div(L.n, 256)), dim3(256), 0, s, grid, L.n, b.box, L.cells, b.keys);
  enqueue_sort(L, b, b.keys, L.nbins_sp, L.nbits_sp, b.perm, nullptr, s);
  hipLaunchKernelGGL(gather_sorted, dim3(ceil_div(L.n, 256)), dim3(256), 0, s, grid, b.perm, L.n, b.sx, b.sy, b.sz);
  hipLaunchKernelGGL(tile_boxes, dim3(L.ntiles), dim3(256), 0, s, b.sx, b.sy, b.sz, L.n, b.boxes);
}

// The loop control of iteration it (diff_control: the sequential normed
// difference, the stop test) runs on the second stream beside iteration
// it + 1's assignment, sort and centroid sums, which do not need its result:
// if iteration it stops the loop, iteration it + 1 (already started) only
// writes buffers the final state does not use, and iteration it + 2 waits for
// it and returns at once.  Hence three centroid buffers (iteration it reads
// it % 3, writes (it+1) % 3, and diff_control(it) still reads both while
// iteration it + 1 writes (it+2) % 3), and change counts / empty-cluster flags
// kept per iteration parity (diff_control(it) reads and clears its parity's
// before iteration it + 2 runs).  With sa == s the same code runs in order on
// one stream.
static int enqueue_kmeans(const Layout& L, const Bufs& b, const double* grid, const double* rho, double threshold,
                          int* interp_idx, int* map_out, double* cen_out, hipStream_t s, hipStream_t sa,
                          const hipEvent_t kev[3]) {
  const long cs = 3L * L.k;
  const bool two = sa != s;
  hipLaunchKernelGGL(init_state, dim3(1), dim3(64), 0, s, b.st, (int)MODE_LATEST);
  enqueue_setup(L, b, grid, s);
  hipLaunchKernelGGL(init_centroids, dim3(ceil_div(L.k, 256)), dim3(256), 0, s, grid, b.idx0, L.k, b.cen3);
  const size_t dyn = (size_t)3 * L.k * sizeof(double);
  for (int it = 0; it < L.max_it; ++it) {
    const int cur = it % 3, nxt = (it + 1) % 3, par = it & 1;
    int* mnew = b.map[par];
    const int* mold = it == 0 ? nullptr : b.map[par ^ 1];
    if (two && it >= 2) THC_HIP(hipStreamWaitEvent(s, kev[1 + par], 0));  // diff_control(it - 2)
    hipLaunchKernelGGL(assign, dim3(L.ntiles), dim3(256), 0, s, b.sx, b.sy, b.sz, b.perm, L.n, b.boxes,
                       b.cen3 + cur * cs, L.k, mnew, mold, b.st, par);
    enqueue_sort(L, b, mnew, L.k, L.nbits_k, b.order, b.st, s);
    hipLaunchKernelGGL(centroid_sums, dim3(ceil_div(L.k, 4)), dim3(256), 0, s, b.order, b.start, b.total, rho,
                       grid, L.k, b.cen3 + nxt * cs, b.st, par);
    if (two) {
      THC_HIP(hipEventRecord(kev[0], s));
      THC_HIP(hipStreamWaitEvent(sa, kev[0], 0));
    }
    hipLaunchKernelGGL(diff_control, dim3(1), dim3(1024), dyn, sa, b.cen3 + nxt * cs, b.cen3 + cur * cs, L.k, L.n,
                       threshold, it, L.max_it, nxt, b.st, b.diffs, b.pcts);
    if (two) THC_HIP(hipEventRecord(kev[1 + par], sa));
  }
  if (two) THC_HIP(hipStreamWaitEvent(s, kev[1 + ((L.max_it - 1) & 1)], 0));  // last loop control
  const int latest = L.max_it % 3;
  hipLaunchKernelGGL(map_to_grid, dim3(L.k), dim3(256), 0, s, b.sx, b.sy, b.sz, b.perm, L.n, b.boxes, L.ntiles,
                     b.cen3, cs, latest, b.st, interp_idx);
  hipLaunchKernelGGL(sort_indices, dim3(1), dim3(1024), 0, s, interp_idx, L.k, b.st);
  if (map_out || cen_out)
    hipLaunchKernelGGL(copy_final, dim3(256), dim3(256), 0, s, b.map[0], b.map[1], L.n, b.cen3, L.k, latest, b.st,
                       map_out, cen_out);
  THC_HIP(hipGetLastError());
  return THC_OK;
}

// ------------------------------------------------------------------ batches
//
// A batch runs the k-means of many systems together: every launch below
// covers one step (setup, assignment, sort, centroid sums, loop control, map
// to grid) of all the systems, so a 20-iteration batch of 32 systems is about
// 160 launches instead of 32 times that.  Each workgroup finds its system in
// a prefix table of per-system block counts and runs the same device code as
// the single-system kernels with that system's sizes and buffers, so every
// system's result is bit-identical to its own thc_kmeans.

struct Job {
  Layout L;
  Bufs b;
  const double* grid;
  const double* rho;
  double threshold;
  int* interp_idx;
};

// block-count tables, one row of nj + 1 prefix sums per kind
enum { JK_NBB, JK_N256, JK_NCS, JK_COLSP, JK_COLK, JK_TILES, JK_K256, JK_K4, JK_K, JK_COUNT };

inline int job_blocks(const Layout& L, int kind) {
  switch (kind) {
    case JK_NBB: return L.nbb;
    case JK_N256: return (int)ceil_div(L.n, 256);
    case JK_NCS: return L.ncs;
    case JK_COLSP: return ceil_div(L.nbins_sp, 64);
    case JK_COLK: return ceil_div(L.k, 64);
    case JK_TILES: return L.ntiles;
    case JK_K256: return ceil_div(L.k, 256);
    case JK_K4: return ceil_div(L.k, 4);
    default: return L.k;
  }
}

// system of workgroup b: off[j] <= b < off[j + 1]
__device__ inline int job_of(const int* off, int nj, int b) {
  int lo = 0, hi = nj;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

#define KM_JOB(OFF)                                 \
  const int j_ = job_of(OFF, nj, (int)blockIdx.x);  \
  const Job& q = jobs[j_];                          \
  const int bid = (int)blockIdx.x - OFF[j_], nblk = OFF[j_ + 1] - OFF[j_];

__global__ void init_state_m(const Job* jobs) { init_state_b(jobs[blockIdx.x].b.st, MODE_LATEST, 0, 1); }

__global__ __launch_bounds__(256) void bbox_partial_m(const Job* jobs, const int* off, int nj) {
  KM_JOB(off)
  bbox_partial_b(q.grid, q.L.n, q.b.part, bid, nblk);
}

__global__ void bbox_final_m(const Job* jobs) {
  const Job& q = jobs[blockIdx.x];
  bbox_final_b(q.b.part, q.L.nbb, q.b.box, 0, 1);
}

__global__ __launch_bounds__(256) void morton_keys_m(const Job* jobs, const int* off, int nj) {
  KM_JOB(off)
  morton_keys_b(q.grid, q.L.n, q.b.box, q.L.cells, q.b.keys, bid, nblk);
}

__global__ __launch_bounds__(256) void gather_sorted_m(const Job* jobs, const int* off, int nj) {
  KM_JOB(off)
  gather_sorted_b(q.grid, q.b.perm, q.L.n, q.b.sx, q.b.sy, q.b.sz, bid, nblk);
}

__global__ __launch_bounds__(256) void tile_boxes_m(const Job* jobs, const int* off, int nj) {
  KM_JOB(off)
  tile_boxes_b(q.b.sx, q.b.sy, q.b.sz, q.L.n, q.b.boxes, bid, nblk);
}

__global__ __launch_bounds__(256) void init_centroids_m(const Job* jobs, const int* off, int nj) {
  KM_JOB(off)
  init_centroids_b(q.grid, q.b.idx0, q.L.k, q.b.cen3, bid, nblk);
}

// Counting sort of one batch: mode 0 sorts the Morton keys into perm (setup),
// mode 1 the iteration's cluster map into order.  Systems past their last
// iteration skip mode 1.
struct SortArgs {
  const int* keys;
  int nbins, nbits;
  int* out;
  const State* st;
};
__device__ inline bool sort_args(const Job& q, int mode, int it, SortArgs& a) {
  if (mode == 0) {
    a = SortArgs{q.b.keys, q.L.nbins_sp, q.L.nbits_sp, q.b.perm, nullptr};
    return true;
  }
  a = SortArgs{q.b.map[it & 1], q.L.k, q.L.nbits_k, q.b.order, q.b.st};
  return it < q.L.max_it;
}

__global__ __launch_bounds__(256) void cs_hist_m(const Job* jobs, const int* off, int nj, int mode, int it) {
  KM_JOB(off)
  SortArgs a;
  if (!sort_args(q, mode, it, a)) return;
  cs_hist_b(a.keys, q.L.n, a.nbins, q.b.tc, a.st, bid, nblk);
}

__global__ __launch_bounds__(64 * CS_SCAN_W) void cs_colscan_m(const Job* jobs, const int* off, int nj, int mode,
                                                                 int it) {
  KM_JOB(off)
  SortArgs a;
  if (!sort_args(q, mode, it, a)) return;
  cs_colscan_b(q.b.tc, q.L.ncs, a.nbins, q.b.total, a.st, bid, nblk);
}

__global__ __launch_bounds__(1024) void cs_totscan_m(const Job* jobs, int mode, int it) {
  const Job& q = jobs[blockIdx.x];
  SortArgs a;
  if (!sort_args(q, mode, it, a)) return;
  cs_totscan_b(q.b.total, a.nbins, q.b.start, a.st, 0, 1);
}

__global__ __launch_bounds__(256) void cs_scatter_m(const Job* jobs, const int* off, int nj, int mode, int it) {
  KM_JOB(off)
  SortArgs a;
  if (!sort_args(q, mode, it, a)) return;
  cs_scatter_b(a.keys, q.L.n, a.nbins, a.nbits, q.b.tc, q.b.start, a.out, a.st, bid, nblk);
}

__global__ __launch_bounds__(256) void assign_m(const Job* jobs, const int* off, int nj, int it) {
  KM_JOB(off)
  if (it >= q.L.max_it) return;
  const int par = it & 1;
  assign_b(q.b.sx, q.b.sy, q.b.sz, q.b.perm, q.L.n, q.b.boxes, q.b.cen3 + (it % 3) * 3L * q.L.k, q.L.k,
           q.b.map[par], it == 0 ? nullptr : q.b.map[par ^ 1], q.b.st, par, bid, nblk);
}

__global__ __launch_bounds__(256) void centroid_sums_m(const Job* jobs, const int* off, int nj, int it) {
  KM_JOB(off)
  if (it >= q.L.max_it) return;
  centroid_sums_b(q.b.order, q.b.start, q.b.total, q.rho, q.grid, q.L.k, q.b.cen3 + ((it + 1) % 3) * 3L * q.L.k,
                  q.b.st, it & 1, bid, nblk);
}

__global__ __launch_bounds__(1024) void diff_control_m(const Job* jobs, int it) {
  const Job& q = jobs[blockIdx.x];
  if (it >= q.L.max_it) return;
  const long cs = 3L * q.L.k;
  diff_control_b(q.b.cen3 + ((it + 1) % 3) * cs, q.b.cen3 + (it % 3) * cs, q.L.k, q.L.n, q.threshold, it,
                 q.L.max_it, (it + 1) % 3, q.b.st, q.b.diffs, q.b.pcts, 0, 1);
}

__global__ __launch_bounds__(256) void map_to_grid_m(const Job* jobs, const int* off, int nj) {
  KM_JOB(off)
  map_to_grid_b(q.b.sx, q.b.sy, q.b.sz, q.b.perm, q.L.n, q.b.boxes, q.L.ntiles, q.b.cen3, 3L * q.L.k,
                q.L.max_it % 3, q.b.st, q.interp_idx, bid, nblk);
}

__global__ __launch_bounds__(1024) void sort_indices_m(const Job* jobs) {
  const Job& q = jobs[blockIdx.x];
  sort_indices_b(q.interp_idx, q.L.k, q.b.st, 0, 1);
}

// every system's State, diffs and pcts into one block for a single copy back:
// system j at out + j * stride bytes
constexpr long STATE_SLOT = (sizeof(State) + 63) / 64 * 64;  // State, then diffs and pcts

__global__ void collect_m(const Job* jobs, char* out, long stride, int max_it) {
  const Job& q = jobs[blockIdx.x];
  char* o = out + blockIdx.x * stride;
  const int t = threadIdx.x;
  if (t == 0) *reinterpret_cast<State*>(o) = *q.b.st;
  double* d = reinterpret_cast<double*>(o + STATE_SLOT);
  for (int i = t; i < max_it; i += blockDim.x) {
    d[i] = i < q.L.max_it ? q.b.diffs[i] : 0.0;
    d[max_it + i] = i < q.L.max_it ? q.b.pcts[i] : 0.0;
  }
}

#undef KM_JOB

}  // namespace km

static std::atomic<int> km_attrs_done{0};  // batched calls run two k-means at once
static int km_set_attrs() {
  if (km_attrs_done) return THC_OK;
  THC_HIP(hipFuncSetAttribute((const void*)km::diff_control, hipFuncAttributeMaxDynamicSharedMemorySize,
                              3 * km::MAXB * (int)sizeof(double)));
  THC_HIP(hipFuncSetAttribute((const void*)km::diff_control_m, hipFuncAttributeMaxDynamicSharedMemorySize,
                              3 * km::MAXB * (int)sizeof(double)));
  km_attrs_done = 1;
  return THC_OK;
}

int kmeans_run(thc_ctx* ctx, const double* grid, const double* rho, long n, int k, int max_it, double threshold,
               const int* host_init_idx, int* interp_idx, int* grid_map_out, double* centroids_out,
               thc_kmeans_stats* stats) {
  using namespace km;
  if (!ctx || !grid || !rho || !interp_idx || !host_init_idx) return THC_ERR_ARG;
  if (n < 1 || k < 1 || k > n || max_it < 1 || n > 0x7fffffffL) return THC_ERR_ARG;
  if (k > MAXB) return THC_ERR_UNSUPPORTED;
  for (int j = 0; j < k; ++j)
    if (host_init_idx[j] < 0 || host_init_idx[j] >= n) {
      set_last_error("a starting index lies outside [0, n)", __FILE__, __LINE__);
      return THC_ERR_ARG;
    }
  THC_HIP(hipSetDevice(ctx->device));
  THC_TRY(km_set_attrs());
  const Layout L = plan(n, k, max_it);
  const size_t need = bytes_for(L, true);
  if (need > ctx->kmws.cap) {
    if (ctx->kmgraph) ctx->kmgraph->reset();
    THC_TRY(ctx->kmws.reserve(need));
  }
  ctx->kmws.reset();
  Bufs b;
  if (!carve(ctx->kmws, L, b, true)) return THC_ERR_NO_MEMORY;
  hipStream_t s = ctx->stream;
  THC_TRY(ensure_pinned(ctx, sizeof(int) * k + sizeof(State) + 2 * sizeof(double) * max_it + 64));
  int* hidx = reinterpret_cast<int*>(ctx->pinned);
  std::copy(host_init_idx, host_init_idx + k, hidx);
  THC_HIP(hipEventRecord(ctx->ev[0], s));
  THC_HIP(hipMemcpyAsync(b.idx0, hidx, sizeof(int) * k, hipMemcpyHostToDevice, s));
  if (ctx->use_graphs) {
    if (!ctx->kmgraph) ctx->kmgraph = new KmGraphCache();
    KmGraphCache& gc = *ctx->kmgraph;
    const void* key[5] = {grid, rho, interp_idx, grid_map_out, centroids_out};
    bool hit = gc.exec && gc.n == n && gc.k == k && gc.max_it == max_it && gc.threshold == threshold &&
               gc.base == ctx->kmws.base && gc.overlap == ctx->km_overlap;
    for (int q = 0; q < 5 && hit; ++q) hit = gc.key_ptr[q] == key[q];
    if (!hit) {
      gc.reset();
      THC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      const int est = enqueue_kmeans(L, b, grid, rho, threshold, interp_idx, grid_map_out, centroids_out, s,
                                     ctx->km_overlap ? ctx->aux : s, ctx->kev);
      hipError_t e = hipStreamEndCapture(s, &gc.graph);
      THC_TRY(est);
      THC_HIP(e);
      THC_HIP(hipGraphInstantiate(&gc.exec, gc.graph, nullptr, nullptr, 0));
      for (int q = 0; q < 5; ++q) gc.key_ptr[q] = key[q];
      gc.n = n; gc.k = k; gc.max_it = max_it; gc.threshold = threshold; gc.base = ctx->kmws.base;
      gc.overlap = ctx->km_overlap;
    }
    THC_HIP(hipGraphLaunch(gc.exec, s));
  } else {
    THC_TRY(enqueue_kmeans(L, b, grid, rho, threshold, interp_idx, grid_map_out, centroids_out, s,
                           ctx->km_overlap ? ctx->aux : s, ctx->kev));
  }
  char* hp = reinterpret_cast<char*>(ctx->pinned);
  State* hst = reinterpret_cast<State*>(hp + align_up(sizeof(int) * k, 64));
  double* hd = reinterpret_cast<double*>(reinterpret_cast<char*>(hst) + align_up(sizeof(State), 64));
  double* hpct = hd + max_it;
  THC_HIP(hipMemcpyAsync(hst, b.st, sizeof(State), hipMemcpyDeviceToHost, s));
  THC_HIP(hipMemcpyAsync(hd, b.diffs, sizeof(double) * max_it, hipMemcpyDeviceToHost, s));
  THC_HIP(hipMemcpyAsync(hpct, b.pcts, sizeof(double) * max_it, hipMemcpyDeviceToHost, s));
  THC_HIP(hipEventRecord(ctx->ev[1], s));
  THC_HIP(hipStreamSynchronize(s));
  if (stats) {
    stats->iterations = hst->iters;
    stats->converged = hst->mode == MODE_FINAL;
    stats->duplicates = hst->dups;
    stats->interp_zeroed = hst->mode == MODE_ZEROS;
    stats->last_diff = hst->last_diff;
    if (stats->diffs) std::copy(hd, hd + hst->iters, stats->diffs);
    if (stats->pcts) std::copy(hpct, hpct + hst->iters, stats->pcts);
    float ms = 0;
    THC_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    stats->elapsed_ms = ms;
  }
  if (hst->error) return THC_ERR_EMPTY_CLUSTER;
  if (hst->dups) return THC_WARN_DUPLICATE_INDICES;
  return THC_OK;
}


extern "C" int thc_seed_indices_impl(long n, int k, unsigned int seed, int* out);

int kmeans_run_many(thc_ctx* ctx, int nj, const KmSpec* spec, int* status, thc_kmeans_stats* stats) {
  using namespace km;
  if (!ctx || nj < 0 || (nj > 0 && (!spec || !status))) return THC_ERR_ARG;
  if (nj == 0) return THC_OK;
  THC_HIP(hipSetDevice(ctx->device));
  THC_TRY(km_set_attrs());
  // host-side checks and layouts; a system that fails them gets its status
  // and drops out of the batch
  std::vector<int> live;
  std::vector<Layout> lay(nj);
  std::vector<std::vector<int>> init(nj);
  size_t need = 0;
  int max_it = 1, maxk = 1;
  long idx_total = 0;
  for (int j = 0; j < nj; ++j) {
    const KmSpec& q = spec[j];
    status[j] = THC_OK;
    if (!q.grid || !q.rho || !q.interp_idx || q.n < 1 || q.k < 1 || q.k > q.n || q.max_it < 1 ||
        q.n > 0x7fffffffL) {
      status[j] = THC_ERR_ARG;
      continue;
    }
    if (q.k > MAXB) {
      status[j] = THC_ERR_UNSUPPORTED;
      continue;
    }
    init[j].resize(q.k);
    if (thc_seed_indices_impl(q.n, q.k, q.seed, init[j].data())) {
      status[j] = THC_ERR_ARG;
      continue;
    }
    lay[j] = plan(q.n, q.k, q.max_it);
    need += bytes_for(lay[j], true);
    max_it = std::max(max_it, q.max_it);
    maxk = std::max(maxk, q.k);
    idx_total += q.k;
    live.push_back(j);
  }
  const int nl = (int)live.size();
  if (nl == 0) return THC_OK;
  // device blob: the job table, the block-count prefix tables, the starting
  // indices; the results block comes back through collect_m
  const size_t job_bytes = align_up(sizeof(Job) * nl, 256);
  const size_t off_bytes = align_up(sizeof(int) * (size_t)JK_COUNT * (nl + 1), 256);
  const size_t idx_bytes = align_up(sizeof(int) * (size_t)idx_total, 256);
  const size_t blob = job_bytes + off_bytes + idx_bytes;
  const long res_stride = (long)align_up(STATE_SLOT + 2 * sizeof(double) * max_it, 64);
  const size_t res_bytes = (size_t)res_stride * nl;
  need += blob + align_up(res_bytes, 256) + 1024;
  if (need > ctx->kmws.cap) {
    if (ctx->kmgraph) ctx->kmgraph->reset();
    THC_TRY(ctx->kmws.reserve(need));
  }
  ctx->kmws.reset();
  char* dblob = ctx->kmws.take<char>(blob);
  char* dres = ctx->kmws.take<char>(res_bytes);
  if (!dblob || !dres) return THC_ERR_NO_MEMORY;
  THC_TRY(ensure_pinned(ctx, std::max(blob, res_bytes)));
  char* hb = reinterpret_cast<char*>(ctx->pinned);
  Job* hj = reinterpret_cast<Job*>(hb);
  int* hoff = reinterpret_cast<int*>(hb + job_bytes);
  int* hidx = reinterpret_cast<int*>(hb + job_bytes + off_bytes);
  int* didx = reinterpret_cast<int*>(dblob + job_bytes + off_bytes);
  long ipos = 0;
  for (int t = 0; t < nl; ++t) {
    const int j = live[t];
    Job& J = hj[t];
    J.L = lay[j];
    if (!carve(ctx->kmws, J.L, J.b, true)) return THC_ERR_NO_MEMORY;
    J.b.idx0 = didx + ipos;  // read by init_centroids only
    std::copy(init[j].begin(), init[j].end(), hidx + ipos);
    ipos += spec[j].k;
    J.grid = spec[j].grid;
    J.rho = spec[j].rho;
    J.threshold = spec[j].threshold;
    J.interp_idx = spec[j].interp_idx;
  }
  for (int kind = 0; kind < JK_COUNT; ++kind) {
    int* o = hoff + kind * (nl + 1);
    o[0] = 0;
    for (int t = 0; t < nl; ++t) o[t + 1] = o[t] + job_blocks(hj[t].L, kind);
  }
  long total_blocks[JK_COUNT];
  for (int kind = 0; kind < JK_COUNT; ++kind) total_blocks[kind] = hoff[kind * (nl + 1) + nl];
  for (int kind = 0; kind < JK_COUNT; ++kind)
    if (total_blocks[kind] > 0x7fffffffL) return THC_ERR_ARG;
  hipStream_t s = ctx->stream;
  const Job* dj = reinterpret_cast<const Job*>(dblob);
  const int* doff = reinterpret_cast<const int*>(dblob + job_bytes);
  auto off = [&](int kind) { return doff + kind * (nl + 1); };
  auto grid_of = [&](int kind) { return dim3((unsigned)total_blocks[kind]); };
  THC_HIP(hipEventRecord(ctx->ev[0], s));
  THC_HIP(hipMemcpyAsync(dblob, hb, blob, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(init_state_m, dim3(nl), dim3(64), 0, s, dj);
  hipLaunchKernelGGL(bbox_partial_m, grid_of(JK_NBB), dim3(256), 0, s, dj, off(JK_NBB), nl);
  hipLaunchKernelGGL(bbox_final_m, dim3(nl), dim3(64), 0, s, dj);
  hipLaunchKernelGGL(morton_keys_m, grid_of(JK_N256), dim3(256), 0, s, dj, off(JK_N256), nl);
  auto sort = [&](int mode, int it) {
    const int col = mode == 0 ? JK_COLSP : JK_COLK;
    hipLaunchKernelGGL(cs_hist_m, grid_of(JK_NCS), dim3(256), 0, s, dj, off(JK_NCS), nl, mode, it);
    hipLaunchKernelGGL(cs_colscan_m, grid_of(col), dim3(64 * CS_SCAN_W), 0, s, dj, off(col), nl, mode, it);
    hipLaunchKernelGGL(cs_totscan_m, dim3(nl), dim3(1024), 0, s, dj, mode, it);
    hipLaunchKernelGGL(cs_scatter_m, grid_of(JK_NCS), dim3(256), 0, s, dj, off(JK_NCS), nl, mode, it);
  };
  sort(0, 0);
  hipLaunchKernelGGL(gather_sorted_m, grid_of(JK_N256), dim3(256), 0, s, dj, off(JK_N256), nl);
  hipLaunchKernelGGL(tile_boxes_m, grid_of(JK_TILES), dim3(256), 0, s, dj, off(JK_TILES), nl);
  hipLaunchKernelGGL(init_centroids_m, grid_of(JK_K256), dim3(256), 0, s, dj, off(JK_K256), nl);
  const size_t dyn = (size_t)3 * maxk * sizeof(double);
  for (int it = 0; it < max_it; ++it) {
    hipLaunchKernelGGL(assign_m, grid_of(JK_TILES), dim3(256), 0, s, dj, off(JK_TILES), nl, it);
    sort(1, it);
    hipLaunchKernelGGL(centroid_sums_m, grid_of(JK_K4), dim3(256), 0, s, dj, off(JK_K4), nl, it);
    hipLaunchKernelGGL(diff_control_m, dim3(nl), dim3(1024), dyn, s, dj, it);
  }
  hipLaunchKernelGGL(map_to_grid_m, grid_of(JK_K), dim3(256), 0, s, dj, off(JK_K), nl);
  hipLaunchKernelGGL(sort_indices_m, dim3(nl), dim3(1024), 0, s, dj);
  hipLaunchKernelGGL(collect_m, dim3(nl), dim3(256), 0, s, dj, dres, res_stride, max_it);
  THC_HIP(hipGetLastError());
  // the pinned buffer still holds the blob being copied: read results back
  // only after that copy (stream order) into the same buffer
  THC_HIP(hipMemcpyAsync(hb, dres, res_bytes, hipMemcpyDeviceToHost, s));
  THC_HIP(hipEventRecord(ctx->ev[1], s));
  THC_HIP(hipStreamSynchronize(s));
  float ms = 0;
  THC_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  for (int t = 0; t < nl; ++t) {
    const int j = live[t];
    const char* r = hb + t * res_stride;
    const State* hst = reinterpret_cast<const State*>(r);
    const double* hd = reinterpret_cast<const double*>(r + STATE_SLOT);
    const double* hp = hd + max_it;
    if (stats) {
      thc_kmeans_stats& o = stats[j];
      o.iterations = hst->iters;
      o.converged = hst->mode == MODE_FINAL;
      o.duplicates = hst->dups;
      o.interp_zeroed = hst->mode == MODE_ZEROS;
      o.last_diff = hst->last_diff;
      if (o.diffs) std::copy(hd, hd + hst->iters, o.diffs);
      if (o.pcts) std::copy(hp, hp + hst->iters, o.pcts);
      o.elapsed_ms = ms;
    }
    status[j] = hst->error ? THC_ERR_EMPTY_CLUSTER : (hst->dups ? THC_WARN_DUPLICATE_INDICES : THC_OK);
  }
  return THC_OK;
}

int kmeans_classify(thc_ctx* ctx, const double* grid, long n, const double* centroids, int k, int* grid_map) {
  using namespace km;
  if (!ctx || !grid || !centroids || !grid_map || n < 1 || k < 1 || n > 0x7fffffffL) return THC_ERR_ARG;
  if (k > MAXB) return THC_ERR_UNSUPPORTED;
  THC_HIP(hipSetDevice(ctx->device));
  const Layout L = plan(n, k, 1);
  THC_TRY(ctx->ws.reserve(bytes_for(L, false)));
  ctx->ws.reset();
  Bufs b;
  if (!carve(ctx->ws, L, b, false)) return THC_ERR_NO_MEMORY;
  hipStream_t s = ctx->stream;
  enqueue_setup(L, b, grid, s);
  hipLaunchKernelGGL(assign, dim3(L.ntiles), dim3(256), 0, s, b.sx, b.sy, b.sz, b.perm, L.n, b.boxes, centroids, k,
                     grid_map, (const int*)nullptr, (State*)nullptr, 0);
  THC_HIP(hipGetLastError());
  THC_HIP(hipStreamSynchronize(s));
  return THC_OK;
}

int kmeans_update(thc_ctx* ctx, const double* rho, const double* grid, long n, const int* grid_map, int k,
                  double* centroids) {
  using namespace km;
  if (!ctx || !rho || !grid || !grid_map || !centroids || n < 1 || k < 1 || n > 0x7fffffffL) return THC_ERR_ARG;
  if (k > MAXB) return THC_ERR_UNSUPPORTED;
  THC_HIP(hipSetDevice(ctx->device));
  const Layout L = plan(n, k, 1);
  THC_TRY(ctx->ws.reserve(bytes_for(L, false) + carve_bytes<int>(n)));
  ctx->ws.reset();
  Bufs b;
  if (!carve(ctx->ws, L, b, false)) return THC_ERR_NO_MEMORY;
  int* order = ctx->ws.take<int>(n);
  if (!order) return THC_ERR_NO_MEMORY;
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(init_state, dim3(1), dim3(64), 0, s, b.st, (int)MODE_LATEST);
  hipLaunchKernelGGL(check_keys, dim3(256), dim3(256), 0, s, grid_map, n, k, b.st);
  enqueue_sort(L, b, grid_map, k, L.nbits_k, order, nullptr, s);
  hipLaunchKernelGGL(centroid_sums, dim3(ceil_div(k, 4)), dim3(256), 0, s, order, b.start, b.total, rho, grid, k,
                     centroids, b.st, 0);
  THC_HIP(hipGetLastError());
  State hst;
  THC_HIP(hipMemcpyAsync(&hst, b.st, sizeof(State), hipMemcpyDeviceToHost, s));
  THC_HIP(hipStreamSynchronize(s));
  if (hst.error & 2) {
    set_last_error("a grid_map entry lies outside [0, k)", __FILE__, __LINE__);
    return THC_ERR_ARG;
  }
  if (hst.serr[0]) return THC_ERR_EMPTY_CLUSTER;
  return THC_OK;
}

int kmeans_map_to_grid(thc_ctx* ctx, const double* grid, long n, const double* centroids, int k, int* interp_idx) {
  using namespace km;
  if (!ctx || !grid || !centroids || !interp_idx || n < 1 || k < 1 || n > 0x7fffffffL) return THC_ERR_ARG;
  if (k > MAXB) return THC_ERR_UNSUPPORTED;
  THC_HIP(hipSetDevice(ctx->device));
  const Layout L = plan(n, k, 1);
  THC_TRY(ctx->ws.reserve(bytes_for(L, false)));
  ctx->ws.reset();
  Bufs b;
  if (!carve(ctx->ws, L, b, false)) return THC_ERR_NO_MEMORY;
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(init_state, dim3(1), dim3(64), 0, s, b.st, (int)MODE_LATEST);
  enqueue_setup(L, b, grid, s);
  hipLaunchKernelGGL(map_to_grid, dim3(k), dim3(256), 0, s, b.sx, b.sy, b.sz, b.perm, L.n, b.boxes, L.ntiles,
                     centroids, 0L, 0, (const State*)nullptr, interp_idx);
  hipLaunchKernelGGL(sort_indices, dim3(1), dim3(1024), 0, s, interp_idx, k, b.st);
  THC_HIP(hipGetLastError());
  State hst;
  THC_HIP(hipMemcpyAsync(&hst, b.st, sizeof(State), hipMemcpyDeviceToHost, s));
  THC_HIP(hipStreamSynchronize(s));
  return hst.dups ? THC_WARN_DUPLICATE_INDICES : THC_OK;
}

namespace km {
// y = sqrt(x) with the square root the distance kernels above use
__global__ __launch_bounds__(256) void device_sqrt(const double* x, long n, double* y) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += 256L * gridDim.x) y[i] = sqrt(x[i]);
}
}  // namespace km

}  // namespace thc

extern "C" int thc_device_sqrt(thc_ctx* ctx, const double* x, long n, double* y) {
  using namespace thc;
  if (!ctx || n < 0 || (n > 0 && (!x || !y))) return THC_ERR_ARG;
  if (n == 0) return THC_OK;
  THC_HIP(hipSetDevice(ctx->device));
  const long blocks = std::min<long>(8192, (n + 255) / 256);
  hipLaunchKernelGGL(km::device_sqrt, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, x, n, y);
  THC_HIP(hipGetLastError());
  THC_HIP(hipStreamSynchronize(ctx->stream));
  return THC_OK;
}
