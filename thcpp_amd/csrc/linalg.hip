// FP64 dense factorisations on the GPU for the K x K matrices of the path:
// Cholesky of CC^T and of V (pzpotrf 'L' in the reference,
// src/matrix_operations.h:259-267) and the SPD inverse used to apply
// CC^T^-1 to ZC^T as one MFMA GEMM.
//
// Cholesky: right-looking, 64-column panels, one fused kernel per panel
// (panel solve by the inverse of the diagonal block, trailing update, and the
// next diagonal block's factor and inverse); layout below.
// Inverse: the lower factor is padded with an identity to 64 * 2^p, its
// 64 x 64 diagonal blocks are inverted in parallel, then recursive doubling
// forms Linv21 = -Linv22 (L21 Linv11) for every pair of blocks of one size in
// two batched GEMMs per level; finally A^-1 = Linv^T Linv.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <utility>

#include "common.h"

namespace thc {
namespace {

constexpr int NB = 64;

__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Lower NB x NB block at (i0, j0) of a, jb x jb valid, as rows of an LDS
// array; outside jb it is the identity.  Load r reads one matrix row (the 64
// lanes read 64 consecutive doubles).  All 64 loads are issued before the
// first LDS store (masked lanes read a[0], which exists), so the block costs
// one memory round trip.  TRANS stores the block transposed.
template <bool TRANS>
__device__ inline void load_lower_block(const double* a, long lda, int i0, int j0, int jb, double (*Ls)[NB + 1]) {
  const int t = threadIdx.x;
  double v[NB];
#pragma unroll
  for (int rr = 0; rr < NB; ++rr) {
    const bool ok = rr < jb && t <= rr;
    v[rr] = a[ok ? (long)(i0 + rr) * lda + j0 + t : 0];
  }
#pragma unroll
  for (int rr = 0; rr < NB; ++rr) {
    const bool ok = rr < jb && t <= rr;
    const double x = ok ? v[rr] : ((rr >= jb && rr == t) ? 1.0 : 0.0);
    if (TRANS) Ls[t][rr] = x;
    else Ls[rr][t] = x;
  }
}

// Forward substitution for one row held in registers: x L^T = b, with L given
// transposed (Lt[c][c2] = L[c2][c], a broadcast read for the whole wave).
// Right-looking: once x[c] is final every later entry drops its x[c] term, so
// the updates of one step are independent of each other.
__device__ inline void solve_row(double (&x)[NB], const double (*Lt)[NB + 1]) {
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    x[c] = x[c] / Lt[c][c];
#pragma unroll
    for (int c2 = c + 1; c2 < NB; ++c2) x[c2] = fma(-x[c], Lt[c][c2], x[c2]);
  }
}

// ---- fused right-looking Cholesky ------------------------------------------
// The matrix is taken in 64 x 64 blocks, extended by the identity beyond k.
// Step j (one kernel) owns panel j: its diagonal block was factored by step
// j - 1, which also left W_j = L_jj^-1.  Every workgroup takes one lower tile
// (I, J) of the trailing matrix (I >= J > j) and
//   1. forms L_I = A_Ij W_j^T and L_J = A_Jj W_j^T (MFMA, from the unsolved
//      panel rows, so the panel itself is never rewritten during the step);
//   2. applies A_IJ -= L_I L_J^T (MFMA);
//   3. for I = J, parks L_I transposed in the unused upper block (j, I);
//   4. tile (j+1, j+1) -- workgroup 0 -- then factors its block: wave 0 runs
//      the pivot steps with lane r owning row r in registers, and wave 1,
//      in lockstep one barrier behind each pivot, forms W_{j+1} column by
//      column (lane c solves L x = e_c, right-looking).
// A last kernel moves the parked panels into the lower triangle and zeroes
// the upper one.  Step -1 only factors block 0.
constexpr int CB = NB;      // block size
constexpr int CT = 4 * 64;  // threads of a step workgroup

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ inline d4 mfma_f64(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

__device__ inline double readlane_d(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// 1/sqrt(d) from the hardware reciprocal square root (about 23 bits) and two
// Newton steps y <- y (3/2 - d/2 y^2) (quadratic: 2^-23 -> 2^-45 -> rounding
// level), sqrt(d) = d / sqrt(d); seven dependent operations after the
// estimate.  Non-positive or NaN d give NaN, as sqrt would.
__device__ inline void sqrt_rsqrt(double d, double& s, double& rs) {
  const double h = 0.5 * d;
  double y = __builtin_amdgcn_rsq(d);
  y = y * fma(-h, y * y, 1.5);
  y = y * fma(-h, y * y, 1.5);
  rs = y;
  s = d * y;
}

// a[r][c] of the matrix extended by the identity beyond k
__device__ inline double ext_at(const double* a, long lda, int k, long r, long c) {
  const bool in = r < k && c < k;
  const double v = a[in ? r * lda + c : 0];
  return in ? v : (r == c ? 1.0 : 0.0);
}

// Block at (r0, c0) into X; each wave reads whole 64-element rows, all loads
// issued before the first LDS store.
__device__ inline void load_tile(const double* a, long lda, int k, long r0, long c0, double (*X)[CB + 1]) {
  const int c = threadIdx.x & 63, rb = threadIdx.x >> 6;
  double v[CB / 4];
#pragma unroll
  for (int i = 0; i < CB / 4; ++i) v[i] = ext_at(a, lda, k, r0 + rb + 4 * i, c0 + c);
#pragma unroll
  for (int i = 0; i < CB / 4; ++i) X[rb + 4 * i][c] = v[i];
}

// Rows 16 w .. 16 w + 15 of X become X W^T (wave w; only its own rows are
// read and then rewritten).  f64 MFMA D layout: lane l holds column l % 16 of
// rows l / 16 + 4 v.
__device__ inline void panel_rows(double (*X)[CB + 1], const double (*W)[CB + 1]) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  d4 acc[4] = {};
#pragma unroll
  for (int q0 = 0; q0 < CB; q0 += 4) {
    const double av = X[16 * w + (l & 15)][q0 + (l >> 4)];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[cb] = mfma_f64(av, W[16 * cb + (l & 15)][q0 + (l >> 4)], acc[cb]);
  }
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int v = 0; v < 4; ++v) X[16 * w + (l >> 4) + 4 * v][16 * cb + (l & 15)] = acc[cb][v];
}

// One pivot step KK of factor_block (after its barrier).  Wave 0 (lane r
// owns row r of the block in x): column KK + 1 takes step KK's update first,
// with L[KK+1][KK] read from lane KK + 1's register, so the next pivot, its
// Newton steps and the scaled column KK + 1 (published as row KK + 1 of Lt)
// follow without an LDS round trip; the other columns read row KK of Lt, up
// to RA entries of it issued before that dependent chain so that the LDS
// latency hides behind it.  Wave 1 (lane c owns column c of W = L^-1 in x)
// takes substitution step KK with row KK of Lt and Rs[KK].
constexpr int RA = 32;

template <int KK>
__device__ __forceinline__ void pivot_step(double (&x)[CB], double& v, int& bad, double (*Lt)[CB + 1], double* Rs,
                                           int w, int l, bool need_w) {
  if (w == 0) {
    if (KK + 1 < CB) {
      constexpr int c0 = KK + 2, na = CB - c0 < RA ? CB - c0 : RA;
      double ra[na > 0 ? na : 1];
#pragma unroll
      for (int i = 0; i < na; ++i) ra[i] = Lt[KK][c0 + i];
      __asm__ volatile("" ::: "memory");
      x[KK + 1] = fma(-v, readlane_d(v, KK + 1), x[KK + 1]);
      const double d = readlane_d(x[KK + 1], KK + 1);
      if (!(d > 0.0) && bad == 0) bad = KK + 2;
      double s, rs;
      sqrt_rsqrt(d, s, rs);
      const double vn = (l == KK + 1) ? s : x[KK + 1] * rs;
#pragma unroll
      for (int i = 0; i < na; ++i) x[c0 + i] = fma(-v, ra[i], x[c0 + i]);
#pragma unroll
      for (int c = c0 + na; c < CB; ++c) x[c] = fma(-v, Lt[KK][c], x[c]);
      x[KK + 1] = vn;
      Lt[KK + 1][l] = vn;
      if (l == KK + 1) Rs[KK + 1] = rs;
      v = vn;
    }
  } else if (w == 1 && need_w) {
    constexpr int c0 = KK + 1, na = CB - c0 < RA ? CB - c0 : RA;
    const double r = Rs[KK];
    double ra[na > 0 ? na : 1];
#pragma unroll
    for (int i = 0; i < na; ++i) ra[i] = Lt[KK][c0 + i];
    __asm__ volatile("" ::: "memory");
    x[KK] *= r;
#pragma unroll
    for (int i = 0; i < na; ++i) x[c0 + i] = fma(-ra[i], x[KK], x[c0 + i]);
#pragma unroll
    for (int c = c0 + na; c < CB; ++c) x[c] = fma(-Lt[KK][c], x[KK], x[c]);
  }
}

template <int... K>
__device__ __forceinline__ void pivot_steps(std::integer_sequence<int, K...>, double (&x)[CB], double& v, int& bad,
                                            double (*Lt)[CB + 1], double* Rs, int w, int l, bool need_w) {
  ((__syncthreads(), pivot_step<K>(x, v, bad, Lt, Rs, w, l, need_w)), ...);
}

// Factor the block in C (LDS, lower part read) at global rows/cols base ..
// base + 63 (jb of them inside the matrix): L (lower) to adiag, W = L^-1 to
// w_out (row-major 64 x 64, zero above the diagonal) when need_w.  Lt (row kk
// = column kk of L) and Rs (1 / L[kk][kk]) are LDS scratch.  All four waves
// call this (the barriers are workgroup-wide); waves 0 and 1 work in
// lockstep, one barrier per pivot.  info: first non-positive pivot (1-based,
// global), kept if already set.
__device__ void factor_block(double (*C)[CB + 1], double (*Lt)[CB + 1], double* Rs, bool need_w, double* adiag,
                             long lda, int jb, double* w_out, int* info, int base) {
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  double x[CB];
  int bad = 0;
  double v = 0.0;
  if (w == 0) {
#pragma unroll
    for (int c = 0; c < CB; ++c) x[c] = C[l][c];
    const double d = readlane_d(x[0], 0);
    if (!(d > 0.0)) bad = 1;
    double s, rs;
    sqrt_rsqrt(d, s, rs);
    v = (l == 0) ? s : x[0] * rs;
    x[0] = v;
    Lt[0][l] = v;
    if (l == 0) Rs[0] = rs;
  } else if (w == 1) {
#pragma unroll
    for (int c = 0; c < CB; ++c) x[c] = (c == l) ? 1.0 : 0.0;
  }
  pivot_steps(std::make_integer_sequence<int, CB>{}, x, v, bad, Lt, Rs, w, l, need_w);
  // after the last barrier only wave 1 reads Lt (row 63, no entries) and Rs
  if (w == 0) {
    if (l == 0 && bad && *info == 0) *info = base + bad;
#pragma unroll
    for (int c = 0; c < CB; ++c) C[l][c] = x[c];
  } else if (w == 1 && need_w) {
#pragma unroll
    for (int c = 0; c < CB; ++c) Lt[c][l] = x[c];
  }
  __syncthreads();
  const int c = l, rb = w;
#pragma unroll
  for (int i = 0; i < CB / 4; ++i) {
    const int r = rb + 4 * i;
    if (r < jb && c <= r) adiag[(long)r * lda + c] = C[r][c];
    if (need_w) w_out[r * CB + c] = Lt[r][c];
  }
}

__device__ inline void lower_tile_of(int t, int& ti, int& tj) {
  int i = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((i + 1) * (i + 2) / 2 <= t) ++i;
  while (i * (i + 1) / 2 > t) --i;
  ti = i;
  tj = t - i * (i + 1) / 2;
}

__global__ __launch_bounds__(CT) void chol_step(double* a, long lda, int k, int nb, int j,
                                                const double* __restrict__ w_in, double* __restrict__ w_out, int* info) {
  __shared__ double Ws[CB][CB + 1];
  __shared__ double Xi[CB][CB + 1];
  __shared__ double Xj[CB][CB + 1];
  __shared__ double Rs[CB];
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  int ti, tj;
  lower_tile_of((int)blockIdx.x, ti, tj);
  const int I = j + 1 + ti, J = j + 1 + tj;
  const long rI = (long)I * CB, rJ = (long)J * CB;
  const bool factor = blockIdx.x == 0;  // tile (j + 1, j + 1)
  if (j >= 0) {
    const long cj = (long)j * CB;
    load_tile(a, lda, k, rI, cj, Xi);
    if (J != I) load_tile(a, lda, k, rJ, cj, Xj);
    {
      double v[CB / 4];
#pragma unroll
      for (int i = 0; i < CB / 4; ++i) v[i] = w_in[(w + 4 * i) * CB + l];
#pragma unroll
      for (int i = 0; i < CB / 4; ++i) Ws[w + 4 * i][l] = v[i];
    }
    __syncthreads();
    panel_rows(Xi, Ws);
    if (J != I) panel_rows(Xj, Ws);
    __syncthreads();
    const double(*Lj)[CB + 1] = (J != I) ? Xj : Xi;
    if (I == J) {
      // L_I^T parked in the upper block (j, I): a[cj + c][rI + r] = L_I[r][c]
#pragma unroll
      for (int i = 0; i < CB / 4; ++i) {
        const int c = w + 4 * i;
        if (rI + l < k) a[(cj + c) * lda + rI + l] = Xi[l][c];
      }
    }
    d4 acc[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[cb][v] = ext_at(a, lda, k, rI + 16 * w + (l >> 4) + 4 * v, rJ + 16 * cb + (l & 15));
#pragma unroll
    for (int q0 = 0; q0 < CB; q0 += 4) {
      const double av = -Xi[16 * w + (l & 15)][q0 + (l >> 4)];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) acc[cb] = mfma_f64(av, Lj[16 * cb + (l & 15)][q0 + (l >> 4)], acc[cb]);
    }
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long r = rI + 16 * w + (l >> 4) + 4 * v, c = rJ + 16 * cb + (l & 15);
        if (r < k && c < k) a[r * lda + c] = acc[cb][v];
      }
    if (!factor) return;
    // I == J here, so Xj is free; the other waves read only Xi
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
      for (int v = 0; v < 4; ++v) Xj[16 * w + (l >> 4) + 4 * v][16 * cb + (l & 15)] = acc[cb][v];
  } else {
    if (!factor) return;  // step -1 is launched as one workgroup
    load_tile(a, lda, k, 0, 0, Xj);
  }
  __syncthreads();
  factor_block(Xj, Ws, Rs, j + 2 < nb, a + rI * lda + rI, lda, (int)std::min<long>(CB, k - rI), w_out, info,
               (int)rI);
}

// Parked panels into the lower triangle, upper triangle zeroed: workgroup
// per lower block (I, p); the upper block (p, I) holds L(I, p) transposed.
__global__ __launch_bounds__(CT) void chol_finish(double* a, long lda, int k) {
  __shared__ double S[CB][CB + 1];
  int I, p;
  lower_tile_of((int)blockIdx.x, I, p);
  const int c = threadIdx.x & 63, rb = threadIdx.x >> 6;
  const long rI = (long)I * CB, rp = (long)p * CB;
  if (I == p) {
#pragma unroll
    for (int i = 0; i < CB / 4; ++i) {
      const int r = rb + 4 * i;
      if (c > r && rI + c < k) a[(rI + r) * lda + rI + c] = 0.0;
    }
    return;
  }
  const bool col_ok = rI + c < k;
  double v[CB / 4];
#pragma unroll
  for (int i = 0; i < CB / 4; ++i) v[i] = a[col_ok ? (rp + rb + 4 * i) * lda + rI + c : 0];
#pragma unroll
  for (int i = 0; i < CB / 4; ++i) {
    S[rb + 4 * i][c] = v[i];
    if (col_ok) a[(rp + rb + 4 * i) * lda + rI + c] = 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < CB / 4; ++i) {
    const int r = rb + 4 * i;
    if (rI + r < k) a[(rI + r) * lda + rp + c] = S[c][r];
  }
}

__global__ __launch_bounds__(256) void zero_upper_kernel(double* a, int k, long lda) {
  const long total = (long)k * k;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += 256L * gridDim.x) {
    const long r = e / k, c = e % k;
    if (c > r) a[r * lda + c] = 0.0;
  }
}

__global__ __launch_bounds__(256) void symmetrize_kernel(double* a, int k, long lda) {
  const long total = (long)k * k;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += 256L * gridDim.x) {
    const long r = e / k, c = e % k;
    if (c > r) a[r * lda + c] = a[c * lda + r];
  }
}

// Lp (kp x kp) = [[L, 0], [0, I]] with the strict upper part zero.
__global__ __launch_bounds__(256) void pad_lower(const double* l, int k, long lda, double* lp, int kp) {
  const long total = (long)kp * kp;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += 256L * gridDim.x) {
    const long r = e / kp, c = e % kp;
    double v = 0.0;
    if (r < k && c < k) v = (c <= r) ? l[r * lda + c] : 0.0;
    else if (r == c) v = 1.0;
    lp[e] = v;
  }
}

// Invert each 64 x 64 lower diagonal block of lp into linv (same positions):
// thread t solves x L^T = e_t, i.e. x = column t of L^-1.
__global__ __launch_bounds__(64) void invert_diag_blocks(const double* lp, int kp, double* linv) {
  __shared__ double Lt[NB][NB + 1];
  const int j0 = blockIdx.x * NB, t = threadIdx.x;
  load_lower_block<true>(lp, kp, j0, j0, NB, Lt);
  wave_sync_lds();
  double x[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) x[c] = (c == t) ? 1.0 : 0.0;
  solve_row(x, Lt);
#pragma unroll
  for (int i = 0; i < NB; ++i) linv[(long)(j0 + i) * kp + j0 + t] = x[i];
}

__global__ __launch_bounds__(256) void fill_zero(double* a, long count) {
  for (long e = blockIdx.x * 256L + threadIdx.x; e < count; e += 256L * gridDim.x) a[e] = 0.0;
}

inline int grid_for(long count) { return (int)std::min<long>(4096, std::max<long>(1, (count + 255) / 256)); }

// ---- symmetric tridiagonal reduction (rank of CC^T) -------------------------
// Householder reduction A = Q T Q^T of a symmetric k x k matrix, one launch
// per column j.  Launch j applies the rank-2 update of reflection j - 1
// (A22 -= v w^T + w v^T, w = p - (beta/2)(p.v) v) to the rows it owns, forms
// the updated row j in every workgroup (row j is read, never written, in this
// launch), takes reflection j from it (v = x - alpha e1, beta = 1/(alpha
// (alpha - x0)), alpha = -sign(x0) |x|), and multiplies its updated rows by v
// (p = beta A22 v).  The diagonal d and off-diagonal e of T go to d, e.
constexpr int TD_MAXK = 4096;

__device__ inline double td_block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void tridiag_step(double* __restrict__ W, int k, int j,
                                                    const double* __restrict__ vp, const double* __restrict__ pp,
                                                    const double* __restrict__ bp, double* __restrict__ vn,
                                                    double* __restrict__ pn, double* __restrict__ bn,
                                                    double* __restrict__ d, double* __restrict__ e) {
  __shared__ double sv[TD_MAXK], sw[TD_MAXK], sx[TD_MAXK];
  __shared__ double red[4];
  const int t = threadIdx.x;
  const bool upd = j > 0;
  if (upd) {
    double g = 0.0;
    for (int c = j + t; c < k; c += 256) g += pp[c] * vp[c];
    g = td_block_sum(g, red);
    const double hb = 0.5 * bp[0] * g;
    for (int c = j + t; c < k; c += 256) {
      sv[c] = vp[c];
      sw[c] = pp[c] - hb * vp[c];
    }
    __syncthreads();
  }
  for (int c = j + t; c < k; c += 256) {
    double x = W[(long)j * k + c];
    if (upd) x -= sv[j] * sw[c] + sw[j] * sv[c];
    sx[c] = x;
  }
  __syncthreads();
  double nrm2 = 0.0;
  for (int c = j + 1 + t; c < k; c += 256) nrm2 += sx[c] * sx[c];
  nrm2 = td_block_sum(nrm2, red);
  double alpha = 0.0, beta = 0.0;
  if (j + 1 < k && nrm2 > 0.0) {
    const double x0 = sx[j + 1];
    alpha = x0 >= 0.0 ? -sqrt(nrm2) : sqrt(nrm2);
    beta = 1.0 / (alpha * (alpha - x0));
  }
  __syncthreads();
  if (t == 0 && j + 1 < k) sx[j + 1] -= alpha;  // v = x - alpha e1
  __syncthreads();
  if (blockIdx.x == 0) {
    if (t == 0) {
      d[j] = sx[j];
      if (j + 1 < k) e[j] = alpha;
      bn[0] = beta;
    }
    for (int c = j + 1 + t; c < k; c += 256) vn[c] = sx[c];
  }
  for (int i = j + 1 + blockIdx.x; i < k; i += gridDim.x) {
    double acc = 0.0;
    double* row = W + (long)i * k;
    for (int c = j + 1 + t; c < k; c += 256) {
      double x = row[c];
      if (upd) x -= sv[i] * sw[c] + sw[i] * sv[c];
      row[c] = x;
      acc += x * sx[c];
    }
    acc = td_block_sum(acc, red);
    if (t == 0) pn[i] = beta * acc;
  }
}

__global__ __launch_bounds__(256) void copy_square(const double* a, long lda, int k, double* w) {
  const long total = (long)k * k;
  for (long q = blockIdx.x * 256L + threadIdx.x; q < total; q += 256L * gridDim.x) w[q] = a[(q / k) * lda + q % k];
}

// number of eigenvalues of the tridiagonal (d, e) below x (Sturm count of the
// LDL^T pivots of T - x I)
int sturm_count(const double* d, const double* e, int k, double x) {
  int neg = 0;
  double q = 1.0;
  for (int i = 0; i < k; ++i) {
    const double off = i > 0 ? e[i - 1] * e[i - 1] : 0.0;
    q = d[i] - x - (i > 0 ? off / q : 0.0);
    if (q == 0.0) q = -1e-300;  // an exact zero pivot counts as negative
    if (q < 0.0) ++neg;
  }
  return neg;
}

// eigenvalue of rank 'idx' (0 = smallest) by bisection within the Gershgorin
// interval, to the resolution of double
double tridiag_eig(const double* d, const double* e, int k, int idx) {
  double lo = 0.0, hi = 0.0;
  for (int i = 0; i < k; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(e[i]) : 0.0);
    lo = i == 0 ? d[i] - r : std::min(lo, d[i] - r);
    hi = i == 0 ? d[i] + r : std::max(hi, d[i] + r);
  }
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (sturm_count(d, e, k, mid) > idx) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

}  // namespace

size_t chol_work_bytes(int) { return 2 * align_up(sizeof(double) * CB * CB, 256); }

int cholesky_f64(double* a, int k, long lda, int* d_info, void* work, hipStream_t s) {
  if (k < 1) return THC_OK;
  // W_j lives in wb[j & 1]: step j - 1 writes it, step j reads it
  double* wb[2] = {reinterpret_cast<double*>(work),
                   reinterpret_cast<double*>(reinterpret_cast<char*>(work) + align_up(sizeof(double) * CB * CB, 256))};
  const int nb = ceil_div(k, CB);
  for (int j = -1; j < nb - 1; ++j) {
    const long tb = j < 0 ? 1 : nb - j - 1;  // step -1 has only the factor of block 0
    hipLaunchKernelGGL(chol_step, dim3((unsigned)(tb * (tb + 1) / 2)), dim3(CT), 0, s, a, lda, k, nb, j,
                       wb[(j + 2) & 1], wb[(j + 1) & 1], d_info);
    THC_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(chol_finish, dim3((unsigned)((long)nb * (nb + 1) / 2)), dim3(CT), 0, s, a, lda, k);
  THC_HIP(hipGetLastError());
  return THC_OK;
}

static int padded(int k) {
  int kp = NB;
  while (kp < k) kp *= 2;
  return kp;
}

size_t spd_inverse_work_bytes(int k) {
  const size_t kp = padded(k);
  return 3 * align_up(sizeof(double) * kp * kp, 256) + 1024;
}

int spd_inverse_from_chol(const double* l, int k, long lda, double* ainv, void* work, hipStream_t s) {
  const int kp = padded(k);
  const size_t mat = align_up(sizeof(double) * (size_t)kp * kp, 256);
  double* lp = reinterpret_cast<double*>(work);
  double* linv = reinterpret_cast<double*>(reinterpret_cast<char*>(work) + mat);
  double* tmp = reinterpret_cast<double*>(reinterpret_cast<char*>(work) + 2 * mat);
  const long kk = (long)kp * kp;
  hipLaunchKernelGGL(pad_lower, dim3(grid_for(kk)), dim3(256), 0, s, l, k, lda, lp, kp);
  hipLaunchKernelGGL(fill_zero, dim3(grid_for(kk)), dim3(256), 0, s, linv, kk);
  hipLaunchKernelGGL(invert_diag_blocks, dim3(kp / NB), dim3(64), 0, s, lp, kp, linv);
  THC_HIP(hipGetLastError());
  for (int sz = NB; sz < kp; sz *= 2) {
    const int pairs = kp / (2 * sz);
    const long step = (long)2 * sz * kp + 2 * sz;  // diagonal stride between pairs
    // tmp_p = L21 Linv11
    GemmArgs g = {};
    g.m = sz; g.n = sz; g.k = sz;
    g.a = lp + (long)sz * kp; g.lda = kp; g.a_trans = 0;
    g.b = linv; g.ldb = kp; g.b_trans = 0;
    g.c = tmp; g.ldc = sz;
    g.alpha = 1.0; g.beta = 0.0; g.batch = pairs;
    g.tri = GEMM_TRI_B_KGE_N;  // Linv11 is lower triangular
    g.stride_a = step; g.stride_b = step; g.stride_c = (long)sz * sz;
    THC_TRY(gemm_f64(g, s));
    // Linv21 = -Linv22 tmp_p
    GemmArgs h = {};
    h.m = sz; h.n = sz; h.k = sz;
    h.a = linv + (long)sz * kp + sz; h.lda = kp; h.a_trans = 0;
    h.b = tmp; h.ldb = sz; h.b_trans = 0;
    h.c = linv + (long)sz * kp; h.ldc = kp;
    h.alpha = -1.0; h.beta = 0.0; h.batch = pairs;
    h.tri = GEMM_TRI_A_KLE_M;  // Linv22 is lower triangular
    h.stride_a = step; h.stride_b = (long)sz * sz; h.stride_c = step;
    THC_TRY(gemm_f64(h, s));
  }
  // A^-1 = Linv^T Linv (top-left k x k): lower tiles only, each over the k
  // range where both triangular operands are nonzero, then mirrored
  GemmArgs f = {};
  f.m = k; f.n = k; f.k = k;
  f.a = linv; f.lda = kp; f.a_trans = 1;
  f.b = linv; f.ldb = kp; f.b_trans = 0;
  f.c = ainv; f.ldc = k;
  f.alpha = 1.0; f.beta = 0.0; f.batch = 1;
  f.lower = 1;
  f.tri = GEMM_TRI_A_KGE_M | GEMM_TRI_B_KGE_N;
  THC_TRY(gemm_f64(f, s));
  hipLaunchKernelGGL(symmetrize_kernel, dim3(grid_for((long)k * k)), dim3(256), 0, s, ainv, k, (long)k);
  THC_HIP(hipGetLastError());
  return THC_OK;
}

size_t tridiag_work_bytes(int k) {
  return align_up(sizeof(double) * (size_t)k * k, 256) + 6 * align_up(sizeof(double) * (size_t)k, 256) +
         2 * align_up(sizeof(double) * 4, 256) + 256;
}

int sym_tridiag(const double* a, int k, long lda, double* d_host, double* e_host, void* work, hipStream_t s) {
  if (k < 1) return THC_OK;
  if (k > TD_MAXK) return THC_ERR_UNSUPPORTED;
  char* w = reinterpret_cast<char*>(work);
  auto take = [&](size_t count) {
    double* p = reinterpret_cast<double*>(w);
    w += align_up(sizeof(double) * count, 256);
    return p;
  };
  double* W = take((size_t)k * k);
  double* vb[2] = {take(k), take(k)};
  double* pb[2] = {take(k), take(k)};
  double* bb[2] = {take(4), take(4)};
  double* dd = take(k);
  double* ee = take(k);
  hipLaunchKernelGGL(copy_square, dim3(grid_for((long)k * k)), dim3(256), 0, s, a, lda, k, W);
  THC_HIP(hipGetLastError());
  for (int j = 0; j < k; ++j) {
    const int cur = j & 1, prev = cur ^ 1;
    const int rows = std::max(1, std::min(128, k - j - 1));
    hipLaunchKernelGGL(tridiag_step, dim3(rows), dim3(256), 0, s, W, k, j, vb[prev], pb[prev], bb[prev], vb[cur],
                       pb[cur], bb[cur], dd, ee);
    THC_HIP(hipGetLastError());
  }
  THC_HIP(hipMemcpyAsync(d_host, dd, sizeof(double) * k, hipMemcpyDeviceToHost, s));
  if (k > 1) THC_HIP(hipMemcpyAsync(e_host, ee, sizeof(double) * (k - 1), hipMemcpyDeviceToHost, s));
  return THC_OK;
}

void tridiag_rank(const double* d, const double* e, int k, int* rank, double* smax) {
  if (k < 1) {
    *rank = 0;
    *smax = 0.0;
    return;
  }
  // singular values of a symmetric matrix are |eigenvalues|; the largest is
  // the larger magnitude of the two extreme eigenvalues
  const double lmin = tridiag_eig(d, e, k, 0), lmax = tridiag_eig(d, e, k, k - 1);
  const double s0 = std::max(std::fabs(lmin), std::fabs(lmax));
  // MatrixOperations::rank: sigma_i < max(m, n) s0 1e-16 counts as zero
  const double rcond = (double)k * s0 * 1e-16;
  const int nul = sturm_count(d, e, k, rcond) - sturm_count(d, e, k, -rcond);
  *rank = k - nul;
  *smax = s0;
}

int zero_upper(double* a, int k, long lda, hipStream_t s) {
  hipLaunchKernelGGL(zero_upper_kernel, dim3(grid_for((long)k * k)), dim3(256), 0, s, a, k, lda);
  THC_HIP(hipGetLastError());
  return THC_OK;
}

int symmetrize_from_lower(double* a, int k, long lda, hipStream_t s) {
  hipLaunchKernelGGL(symmetrize_kernel, dim3(grid_for((long)k * k)), dim3(256), 0, s, a, k, lda);
  THC_HIP(hipGetLastError());
  return THC_OK;
}

}  // namespace thc

// ---------------------------------------------------------------- no-allocation Cholesky
//
// thc_cholesky_noalloc's scratch: the two panel-inverse buffers cholesky_f64
// works in, then its info word.  A file-scope device array: each GPU the
// library runs on holds one copy, for the life of the process.  In host code
// the name 'scratch' is a host-side stand-in, never a device address; the
// device address of each GPU's copy comes from hipGetSymbolAddress, taken once
// per device (from thc_ctx_create, with that device current) and kept below.
namespace thc {
namespace {
constexpr size_t kCholScratchWork = 2 * 64 * 64 * sizeof(double);  // chol_work_bytes(): two 64 x 64 panels
constexpr size_t kCholScratchBytes = kCholScratchWork + 256;        // + the info word (256-byte slot)
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace
}  // namespace thc

static __device__ __attribute__((aligned(256))) unsigned char scratch[thc::kCholScratchBytes];

namespace thc {

int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return THC_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return THC_OK;
  void* p = nullptr;
  THC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = p;
  return THC_OK;
}

namespace {
__global__ void zero_info(int* p) {
  if (threadIdx.x == 0) *p = 0;
}
}  // namespace

}  // namespace thc

extern "C" int thc_cholesky_noalloc(thc_ctx* ctx, double* a, int k, int* info) {
  using namespace thc;
  if (!ctx || !a || !info || k < 1) return THC_ERR_ARG;
  if (chol_work_bytes(k) > kCholScratchWork) {
    set_last_error("thc_cholesky_noalloc: the Cholesky panels need more than the static scratch holds", __FILE__,
                   __LINE__);
    return THC_ERR_UNSUPPORTED;
  }
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return THC_ERR_UNSUPPORTED;
  if (!ctx->pinned || ctx->pinned_bytes < sizeof(int)) return THC_ERR_ARG;  // thc_ctx_create provides it
  THC_HIP(hipSetDevice(ctx->device));
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = g_scratch_ptr[ctx->device];
  }
  if (!base) THC_TRY(noalloc_scratch_init(ctx->device));  // a device the library was not set up on
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = g_scratch_ptr[ctx->device];
  }
  double* work = reinterpret_cast<double*>(base);
  int* d_info = reinterpret_cast<int*>(reinterpret_cast<char*>(base) + kCholScratchWork);
  hipLaunchKernelGGL(zero_info, dim3(1), dim3(64), 0, ctx->stream, d_info);
  THC_HIP(hipGetLastError());
  THC_TRY(cholesky_f64(a, k, k, d_info, work, ctx->stream));
  int* h = reinterpret_cast<int*>(ctx->pinned);
  THC_HIP(hipMemcpyAsync(h, d_info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  THC_HIP(hipStreamSynchronize(ctx->stream));
  *info = *h;
  return THC_OK;
}
