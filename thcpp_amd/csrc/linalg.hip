// FP64 dense factorisations on the GPU for the K x K matrices of the path:
// Cholesky of CC^T and of V (pzpotrf 'L' in the reference,
// src/matrix_operations.h:259-267) and the SPD inverse used to apply
// CC^T^-1 to ZC^T as one MFMA GEMM.
//
// Cholesky: right-looking, 64-column panels.  Per panel: one wave factors the
// 64 x 64 diagonal block (row per lane, in registers); the panel below is
// solved against it row by row (L21 = A21 L11^-T, one row per thread); the
// trailing lower triangle takes the rank-64 update (MFMA GEMM, lower tiles
// only).
// Inverse: the lower factor is padded with an identity to 64 * 2^p, its
// 64 x 64 diagonal blocks are inverted in parallel, then recursive doubling
// forms Linv21 = -Linv22 (L21 Linv11) for every pair of blocks of one size in
// two batched GEMMs per level; finally A^-1 = Linv^T Linv.
#include <algorithm>

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

// Rows r0 .. r0 + NB - 1 (below 'rows': zero) of the rows x jb panel p (row
// stride lda) into Xs, columns >= jb zero; batched loads as above.
__device__ inline void load_panel(const double* p, long lda, long r0, long rows, int jb, double (*Xs)[NB + 1]) {
  const int t = threadIdx.x;
  double v[NB];
#pragma unroll
  for (int rr = 0; rr < NB; ++rr) {
    const bool ok = r0 + rr < rows && t < jb;
    v[rr] = p[ok ? (r0 + rr) * lda + t : 0];
  }
#pragma unroll
  for (int rr = 0; rr < NB; ++rr) Xs[rr][t] = (r0 + rr < rows && t < jb) ? v[rr] : 0.0;
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

__device__ inline double readlane_d(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Factor the jb x jb diagonal block at (j0, j0) in place (lower part only).
// Right-looking, lane r owning row r in registers: at step k the pivot
// L[k][k] = sqrt(A[k][k]) is read from lane k, column k is scaled by its
// reciprocal and published in LDS as row k of Lt, and every lane drops
// L[r][k] L[c][k] from its entries c > k (broadcast reads of Lt[k][c]).
// Rows/columns >= jb are an identity extension.  Also writes lt (NB x NB,
// lt[c][r] = L[r][c], zero above the diagonal of L) and rinv[c] = 1 / L[c][c]
// for the panel solve.  info: first failing leading minor (1-based, global),
// kept if already set.  (A rolled form with the rows in LDS is smaller code
// but 2.3x slower: its read-modify-write chains are LDS-latency bound.)
__global__ __launch_bounds__(64) void potrf_diag(double* a, long lda, int j0, int jb, int* info, double* lt,
                                                 double* rinv) {
  __shared__ double Ls[NB][NB + 1];
  __shared__ __attribute__((aligned(16))) double Lt[NB][NB];
  const int r = threadIdx.x;
  load_lower_block<false>(a, lda, j0, j0, jb, Ls);
  wave_sync_lds();
  double row[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) row[c] = Ls[r][c];
  int bad = 0;
  double my_rp = 1.0;
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const double d = readlane_d(row[k], k);
    if (!(d > 0.0) && bad == 0 && k < jb) bad = k + 1;
    const double piv = sqrt(d);
    const double rp = 1.0 / piv;
    if (r == k) my_rp = rp;
    const double v = (r == k) ? piv : row[k] * rp;
    row[k] = v;
    Lt[k][r] = v;
    wave_sync_lds();
#pragma unroll
    for (int c = k + 1; c < NB; ++c) row[c] = fma(-v, Lt[k][c], row[c]);
  }
  if (r == 0 && bad && *info == 0) *info = j0 + bad;
  rinv[r] = my_rp;
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    lt[c * NB + r] = (c <= r) ? row[c] : 0.0;
    Ls[r][c] = row[c];
  }
  wave_sync_lds();
  for (int rr = 0; rr < jb; ++rr)
    if (r <= rr) a[(long)(j0 + rr) * lda + j0 + r] = Ls[rr][r];
}

// L21 = A21 L11^-T in place for the 'rows' rows below the diagonal block at
// (j0, j0): one row per thread (x L11^T = a, forward substitution,
// right-looking, each x[c] scaled by the stored reciprocal of L[c][c]), 64 rows
// per workgroup.  L11^T and the reciprocals come from potrf_diag's output.
__global__ __launch_bounds__(64) void trsm_panel(double* a, long lda, int j0, int jb, long rows,
                                                 const double* __restrict__ lt, const double* __restrict__ rinv) {
  __shared__ double Lt[NB][NB + 1];
  __shared__ double Xs[NB][NB + 1];
  __shared__ double Ri[NB];
  const int t = threadIdx.x;
  const long r0 = (long)blockIdx.x * NB;
  double* p = a + (long)(j0 + jb) * lda + j0;
  {
    double v[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) v[c] = lt[c * NB + t];
    const double ri = rinv[t];
    load_panel(p, lda, r0, rows, jb, Xs);
#pragma unroll
    for (int c = 0; c < NB; ++c) Lt[c][t] = v[c];
    Ri[t] = ri;
  }
  wave_sync_lds();
  double x[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) x[c] = Xs[t][c];
  // col holds step c's factors: col[c] = 1 / L[c][c], col[c2] = L[c2][c] (c2 > c).
  // The next step's column is read before this step's FMAs, behind a compiler
  // fence, so the LDS latency overlaps the arithmetic.
  double col[NB];
  col[0] = Ri[0];
#pragma unroll
  for (int c2 = 1; c2 < NB; ++c2) col[c2] = Lt[0][c2];
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    double nxt[NB];
    if (c + 1 < NB) {
      nxt[c + 1] = Ri[c + 1];
#pragma unroll
      for (int c2 = c + 2; c2 < NB; ++c2) nxt[c2] = Lt[c + 1][c2];
    }
    __asm__ volatile("" ::: "memory");
    x[c] *= col[c];
#pragma unroll
    for (int c2 = c + 1; c2 < NB; ++c2) x[c2] = fma(-x[c], col[c2], x[c2]);
    if (c + 1 < NB) {
#pragma unroll
      for (int c2 = c + 1; c2 < NB; ++c2) col[c2] = nxt[c2];
    }
  }
#pragma unroll
  for (int c = 0; c < NB; ++c) Xs[t][c] = x[c];
  wave_sync_lds();
  for (int rr = 0; rr < NB; ++rr) {
    const long pr = r0 + rr;
    if (pr < rows && t < jb) p[pr * lda + t] = Xs[rr][t];
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

}  // namespace

size_t chol_work_bytes(int) { return align_up(sizeof(double) * NB * NB, 256) + align_up(sizeof(double) * NB, 256); }

int cholesky_f64(double* a, int k, long lda, int* d_info, void* work, hipStream_t s) {
  double* lt = reinterpret_cast<double*>(work);
  double* rinv = reinterpret_cast<double*>(reinterpret_cast<char*>(work) + align_up(sizeof(double) * NB * NB, 256));
  for (int j0 = 0; j0 < k; j0 += NB) {
    const int jb = std::min(NB, k - j0);
    hipLaunchKernelGGL(potrf_diag, dim3(1), dim3(64), 0, s, a, lda, j0, jb, d_info, lt, rinv);
    THC_HIP(hipGetLastError());
    const long rows = k - j0 - jb;
    if (rows <= 0) continue;
    hipLaunchKernelGGL(trsm_panel, dim3(ceil_div(rows, NB)), dim3(64), 0, s, a, lda, j0, jb, rows, lt, rinv);
    THC_HIP(hipGetLastError());
    // trailing lower update A22 -= L21 L21^T
    GemmArgs u = {};
    u.m = rows; u.n = rows; u.k = jb;
    u.a = a + (long)(j0 + jb) * lda + j0; u.lda = lda; u.a_trans = 0;
    u.b = u.a; u.ldb = lda; u.b_trans = 1;
    u.c = a + (long)(j0 + jb) * lda + (j0 + jb); u.ldc = lda;
    u.alpha = -1.0; u.beta = 1.0; u.lower = 1; u.batch = 1;
    THC_TRY(gemm_f64(u, s));
  }
  hipLaunchKernelGGL(zero_upper_kernel, dim3(grid_for((long)k * k)), dim3(256), 0, s, a, k, lda);
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
    g.stride_a = step; g.stride_b = step; g.stride_c = (long)sz * sz;
    THC_TRY(gemm_f64(g, s));
    // Linv21 = -Linv22 tmp_p
    GemmArgs h = {};
    h.m = sz; h.n = sz; h.k = sz;
    h.a = linv + (long)sz * kp + sz; h.lda = kp; h.a_trans = 0;
    h.b = tmp; h.ldb = sz; h.b_trans = 0;
    h.c = linv + (long)sz * kp; h.ldc = kp;
    h.alpha = -1.0; h.beta = 0.0; h.batch = pairs;
    h.stride_a = step; h.stride_b = (long)sz * sz; h.stride_c = step;
    THC_TRY(gemm_f64(h, s));
  }
  // A^-1 = Linv^T Linv (top-left k x k)
  GemmArgs f = {};
  f.m = k; f.n = k; f.k = k;
  f.a = linv; f.lda = kp; f.a_trans = 1;
  f.b = linv; f.ldb = kp; f.b_trans = 0;
  f.c = ainv; f.ldc = k;
  f.alpha = 1.0; f.beta = 0.0; f.batch = 1;
  THC_TRY(gemm_f64(f, s));
  return THC_OK;
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
