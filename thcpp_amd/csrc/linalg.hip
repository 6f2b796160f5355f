// FP64 dense factorisations on the GPU for the K x K matrices of the path:
// Cholesky of CC^T and of V (pzpotrf 'L' in the reference,
// src/matrix_operations.h:259-267) and the SPD inverse used to apply
// CC^T^-1 to ZC^T as one MFMA GEMM.
//
// Cholesky: right-looking, 64-column panels.  Per panel: one workgroup
// factors the 64 x 64 diagonal block in LDS and inverts it; the panel below is
// multiplied by that inverse (GEMM); the trailing lower triangle takes the
// rank-64 update (GEMM, lower tiles only).
// Inverse: the lower factor is padded with an identity to 64 * 2^p, its
// 64 x 64 diagonal blocks are inverted in parallel, then recursive doubling
// forms Linv21 = -Linv22 (L21 Linv11) for every pair of blocks of one size in
// two batched GEMMs per level; finally A^-1 = Linv^T Linv.
#include <algorithm>

#include "common.h"

namespace thc {
namespace {

constexpr int NB = 64;

// In-LDS factor and inverse of one NB x NB lower block, one wave.
// Factor (Cholesky-Crout): lane r owns row r and computes
// L[r][k] = (A[r][k] - sum_{q<k} L[r][q] L[k][q]) / L[k][k]; row k is read as
// broadcasts, the lane's own row with the (NB + 1) stride that keeps the 64
// lanes on distinct banks.  Inverse: lane c owns column c of L^-1,
// x[i] = (delta_ic - sum_{q<i} L[i][q] x[q]) / L[i][i].  Dot products use
// four partial sums so loads pipeline and the dependent chain is short.
// Rows and columns >= jb are an identity extension.  Returns the 1-based first
// failing pivot (0 if none).
__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ inline double dot_lds(const double* a, int astride, const double* b, int bstride, int n) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int q = 0;
  for (; q + 4 <= n; q += 4) {
    s0 += a[q * astride] * b[q * bstride];
    s1 += a[(q + 1) * astride] * b[(q + 1) * bstride];
    s2 += a[(q + 2) * astride] * b[(q + 2) * bstride];
    s3 += a[(q + 3) * astride] * b[(q + 3) * bstride];
  }
  for (; q < n; ++q) s0 += a[q * astride] * b[q * bstride];
  return (s0 + s1) + (s2 + s3);
}

__device__ int block_chol_inv(double (*Ls)[NB + 1], double (*Ds)[NB + 1], int jb) {
  const int r = threadIdx.x;
  int bad = 0;
  for (int k = 0; k < NB; ++k) {
    const double s = Ls[r][k] - dot_lds(&Ls[r][0], 1, &Ls[k][0], 1, k);
    const double d = __shfl(s, k);
    if (!(d > 0.0) && bad == 0 && k < jb) bad = k + 1;
    const double piv = sqrt(d);
    if (r >= k) Ls[r][k] = (r == k) ? piv : s / piv;
    wave_sync_lds();
  }
  // column r of L^-1, kept in Ds[.][r]
  for (int i = 0; i < NB; ++i) {
    double s = (i == r) ? 1.0 : 0.0;
    if (i > r) s = -dot_lds(&Ls[i][r], 1, &Ds[r][r], NB + 1, i - r);
    Ds[i][r] = (i < r) ? 0.0 : s / Ls[i][i];
  }
  wave_sync_lds();
  return bad;
}

// Factor the jb x jb diagonal block at (j0, j0) in place and write its lower
// inverse to dinv (NB x NB, row-major, zero upper).  info: first failing
// leading minor (1-based, global), kept if already set.
__global__ __launch_bounds__(64) void potrf_diag(double* a, long lda, int j0, int jb, int* info, double* dinv) {
  __shared__ double L[NB][NB + 1];
  __shared__ double D[NB][NB + 1];
  const int r = threadIdx.x;
  for (int c = 0; c < NB; ++c) {
    double v = 0.0;
    if (r < jb && c < jb) v = c <= r ? a[(long)(j0 + r) * lda + j0 + c] : 0.0;
    else if (r == c) v = 1.0;
    L[r][c] = v;
  }
  __syncthreads();
  const int bad = block_chol_inv(L, D, jb);
  if (r == 0 && bad && *info == 0) *info = j0 + bad;
  if (r < jb)
    for (int c = 0; c <= r; ++c) a[(long)(j0 + r) * lda + j0 + c] = L[r][c];
  for (int c = 0; c < NB; ++c) dinv[r * NB + c] = (r < jb && c < jb) ? D[r][c] : 0.0;
}

__global__ __launch_bounds__(256) void copy_block(const double* src, long lds, double* dst, long ldd, long rows,
                                                  int cols) {
  const long total = rows * cols;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += 256L * gridDim.x) {
    const long r = e / cols;
    const int c = (int)(e % cols);
    dst[r * ldd + c] = src[r * lds + c];
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

// Invert each 64 x 64 lower diagonal block of lp into linv (same positions).
__global__ __launch_bounds__(64) void invert_diag_blocks(const double* lp, int kp, double* linv) {
  __shared__ double L[NB][NB + 1];
  __shared__ double D[NB][NB + 1];
  const int j0 = blockIdx.x * NB, r = threadIdx.x;
  for (int c = 0; c < NB; ++c) L[r][c] = lp[(long)(j0 + r) * kp + j0 + c];
  __syncthreads();
  for (int i = 0; i < NB; ++i) {
    double s = (i == r) ? 1.0 : 0.0;
    if (i > r) s = -dot_lds(&L[i][r], 1, &D[r][r], NB + 1, i - r);
    D[i][r] = (i < r) ? 0.0 : s / L[i][i];
  }
  __syncthreads();
  for (int c = 0; c < NB; ++c) linv[(long)(j0 + r) * kp + j0 + c] = D[r][c];
}

__global__ __launch_bounds__(256) void fill_zero(double* a, long count) {
  for (long e = blockIdx.x * 256L + threadIdx.x; e < count; e += 256L * gridDim.x) a[e] = 0.0;
}

inline int grid_for(long count) { return (int)std::min<long>(4096, std::max<long>(1, (count + 255) / 256)); }

}  // namespace

size_t chol_work_bytes(int k) {
  return align_up(sizeof(double) * NB * NB, 256) + align_up(sizeof(double) * (size_t)k * NB, 256) + 512;
}

int cholesky_f64(double* a, int k, long lda, int* d_info, void* work, hipStream_t s) {
  double* dinv = reinterpret_cast<double*>(work);
  double* panel = reinterpret_cast<double*>(reinterpret_cast<char*>(work) + align_up(sizeof(double) * NB * NB, 256));
  for (int j0 = 0; j0 < k; j0 += NB) {
    const int jb = std::min(NB, k - j0);
    hipLaunchKernelGGL(potrf_diag, dim3(1), dim3(64), 0, s, a, lda, j0, jb, d_info, dinv);
    THC_HIP(hipGetLastError());
    const long rows = k - j0 - jb;
    if (rows <= 0) continue;
    // panel = A[j0+jb:, j0:j0+jb] * Dinv^T
    GemmArgs g = {};
    g.m = rows; g.n = jb; g.k = jb;
    g.a = a + (long)(j0 + jb) * lda + j0; g.lda = lda; g.a_trans = 0;
    g.b = dinv; g.ldb = NB; g.b_trans = 1;
    g.c = panel; g.ldc = NB;
    g.alpha = 1.0; g.beta = 0.0; g.batch = 1;
    THC_TRY(gemm_f64(g, s));
    hipLaunchKernelGGL(copy_block, dim3(grid_for(rows * jb)), dim3(256), 0, s, panel, (long)NB,
                       a + (long)(j0 + jb) * lda + j0, lda, rows, jb);
    // trailing lower update A22 -= panel panel^T
    GemmArgs u = {};
    u.m = rows; u.n = rows; u.k = jb;
    u.a = panel; u.lda = NB; u.a_trans = 0;
    u.b = panel; u.ldb = NB; u.b_trans = 1;
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
