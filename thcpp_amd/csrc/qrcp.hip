// Interpolation points by QR with column pivoting (QRCPSolver::kernel,
// src/qrcp.cpp:100-173; pzgeqpf at src/matrix_operations.h:756-788) without
// forming the orbital-product matrix.
//
// The reference factors ZT[(i k), a] = phi_i(r_a) phi_k(r_a) (M^2 x N; for the
// half-rotated set, i runs over the occupied orbitals) and keeps the first K
// pivot columns.  Column pivoting only ever looks at inner products of
// columns, and those are
//   <Z_a, Z_b> = (sum_i phi_i(a) phi_i(b)) (sum_k phi_k(a) phi_k(b)),
// i.e. the Gram matrix G = P o Pocc (P o P without half rotation).  Pivoted
// Cholesky of G picks, step by step, the column of largest residual norm and
// projects it out -- the same pivots as QR with column pivoting of ZT in exact
// arithmetic, with the pivot's residual norm equal to |R_jj|.  Each step is
// one kernel over the N grid points:
//   p      = argmax residual d (first index on ties), from the previous step's
//            per-block maxima;
//   g_a    = G[a][p] from two M-length dot products with the pivot's row;
//   l_a    = (g_a - sum_{q<j} L[q][a] L[q][p]) / sqrt(d_p);
//   d_a   -= l_a^2  (the pivot itself is retired);
//   new per-block maxima of d.
// Cost per step: one read of phi (N M) and of the j previous columns of L.
//
// Complex orbitals (ZT[(i k), a] = conj(phi_i(r_a)) phi_k(r_a), src/qrcp.cpp:
// 33-48): with Q[a][b] = sum_k conj(phi_k(a)) phi_k(b) the Gram matrix is
// conj(Q) o Q = |Q|^2, real, and the steps above apply with g_a = |Q[a][p]|^2.
// Half-rotated (i over the occupied set) it is conj(Qocc) o Q, Hermitian: the
// columns of L are complex, l_a = (G[a][p] - sum_q L[a][q] conj(L[p][q])) /
// sqrt(d_p) and d_a -= |l_a|^2; the residual norms, hence the pivots, are
// those of QR with column pivoting of the complex ZT.
#include <algorithm>
#include <cfloat>
#include <vector>

#include "common.h"

namespace thc {
namespace {

constexpr int QB = 256;          // threads per block
constexpr double RETIRED = -DBL_MAX;  // residual of a chosen point (no arithmetic reaches it)

__device__ inline bool wins(double v, int i, double bv, int bi) { return v > bv || (v == bv && i < bi); }

// (max value, first index) over the block into pv/pi[blockIdx.x]
__device__ void block_argmax(double v, int i, double* pv, int* pi) {
  __shared__ double sv[QB / 64];
  __shared__ int si[QB / 64];
  for (int o = 32; o > 0; o >>= 1) {
    const double ov = __shfl_xor(v, o);
    const int oi = __shfl_xor(i, o);
    if (wins(ov, oi, v, i)) { v = ov; i = oi; }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { sv[w] = v; si[w] = i; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < QB / 64; ++q)
      if (wins(sv[q], si[q], v, i)) { v = sv[q]; i = si[q]; }
    pv[blockIdx.x] = v;
    pi[blockIdx.x] = i;
  }
}

__device__ inline double dot_row(const double* __restrict__ x, const double* __restrict__ y, int m) {
  double s0 = 0.0, s1 = 0.0;
  int q = 0;
  for (; q + 2 <= m; q += 2) {
    s0 = fma(x[q], y[q], s0);
    s1 = fma(x[q + 1], y[q + 1], s1);
  }
  if (q < m) s0 = fma(x[q], y[q], s0);
  return s0 + s1;
}

// sum_q conj(x_q) y_q over interleaved complex rows (m2 doubles)
__device__ inline void cdot_row(const double* __restrict__ x, const double* __restrict__ y, int m2, double& re,
                                double& im) {
  double r = 0.0, i = 0.0;
  for (int q = 0; q + 2 <= m2; q += 2) {
    r = fma(x[q], y[q], r);
    r = fma(x[q + 1], y[q + 1], r);
    i = fma(x[q], y[q + 1], i);
    i = fma(-x[q + 1], y[q], i);
  }
  re = r;
  im = i;
}

// d_a = G[a][a] and the first per-block maxima (sum |phi|^2 over interleaved
// complex rows is the real row's dot product, so one kernel serves both)
__global__ __launch_bounds__(QB) void qr_init(const double* __restrict__ phi, const double* __restrict__ occ, long n,
                                              int m, int mo, double* __restrict__ d, double* pv, int* pi) {
  const long a = (long)blockIdx.x * QB + threadIdx.x;
  double v = -DBL_MAX;
  int vi = INT_MAX;
  if (a < n) {
    const double* x = phi + a * m;
    const double s = dot_row(x, x, m);
    double g = s * s;
    if (occ) {
      const double* y = occ + a * mo;
      g = s * dot_row(y, y, mo);
    }
    d[a] = g;
    v = g;
    vi = (int)a;
  }
  block_argmax(v, vi, pv, pi);
}

// Pivot step j; partial maxima are read from (pv_in, pi_in) and written to
// (pv_out, pi_out), which the host alternates between steps.
// MODE 0: real orbitals; 1: complex orbitals, all on both sides (real G);
// 2: complex half-rotated (complex G and L: L[2q] real, L[2q + 1] imaginary
// parts of column q).
template <int MODE>
__global__ __launch_bounds__(QB) void qr_step(const double* __restrict__ phi, const double* __restrict__ occ, long n,
                                              int m, int mo, int j, int nparts, double* __restrict__ d,
                                              double* __restrict__ L, const double* pv_in, const int* pi_in,
                                              double* pv_out, int* pi_out, int* piv, double* rdiag, int* fail) {
  extern __shared__ double sh[];  // pivot rows: phi (m), occ (mo), then L[q][p] for q < j (2j values in MODE 2)
  __shared__ double gv;
  __shared__ int gi;
  // global argmax of the previous residuals (every block, redundantly)
  {
    double v = -DBL_MAX;
    int vi = INT_MAX;
    for (int q = threadIdx.x; q < nparts; q += QB)
      if (wins(pv_in[q], pi_in[q], v, vi)) { v = pv_in[q]; vi = pi_in[q]; }
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(v, o);
      const int oi = __shfl_xor(vi, o);
      if (wins(ov, oi, v, vi)) { v = ov; vi = oi; }
    }
    __shared__ double wv[QB / 64];
    __shared__ int wi[QB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { wv[w] = v; wi[w] = vi; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int q = 1; q < QB / 64; ++q)
        if (wins(wv[q], wi[q], v, vi)) { v = wv[q]; vi = wi[q]; }
      gv = v;
      gi = vi;
    }
    __syncthreads();
  }
  const double pval = gv;
  const long p = gi;
  // residual exhausted (K beyond the numerical rank): below 1e-13 of the first
  // pivot the squared residuals are rounding noise of the Gram downdate
  const double d0 = j > 0 ? rdiag[0] * rdiag[0] : pval;
  if (!(pval > 1e-13 * d0) || p < 0 || p >= n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *fail = j + 1;
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    piv[j] = (int)p;
    rdiag[j] = sqrt(pval);
  }
  double* prow = sh;
  double* orow = sh + m;
  double* lp = orow + (occ ? mo : 0);
  for (int q = threadIdx.x; q < m; q += QB) prow[q] = phi[p * m + q];
  if (occ)
    for (int q = threadIdx.x; q < mo; q += QB) orow[q] = occ[p * mo + q];
  constexpr int LW = MODE == 2 ? 2 : 1;  // doubles per entry of L
  for (int q = threadIdx.x; q < LW * j; q += QB) lp[q] = L[(long)q * n + p];
  __syncthreads();
  const double rs = 1.0 / sqrt(pval);
  const long a = (long)blockIdx.x * QB + threadIdx.x;
  double v = -DBL_MAX;
  int vi = INT_MAX;
  if (a < n) {
    double da = d[a];
    if (MODE == 2) {
      double qr, qi, o_r, o_i;
      cdot_row(phi + a * m, prow, m, qr, qi);
      cdot_row(occ + a * mo, orow, mo, o_r, o_i);
      const double gr = o_r * qr + o_i * qi, gi = o_r * qi - o_i * qr;  // conj(Qocc) Q
      double ar = 0.0, ai = 0.0;
      for (int q = 0; q < j; ++q) {
        const double xr = L[(long)(2 * q) * n + a], xi = L[(long)(2 * q + 1) * n + a];
        const double yr = lp[2 * q], yi = lp[2 * q + 1];
        ar = fma(xr, yr, fma(xi, yi, ar));
        ai = fma(xi, yr, fma(-xr, yi, ai));
      }
      const double lr = (gr - ar) * rs, li = (gi - ai) * rs;
      L[(long)(2 * j) * n + a] = lr;
      L[(long)(2 * j + 1) * n + a] = li;
      if (a == p || da == RETIRED) da = RETIRED;
      else da -= lr * lr + li * li;
    } else {
      double g;
      if (MODE == 1) {
        double qr, qi;
        cdot_row(phi + a * m, prow, m, qr, qi);
        g = qr * qr + qi * qi;  // |Q[a][p]|^2
      } else {
        const double s = dot_row(phi + a * m, prow, m);
        g = occ ? s * dot_row(occ + a * mo, orow, mo) : s * s;
      }
      double acc0 = 0.0, acc1 = 0.0;
      int q = 0;
      for (; q + 2 <= j; q += 2) {
        acc0 = fma(L[(long)q * n + a], lp[q], acc0);
        acc1 = fma(L[(long)(q + 1) * n + a], lp[q + 1], acc1);
      }
      if (q < j) acc0 = fma(L[(long)q * n + a], lp[q], acc0);
      const double l = (g - (acc0 + acc1)) * rs;
      L[(long)j * n + a] = l;
      if (a == p || da == RETIRED) da = RETIRED;
      else da -= l * l;
    }
    d[a] = da;
    v = da;
    vi = (int)a;
  }
  block_argmax(v, vi, pv_out, pi_out);
}

}  // namespace

size_t qrcp_work_bytes(long n, int k, bool cplx) {
  const long nparts = (n + QB - 1) / QB;
  return carve_bytes<double>((size_t)n) + carve_bytes<double>((size_t)n * k * (cplx ? 2 : 1)) +
         2 * carve_bytes<double>(nparts) + 2 * carve_bytes<int>(nparts) + carve_bytes<int>(k) +
         carve_bytes<double>(k) + carve_bytes<int>(1);
}

int qrcp_points(const double* phi, const double* occ, long n, int m, int mo, int k, int* piv_out, double* rdiag_out,
                Arena& ws, hipStream_t s, bool cplx) {
  const long nparts = (n + QB - 1) / QB;
  if (nparts > 0x7fffffffL) return THC_ERR_ARG;
  if (cplx && (m % 2 || (occ && mo % 2))) return THC_ERR_ARG;
  const int mode = !cplx ? 0 : (occ ? 2 : 1);
  double* d = ws.take<double>((size_t)n);
  double* L = ws.take<double>((size_t)n * k * (mode == 2 ? 2 : 1));
  double* pv[2] = {ws.take<double>(nparts), ws.take<double>(nparts)};
  int* pi[2] = {ws.take<int>(nparts), ws.take<int>(nparts)};
  int* piv = ws.take<int>(k);
  double* rd = ws.take<double>(k);
  int* fail = ws.take<int>(1);
  if (!d || !L || !pv[0] || !pv[1] || !pi[0] || !pi[1] || !piv || !rd || !fail) return THC_ERR_NO_MEMORY;
  const size_t lds = sizeof(double) * ((size_t)m + (occ ? mo : 0) + (size_t)k * (mode == 2 ? 2 : 1));
  const void* step_fn = mode == 0 ? (const void*)qr_step<0> : (mode == 1 ? (const void*)qr_step<1>
                                                                          : (const void*)qr_step<2>);
  if (lds > 160 * 1024) return THC_ERR_UNSUPPORTED;
  if (lds > 64 * 1024) THC_HIP(hipFuncSetAttribute(step_fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  THC_HIP(hipMemsetAsync(fail, 0, sizeof(int), s));
  hipLaunchKernelGGL(qr_init, dim3((unsigned)nparts), dim3(QB), 0, s, phi, occ, n, m, mo, d, pv[0], pi[0]);
  THC_HIP(hipGetLastError());
  for (int j = 0; j < k; ++j) {
    const int in = j & 1;
    if (mode == 0)
      hipLaunchKernelGGL(qr_step<0>, dim3((unsigned)nparts), dim3(QB), lds, s, phi, occ, n, m, mo, j, (int)nparts, d,
                         L, pv[in], pi[in], pv[in ^ 1], pi[in ^ 1], piv, rd, fail);
    else if (mode == 1)
      hipLaunchKernelGGL(qr_step<1>, dim3((unsigned)nparts), dim3(QB), lds, s, phi, occ, n, m, mo, j, (int)nparts, d,
                         L, pv[in], pi[in], pv[in ^ 1], pi[in ^ 1], piv, rd, fail);
    else
      hipLaunchKernelGGL(qr_step<2>, dim3((unsigned)nparts), dim3(QB), lds, s, phi, occ, n, m, mo, j, (int)nparts, d,
                         L, pv[in], pi[in], pv[in ^ 1], pi[in ^ 1], piv, rd, fail);
  }
  THC_HIP(hipGetLastError());
  int h_fail = 0;
  std::vector<int> h(k);
  THC_HIP(hipMemcpyAsync(&h_fail, fail, sizeof(int), hipMemcpyDeviceToHost, s));
  THC_HIP(hipMemcpyAsync(h.data(), piv, sizeof(int) * k, hipMemcpyDeviceToHost, s));
  if (rdiag_out) THC_HIP(hipMemcpyAsync(rdiag_out, rd, sizeof(double) * k, hipMemcpyDeviceToHost, s));
  THC_HIP(hipStreamSynchronize(s));
  if (h_fail) return THC_ERR_NOT_POSDEF;
  // the reference keeps the first K pivots in ascending grid order
  std::sort(h.begin(), h.end());
  THC_HIP(hipMemcpyAsync(piv_out, h.data(), sizeof(int) * k, hipMemcpyHostToDevice, s));
  THC_HIP(hipStreamSynchronize(s));
  return THC_OK;
}

}  // namespace thc
