// C ABI of libthcpp_amd: context, workspace, the ISDF pipeline (IVecs ctor +
// kernel at one rank, src/interpolating_vectors.cpp:17-57 and 406-455) and the
// k-means entry points (kmeans.hip).
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <future>
#include <string>
#include <thread>

#include "common.h"
#include "kmeans.h"

extern "C" int thc_seed_indices_impl(long n, int k, unsigned int seed, int* out);

namespace thc {

static thread_local std::string g_last_error;

void set_last_error(const char* what, const char* file, int line) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s (%s:%d)", what, file, line);
  g_last_error = buf;
}

int Arena::reserve(size_t bytes) {
  if (bytes <= cap) return THC_OK;
  release();
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_last_error(hipGetErrorString(e), __FILE__, __LINE__);
    return THC_ERR_NO_MEMORY;
  }
  base = reinterpret_cast<char*>(p);
  cap = bytes;
  used = 0;
  return THC_OK;
}

void Arena::release() {
  if (base) (void)hipFree(base);
  base = nullptr;
  cap = 0;
  used = 0;
}

int ensure_pinned(thc_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->pinned_bytes) return THC_OK;
  if (ctx->pinned) THC_HIP(hipHostFree(ctx->pinned));
  ctx->pinned = nullptr;
  ctx->pinned_bytes = 0;
  bytes = align_up(bytes, 4096);
  THC_HIP(hipHostMalloc(&ctx->pinned, bytes, hipHostMallocDefault));
  ctx->pinned_bytes = bytes;
  return THC_OK;
}

namespace {

// C[mu][nu] = Z[mu][idx[nu]]
__global__ __launch_bounds__(256) void gather_cols(const double* z, long ldz, const int* idx, int k, double* c) {
  const long total = (long)k * k;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += 256L * gridDim.x) {
    const long r = e / k, q = e % k;
    const long col = idx[q];
    c[e] = (col >= 0 && col < ldz) ? z[r * ldz + col] : 0.0;
  }
}

// rows idx of a (n x m) into out (k x m); indices outside [0, n) give zero rows
__global__ __launch_bounds__(256) void gather_rows(const double* a, long n, int m, const int* idx, int k,
                                                   double* out) {
  const long total = (long)k * m;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += 256L * gridDim.x) {
    const long row = idx[e / m];
    out[e] = (row >= 0 && row < n) ? a[row * m + e % m] : 0.0;
  }
}

// w = i * phi for interleaved complex values (re, im) -> (-im, re)
__global__ __launch_bounds__(256) void times_i(const double* phi, long count, double* w) {
  for (long e = blockIdx.x * 256L + threadIdx.x; e < count; e += 256L * gridDim.x) {
    const double re = phi[2 * e], im = phi[2 * e + 1];
    w[2 * e] = -im;
    w[2 * e + 1] = re;
  }
}

// Half-rotated ZC^T of complex orbitals in the stacked real layout zt (n x 2k
// row-major: zt[r][mu] = Re CZt[mu][r], zt[r][k + mu] = Im CZt[mu][r]).  On
// entry zt holds (Re P, Im P) and t (same layout) (Re Pocc, Im Pocc); CZt =
// Pocc o conj(P) (setup_CZt_half, src/interpolating_vectors.cpp:152-155).
__global__ __launch_bounds__(256) void half_complex_combine(double* zt, const double* t, long n, int k) {
  const long k2 = 2L * k, total = n * k;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += 256L * gridDim.x) {
    const long r = e / k, o = r * k2 + (e - r * k);
    const double rp = zt[o], ip = zt[o + k], ro = t[o], io = t[o + k];
    zt[o] = ro * rp + io * ip;
    zt[o + k] = io * rp - ro * ip;
  }
}

// The real 2k x 2k embedding [[Re C, -Im C], [Im C, Re C]] of the Hermitian
// CC^T = CZt[:, idx], from the stacked zt: Re C[i][j] = zt[idx[j]][i],
// Im C[i][j] = zt[idx[j]][k + i]; indices outside [0, n) give zeros (the
// caller reports them).
__global__ __launch_bounds__(256) void half_complex_cct(const double* zt, long n, const int* idx, int k,
                                                        double* c2) {
  const long k2 = 2L * k, total = k2 * k2;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += 256L * gridDim.x) {
    const int i = (int)(e / k2), j = (int)(e - (long)i * k2);
    const int bi = i >= k, bj = j >= k, ii = i - bi * k, jj = j - bj * k;
    const long q = idx[jj];
    double cr = 0.0, ci = 0.0;
    if (q >= 0 && q < n) {
      cr = zt[q * k2 + ii];
      ci = zt[q * k2 + k + ii];
    }
    c2[e] = bi == bj ? cr : (bi ? ci : -ci);
  }
}

// Muv of the half-rotated pass from the Gram matrix g (2k x 2k, full) of the
// weighted spectra of Re X (rows 0..k-1) and Im X (rows k..2k-1):
// Muv = G_rr - G_ii - i (G_ri + G_ir), written (re, im) interleaved.
__global__ __launch_bounds__(256) void half_complex_muv(const double* g, int k, double* muv) {
  const long k2 = 2L * k, total = (long)k * k;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += 256L * gridDim.x) {
    const long mu = e / k, nu = e - mu * k;
    muv[2 * e] = g[mu * k2 + nu] - g[(k + mu) * k2 + k + nu];
    muv[2 * e + 1] = -(g[mu * k2 + k + nu] + g[(k + mu) * k2 + nu]);
  }
}

// out[mu][r] = (x[mu][r], x[k + mu][r]) interleaved, x 2k x n
__global__ __launch_bounds__(256) void interleave_halves(const double* x, long n, int k, double* out) {
  const long total = (long)k * n, half = (long)k * n;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += 256L * gridDim.x) {
    out[2 * e] = x[e];
    out[2 * e + 1] = x[half + e];
  }
}

// out[mu][r] = (zt[r][mu], zt[r][k + mu]) interleaved, zt n x 2k
__global__ __launch_bounds__(256) void untranspose_stacked(const double* zt, long n, int k, double* out) {
  const long total = (long)k * n, k2 = 2L * k;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += 256L * gridDim.x) {
    const long mu = e / n, r = e - mu * n;
    out[2 * e] = zt[r * k2 + mu];
    out[2 * e + 1] = zt[r * k2 + k + mu];
  }
}

__global__ void repeat_idx(const int* idx, int k, int* idx2) {
  for (int j = blockIdx.x * 256 + threadIdx.x; j < 2 * k; j += 256 * gridDim.x) idx2[j] = idx[j < k ? j : j - k];
}

__global__ __launch_bounds__(256) void copy_f64(const double* x, double* y, long count) {
  for (long e = blockIdx.x * 256L + threadIdx.x; e < count; e += 256L * gridDim.x) y[e] = x[e];
}

__global__ void check_idx(const int* idx, int k, long n, int* bad) {
  for (int j = blockIdx.x * 256 + threadIdx.x; j < k; j += 256 * gridDim.x)
    if (idx[j] < 0 || idx[j] >= n) atomicOr(bad, 1);
}

inline int grid_for(long count) { return (int)std::min<long>(8192, std::max<long>(1, (count + 255) / 256)); }

// Stacked half-rotated ZC^T of complex orbitals (half_complex_combine) for m
// orbitals (aoR: n x 2m interleaved) and m_occ occupied ones (aoR_half: n x
// 2 m_occ).  P[mu][r] = sum_i conj(phi_i(r_mu)) phi_i(r): Re P is Phi against
// Phi over the 2m interleaved columns, Im P is Phi against i Phi (w = i Phi,
// wo = i Phi_occ, filled here), rows of the second operand gathered at the
// points.  t: n x 2k scratch.
int half_complex_zt(const double* aoR, const double* aoR_half, long n, int m, int m_occ, const int* idx, int k,
                    double* zt, double* t, double* w, double* wo, hipStream_t s) {
  hipLaunchKernelGGL(times_i, dim3(grid_for(n * m)), dim3(256), 0, s, aoR, n * m, w);
  hipLaunchKernelGGL(times_i, dim3(grid_for(n * m_occ)), dim3(256), 0, s, aoR_half, n * m_occ, wo);
  THC_HIP(hipGetLastError());
  const struct { const double* a; const double* b; int width; double* c; } prods[4] = {
      {aoR, aoR, 2 * m, zt}, {aoR, w, 2 * m, zt + k}, {aoR_half, aoR_half, 2 * m_occ, t}, {aoR_half, wo, 2 * m_occ, t + k}};
  for (const auto& q : prods) {
    GemmArgs g = {};
    g.m = n; g.n = k; g.k = q.width;
    g.a = q.a; g.lda = q.width; g.a_trans = 0;
    g.b = q.b; g.ldb = q.width; g.b_trans = 1; g.b_rows = idx; g.b_nrows = n;
    g.c = q.c; g.ldc = 2L * k; g.c_trans = 0;
    g.alpha = 1.0; g.beta = 0.0; g.batch = 1;
    THC_TRY(gemm_f64(g, s));
  }
  hipLaunchKernelGGL(half_complex_combine, dim3(grid_for(n * k)), dim3(256), 0, s, zt, t, n, k);
  THC_HIP(hipGetLastError());
  return THC_OK;
}

// deterministic operands of the GEMM timing hook: values in [-0.5, 0.5),
// gathered rows spread over [0, rows)
__global__ void fill_pattern(double* a, long count) {
  for (long e = blockIdx.x * 256L + threadIdx.x; e < count; e += 256L * gridDim.x)
    a[e] = (double)((unsigned long)(e * 2654435761UL) % 1000UL) * 1e-3 - 0.5;
}
__global__ void fill_spread(int* idx, int count, long rows) {
  for (int e = blockIdx.x * 256 + threadIdx.x; e < count; e += 256 * gridDim.x)
    idx[e] = (int)(((long)e * rows) / count);
}

struct IsdfBufs {
  double *z, *x, *c, *l, *ci, *r, *v;
  double *phim, *phio;  // Phi rows at the points (k x m), occupied set (k x m_occ)
  double* w;            // i Phi for complex orbitals (n x 2m), else null
  double* wo;           // i Phi_occ for complex half-rotated runs, else null
  int* idx2;            // the points twice (columns of the 2k-row real embedding), else null
  void* b;      // weighted half spectrum (f64 or f32)
  void *chol_w, *inv_w, *fft_w, *syrk_w;
  void* solve_w;        // digit planes of the int8 solve (mixed precision)
  void* rank_w;         // tridiagonal reduction of CC^T (rank diagnostic)
  int* info;
};

// m is the real row width of Phi (2 x orbitals for interleaved complex ones),
// m_occ that of the occupied set; k the rows of the real system solved (2 N_mu
// for complex half-rotated runs, ch)
size_t isdf_bytes(long n, int k, int m, int m_occ, bool cplx, bool ch, const int fft[3], int prec, int refine,
                  bool own_x, bool rank) {
  const long g2 = half_spectrum_len(fft);
  const size_t kk = (size_t)k * k;
  size_t b = carve_bytes<double>((size_t)k * m) + carve_bytes<double>((size_t)k * std::max(m_occ, 1));
  if (cplx) b += carve_bytes<double>((size_t)n * m);
  if (ch) b += carve_bytes<double>((size_t)n * m_occ) + carve_bytes<int>((size_t)k);
  b += carve_bytes<double>((size_t)k * n);                   // z
  if (own_x) b += carve_bytes<double>((size_t)k * n);        // x
  b += 4 * carve_bytes<double>(kk);                          // c, l, ci, v
  if (refine > 0) b += carve_bytes<double>((size_t)k * n);   // r
  // the weighted spectrum (the mixed path's FFT writes V's digit planes instead)
  if (prec == 0) b += carve_bytes<double>((size_t)k * g2);
  else if (prec != THC_PREC_MIXED) b += carve_bytes<float>((size_t)k * g2);
  const int fprec = prec == 0 ? 0 : 1;  // FFT / spectrum precision
  b += chol_work_bytes(k) + 256 + spd_inverse_work_bytes(k) + 256 + fft_work_bytes(k, fft, fprec) + 256 +
       (prec == THC_PREC_MIXED ? oz_lines_work_bytes(k, fft) : syrk_work_bytes(k, g2, fprec)) + 256;
  if (prec == THC_PREC_MIXED) b += oz_solve_work_bytes(k, n) + 256;
  if (rank) b += tridiag_work_bytes(k) + 256;
  b += carve_bytes<int>(16);
  return b;
}

bool isdf_carve(Arena& a, long n, int k, int m, int m_occ, bool cplx, bool ch, const int fft[3], int prec,
                int refine, double* xi, bool rank, IsdfBufs& B) {
  const long g2 = half_spectrum_len(fft);
  const size_t kk = (size_t)k * k;
  B = IsdfBufs{};
  B.phim = a.take<double>((size_t)k * m);
  B.phio = a.take<double>((size_t)k * std::max(m_occ, 1));
  B.w = cplx ? a.take<double>((size_t)n * m) : nullptr;
  if (cplx && !B.w) return false;
  B.wo = ch ? a.take<double>((size_t)n * m_occ) : nullptr;
  B.idx2 = ch ? a.take<int>((size_t)k) : nullptr;
  if (ch && (!B.wo || !B.idx2)) return false;
  B.z = a.take<double>((size_t)k * n);
  B.x = xi ? xi : a.take<double>((size_t)k * n);
  B.c = a.take<double>(kk);
  B.l = a.take<double>(kk);
  B.ci = a.take<double>(kk);
  B.v = a.take<double>(kk);
  B.r = refine > 0 ? a.take<double>((size_t)k * n) : nullptr;
  if (prec == 0) B.b = a.take<double>((size_t)k * g2);
  else if (prec != THC_PREC_MIXED) B.b = a.take<float>((size_t)k * g2);
  B.chol_w = a.take<char>(chol_work_bytes(k));
  B.inv_w = a.take<char>(spd_inverse_work_bytes(k));
  const int fprec = prec == 0 ? 0 : 1;
  B.fft_w = a.take<char>(fft_work_bytes(k, fft, fprec));
  B.syrk_w = a.take<char>(prec == THC_PREC_MIXED ? oz_lines_work_bytes(k, fft) : syrk_work_bytes(k, g2, fprec));
  B.solve_w = prec == THC_PREC_MIXED ? a.take<char>(oz_solve_work_bytes(k, n)) : nullptr;
  B.rank_w = rank ? a.take<char>(tridiag_work_bytes(k)) : nullptr;
  B.info = a.take<int>(16);
  return B.phim && B.phio && B.z && B.x && B.c && B.l && B.ci && B.v && (refine <= 0 || B.r) &&
         (prec == THC_PREC_MIXED || B.b) &&
         B.chol_w && B.inv_w && B.fft_w && B.syrk_w && (prec != THC_PREC_MIXED || B.solve_w) &&
         (!rank || B.rank_w) && B.info;
}

__global__ void zero_ints(int* p, int count) {
  if (threadIdx.x < count) p[threadIdx.x] = 0;
}

}  // namespace

// IVecs ctor + kernel at one rank.
// aoR_half (n x m_occ, nullable) selects the half-rotated ZC^T (Pocc o P).
// cplx: aoR holds m complex orbitals interleaved (n x 2m doubles); then
// ZC^T = |P|^2 = (Re P)^2 + (Im P)^2 with Re P from Phi against Phi and Im P
// from Phi against i Phi, both real GEMMs over the 2m interleaved columns.
// cplx with aoR_half (complex half-rotated, ch): ZC^T = Pocc o conj(P) and the
// Hermitian CC^T are complex; the solve runs on the real embedding (2k rows:
// Re X then Im X, against [[Re C, -Im C], [Im C, Re C]]), the FFT and V on
// those 2k real rows, and Muv is assembled from the 2k x 2k Gram matrix
// (half_complex_muv).  v_out and xi_out are then complex, interleaved.
int isdf_run(thc_ctx* ctx, const double* aoR, const double* aoR_half, long n, int m_orb, int m_occ, bool cplx,
             const int* idx, int k, const int fft[3], const double* coulG, const thc_isdf_opts* opts,
             double* v_out, double* luv_out, double* xi_out, thc_isdf_stats* stats) {
  if (!ctx || !aoR || !idx || !fft || !coulG || !v_out || n < 1 || m_orb < 1 || k < 1 || k > n) return THC_ERR_ARG;
  if (aoR_half && m_occ < 1) return THC_ERR_ARG;
  const bool ch = cplx && aoR_half;
  const int m = cplx ? 2 * m_orb : m_orb;           // real row width of Phi
  const int mo = cplx ? 2 * m_occ : m_occ;          // real row width of Phi_occ
  const int kr = ch ? 2 * k : k;                    // rows of the real system solved
  const int prec = opts ? opts->precision : THC_PREC_FP64;
  const int refine = opts ? std::max(0, opts->refine_steps) : 0;
  const int want_luv = opts ? opts->want_luv : 1;
  const bool want_rank = opts && opts->check_rank;
  if (prec != THC_PREC_FP64 && prec != THC_PREC_FP32 && prec != THC_PREC_MIXED) return THC_ERR_UNSUPPORTED;
  const long nfft = (long)fft[0] * fft[1] * fft[2];
  if (fft[0] < 1 || fft[1] < 1 || fft[2] < 1) return THC_ERR_ARG;
  // The reference only warns when prod(fft_grid) != N (interpolating_vectors.cpp:368-370) and then
  // transforms a prefix of each column (or runs past it); here that case is refused.
  if (nfft != n) {
    set_last_error("prod(fft_grid) differs from the number of grid points", __FILE__, __LINE__);
    return THC_ERR_ARG;
  }
  THC_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  if (want_rank && kr > THC_MAX_CLUSTERS) return THC_ERR_UNSUPPORTED;
  // int32 digit-group sums of the int8 solve are exact up to 4096 rows
  if (prec == THC_PREC_MIXED && kr > 4096) return THC_ERR_UNSUPPORTED;
  THC_TRY(ctx->ws.reserve(isdf_bytes(n, kr, m, mo, cplx, ch, fft, prec, refine, ch || xi_out == nullptr, want_rank)));
  ctx->ws.reset();
  IsdfBufs B;
  if (!isdf_carve(ctx->ws, n, kr, m, mo, cplx, ch, fft, prec, refine, ch ? nullptr : xi_out, want_rank, B))
    return THC_ERR_NO_MEMORY;
  if (want_rank) THC_TRY(ensure_pinned(ctx, sizeof(double) * 2 * (size_t)kr));
  const long g2 = half_spectrum_len(fft);
  hipEvent_t* ev = ctx->ev;

  hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, B.info, 16);
  hipLaunchKernelGGL(check_idx, dim3(grid_for(k)), dim3(256), 0, s, idx, k, n, B.info + 2);
  THC_HIP(hipEventRecord(ev[0], s));
  // 2 (on the second stream, beside stage 1). CC^T = ZC^T[:, idx] formed
  //    directly from Phi_mu: the GEMM below sees the same operand values in
  //    the same roles and k order as stage 1 does at those entries, so CC^T
  //    equals the gathered columns of ZC^T bit for bit.  Then the Cholesky
  //    factor and the explicit inverse.
  hipStream_t sa = ctx->aux;
  if (cplx && !ch) {
    hipLaunchKernelGGL(times_i, dim3(grid_for((long)n * m_orb)), dim3(256), 0, s, aoR, (long)n * m_orb, B.w);
    THC_HIP(hipGetLastError());
  }
  THC_HIP(hipEventRecord(ctx->fork, s));
  THC_HIP(hipStreamWaitEvent(sa, ctx->fork, 0));
  double* rk_d = reinterpret_cast<double*>(ctx->pinned);
  // The digit epilogue runs on the one-workgroup-per-CU 128 x 128 x 32 tile, the
  // tile the ZC^T GEMM takes anyway for inner dimensions up to 256; deeper
  // products (M > 256) run faster on the two-per-CU tile with a separate split.
  const bool fuse_split = prec == THC_PREC_MIXED && !aoR_half && !cplx && refine == 0 && m <= 256;
  if (ch) {
    // 1 + 2, complex half-rotated: the stacked ZC^T (the X buffer as scratch),
    // then CC^T's real embedding gathered from it, its Cholesky factor and
    // inverse, all on the main stream
    hipLaunchKernelGGL(repeat_idx, dim3(grid_for(2L * k)), dim3(256), 0, s, idx, k, B.idx2);
    THC_TRY(half_complex_zt(aoR, aoR_half, n, m_orb, m_occ, idx, k, B.z, B.x, B.w, B.wo, s));
    THC_HIP(hipEventRecord(ev[1], s));
    hipLaunchKernelGGL(half_complex_cct, dim3(grid_for((long)kr * kr)), dim3(256), 0, s, B.z, n, idx, k, B.c);
    hipLaunchKernelGGL(copy_f64, dim3(grid_for((long)kr * kr)), dim3(256), 0, s, B.c, B.l, (long)kr * kr);
    THC_HIP(hipGetLastError());
    THC_TRY(cholesky_f64(B.l, kr, kr, B.info, B.chol_w, s));
    THC_TRY(spd_inverse_from_chol(B.l, kr, kr, B.ci, B.inv_w, s));
    if (want_rank) THC_TRY(sym_tridiag(B.c, kr, kr, rk_d, rk_d + kr, B.rank_w, s));
    THC_HIP(hipEventRecord(ctx->join, sa));
  } else {
    hipLaunchKernelGGL(gather_rows, dim3(grid_for((long)k * m)), dim3(256), 0, sa, aoR, n, m, idx, k, B.phim);
    if (aoR_half)
      hipLaunchKernelGGL(gather_rows, dim3(grid_for((long)k * m_occ)), dim3(256), 0, sa, aoR_half, n, m_occ, idx, k,
                         B.phio);
    THC_HIP(hipGetLastError());
    {
      GemmArgs g = {};
      g.m = k; g.n = k; g.k = m;
      g.a = B.phim; g.lda = m; g.a_trans = 0;
      g.b = aoR; g.ldb = m; g.b_trans = 1; g.b_rows = idx; g.b_nrows = n;
      g.c = B.c; g.ldc = k; g.c_trans = 0;
      g.alpha = 1.0; g.beta = 0.0; g.epi_square = aoR_half ? 0 : 1; g.batch = 1;
      THC_TRY(gemm_f64(g, sa));
      if (aoR_half) {
        GemmArgs h = g;
        h.k = m_occ;
        h.a = B.phio; h.lda = m_occ;
        h.b = aoR_half; h.ldb = m_occ;
        h.epi_hadamard = 1;
        THC_TRY(gemm_f64(h, sa));
      }
      if (cplx) {
        GemmArgs h = g;
        h.b = B.w;
        h.epi_square = 0; h.epi_addsquare = 1;
        THC_TRY(gemm_f64(h, sa));
      }
    }
    hipLaunchKernelGGL(copy_f64, dim3(grid_for((long)k * k)), dim3(256), 0, sa, B.c, B.l, (long)k * k);
    THC_HIP(hipGetLastError());
    THC_TRY(cholesky_f64(B.l, k, k, B.info, B.chol_w, sa));
    THC_TRY(spd_inverse_from_chol(B.l, k, k, B.ci, B.inv_w, sa));
    THC_HIP(hipEventRecord(ctx->join, sa));
    // rank diagnostic (IVecs::check_rank) after the join: nothing on the main
    // stream waits for it; its tridiagonal form lands in pinned host memory
    if (want_rank) THC_TRY(sym_tridiag(B.c, k, k, rk_d, rk_d + k, B.rank_w, sa));
    // 1. Zt = (Phi Phi_mu^T)^2, N x K row-major (Zt[r][mu] = ZC^T[mu][r]): B rows
    //    gathered by idx, epilogue squares; this layout lets the solve read both
    //    operands along k and keeps the 4 GB store row-contiguous
    //    Half-rotated (setup_CZt_half): Zt = P o Pocc, P stored first, then the
    //    occupied-orbital product multiplied into it in the epilogue.
    //    Mixed path without refinement: nothing but the solve reads Zt, so the
    //    GEMM's epilogue writes the solve's int8 digit planes of Zt directly
    //    (the same values as splitting the FP64 Zt afterwards) and Zt itself is
    //    never stored.
    {
      GemmArgs g = {};
      g.m = n; g.n = k; g.k = m;
      g.a = aoR; g.lda = m; g.a_trans = 0;
      g.b = aoR; g.ldb = m; g.b_trans = 1; g.b_rows = idx; g.b_nrows = n;
      g.c = B.z; g.ldc = k; g.c_trans = 0;
      g.alpha = 1.0; g.beta = 0.0; g.epi_square = aoR_half ? 0 : 1; g.batch = 1;
      if (fuse_split) {
        THC_TRY(oz_solve_prepare(k, n, idx, aoR, m, B.solve_w, g, s));
        g.epi_square = 0;
      }
      THC_TRY(gemm_f64(g, s));
      if (aoR_half) {
        GemmArgs h = g;
        h.k = m_occ;
        h.a = aoR_half; h.lda = m_occ;
        h.b = aoR_half; h.ldb = m_occ;
        h.epi_hadamard = 1;
        THC_TRY(gemm_f64(h, s));
      }
      if (cplx) {  // Zt += (Im P)^2
        GemmArgs h = g;
        h.b = B.w;
        h.epi_square = 0; h.epi_addsquare = 1;
        THC_TRY(gemm_f64(h, s));
      }
    }
    THC_HIP(hipEventRecord(ev[1], s));
  }
  // join: the solve needs (CC^T)^-1
  THC_HIP(hipStreamWaitEvent(s, ctx->join, 0));
  THC_HIP(hipEventRecord(ev[2], s));
  // 3. X = CC^T^-1 ZC^T.  Below FP64 the FFT rounds X to FP32 as it reads it,
  // so when nothing else reads X (no xi output, no refinement) the solve
  // stores it as FP32 directly: the same values, half the bytes.
  const int x_f32 = prec != THC_PREC_FP64 && xi_out == nullptr && refine == 0;
  if (prec == THC_PREC_MIXED) {
    // exact int8 digit products of the scaled operands (ozaki.hip); the point
    // scales of the 2k embedded columns are those of idx, twice
    THC_TRY(oz_solve(B.ci, kr, B.z, kr, n, ch ? B.idx2 : idx, aoR, m, aoR_half, aoR_half ? mo : 0, B.x, x_f32, n,
                     B.solve_w, s, ev[8], ev[9], fuse_split ? 1 : 0));
  } else {
    GemmArgs g = {};
    g.m = kr; g.n = n; g.k = kr;
    g.a = B.ci; g.lda = kr; g.a_trans = 0;
    g.b = B.z; g.ldb = kr; g.b_trans = 1;
    g.c = B.x; g.ldc = n;
    g.alpha = 1.0; g.beta = 0.0; g.batch = 1;
    g.c_f32 = x_f32;
    THC_HIP(hipEventRecord(ev[8], s));
    THC_TRY(gemm_f64(g, s));
    THC_HIP(hipEventRecord(ev[9], s));
  }
  THC_HIP(hipEventRecord(ev[3], s));
  // 3b. FP64 iterative refinement in the transposed layout:
  //     Rt = Zt - X^T C^T;  X += C^-1 Rt^T
  for (int it = 0; it < refine; ++it) {
    hipLaunchKernelGGL(copy_f64, dim3(grid_for((long)kr * n)), dim3(256), 0, s, B.z, B.r, (long)kr * n);
    GemmArgs g = {};
    g.m = n; g.n = kr; g.k = kr;
    g.a = B.x; g.lda = n; g.a_trans = 1;
    g.b = B.c; g.ldb = kr; g.b_trans = 1;
    g.c = B.r; g.ldc = kr;
    g.alpha = -1.0; g.beta = 1.0; g.batch = 1;
    THC_TRY(gemm_f64(g, s));
    GemmArgs h = {};
    h.m = kr; h.n = n; h.k = kr;
    h.a = B.ci; h.lda = kr; h.a_trans = 0;
    h.b = B.r; h.ldb = kr; h.b_trans = 1;
    h.c = B.x; h.ldc = n;
    h.alpha = 1.0; h.beta = 1.0; h.batch = 1;
    THC_TRY(gemm_f64(h, s));
  }
  THC_HIP(hipEventRecord(ev[4], s));
  // 4. weighted half spectrum
  //    (mixed: pass B writes V's int8 digit planes, one exponent per y-line)
  FftDigits dig = {};
  if (prec == THC_PREC_MIXED) dig = oz_lines_layout(kr, fft, B.syrk_w);
  THC_TRY(fft_coulomb(B.x, x_f32, kr, fft, coulG, B.b, prec == THC_PREC_FP64 ? 0 : 1, B.fft_w, s, ev + 12,
                      prec == THC_PREC_MIXED ? &dig : nullptr));
  THC_HIP(hipEventRecord(ev[5], s));
  // 5. V = B B^T, FP64 sum of the split-K partials (exact int8 digit products on the mixed path);
  //    complex half-rotated: the 2k x 2k Gram matrix, then Muv from its blocks
  double* gram = ch ? B.v : v_out;
  if (prec == THC_PREC_MIXED) {
    THC_TRY(oz_syrk_lines(dig, kr, fft, gram, kr, B.syrk_w, s, ev[10], ev[11]));
  } else {
    THC_TRY(syrk_splitk(B.b, kr, g2, g2, prec == THC_PREC_FP64 ? 0 : 1, gram, kr, B.syrk_w, s, ev[10], ev[11]));
  }
  if (ch) {
    hipLaunchKernelGGL(half_complex_muv, dim3(grid_for((long)k * k)), dim3(256), 0, s, B.v, k, v_out);
    if (xi_out)
      hipLaunchKernelGGL(interleave_halves, dim3(grid_for((long)k * n)), dim3(256), 0, s, B.x, n, k, xi_out);
    THC_HIP(hipGetLastError());
  }
  THC_HIP(hipEventRecord(ev[6], s));
  // 6. Luv = chol(V)
  if (want_luv && luv_out && !ch) {
    hipLaunchKernelGGL(copy_f64, dim3(grid_for((long)k * k)), dim3(256), 0, s, v_out, luv_out, (long)k * k);
    THC_TRY(cholesky_f64(luv_out, k, k, B.info + 1, B.chol_w, s));
  }
  THC_HIP(hipEventRecord(ev[7], s));
  int info[4] = {0, 0, 0, 0};
  THC_HIP(hipMemcpyAsync(info, B.info, sizeof(info), hipMemcpyDeviceToHost, s));
  THC_HIP(hipStreamSynchronize(s));
  int cct_rank = -1;
  double cct_smax = 0.0;
  if (want_rank) {
    THC_HIP(hipStreamSynchronize(sa));
    tridiag_rank(rk_d, rk_d + kr, kr, &cct_rank, &cct_smax);
    if (ch) cct_rank /= 2;  // the embedding holds every singular value of CC^T twice
  }
  if (stats) {
    stats->cct_chol_info = info[0];
    stats->luv_chol_info = info[1];
    stats->cct_rank = cct_rank;
    stats->cct_smax = cct_smax;
    float t;
    const int map[6][2] = {{0, 1}, {1, 2}, {2, 3}, {4, 5}, {5, 6}, {6, 7}};
    for (int q = 0; q < 6; ++q) {
      THC_HIP(hipEventElapsedTime(&t, ev[map[q][0]], ev[map[q][1]]));
      stats->stage_ms[q] = t;
    }
    THC_HIP(hipEventElapsedTime(&t, ev[0], ev[7]));
    stats->stage_ms[6] = t;
    THC_HIP(hipEventElapsedTime(&t, ev[3], ev[4]));
    stats->stage_ms[7] = t;
    THC_HIP(hipEventElapsedTime(&t, ev[8], ev[9]));
    stats->stage_ms[8] = t;
    THC_HIP(hipEventElapsedTime(&t, ev[10], ev[11]));
    stats->stage_ms[9] = t;
    THC_HIP(hipEventElapsedTime(&t, ev[12], ev[13]));
    stats->stage_ms[10] = t;
    THC_HIP(hipEventElapsedTime(&t, ev[13], ev[14]));
    stats->stage_ms[11] = t;
  }
  if (info[2]) {
    set_last_error("an interpolating-point index lies outside [0, n)", __FILE__, __LINE__);
    return THC_ERR_ARG;
  }
  if (info[0]) return THC_ERR_NOT_POSDEF;
  if (info[1] && want_luv && luv_out) return THC_WARN_LUV_NOT_POSDEF;
  if (want_rank && cct_rank < k) return THC_WARN_RANK_DEFICIENT;
  return THC_OK;
}

}  // namespace thc

// ======================================================================== C ABI

using namespace thc;

extern "C" {

const char* thc_version(void) { return "thcpp_amd 0.1 (gfx950)"; }

const char* thc_last_error(void) { return g_last_error.c_str(); }

const char* thc_status_string(int st) {
  switch (st) {
    case THC_OK: return "ok";
    case THC_ERR_ARG: return "bad argument";
    case THC_ERR_HIP: return "HIP runtime error";
    case THC_ERR_EMPTY_CLUSTER: return "CRITICAL ERROR: Found centroid with weight ~ 0";
    case THC_ERR_NOT_POSDEF: return "CC^T is not numerically positive definite";
    case THC_ERR_NO_MEMORY: return "device workspace allocation failed";
    case THC_ERR_UNSUPPORTED: return "unsupported size";
    case THC_WARN_DUPLICATE_INDICES: return "WARNING: Found repeated interpolation indices";
    case THC_WARN_LUV_NOT_POSDEF: return "WARNING: Cholesky decomposition of Muv failed";
    case THC_WARN_RANK_DEFICIENT: return "WARNING: CCt matrix is rank deficient";
    default: return "unknown status";
  }
}

int thc_ctx_create(int device, void* stream, thc_ctx** out) {
  if (!out) return THC_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  THC_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return THC_ERR_ARG;
  THC_HIP(hipSetDevice(device));
  thc_ctx* c = new thc_ctx();
  c->device = device;
  if (stream) {
    c->stream = reinterpret_cast<hipStream_t>(stream);
  } else {
    // blocking: ordered after earlier work on the legacy default stream, where
    // hosts that pass no stream (torch among them) produce the inputs
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamDefault);
    if (e != hipSuccess) {
      delete c;
      set_last_error(hipGetErrorString(e), __FILE__, __LINE__);
      return THC_ERR_HIP;
    }
    c->own_stream = true;
  }
  for (auto& e : c->ev) {
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) {
      set_last_error(hipGetErrorString(r), __FILE__, __LINE__);
      thc_ctx_destroy(c);
      return THC_ERR_HIP;
    }
  }
  // set-up that thc_cholesky_noalloc relies on: the device address of its
  // static scratch on this device, and a small pinned buffer for its status
  int st = noalloc_scratch_init(device);
  if (st == THC_OK) st = ensure_pinned(c, 4096);
  if (st != THC_OK) {
    thc_ctx_destroy(c);
    return st;
  }
  hipError_t r = hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking);
  if (r == hipSuccess) r = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
  if (r == hipSuccess) r = hipEventCreateWithFlags(&c->join, hipEventDisableTiming);
  for (int q = 0; q < 3 && r == hipSuccess; ++q) r = hipEventCreateWithFlags(&c->kev[q], hipEventDisableTiming);
  if (r != hipSuccess) {
    set_last_error(hipGetErrorString(r), __FILE__, __LINE__);
    thc_ctx_destroy(c);
    return THC_ERR_HIP;
  }
  *out = c;
  return THC_OK;
}

int thc_ctx_destroy(thc_ctx* c) {
  if (!c) return THC_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->kmgraph) {
    c->kmgraph->reset();
    delete c->kmgraph;
  }
  c->ws.release();
  c->kmws.release();
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->aux) {
    (void)hipStreamSynchronize(c->aux);
    (void)hipStreamDestroy(c->aux);
  }
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->join) (void)hipEventDestroy(c->join);
  for (hipEvent_t e : c->kev)
    if (e) (void)hipEventDestroy(e);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  for (thc_ctx* l : c->lanes) thc_ctx_destroy(l);
  if (c->kaux) thc_ctx_destroy(c->kaux);
  delete c;
  return THC_OK;
}

int thc_ctx_set_graphs(thc_ctx* c, int enable) {
  if (!c) return THC_ERR_ARG;
  c->use_graphs = enable != 0;
  return THC_OK;
}

int thc_kmeans_initial_indices(long n, int k, unsigned int seed, int* host_idx_out) {
  if (thc_seed_indices_impl(n, k, seed, host_idx_out)) return THC_ERR_ARG;
  return THC_OK;
}

int thc_classify_grid_points(thc_ctx* ctx, const double* grid, long n, const double* centroids, int k,
                             int* grid_map) {
  return kmeans_classify(ctx, grid, n, centroids, k, grid_map);
}

int thc_update_centroids(thc_ctx* ctx, const double* rho, const double* grid, long n, const int* grid_map, int k,
                         double* centroids) {
  return kmeans_update(ctx, rho, grid, n, grid_map, k, centroids);
}

int thc_map_to_grid(thc_ctx* ctx, const double* grid, long n, const double* centroids, int k, int* interp_idx) {
  return kmeans_map_to_grid(ctx, grid, n, centroids, k, interp_idx);
}

int thc_kmeans(thc_ctx* ctx, const double* grid, const double* rho, long n, int k, int max_it, double threshold,
               unsigned int seed, const int* host_init_idx, int* interp_idx, int* grid_map, double* centroids,
               thc_kmeans_stats* stats) {
  if (!ctx || n < 1 || k < 1 || k > n) return THC_ERR_ARG;
  const int* init = host_init_idx;
  std::vector<int> tmp;
  if (!init) {
    tmp.resize(k);
    if (thc_seed_indices_impl(n, k, seed, tmp.data())) return THC_ERR_ARG;
    init = tmp.data();
  }
  return kmeans_run(ctx, grid, rho, n, k, max_it, threshold, init, interp_idx, grid_map, centroids, stats);
}

int thc_czt(thc_ctx* ctx, const double* aoR, long n, int m, const int* idx, int k, double* czt) {
  if (!ctx || !aoR || !idx || !czt || n < 1 || m < 1 || k < 1) return THC_ERR_ARG;
  THC_HIP(hipSetDevice(ctx->device));
  GemmArgs g = {};
  g.m = k; g.n = n; g.k = m;
  g.a = aoR; g.lda = m; g.a_trans = 0; g.a_rows = idx; g.a_nrows = n;
  g.b = aoR; g.ldb = m; g.b_trans = 1;
  g.c = czt; g.ldc = n;
  g.alpha = 1.0; g.beta = 0.0; g.epi_square = 1; g.batch = 1;
  THC_TRY(gemm_f64(g, ctx->stream));
  THC_HIP(hipStreamSynchronize(ctx->stream));
  return THC_OK;
}

int thc_czt_complex(thc_ctx* ctx, const double* aoR_ri, long n, int m, const int* idx, int k, double* czt) {
  if (!ctx || !aoR_ri || !idx || !czt || n < 1 || m < 1 || k < 1) return THC_ERR_ARG;
  THC_HIP(hipSetDevice(ctx->device));
  THC_TRY(ctx->ws.reserve(carve_bytes<double>((size_t)n * 2 * m)));
  ctx->ws.reset();
  double* w = ctx->ws.take<double>((size_t)n * 2 * m);
  if (!w) return THC_ERR_NO_MEMORY;
  hipLaunchKernelGGL(times_i, dim3(grid_for((long)n * m)), dim3(256), 0, ctx->stream, aoR_ri, (long)n * m, w);
  THC_HIP(hipGetLastError());
  GemmArgs g = {};
  g.m = k; g.n = n; g.k = 2 * m;
  g.a = aoR_ri; g.lda = 2 * m; g.a_trans = 0; g.a_rows = idx; g.a_nrows = n;
  g.b = aoR_ri; g.ldb = 2 * m; g.b_trans = 1;
  g.c = czt; g.ldc = n;
  g.alpha = 1.0; g.beta = 0.0; g.epi_square = 1; g.batch = 1;
  THC_TRY(gemm_f64(g, ctx->stream));
  GemmArgs h = g;
  h.b = w;
  h.epi_square = 0; h.epi_addsquare = 1;
  THC_TRY(gemm_f64(h, ctx->stream));
  THC_HIP(hipStreamSynchronize(ctx->stream));
  return THC_OK;
}

int thc_isdf_complex(thc_ctx* ctx, const double* aoR_ri, long n, int m, const int* idx, int k,
                     const int fft_grid[3], const double* coulG, const thc_isdf_opts* opts, double* v, double* luv,
                     double* xi, thc_isdf_stats* stats) {
  return isdf_run(ctx, aoR_ri, nullptr, n, m, 0, true, idx, k, fft_grid, coulG, opts, v, luv, xi, stats);
}

int thc_qrcp_points(thc_ctx* ctx, const double* aoR, const double* aoR_half, long n, int m, int m_occ, int k,
                    int* interp_idx, double* rdiag) {
  if (!ctx || !aoR || !interp_idx || n < 1 || m < 1 || k < 1 || k > n) return THC_ERR_ARG;
  if (aoR_half && m_occ < 1) return THC_ERR_ARG;
  THC_HIP(hipSetDevice(ctx->device));
  THC_TRY(ctx->ws.reserve(qrcp_work_bytes(n, k)));
  ctx->ws.reset();
  return qrcp_points(aoR, aoR_half, n, m, aoR_half ? m_occ : 0, k, interp_idx, rdiag, ctx->ws, ctx->stream);
}

int thc_qrcp_points_complex(thc_ctx* ctx, const double* aoR_ri, const double* aoR_half_ri, long n, int m, int m_occ,
                            int k, int* interp_idx, double* rdiag) {
  if (!ctx || !aoR_ri || !interp_idx || n < 1 || m < 1 || k < 1 || k > n) return THC_ERR_ARG;
  if (aoR_half_ri && m_occ < 1) return THC_ERR_ARG;
  THC_HIP(hipSetDevice(ctx->device));
  THC_TRY(ctx->ws.reserve(qrcp_work_bytes(n, k, true)));
  ctx->ws.reset();
  return qrcp_points(aoR_ri, aoR_half_ri, n, 2 * m, aoR_half_ri ? 2 * m_occ : 0, k, interp_idx, rdiag, ctx->ws,
                     ctx->stream, true);
}

int thc_czt_half(thc_ctx* ctx, const double* aoR, const double* aoR_half, long n, int m, int m_occ, const int* idx,
                 int k, double* czt) {
  if (!ctx || !aoR || !aoR_half || !idx || !czt || n < 1 || m < 1 || m_occ < 1 || k < 1) return THC_ERR_ARG;
  THC_HIP(hipSetDevice(ctx->device));
  GemmArgs g = {};
  g.m = k; g.n = n; g.k = m;
  g.a = aoR; g.lda = m; g.a_trans = 0; g.a_rows = idx; g.a_nrows = n;
  g.b = aoR; g.ldb = m; g.b_trans = 1;
  g.c = czt; g.ldc = n;
  g.alpha = 1.0; g.beta = 0.0; g.batch = 1;
  THC_TRY(gemm_f64(g, ctx->stream));
  GemmArgs h = g;
  h.k = m_occ;
  h.a = aoR_half; h.lda = m_occ;
  h.b = aoR_half; h.ldb = m_occ;
  h.epi_hadamard = 1;
  THC_TRY(gemm_f64(h, ctx->stream));
  THC_HIP(hipStreamSynchronize(ctx->stream));
  return THC_OK;
}

int thc_czt_half_complex(thc_ctx* ctx, const double* aoR_ri, const double* aoR_half_ri, long n, int m, int m_occ,
                         const int* idx, int k, double* czt_ri) {
  if (!ctx || !aoR_ri || !aoR_half_ri || !idx || !czt_ri || n < 1 || m < 1 || m_occ < 1 || k < 1 || k > n)
    return THC_ERR_ARG;
  THC_HIP(hipSetDevice(ctx->device));
  const size_t zb = (size_t)n * 2 * k;
  THC_TRY(ctx->ws.reserve(2 * carve_bytes<double>(zb) + carve_bytes<double>((size_t)n * 2 * m) +
                          carve_bytes<double>((size_t)n * 2 * m_occ)));
  ctx->ws.reset();
  double* zt = ctx->ws.take<double>(zb);
  double* t = ctx->ws.take<double>(zb);
  double* w = ctx->ws.take<double>((size_t)n * 2 * m);
  double* wo = ctx->ws.take<double>((size_t)n * 2 * m_occ);
  if (!zt || !t || !w || !wo) return THC_ERR_NO_MEMORY;
  THC_TRY(half_complex_zt(aoR_ri, aoR_half_ri, n, m, m_occ, idx, k, zt, t, w, wo, ctx->stream));
  hipLaunchKernelGGL(untranspose_stacked, dim3(grid_for((long)k * n)), dim3(256), 0, ctx->stream, zt, n, k, czt_ri);
  THC_HIP(hipGetLastError());
  THC_HIP(hipStreamSynchronize(ctx->stream));
  return THC_OK;
}

int thc_cct(thc_ctx* ctx, const double* czt, long n, const int* idx, int k, double* cct) {
  if (!ctx || !czt || !idx || !cct || n < 1 || k < 1) return THC_ERR_ARG;
  THC_HIP(hipSetDevice(ctx->device));
  hipLaunchKernelGGL(gather_cols, dim3(grid_for((long)k * k)), dim3(256), 0, ctx->stream, czt, n, idx, k, cct);
  THC_HIP(hipGetLastError());
  THC_HIP(hipStreamSynchronize(ctx->stream));
  return THC_OK;
}

int thc_cholesky(thc_ctx* ctx, double* a, int k, int* info) {
  if (!ctx || !a || k < 1) return THC_ERR_ARG;
  THC_HIP(hipSetDevice(ctx->device));
  THC_TRY(ctx->ws.reserve(chol_work_bytes(k) + carve_bytes<int>(4)));
  ctx->ws.reset();
  void* w = ctx->ws.take<char>(chol_work_bytes(k));
  int* d_info = ctx->ws.take<int>(4);
  if (!w || !d_info) return THC_ERR_NO_MEMORY;
  hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, ctx->stream, d_info, 4);
  THC_TRY(cholesky_f64(a, k, k, d_info, w, ctx->stream));
  int h = 0;
  THC_HIP(hipMemcpyAsync(&h, d_info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  THC_HIP(hipStreamSynchronize(ctx->stream));
  if (info) *info = h;
  return THC_OK;
}

int thc_gemm_tuning(const int* entries, int n) { return gemm_set_tuning(entries, n); }

int thc_gemm_time(thc_ctx* ctx, long m, long n, long k, int kind, int cfg, int reps, double* ms) {
  if (!ctx || !ms || m < 1 || n < 1 || k < 1 || (kind != 0 && kind != 1) || cfg < 0 || cfg >= GEMM_CFG_COUNT ||
      reps < 1 || reps > 1000)
    return THC_ERR_ARG;
  if (kind == 1 && (n > 0x7fffffffL || m > 0x7fffffffL)) return THC_ERR_ARG;
  THC_HIP(hipSetDevice(ctx->device));
  const long nb = kind == 0 ? n : 0;
  const size_t need = carve_bytes<double>((size_t)m * k) + carve_bytes<double>((size_t)std::max(nb, 1L) * k) +
                      carve_bytes<double>((size_t)m * n) + carve_bytes<int>((size_t)n) + 1024;
  THC_TRY(ctx->ws.reserve(need));
  ctx->ws.reset();
  double* a = ctx->ws.take<double>((size_t)m * k);
  double* b = ctx->ws.take<double>((size_t)std::max(nb, 1L) * k);
  double* c = ctx->ws.take<double>((size_t)m * n);
  int* idx = ctx->ws.take<int>((size_t)n);
  if (!a || !b || !c || !idx) return THC_ERR_NO_MEMORY;
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(fill_pattern, dim3(grid_for(m * k)), dim3(256), 0, s, a, m * k);
  if (kind == 0) hipLaunchKernelGGL(fill_pattern, dim3(grid_for(n * k)), dim3(256), 0, s, b, n * k);
  else hipLaunchKernelGGL(fill_spread, dim3(grid_for(n)), dim3(256), 0, s, idx, (int)n, m);
  THC_HIP(hipGetLastError());
  GemmArgs g = {};
  g.m = m; g.n = n; g.k = k;
  g.a = a; g.lda = k; g.a_trans = 0;
  if (kind == 0) {
    g.b = b; g.ldb = k; g.b_trans = 1;
  } else {  // the ZC^T GEMM: B rows gathered from A, square epilogue
    g.b = a; g.ldb = k; g.b_trans = 1; g.b_rows = idx; g.b_nrows = m;
    g.epi_square = 1;
  }
  g.c = c; g.ldc = n;
  g.alpha = 1.0; g.beta = 0.0; g.batch = 1;
  g.cfg = cfg;
  THC_TRY(gemm_f64(g, s));  // warm-up; refuses a configuration that does not apply
  THC_HIP(hipEventRecord(ctx->ev[0], s));
  for (int r = 0; r < reps; ++r) THC_TRY(gemm_f64(g, s));
  THC_HIP(hipEventRecord(ctx->ev[1], s));
  THC_HIP(hipStreamSynchronize(s));
  float t = 0;
  THC_HIP(hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]));
  *ms = t / reps;
  return THC_OK;
}

int thc_cct_rank(thc_ctx* ctx, const double* cct, int k, int* rank, double* smax) {
  if (!ctx || !cct || !rank || k < 1) return THC_ERR_ARG;
  if (k > THC_MAX_CLUSTERS) return THC_ERR_UNSUPPORTED;
  THC_HIP(hipSetDevice(ctx->device));
  THC_TRY(ctx->ws.reserve(tridiag_work_bytes(k) + 256));
  ctx->ws.reset();
  void* w = ctx->ws.take<char>(tridiag_work_bytes(k));
  if (!w) return THC_ERR_NO_MEMORY;
  THC_TRY(ensure_pinned(ctx, sizeof(double) * 2 * (size_t)k));
  double* d = reinterpret_cast<double*>(ctx->pinned);
  THC_TRY(sym_tridiag(cct, k, k, d, d + k, w, ctx->stream));
  THC_HIP(hipStreamSynchronize(ctx->stream));
  double s0 = 0.0;
  tridiag_rank(d, d + k, k, rank, &s0);
  if (smax) *smax = s0;
  return *rank < k ? THC_WARN_RANK_DEFICIENT : THC_OK;
}

int thc_least_squares(thc_ctx* ctx, const double* cct, const double* czt, long n, int k, int refine_steps,
                      double* x, int* info) {
  if (!ctx || !cct || !czt || !x || n < 1 || k < 1) return THC_ERR_ARG;
  THC_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const size_t kk = (size_t)k * k;
  size_t need = 2 * carve_bytes<double>(kk) + chol_work_bytes(k) + spd_inverse_work_bytes(k) + 1024 +
                carve_bytes<int>(4);
  if (refine_steps > 0) need += carve_bytes<double>((size_t)k * n);
  THC_TRY(ctx->ws.reserve(need));
  ctx->ws.reset();
  double* l = ctx->ws.take<double>(kk);
  double* ci = ctx->ws.take<double>(kk);
  void* cw = ctx->ws.take<char>(chol_work_bytes(k));
  void* iw = ctx->ws.take<char>(spd_inverse_work_bytes(k));
  int* d_info = ctx->ws.take<int>(4);
  double* r = refine_steps > 0 ? ctx->ws.take<double>((size_t)k * n) : nullptr;
  if (!l || !ci || !cw || !iw || !d_info || (refine_steps > 0 && !r)) return THC_ERR_NO_MEMORY;
  hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, d_info, 4);
  hipLaunchKernelGGL(copy_f64, dim3(grid_for((long)kk)), dim3(256), 0, s, cct, l, (long)kk);
  THC_TRY(cholesky_f64(l, k, k, d_info, cw, s));
  THC_TRY(spd_inverse_from_chol(l, k, k, ci, iw, s));
  GemmArgs g = {};
  g.m = k; g.n = n; g.k = k;
  g.a = ci; g.lda = k; g.b = czt; g.ldb = n; g.c = x; g.ldc = n;
  g.alpha = 1.0; g.beta = 0.0; g.batch = 1;
  THC_TRY(gemm_f64(g, s));
  for (int it = 0; it < refine_steps; ++it) {
    hipLaunchKernelGGL(copy_f64, dim3(grid_for((long)k * n)), dim3(256), 0, s, czt, r, (long)k * n);
    GemmArgs q = g;
    q.a = cct; q.b = x; q.c = r; q.alpha = -1.0; q.beta = 1.0;
    THC_TRY(gemm_f64(q, s));
    GemmArgs u = g;
    u.a = ci; u.b = r; u.c = x; u.alpha = 1.0; u.beta = 1.0;
    THC_TRY(gemm_f64(u, s));
  }
  int h = 0;
  THC_HIP(hipMemcpyAsync(&h, d_info, sizeof(int), hipMemcpyDeviceToHost, s));
  THC_HIP(hipStreamSynchronize(s));
  if (info) *info = h;
  return h ? THC_ERR_NOT_POSDEF : THC_OK;
}

int thc_fft_vectors(thc_ctx* ctx, const double* x, int k, const int fft_grid[3], const double* coulG,
                    double* out) {
  if (!ctx || !x || !fft_grid || !coulG || !out || k < 1) return THC_ERR_ARG;
  THC_HIP(hipSetDevice(ctx->device));
  THC_TRY(ctx->ws.reserve(fft_work_bytes(k, fft_grid, 0) + 256));
  ctx->ws.reset();
  void* w = ctx->ws.take<char>(fft_work_bytes(k, fft_grid, 0));
  if (!w) return THC_ERR_NO_MEMORY;
  THC_TRY(fft_coulomb(x, 0, k, fft_grid, coulG, out, 0, w, ctx->stream));
  THC_HIP(hipStreamSynchronize(ctx->stream));
  return THC_OK;
}

int thc_construct_muv(thc_ctx* ctx, const double* b, int k, long g2, double* v) {
  if (!ctx || !b || !v || k < 1 || g2 < 1) return THC_ERR_ARG;
  THC_HIP(hipSetDevice(ctx->device));
  THC_TRY(ctx->ws.reserve(syrk_work_bytes(k, g2, 0) + 256));
  ctx->ws.reset();
  void* w = ctx->ws.take<char>(syrk_work_bytes(k, g2, 0));
  if (!w) return THC_ERR_NO_MEMORY;
  THC_TRY(syrk_splitk(b, k, g2, g2, 0, v, k, w, ctx->stream));
  THC_HIP(hipStreamSynchronize(ctx->stream));
  return THC_OK;
}

int thc_isdf(thc_ctx* ctx, const double* aoR, long n, int m, const int* idx, int k, const int fft_grid[3],
             const double* coulG, const thc_isdf_opts* opts, double* v, double* luv, double* xi,
             thc_isdf_stats* stats) {
  return isdf_run(ctx, aoR, nullptr, n, m, 0, false, idx, k, fft_grid, coulG, opts, v, luv, xi, stats);
}

int thc_isdf_half(thc_ctx* ctx, const double* aoR, const double* aoR_half, long n, int m, int m_occ,
                  const int* idx, int k, const int fft_grid[3], const double* coulG, const thc_isdf_opts* opts,
                  double* muv, double* xi, thc_isdf_stats* stats) {
  if (!aoR_half) return THC_ERR_ARG;
  return isdf_run(ctx, aoR, aoR_half, n, m, m_occ, false, idx, k, fft_grid, coulG, opts, muv, nullptr, xi, stats);
}

int thc_isdf_half_complex(thc_ctx* ctx, const double* aoR_ri, const double* aoR_half_ri, long n, int m, int m_occ,
                          const int* idx, int k, const int fft_grid[3], const double* coulG,
                          const thc_isdf_opts* opts, double* muv_ri, double* xi_ri, thc_isdf_stats* stats) {
  if (!aoR_half_ri) return THC_ERR_ARG;
  return isdf_run(ctx, aoR_ri, aoR_half_ri, n, m, m_occ, true, idx, k, fft_grid, coulG, opts, muv_ri, nullptr, xi_ri,
                  stats);
}

int thc_factorize(thc_ctx* ctx, const double* grid, const double* rho, const double* aoR, long n, int m,
                  int thc_cfac, int max_it, double threshold, unsigned int seed, const int fft_grid[3],
                  const double* coulG, const thc_isdf_opts* opts, int* interp_idx, double* v, double* luv,
                  thc_kmeans_stats* kstats, thc_isdf_stats* istats) {
  if (!ctx || thc_cfac < 1 || m < 1) return THC_ERR_ARG;
  const long k = (long)thc_cfac * m;
  if (k > n || k > 0x7fffffffL) return THC_ERR_ARG;
  const int st = thc_kmeans(ctx, grid, rho, n, (int)k, max_it, threshold, seed, nullptr, interp_idx, nullptr,
                            nullptr, kstats);
  if (st != THC_OK && st < 100) return st;
  const int st2 = isdf_run(ctx, aoR, nullptr, n, m, 0, false, interp_idx, (int)k, fft_grid, coulG, opts, v, luv,
                           nullptr, istats);
  if (st2 != THC_OK) return st2;
  return st;
}

int thc_factorize_batched(thc_ctx* ctx, int nsys, thc_system* sys, const thc_isdf_opts* opts, int max_streams) {
  if (!ctx || nsys < 0 || (nsys > 0 && !sys)) return THC_ERR_ARG;
  if (nsys == 0) return THC_OK;
  THC_HIP(hipSetDevice(ctx->device));
  std::vector<KmSpec> spec(nsys);
  std::vector<thc_kmeans_stats> kstats(nsys);
  std::vector<double> cost(nsys, 0.0);
  for (int i = 0; i < nsys; ++i) {
    const thc_system& q = sys[i];
    const long k = (long)q.thc_cfac * q.m;
    spec[i] = KmSpec{q.grid, q.rho, q.n, (int)std::min<long>(k, 0x7fffffffL), q.max_it, q.threshold, q.seed,
                     q.interp_idx};
    if (q.thc_cfac < 1 || q.m < 1 || k > q.n || k > 0x7fffffffL) spec[i].k = -1;  // refused by the k-means
    kstats[i] = q.kstats;  // keeps the caller's diffs / pcts arrays
    if (spec[i].k > 0) cost[i] = (double)q.n * (double)k * (double)(k + q.m);
  }
  // The ISDF stages run on up to max_streams lanes (the context and up to 15
  // more it keeps), each driven by its own host thread.  The systems are
  // ordered largest first and a lane takes the next one whenever its last one
  // is done (list scheduling: the small systems, whose Cholesky and FFT chains
  // are latency-bound, fill in where the lanes would otherwise end unevenly).
  std::vector<int> order(nsys);
  for (int i = 0; i < nsys; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  const int nl = std::min(std::min(max_streams > 0 ? max_streams : 4, nsys), 16);
  while ((int)ctx->lanes.size() < nl - 1) {
    thc_ctx* c = nullptr;
    THC_TRY(thc_ctx_create(ctx->device, nullptr, &c));
    c->use_graphs = ctx->use_graphs;
    ctx->lanes.push_back(c);
  }
  // The k-means run as two batched runs (each launch covering one step of
  // every system of its run): the nl largest systems on the context's stream,
  // the rest on an auxiliary context from a thread of its own.  The lanes
  // start on the largest systems as soon as their run is done; a lane reaching
  // the rest waits for the second run, which has meanwhile gone on beside the
  // first systems' ISDF stages.  Each system's k-means is the same either way.
  const int na = nsys > nl ? nl : nsys;
  std::vector<int> ga(order.begin(), order.begin() + na), gb(order.begin() + na, order.end());
  auto run_km = [&](thc_ctx* c, const std::vector<int>& ids) -> int {
    std::vector<KmSpec> sp;
    std::vector<thc_kmeans_stats> ks;
    for (int i : ids) {
      sp.push_back(spec[i]);
      ks.push_back(kstats[i]);
    }
    std::vector<int> st(ids.size(), THC_OK);
    const int r = kmeans_run_many(c, (int)ids.size(), sp.data(), st.data(), ks.data());
    for (size_t t = 0; t < ids.size(); ++t) {
      sys[ids[t]].kstats = ks[t];
      sys[ids[t]].status = r == THC_OK ? st[t] : r;
    }
    return r;
  };
  if (!gb.empty() && !ctx->kaux) {
    THC_TRY(thc_ctx_create(ctx->device, nullptr, &ctx->kaux));
    ctx->kaux->use_graphs = ctx->use_graphs;
  }
  std::vector<std::string> errs(nsys);
  std::promise<int> b_done;
  std::shared_future<int> b_ready = b_done.get_future().share();
  auto run_b = [&] {
    const bool ok = hipSetDevice(ctx->device) == hipSuccess;
    const int r = ok ? run_km(ctx->kaux, gb) : THC_ERR_HIP;
    if (r != THC_OK)
      for (int i : gb) {
        sys[i].status = r;
        errs[i] = ok ? g_last_error : std::string("hipSetDevice failed for the second k-means run");
      }
    b_done.set_value(r);
  };
  std::thread kb;
  if (gb.empty()) {
    b_done.set_value(THC_OK);
  } else {
    try {
      kb = std::thread(run_b);
    } catch (...) {
      run_b();  // no thread available: the second run goes first, in this thread
    }
  }
  const int ra = run_km(ctx, ga);
  if (ra != THC_OK) {
    if (kb.joinable()) kb.join();
    return ra;
  }
  // the largest workspace any system of the batch needs: every lane reserves
  // it up front where it fits, so that no lane frees and reallocates its arena
  // in the middle of the batch (hipFree waits for every kernel on the device)
  size_t maxb = 0;
  const int prec = opts ? opts->precision : THC_PREC_FP64;
  const int refine = opts ? std::max(0, opts->refine_steps) : 0;
  const bool rank = opts && opts->check_rank;
  for (int i : order) {
    const thc_system& q = sys[i];
    const int k = spec[i].k;
    if (k < 1 || (prec == THC_PREC_MIXED && k > 4096)) continue;  // refused by the k-means or by isdf_run
    maxb = std::max(maxb, isdf_bytes(q.n, k, q.m, 0, false, false, q.fft_grid, prec, refine, true, rank));
  }
  std::atomic<size_t> next{0};
  auto work = [&](int lane) {
    thc_ctx* c = lane == 0 ? ctx : ctx->lanes[lane - 1];
    if (hipSetDevice(c->device) != hipSuccess) {
      (void)hipGetLastError();
      return;  // this lane takes no system; the others run them
    }
    // a reservation that does not fit leaves the lane to reserve per system
    if (c->ws.reserve(maxb) != THC_OK) (void)hipGetLastError();
    for (;;) {
      const size_t t = next.fetch_add(1);
      if (t >= order.size()) break;
      const int i = order[t];
      if ((int)t >= na) b_ready.wait();
      thc_system& q = sys[i];
      if (q.status != THC_OK && q.status < 100) continue;  // its k-means failed: no ISDF
      const int st2 = isdf_run(c, q.aoR, nullptr, q.n, q.m, 0, false, q.interp_idx, spec[i].k, q.fft_grid,
                               q.coulG, opts, q.v, q.luv, nullptr, &q.istats);
      if (st2 != THC_OK) q.status = st2;  // else keep the k-means status (OK or a warning)
      if (q.status != THC_OK && q.status < 100) errs[i] = g_last_error;
    }
  };
  std::vector<std::thread> threads;
  for (int lane = 1; lane < nl; ++lane) {
    try {
      threads.emplace_back(work, lane);
    } catch (...) {
      work(lane);  // no thread available: this lane takes systems from the list here
    }
  }
  work(0);
  for (auto& t : threads) t.join();
  if (kb.joinable()) kb.join();
  b_ready.wait();
  // systems no lane took (every lane failed to select the device)
  for (size_t t = next.load(); t < order.size(); ++t) {
    thc_system& q = sys[order[t]];
    if (q.status != THC_OK && q.status < 100) continue;
    q.status = THC_ERR_HIP;
    errs[order[t]] = "no batch worker could select the device";
  }
  int first_warn = THC_OK;
  for (int i = 0; i < nsys; ++i) {
    if (sys[i].status != THC_OK && sys[i].status < 100) {
      g_last_error = errs[i].empty() ? std::string("k-means of this system failed") : errs[i];
      return sys[i].status;
    }
    if (sys[i].status != THC_OK && first_warn == THC_OK) first_warn = sys[i].status;
  }
  return first_warn;
}

}  // extern "C"
