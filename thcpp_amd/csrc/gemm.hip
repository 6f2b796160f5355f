// LDS-tiled MFMA GEMM for gfx950 (FP64: v_mfma_f64_16x16x4f64, FP32:
// v_mfma_f32_16x16x4f32), 256 threads = 4 wave64 in a 2 x 2 layout.
//
// Operand fragments (checked on the device by tests/test_gpu_gemm.py):
//   A 16x4 block: lane l holds A[l % 16][l / 16]
//   B 4x16 block: lane l holds B[l / 16][l % 16]
//   D 16x16, 4 values per lane, column l % 16, row
//       f64: (l / 16) + 4 v        f32: 4 (l / 16) + v
// Tiles are staged k-major in LDS (As[k][m], Bs[k][n]) with one element of
// padding per row so the transposing stores do not all hit one bank.
#include "common.h"

namespace thc {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <class T> struct Mfma;
template <> struct Mfma<double> {
  typedef d4 acc_t;
  __device__ static inline acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct Mfma<float> {
  typedef f4 acc_t;
  __device__ static inline acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

enum { EPI_STORE = 0, EPI_SQUARE = 1, EPI_PARTIAL = 2 };

struct KArgs {
  long m, n, k;
  const void* a; long lda; int a_trans; const int* a_rows; long a_nrows;
  const void* b; long ldb; int b_trans;
  void* c; long ldc;
  double alpha, beta;
  int lower;
  long stride_a, stride_b, stride_c;
  long kchunk;  // split-K chunk (EPI_PARTIAL)
};

// Decode the tile handled by this workgroup.  Tiles are walked in groups of 8
// tile-rows so neighbouring workgroups share A rows, and workgroup ids are
// regrouped so each of the 8 XCDs (which receive ids round-robin) gets a
// contiguous run of tiles and therefore reuses operands in its own L2.
__device__ inline bool tile_coords(const KArgs& p, int BM, int BN, int& tm, int& tn) {
  const int nbm = (int)((p.m + BM - 1) / BM);
  const int nbn = (int)((p.n + BN - 1) / BN);
  int id = blockIdx.x;
  const int nb = gridDim.x;
  if ((nb & 7) == 0) id = (id & 7) * (nb >> 3) + (id >> 3);
  if (p.lower) {
    // triangular enumeration tn <= tm
    int t = (int)((sqrt(8.0 * id + 1.0) - 1.0) * 0.5);
    while ((long)(t + 1) * (t + 2) / 2 <= id) ++t;
    while ((long)t * (t + 1) / 2 > id) --t;
    tm = t;
    tn = id - t * (t + 1) / 2;
    return tm < nbm;
  }
  const int per_group = 8 * nbn;
  const int g = id / per_group;
  const int first = g * 8;
  const int gsize = min(nbm - first, 8);
  const int r = id % per_group;
  tm = first + r % gsize;
  tn = r / gsize;
  return tm < nbm && tn < nbn;
}

template <class T, int BM, int BN, int BK, int TM, int TN, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(KArgs p) {
  constexpr int WTM = TM * 16, WTN = TN * 16;
  static_assert(2 * WTM == BM && 2 * WTN == BN, "2x2 waves");
  constexpr int LDA_S = BM + 1, LDB_S = BN + 1;
  __shared__ T As[2][BK * LDA_S];
  __shared__ T Bs[2][BK * LDB_S];
  typedef typename Mfma<T>::acc_t acc_t;

  int tm, tn;
  if (!tile_coords(p, BM, BN, tm, tn)) return;
  const long m0 = (long)tm * BM, n0 = (long)tn * BN;
  const int bz = blockIdx.z;
  const T* A = reinterpret_cast<const T*>(p.a) + (long)bz * p.stride_a;
  const T* B = reinterpret_cast<const T*>(p.b) + (long)bz * p.stride_b;
  T* C = reinterpret_cast<T*>(p.c);

  long kbeg = 0, kend = p.k;
  if (EPI == EPI_PARTIAL) {
    kbeg = (long)blockIdx.y * p.kchunk;
    kend = min(p.k, kbeg + p.kchunk);
  }

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  constexpr int A_ELEMS = BM * BK / 256;
  constexpr int B_ELEMS = BN * BK / 256;
  T ra[A_ELEMS], rb[B_ELEMS];

  // Global -> registers for the k-tile starting at k0.
  auto load_tile = [&](long k0) {
#pragma unroll
    for (int e = 0; e < A_ELEMS; ++e) {
      const int lin = e * 256 + tid;
      int mm, kk;
      if (p.a_trans == 0) { kk = lin % BK; mm = lin / BK; }
      else { mm = lin % BM; kk = lin / BM; }
      const long gm = m0 + mm, gk = k0 + kk;
      T v = T(0);
      if (gm < p.m && gk < kend) {
        if (p.a_trans == 0) {
          if (p.a_rows) {
            const long row = p.a_rows[gm];
            if (row >= 0 && row < p.a_nrows) v = A[row * p.lda + gk];
          } else {
            v = A[gm * p.lda + gk];
          }
        } else {
          v = A[gk * p.lda + gm];
        }
      }
      ra[e] = v;
    }
#pragma unroll
    for (int e = 0; e < B_ELEMS; ++e) {
      const int lin = e * 256 + tid;
      int nn, kk;
      if (p.b_trans == 0) { nn = lin % BN; kk = lin / BN; }
      else { kk = lin % BK; nn = lin / BK; }
      const long gn = n0 + nn, gk = k0 + kk;
      T v = T(0);
      if (gn < p.n && gk < kend) {
        v = (p.b_trans == 0) ? B[gk * p.ldb + gn] : B[gn * p.ldb + gk];
      }
      rb[e] = v;
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int e = 0; e < A_ELEMS; ++e) {
      const int lin = e * 256 + tid;
      int mm, kk;
      if (p.a_trans == 0) { kk = lin % BK; mm = lin / BK; }
      else { mm = lin % BM; kk = lin / BM; }
      As[buf][kk * LDA_S + mm] = ra[e];
    }
#pragma unroll
    for (int e = 0; e < B_ELEMS; ++e) {
      const int lin = e * 256 + tid;
      int nn, kk;
      if (p.b_trans == 0) { nn = lin % BN; kk = lin / BN; }
      else { kk = lin % BK; nn = lin / BK; }
      Bs[buf][kk * LDB_S + nn] = rb[e];
    }
  };

  acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[i][j][v] = T(0);

  const long nk = (kend - kbeg + BK - 1) / BK;
  if (nk > 0) {
    load_tile(kbeg);
    store_tile(0);
    __syncthreads();
    for (long t = 0; t < nk; ++t) {
      const int buf = (int)(t & 1);
      if (t + 1 < nk) load_tile(kbeg + (t + 1) * BK);
#pragma unroll
      for (int k4 = 0; k4 < BK / 4; ++k4) {
        const int kr = k4 * 4 + (lane >> 4);
        T af[TM], bf[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) af[i] = As[buf][kr * LDA_S + wm * WTM + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < TN; ++j) bf[j] = Bs[buf][kr * LDB_S + wn * WTN + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<T>::op(af[i], bf[j], acc[i][j]);
      }
      if (t + 1 < nk) store_tile(buf ^ 1);
      __syncthreads();
    }
  }

  // Epilogue.
  if (EPI == EPI_PARTIAL) {
    // partial tile, tile-local row-major BM x BN, indexed [split][tile]
    const long tile_lin = (long)tm * (tm + 1) / 2 + tn;
    const long ntiles = gridDim.x;
    T* W = C + ((long)blockIdx.y * ntiles + tile_lin) * (BM * BN);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = wm * WTM + i * 16 + Mfma<T>::row(lane, v);
          const int cc = wn * WTN + j * 16 + (lane & 15);
          W[r * BN + cc] = acc[i][j][v];
        }
    return;
  }
  T* Cb = C + (long)bz * p.stride_c;
  const T alpha = (T)p.alpha, beta = (T)p.beta;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long r = m0 + wm * WTM + i * 16 + Mfma<T>::row(lane, v);
        const long cc = n0 + wn * WTN + j * 16 + (lane & 15);
        if (r < p.m && cc < p.n && (!p.lower || cc <= r)) {
          T* dst = Cb + r * p.ldc + cc;
          if (EPI == EPI_SQUARE) {
            const T x = alpha * acc[i][j][v];
            *dst = x * x;
          } else {
            T x = alpha * acc[i][j][v];
            if (beta != T(0)) x += beta * *dst;
            *dst = x;
          }
        }
      }
}

// Sum split-K partial tiles in split order (FP64) into the lower triangle of V
// and mirror it to the upper triangle.
template <class T, int BM, int BN>
__global__ __launch_bounds__(256) void syrk_reduce_kernel(const T* work, int nsplit, long ntiles, long m,
                                                          double* v, long ldv) {
  const long tile_lin = blockIdx.x;
  int tm = (int)((sqrt(8.0 * tile_lin + 1.0) - 1.0) * 0.5);
  while ((long)(tm + 1) * (tm + 2) / 2 <= tile_lin) ++tm;
  while ((long)tm * (tm + 1) / 2 > tile_lin) --tm;
  const int tn = (int)(tile_lin - (long)tm * (tm + 1) / 2);
  for (int e = threadIdx.x; e < BM * BN; e += 256) {
    const int r = e / BN, c = e % BN;
    const long gr = (long)tm * BM + r, gc = (long)tn * BN + c;
    if (gr >= m || gc >= m || gc > gr) continue;
    double s = 0.0;
    for (int sp = 0; sp < nsplit; ++sp) s += (double)work[((long)sp * ntiles + tile_lin) * (BM * BN) + e];
    v[gr * ldv + gc] = s;
    v[gc * ldv + gr] = s;
  }
}

template <class T, int BM, int BN, int BK, int TM, int TN>
int launch(const GemmArgs& g, hipStream_t s) {
  KArgs p;
  p.m = g.m; p.n = g.n; p.k = g.k;
  p.a = g.a; p.lda = g.lda; p.a_trans = g.a_trans; p.a_rows = g.a_rows; p.a_nrows = g.a_nrows;
  p.b = g.b; p.ldb = g.ldb; p.b_trans = g.b_trans;
  p.c = g.c; p.ldc = g.ldc;
  p.alpha = g.alpha; p.beta = g.beta;
  p.lower = g.lower;
  p.stride_a = g.stride_a; p.stride_b = g.stride_b; p.stride_c = g.stride_c;
  p.kchunk = 0;
  const long nbm = (g.m + BM - 1) / BM, nbn = (g.n + BN - 1) / BN;
  long nblk;
  if (g.lower) nblk = nbm * (nbm + 1) / 2;
  else nblk = nbm * nbn;
  if (nblk <= 0 || g.batch <= 0) return THC_OK;
  if (nblk > 0x7fffffffL) return THC_ERR_ARG;
  dim3 grid((unsigned)nblk, 1, (unsigned)g.batch);
  if (g.epi_square)
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, TM, TN, EPI_SQUARE>), grid, dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, TM, TN, EPI_STORE>), grid, dim3(256), 0, s, p);
  THC_HIP(hipGetLastError());
  return THC_OK;
}

template <class T>
int gemm_dispatch(const GemmArgs& g, hipStream_t s) {
  if (g.m <= 0 || g.n <= 0) return THC_OK;
  if (g.lower && g.m != g.n) return THC_ERR_ARG;
  // Small problems (factorisation panels, batched triangular blocks) use
  // 64 x 64 tiles so more workgroups exist.
  const long tiles128 = ((g.m + 127) / 128) * ((g.n + 127) / 128) * g.batch;
  if (tiles128 < 512) return launch<T, 64, 64, 16, 2, 2>(g, s);
  return launch<T, 128, 128, 16, 4, 4>(g, s);
}

constexpr int SY_BM = 128, SY_BN = 128, SY_BK = 16;

}  // namespace

int gemm_f64(const GemmArgs& g, hipStream_t s) { return gemm_dispatch<double>(g, s); }
int gemm_f32(const GemmArgs& g, hipStream_t s) { return gemm_dispatch<float>(g, s); }

static void syrk_plan(long m, long k, long& ntiles, int& nsplit, long& kchunk) {
  const long nb = (m + SY_BM - 1) / SY_BM;
  ntiles = nb * (nb + 1) / 2;
  // aim for >= 2048 workgroups; chunks are whole BK multiples
  long want = (2048 + ntiles - 1) / ntiles;
  long maxsplit = (k + 4 * SY_BK - 1) / (4 * SY_BK);
  if (want > maxsplit) want = maxsplit;
  if (want < 1) want = 1;
  kchunk = (k + want - 1) / want;
  kchunk = (kchunk + SY_BK - 1) / SY_BK * SY_BK;
  nsplit = (int)((k + kchunk - 1) / kchunk);
}

size_t syrk_work_bytes(long m, long k, int prec) {
  long ntiles, kchunk;
  int nsplit;
  syrk_plan(m, k, ntiles, nsplit, kchunk);
  return (size_t)nsplit * ntiles * SY_BM * SY_BN * (prec == 0 ? 8 : 4);
}

int syrk_splitk(const void* a, long m, long k, long lda, int prec, double* v, long ldv, void* work,
                hipStream_t s) {
  if (m <= 0) return THC_OK;
  long ntiles, kchunk;
  int nsplit;
  syrk_plan(m, k, ntiles, nsplit, kchunk);
  KArgs p = {};
  p.m = m; p.n = m; p.k = k;
  p.a = a; p.lda = lda; p.a_trans = 0; p.a_rows = nullptr;
  p.b = a; p.ldb = lda; p.b_trans = 1;
  p.c = work; p.ldc = 0;
  p.alpha = 1.0; p.beta = 0.0;
  p.lower = 1;
  p.kchunk = kchunk;
  dim3 grid((unsigned)ntiles, (unsigned)nsplit, 1);
  if (prec == 0) {
    hipLaunchKernelGGL((gemm_kernel<double, SY_BM, SY_BN, SY_BK, 4, 4, EPI_PARTIAL>), grid, dim3(256), 0, s, p);
    THC_HIP(hipGetLastError());
    hipLaunchKernelGGL((syrk_reduce_kernel<double, SY_BM, SY_BN>), dim3((unsigned)ntiles), dim3(256), 0, s,
                       (const double*)work, nsplit, ntiles, m, v, ldv);
  } else {
    hipLaunchKernelGGL((gemm_kernel<float, SY_BM, SY_BN, SY_BK, 4, 4, EPI_PARTIAL>), grid, dim3(256), 0, s, p);
    THC_HIP(hipGetLastError());
    hipLaunchKernelGGL((syrk_reduce_kernel<float, SY_BM, SY_BN>), dim3((unsigned)ntiles), dim3(256), 0, s,
                       (const float*)work, nsplit, ntiles, m, v, ldv);
  }
  THC_HIP(hipGetLastError());
  return THC_OK;
}

}  // namespace thc
