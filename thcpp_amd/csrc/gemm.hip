// LDS-tiled MFMA GEMM for gfx950 (FP64: v_mfma_f64_16x16x4f64, FP32:
// v_mfma_f32_16x16x4f32).  A workgroup is a WM x WN grid of wave64s; the
// large-matrix configuration is a 128 x 128 x 16 tile on 8 waves (4 x 2, each
// wave 32 x 64), which keeps a wave under 128 VGPRs so that four waves share
// every SIMD (a 2 x 2 grid of 64 x 64 wave tiles needs 340 VGPRs and runs one
// wave per SIMD at ~60% of the rate).
//
// Operand fragments (checked on the device by the GPU tests):
//   A 16x4 block: lane l holds A[l % 16][l / 16]
//   B 4x16 block: lane l holds B[l / 16][l % 16]
//   D 16x16, 4 values per lane, column l % 16, row
//       f64: (l / 16) + 4 v        f32: 4 (l / 16) + v
//
// Operand layouts are template parameters.  An operand that is contiguous
// along k (A row-major, or B given as n x k) is staged in LDS k-contiguous
// (row stride BK + 1), one that is contiguous along m / n is staged k-major
// (row stride BM + 1 / BN + 1); both make the 16-lane fragment reads hit 16
// distinct bank pairs.  Each thread fetches a run of 8 (or 4) contiguous
// elements per tile.  Interior tiles take an unpredicated path (vector loads
// when the operands are 16-byte aligned); only edge tiles test bounds.
// Rows gathered through an index list (Phi_mu, on either operand) are
// resolved once per tile.
#include <mutex>
#include <type_traits>
#include <vector>

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

enum { EPI_STORE = 0, EPI_SQUARE = 1, EPI_PARTIAL = 2, EPI_HADAMARD = 3, EPI_ADDSQUARE = 4, EPI_DIGITS = 5 };
constexpr int DIG_PLANES = 7;  // digits of the mixed solve's ZC^T operand (ozaki.hip)

struct KArgs {
  long m, n, k;
  const void* a; long lda; const int* a_rows; long a_nrows;
  const void* b; long ldb; const int* b_rows; long b_nrows;
  void* c; long ldc; int c_trans;
  double alpha, beta;
  int lower;
  int tri;
  int c_f32;
  long stride_a, stride_b, stride_c;
  long kchunk;  // split-K chunk (EPI_PARTIAL)
  signed char* dig_out;  // EPI_DIGITS (see GemmArgs)
  const double* dig_s;
  const double* dig_inv2;
  long dig_nr;
  int dig_kp;
};

__device__ inline bool tile_coords(const KArgs& p, int BM, int BN, int& tm, int& tn) {
  const int nbm = (int)((p.m + BM - 1) / BM);
  const int nbn = (int)((p.n + BN - 1) / BN);
  int id = blockIdx.x;
  const int nb = gridDim.x;
  // ids reach the 8 XCDs round-robin: give each XCD a contiguous run of tiles
  if ((nb & 7) == 0) id = (id & 7) * (nb >> 3) + (id >> 3);
  if (p.lower) {
    int t = (int)((sqrt(8.0 * id + 1.0) - 1.0) * 0.5);
    while ((long)(t + 1) * (t + 2) / 2 <= id) ++t;
    while ((long)t * (t + 1) / 2 > id) --t;
    tm = t;
    tn = id - t * (t + 1) / 2;
    return tm < nbm;
  }
  // walk groups of up to 4 tile rows column by column: the B column block a
  // group needs is read once and the group's A panels stay in cache (the 64
  // workgroups resident on one XCD then span 4 x 16 tiles; on the 2048 x
  // 262144 x 2048 solve 4-row groups beat 16-row groups by 1%, 8 by 0.9%)
  const int G = 4;
  const int per_group = G * nbn;
  const int g = id / per_group;
  const int first = g * G;
  const int gsize = min(nbm - first, G);
  const int r = id % per_group;
  tm = first + r % gsize;
  tn = r / gsize;
  return tm < nbm && tn < nbn;
}

// Per-thread share of one operand tile: RUN contiguous elements of one line.
//   KCONTIG: the operand is contiguous along k (lines are m or n indices)
//   otherwise contiguous along m / n (lines are k indices)
template <class T, int BMN, int BK, bool KCONTIG, int NT = 256>
struct TileMap {
  static constexpr int ELEMS = BMN * BK / NT;
  static constexpr int RUN = ELEMS;  // one run per thread
  __device__ static inline void coords(int tid, int& line, int& off) {
    if (KCONTIG) {  // BMN lines of BK
      constexpr int TPL = BK / RUN;
      line = tid / TPL;
      off = (tid % TPL) * RUN;
    } else {        // BK lines of BMN
      constexpr int TPL = BMN / RUN;
      line = tid / TPL;
      off = (tid % TPL) * RUN;
    }
  }
};

template <class T, int N>
__device__ inline void load_run(const T* __restrict__ src, T* dst, bool vec) {
  if (vec) {
    if (sizeof(T) == 8) {
      typedef double v2 __attribute__((ext_vector_type(2)));
      const v2* s = reinterpret_cast<const v2*>(src);
#pragma unroll
      for (int e = 0; e < N / 2; ++e) {
        v2 x = s[e];
        dst[2 * e] = x[0];
        dst[2 * e + 1] = x[1];
      }
    } else {
      typedef float v4 __attribute__((ext_vector_type(4)));
      const v4* s = reinterpret_cast<const v4*>(src);
#pragma unroll
      for (int e = 0; e < N / 4; ++e) {
        v4 x = s[e];
        dst[4 * e] = x[0];
        dst[4 * e + 1] = x[1];
        dst[4 * e + 2] = x[2];
        dst[4 * e + 3] = x[3];
      }
    }
  } else {
#pragma unroll
    for (int e = 0; e < N; ++e) dst[e] = src[e];
  }
}

template <class T, int BM, int BN, int BK, int TM, int TN, int WM, int WN, int EPI, bool AK, bool BK_, bool GA,
          bool GB, bool VEC>
__global__ __launch_bounds__(64 * WM * WN) void gemm_kernel(KArgs p) {
  // AK: A contiguous along k (A[m * lda + k]); else A[k * lda + m]
  // BK_: B contiguous along k (B[n * ldb + k]); else B[k * ldb + n]
  constexpr int WTM = TM * 16, WTN = TN * 16, NT = 64 * WM * WN;
  static_assert(WM * WTM == BM && WN * WTN == BN, "wave grid must cover the tile");
  constexpr int A_LINES = AK ? BM : BK, A_LEN = AK ? BK + 1 : BM + 1;
  constexpr int B_LINES = BK_ ? BN : BK, B_LEN = BK_ ? BK + 1 : BN + 1;
  __shared__ __attribute__((aligned(16))) T As[2][A_LINES * A_LEN];
  __shared__ __attribute__((aligned(16))) T Bs[2][B_LINES * B_LEN];
  typedef typename Mfma<T>::acc_t acc_t;
  typedef TileMap<T, BM, BK, AK, NT> MA;
  typedef TileMap<T, BN, BK, BK_, NT> MB;
  constexpr int RA = MA::RUN, RB = MB::RUN;
  static_assert(MA::ELEMS * NT == BM * BK && MB::ELEMS * NT == BN * BK, "tile must split evenly over threads");
  static_assert(!VEC || (RA * sizeof(T) % 16 == 0 && RB * sizeof(T) % 16 == 0), "vector runs are 16-byte multiples");

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
  if (p.tri) {
    if (p.tri & GEMM_TRI_A_KGE_M) kbeg = max(kbeg, m0);
    if (p.tri & GEMM_TRI_B_KGE_N) kbeg = max(kbeg, n0);
    if (p.tri & GEMM_TRI_A_KLE_M) kend = min(kend, m0 + BM);
    kbeg = kbeg / BK * BK;
  }

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;

  int a_line, a_off, b_line, b_off;
  MA::coords(tid, a_line, a_off);
  MB::coords(tid, b_line, b_off);

  // A: fixed line pointer (row m for AK, resolved through the gather list)
  const T* a_ptr = nullptr;
  bool a_row_ok = true;
  if (AK) {
    const long gm = m0 + a_line;
    a_row_ok = gm < p.m;
    long row = gm;
    if (GA) {
      row = a_row_ok ? (long)p.a_rows[gm] : 0;
      a_row_ok = a_row_ok && row >= 0 && row < p.a_nrows;
    }
    a_ptr = A + (a_row_ok ? row : 0) * p.lda + a_off;
  } else {
    a_ptr = A + (long)a_line * p.lda + m0 + a_off;
  }
  const T* b_ptr;
  bool b_row_ok = true;
  if (BK_) {
    const long gn = n0 + b_line;
    b_row_ok = gn < p.n;
    long row = gn;
    if (GB) {
      row = b_row_ok ? (long)p.b_rows[gn] : 0;
      b_row_ok = b_row_ok && row >= 0 && row < p.b_nrows;
    }
    b_ptr = B + (b_row_ok ? row : 0) * p.ldb + b_off;
  } else {
    b_ptr = B + (long)b_line * p.ldb + n0 + b_off;
  }
  // operand rows outside the matrix read a clamped valid row and are zeroed
  const bool a_mn_full = AK || (m0 + BM <= p.m);
  const bool b_mn_full = BK_ || (n0 + BN <= p.n);

  T ra[RA], rb[RB];

  auto load_tile = [&](long k0) {
    const bool full_k = k0 + BK <= kend;
    if (full_k && a_mn_full) {
      if (AK) {
        load_run<T, RA>(a_ptr + k0, ra, VEC);  // invalid rows are zeroed in store_tile
      } else {
        load_run<T, RA>(a_ptr + k0 * p.lda, ra, VEC);
      }
    } else {
#pragma unroll
      for (int e = 0; e < RA; ++e) {
        bool ok;
        long off;
        if (AK) { ok = a_row_ok && (k0 + a_off + e < kend); off = k0 + e; }
        else { ok = (k0 + a_line < kend) && (m0 + a_off + e < p.m); off = k0 * p.lda + e; }
        ra[e] = ok ? a_ptr[off] : T(0);
      }
    }
    if (full_k && b_mn_full) {
      if (BK_) {
        load_run<T, RB>(b_ptr + k0, rb, VEC);
      } else {
        load_run<T, RB>(b_ptr + k0 * p.ldb, rb, VEC);
      }
    } else {
#pragma unroll
      for (int e = 0; e < RB; ++e) {
        bool ok;
        long off;
        if (BK_) { ok = b_row_ok && (k0 + b_off + e < kend); off = k0 + e; }
        else { ok = (k0 + b_line < kend) && (n0 + b_off + e < p.n); off = k0 * p.ldb + e; }
        rb[e] = ok ? b_ptr[off] : T(0);
      }
    }
  };
  auto store_tile = [&](int buf) {
    T* as = As[buf] + a_line * A_LEN + a_off;
    T* bs = Bs[buf] + b_line * B_LEN + b_off;
#pragma unroll
    for (int e = 0; e < RA; ++e) as[e] = a_row_ok ? ra[e] : T(0);
#pragma unroll
    for (int e = 0; e < RB; ++e) bs[e] = b_row_ok ? rb[e] : T(0);
  };

  acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[i][j][v] = T(0);

  const int l16 = lane & 15, l4 = lane >> 4;
  const long nk = (kend - kbeg + BK - 1) / BK;
  if (nk > 0) {
    load_tile(kbeg);
    store_tile(0);
    __syncthreads();
    for (long t = 0; t < nk; ++t) {
      const int buf = (int)(t & 1);
      if (t + 1 < nk) load_tile(kbeg + (t + 1) * BK);
      const T* as = As[buf];
      const T* bs = Bs[buf];
#pragma unroll
      for (int k4 = 0; k4 < BK / 4; ++k4) {
        const int kr = k4 * 4 + l4;
        T af[TM], bf[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          const int mi = wm * WTM + i * 16 + l16;
          af[i] = AK ? as[mi * A_LEN + kr] : as[kr * A_LEN + mi];
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const int nj = wn * WTN + j * 16 + l16;
          bf[j] = BK_ ? bs[nj * B_LEN + kr] : bs[kr * B_LEN + nj];
        }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<T>::op(af[i], bf[j], acc[i][j]);
      }
      if (t + 1 < nk) store_tile(buf ^ 1);
      __syncthreads();
    }
  }

  if constexpr (EPI == EPI_DIGITS) {
    // x = (acc^2 * inv2[col]) / s[row] cut into DIG_PLANES digits (x *= 128,
    // d = trunc(x), x -= d: exact in FP64), the same arithmetic as the
    // standalone split of ZC^T.  All planes of the tile are staged in the
    // free operand LDS (planes 0-3 in As, 4-6 in Bs) as [32-column
    // block][row][32 bytes], then, after one barrier, stored as whole 16-byte
    // pieces of the k-blocked planes: a 4 KB contiguous run per block.
    static_assert(sizeof(T) == 8 && BM == 128 && BN == 128 && DIG_PLANES == 7, "digit epilogue: FP64 128 x 128");
    static_assert(sizeof(As) >= 4 * BM * BN && sizeof(Bs) >= 3 * BM * BN, "seven staged planes fit the operand LDS");
    signed char* stg_a = reinterpret_cast<signed char*>(&As[0][0]);
    signed char* stg_b = reinterpret_cast<signed char*>(&Bs[0][0]);
    double x[TM][TN][4];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const long r = m0 + wm * WTM + i * 16 + Mfma<T>::row(lane, v);
          const long cc = n0 + wn * WTN + j * 16 + l16;
          double xv = 0.0;
          if (r < p.m && cc < p.n) {
            const double sr = p.dig_s[r];
            const double isr = sr > 0.0 ? 1.0 / sr : 0.0;
            const double a = p.alpha * acc[i][j][v];
            xv = a * a * p.dig_inv2[cc] * isr;
          }
          x[i][j][v] = xv;
        }
    __syncthreads();  // every wave is past its last fragment read of As / Bs
#pragma unroll
    for (int q = 0; q < DIG_PLANES; ++q) {
      signed char* stg = q < 4 ? stg_a + q * (BM * BN) : stg_b + (q - 4) * (BM * BN);
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int rl = wm * WTM + i * 16 + Mfma<T>::row(lane, v);
            const int cl = wn * WTN + j * 16 + l16;
            x[i][j][v] *= 128.0;
            const double d = trunc(x[i][j][v]);
            x[i][j][v] -= d;
            stg[(cl >> 5) * (BM * 32) + rl * 32 + (cl & 31)] = (signed char)(int)d;
          }
    }
    __syncthreads();
    const long plane = p.dig_nr * (long)p.dig_kp;
#pragma unroll
    for (int q = 0; q < DIG_PLANES; ++q) {
      const signed char* stg = q < 4 ? stg_a + q * (BM * BN) : stg_b + (q - 4) * (BM * BN);
      for (int t = tid; t < BM * BN / 16; t += NT) {
        const int cb = t / (BM * 2), rem = t % (BM * 2), rl = rem >> 1, h = rem & 1;
        const long gr = m0 + rl, gc = n0 + cb * 32;
        if (gr < p.dig_nr && gc < p.dig_kp)
          *reinterpret_cast<int4*>(p.dig_out + q * plane + (gc >> 5) * p.dig_nr * 32 + gr * 32 + h * 16) =
              *reinterpret_cast<const int4*>(stg + cb * (BM * 32) + rl * 32 + h * 16);
      }
    }
    return;
  }
  if (EPI == EPI_PARTIAL) {
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
          const int cc = wn * WTN + j * 16 + l16;
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
        const long cc = n0 + wn * WTN + j * 16 + l16;
        if (r < p.m && cc < p.n && (!p.lower || cc <= r)) {
          const long off = p.c_trans ? cc * p.ldc + r : r * p.ldc + cc;
          T* dst = Cb + off;
          if (EPI == EPI_STORE && p.c_f32) {
            reinterpret_cast<float*>(p.c)[(long)bz * p.stride_c + off] = (float)(alpha * acc[i][j][v]);
          } else if (EPI == EPI_SQUARE) {
            const T x = alpha * acc[i][j][v];
            *dst = x * x;
          } else if (EPI == EPI_HADAMARD) {
            const T x = alpha * acc[i][j][v];
            *dst = x * *dst;
          } else if (EPI == EPI_ADDSQUARE) {
            const T x = alpha * acc[i][j][v];
            *dst = *dst + x * x;
          } else {
            T x = alpha * acc[i][j][v];
            if (beta != T(0)) x += beta * *dst;
            *dst = x;
          }
        }
      }
}

// FP64 128 x 128 x 16 tile for A (m x k) and B (n x k) both contiguous along
// k, k a multiple of 16: the tiles reach LDS by direct global-to-LDS loads
// (global_load_lds_dwordx4: lane L of a wave writes bytes 16 L .. 16 L + 15 of
// the wave's 1 KB block), so no registers stage them and no LDS stores wait on
// the global loads.  An LDS tile row is 16 doubles (128 bytes, eight 16-byte
// chunks, no padding); logical chunk c of row r sits at chunk c ^ ((r >> 1) & 7),
// which keeps each 16-row fragment read free of bank conflicts.  The MFMA k
// order is permuted so one 16-byte read feeds two k-steps: in half h, k-step s,
// lane group l4 contributes k = 8 h + 2 l4 + s of both operands.  Rows past m
// (n) read row m - 1 (n - 1); they only feed outputs that are never stored.
// PARTIAL: split-K partial tiles of the lower triangle (blockIdx.y = split,
// k range [y kchunk, min(k, (y + 1) kchunk))), stored whole as in gemm_kernel.
typedef __attribute__((address_space(3))) void lds_void;
constexpr int DTL_BM = 128, DTL_BN = 128, DTL_BK = 16, DTL_WN = 2, DTL_THREADS = 512;

// one lane's 16 bytes from global memory into LDS (a plain function: with
// template-dependent operands the builtin is not instantiated for the host
// pass, and the kernel's launch stub goes missing)
__device__ inline void load_lds16(const double* src, double* dst) {
  __builtin_amdgcn_global_load_lds(src, (lds_void*)dst, 16, 0, 0);
}

template <bool PARTIAL>
__global__ __launch_bounds__(DTL_THREADS) void gemm_dtl_kernel(KArgs p) {
  constexpr int BM = DTL_BM, BN = DTL_BN, BK = DTL_BK, TM = 2, TN = 4, WN = DTL_WN;
  constexpr int WTM = TM * 16, WTN = TN * 16;
  static_assert((BM / 8) % (DTL_THREADS / 64) == 0 && BM == BN, "8-row blocks must split evenly over the waves");
  // one variable per buffer: the compiler then knows that the direct loads
  // into one buffer do not alias the fragment reads of the other, and does not
  // wait for the loads before reading
  __shared__ __attribute__((aligned(16))) double As0[BM * BK], As1[BM * BK];
  __shared__ __attribute__((aligned(16))) double Bs0[BN * BK], Bs1[BN * BK];
  int tm, tn;
  if (!tile_coords(p, BM, BN, tm, tn)) return;
  const long m0 = (long)tm * BM, n0 = (long)tn * BN;
  const int bz = blockIdx.z;
  const double* A = reinterpret_cast<const double*>(p.a) + (long)bz * p.stride_a;
  const double* B = reinterpret_cast<const double*>(p.b) + (long)bz * p.stride_b;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int l16 = lane & 15, l4 = lane >> 4;

  // wave w loads 8-row blocks w and w + 8 of each operand; lane L fills row
  // 8 b + L / 8, physical chunk L % 8
  constexpr int NQ = BM / 8 / (DTL_THREADS / 64);
  const double* ap[NQ];
  const double* bp[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int r = 8 * (wave + 8 * q) + (lane >> 3);
    const int lc = (lane & 7) ^ ((r >> 1) & 7);
    ap[q] = A + min(m0 + r, p.m - 1) * p.lda + 2 * lc;
    bp[q] = B + min(n0 + r, p.n - 1) * p.ldb + 2 * lc;
  }
  auto issue = [&](long k0, double* as, double* bs) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int b = wave + 8 * q;
      load_lds16(ap[q] + k0, as + b * 8 * BK);
      load_lds16(bp[q] + k0, bs + b * 8 * BK);
    }
  };

  d4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  long kbeg = 0, kend = p.k;
  if (PARTIAL) {
    kbeg = (long)blockIdx.y * p.kchunk;
    kend = min(p.k, kbeg + p.kchunk);
  }
  const long nk = (kend - kbeg) / BK;
  // one tile: fragments from (as, bs) while the next tile streams into (na, nb)
  auto step = [&](long t, const double* as, const double* bs, double* na, double* nb) {
    if (t + 1 < nk) issue(kbeg + (t + 1) * BK, na, nb);
    __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of this tile's MFMAs
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      double2 af[TM], bf[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int r = wm * WTM + i * 16 + l16;
        af[i] = *reinterpret_cast<const double2*>(as + r * BK + 2 * ((4 * h + l4) ^ ((r >> 1) & 7)));
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int r = wn * WTN + j * 16 + l16;
        bf[j] = *reinterpret_cast<const double2*>(bs + r * BK + 2 * ((4 * h + l4) ^ ((r >> 1) & 7)));
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          acc[i][j] = Mfma<double>::op(af[i].x, bf[j].x, acc[i][j]);
          acc[i][j] = Mfma<double>::op(af[i].y, bf[j].y, acc[i][j]);
        }
    }
    // this wave's next tiles have landed, and every wave is done reading the
    // current ones before they are refilled
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  };
  if (nk > 0) issue(kbeg, As0, Bs0);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (long t = 0; t < nk; t += 2) {
    step(t, As0, Bs0, As1, Bs1);
    if (t + 1 < nk) step(t + 1, As1, Bs1, As0, Bs0);
  }

  if (PARTIAL) {
    const long tile_lin = (long)tm * (tm + 1) / 2 + tn;
    double* W = reinterpret_cast<double*>(p.c) + ((long)blockIdx.y * gridDim.x + tile_lin) * (BM * BN);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = wm * WTM + i * 16 + Mfma<double>::row(lane, v);
          const int cc = wn * WTN + j * 16 + l16;
          W[r * BN + cc] = acc[i][j][v];
        }
    return;
  }
  double* Cb = reinterpret_cast<double*>(p.c) + (long)bz * p.stride_c;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long r = m0 + wm * WTM + i * 16 + Mfma<double>::row(lane, v);
        const long cc = n0 + wn * WTN + j * 16 + l16;
        if (r < p.m && cc < p.n) {
          const long off = p.c_trans ? cc * p.ldc + r : r * p.ldc + cc;
          double x = p.alpha * acc[i][j][v];
          if (p.c_f32) {
            reinterpret_cast<float*>(p.c)[(long)bz * p.stride_c + off] = (float)x;
          } else {
            if (p.beta != 0.0) x += p.beta * Cb[off];
            Cb[off] = x;
          }
        }
      }
}

// V = FP64 sum over the split-K partials (in split order), mirrored to the
// upper triangle: one workgroup per 32 x 32 block of a lower tile, thread t
// owning row t / 8 and columns 4 (t % 8) .. + 3; the mirror image goes
// through an LDS transpose so both stores are row-contiguous.
template <class T, int BM, int BN>
__global__ __launch_bounds__(256) void syrk_reduce_kernel(const T* __restrict__ work, int nsplit, long ntiles, long m,
                                                          double* __restrict__ v, long ldv) {
  constexpr int RB = 32, SUBM = BM / RB, SUBN = BN / RB;
  __shared__ double S[RB][RB + 1];
  const long tile_lin = blockIdx.x / (SUBM * SUBN);
  const int sb = blockIdx.x % (SUBM * SUBN), sr = sb / SUBN, sc = sb % SUBN;
  int tm = (int)((sqrt(8.0 * tile_lin + 1.0) - 1.0) * 0.5);
  while ((long)(tm + 1) * (tm + 2) / 2 <= tile_lin) ++tm;
  while ((long)tm * (tm + 1) / 2 > tile_lin) --tm;
  const int tn = (int)(tile_lin - (long)tm * (tm + 1) / 2);
  const int r = threadIdx.x >> 3, c0 = (threadIdx.x & 7) * 4;
  const int rl = sr * RB + r, cl = sc * RB + c0;
  const T* src = work + tile_lin * (BM * BN) + rl * BN + cl;
  const long sstride = ntiles * (BM * BN);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int sp = 0; sp < nsplit; ++sp) {
    const T* p = src + (long)sp * sstride;
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u] += (double)p[u];
  }
  const long gr = (long)tm * BM + rl, gc = (long)tn * BN + cl;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    S[r][c0 + u] = acc[u];
    if (gr < m && gc + u < m && gc + u <= gr) v[gr * ldv + gc + u] = acc[u];
  }
  __syncthreads();
  const long mr = (long)tn * BN + sc * RB + r, mc = (long)tm * BM + sr * RB + c0;
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (mr < m && mc + u < m && mr < mc + u) v[mr * ldv + mc + u] = S[c0 + u][r];
}

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

KArgs make_args(const GemmArgs& g) {
  KArgs p = {};
  p.m = g.m; p.n = g.n; p.k = g.k;
  p.a = g.a; p.lda = g.lda; p.a_rows = g.a_rows; p.a_nrows = g.a_nrows;
  p.b = g.b; p.ldb = g.ldb; p.b_rows = g.b_rows; p.b_nrows = g.b_nrows;
  p.c = g.c; p.ldc = g.ldc; p.c_trans = g.c_trans;
  p.alpha = g.alpha; p.beta = g.beta;
  p.lower = g.lower;
  p.tri = g.tri;
  p.c_f32 = g.c_f32;
  p.stride_a = g.stride_a; p.stride_b = g.stride_b; p.stride_c = g.stride_c;
  p.dig_out = g.dig_out; p.dig_s = g.dig_s; p.dig_inv2 = g.dig_inv2; p.dig_nr = g.dig_nr; p.dig_kp = g.dig_kp;
  return p;
}

// Tile configuration: BM x BN x BK on a WM x WN wave grid, TM x TN MFMA
// blocks per wave.
template <int BM_, int BN_, int BK_, int TM_, int TN_, int WM_, int WN_>
struct Cfg {
  static constexpr int BM = BM_, BN = BN_, BK = BK_, TM = TM_, TN = TN_, WM = WM_, WN = WN_;
  static constexpr int THREADS = 64 * WM * WN;
};
typedef Cfg<128, 128, 16, 2, 4, 4, 2> CfgLarge;  // 8 waves of 32 x 64
typedef Cfg<64, 64, 16, 2, 2, 2, 2> CfgSmall;    // 4 waves of 32 x 32
typedef Cfg<32, 32, 16, 2, 2, 1, 1> CfgTiny;     // 1 wave of 32 x 32: four times the workgroups
typedef Cfg<128, 128, 32, 2, 4, 4, 2> CfgDeepK;  // 135 KB of LDS: one workgroup per CU

template <class T, class C, int EPI, bool AK, bool BKc, bool GA, bool GB>
void launch_vec(const KArgs& p, bool vec, dim3 grid, hipStream_t s) {
  if (vec)
    hipLaunchKernelGGL((gemm_kernel<T, C::BM, C::BN, C::BK, C::TM, C::TN, C::WM, C::WN, EPI, AK, BKc, GA, GB, true>),
                       grid, dim3(C::THREADS), 0, s, p);
  else
    hipLaunchKernelGGL((gemm_kernel<T, C::BM, C::BN, C::BK, C::TM, C::TN, C::WM, C::WN, EPI, AK, BKc, GA, GB, false>),
                       grid, dim3(C::THREADS), 0, s, p);
}

// Every operand layout the path uses: (A k-contiguous?, B k-contiguous?,
// gather on A?, gather on B?); gathers need the gathered operand k-contiguous
// and only one operand gathered at a time.
template <class T, class C, int EPI>
int launch_layout(const GemmArgs& g, const KArgs& p, dim3 grid, hipStream_t s) {
  const int vw = 16 / (int)sizeof(T);  // elements per 16 bytes
  const bool vec = aligned16(g.a) && aligned16(g.b) && g.lda % vw == 0 && g.ldb % vw == 0 &&
                   g.stride_a % vw == 0 && g.stride_b % vw == 0;
  const bool ak = g.a_trans == 0, bk = g.b_trans == 1, ga = g.a_rows != nullptr, gb = g.b_rows != nullptr;
  if ((ga && !ak) || (gb && !bk) || (ga && gb)) return THC_ERR_ARG;
  if (ak && bk) {
    if (ga) launch_vec<T, C, EPI, true, true, true, false>(p, vec, grid, s);
    else if (gb) launch_vec<T, C, EPI, true, true, false, true>(p, vec, grid, s);
    else launch_vec<T, C, EPI, true, true, false, false>(p, vec, grid, s);
  } else if (ak && !bk) {
    launch_vec<T, C, EPI, true, false, false, false>(p, vec, grid, s);
  } else if (!ak && !bk) {
    launch_vec<T, C, EPI, false, false, false, false>(p, vec, grid, s);
  } else {
    launch_vec<T, C, EPI, false, true, false, false>(p, vec, grid, s);
  }
  THC_HIP(hipGetLastError());
  return THC_OK;
}

template <class T, class C>
int launch(const GemmArgs& g, hipStream_t s) {
  const KArgs p = make_args(g);
  if (g.epi_digits) {
    // B rows gathered, both operands k-contiguous, the deep-k tile (its
    // operand LDS holds one staged digit plane)
    if constexpr (std::is_same<T, double>::value && std::is_same<C, CfgDeepK>::value) {
      if (g.lower || g.tri || g.a_rows || !g.b_rows || g.a_trans != 0 || g.b_trans != 1 || g.batch != 1 ||
          g.epi_square || g.epi_hadamard || g.epi_addsquare || !g.dig_out || !g.dig_s || !g.dig_inv2 ||
          g.dig_kp % 32 != 0 || g.dig_nr % 64 != 0)
        return THC_ERR_ARG;
      const long nblk = ((g.m + C::BM - 1) / C::BM) * ((g.n + C::BN - 1) / C::BN);
      if (nblk <= 0) return THC_OK;
      if (nblk > 0x7fffffffL) return THC_ERR_ARG;
      const int vw = 16 / (int)sizeof(T);
      const bool vec = aligned16(g.a) && aligned16(g.b) && g.lda % vw == 0 && g.ldb % vw == 0;
      launch_vec<T, C, EPI_DIGITS, true, true, false, true>(p, vec, dim3((unsigned)nblk), s);
      THC_HIP(hipGetLastError());
      return THC_OK;
    } else {
      return THC_ERR_ARG;
    }
  }
  const long nbm = (g.m + C::BM - 1) / C::BM, nbn = (g.n + C::BN - 1) / C::BN;
  const long nblk = g.lower ? nbm * (nbm + 1) / 2 : nbm * nbn;
  if (nblk <= 0 || g.batch <= 0) return THC_OK;
  if (nblk > 0x7fffffffL) return THC_ERR_ARG;
  if ((g.a_rows || g.b_rows) && g.lower) return THC_ERR_ARG;
  dim3 grid((unsigned)nblk, 1, (unsigned)g.batch);
  if (g.epi_hadamard || g.epi_addsquare) {
    // only the gathered, k-contiguous layouts of the orbital products
    if (g.epi_square || (g.epi_hadamard && g.epi_addsquare) || g.a_trans != 0 || g.b_trans != 1 ||
        (!g.a_rows == !g.b_rows))
      return THC_ERR_ARG;
    const int vw = 16 / (int)sizeof(T);
    const bool vec = aligned16(g.a) && aligned16(g.b) && g.lda % vw == 0 && g.ldb % vw == 0 &&
                     g.stride_a % vw == 0 && g.stride_b % vw == 0;
    if (g.epi_hadamard) {
      if (g.a_rows) launch_vec<T, C, EPI_HADAMARD, true, true, true, false>(p, vec, grid, s);
      else launch_vec<T, C, EPI_HADAMARD, true, true, false, true>(p, vec, grid, s);
    } else {
      if (g.a_rows) launch_vec<T, C, EPI_ADDSQUARE, true, true, true, false>(p, vec, grid, s);
      else launch_vec<T, C, EPI_ADDSQUARE, true, true, false, true>(p, vec, grid, s);
    }
    THC_HIP(hipGetLastError());
    return THC_OK;
  }
  if (g.epi_square) return launch_layout<T, C, EPI_SQUARE>(g, p, grid, s);
  return launch_layout<T, C, EPI_STORE>(g, p, grid, s);
}

// The direct-to-LDS FP64 tile: plain store epilogue, A and B contiguous along
// k with 16-byte aligned rows, k a multiple of the 16-deep tile.
bool dtl_fits(const GemmArgs& g) {
  return !g.lower && !g.tri && g.a_trans == 0 && g.b_trans == 1 && !g.a_rows && !g.b_rows && !g.epi_square &&
         !g.epi_hadamard && !g.epi_addsquare && g.k > 0 && g.k % DTL_BK == 0 && g.lda % 2 == 0 &&
         g.ldb % 2 == 0 && g.stride_a % 2 == 0 && g.stride_b % 2 == 0 && aligned16(g.a) && aligned16(g.b);
}

int launch_dtl(const GemmArgs& g, hipStream_t s) {
  const KArgs p = make_args(g);
  const long nblk = ((g.m + DTL_BM - 1) / DTL_BM) * ((g.n + DTL_BN - 1) / DTL_BN);
  if (nblk <= 0 || g.batch <= 0) return THC_OK;
  if (nblk > 0x7fffffffL) return THC_ERR_ARG;
  hipLaunchKernelGGL(gemm_dtl_kernel<false>, dim3((unsigned)nblk, 1, (unsigned)g.batch), dim3(DTL_THREADS), 0, s,
                     p);
  THC_HIP(hipGetLastError());
  return THC_OK;
}

// ---- tuning table (written by the autotuner, thcpp_amd/autotune.py) --------

struct TuneEntry {
  int mb, nb, kb, flags, cfg;
};
std::mutex tune_mu;
std::vector<TuneEntry> tune_tab;

int ceil_log2(long v) {
  int b = 0;
  while ((1L << b) < v) ++b;
  return b;
}

int tuned_cfg(const GemmArgs& g) {
  const int mb = ceil_log2(g.m), nb = ceil_log2(g.n), kb = ceil_log2(g.k), fl = gemm_tune_flags(g);
  std::lock_guard<std::mutex> lock(tune_mu);
  for (const TuneEntry& e : tune_tab)
    if (e.mb == mb && e.nb == nb && e.kb == kb && e.flags == fl) return e.cfg;
  return GEMM_CFG_AUTO;
}

// launch on configuration cfg; 1 when cfg does not apply to these arguments
template <class T>
int launch_cfg(const GemmArgs& g, int cfg, hipStream_t s) {
  switch (cfg) {
    case GEMM_CFG_TINY: return launch<T, CfgTiny>(g, s);
    case GEMM_CFG_SMALL: return launch<T, CfgSmall>(g, s);
    case GEMM_CFG_LARGE: return launch<T, CfgLarge>(g, s);
    case GEMM_CFG_DTL:
      if (sizeof(T) == 8 && dtl_fits(g)) return launch_dtl(g, s);
      return 1;
    case GEMM_CFG_DEEPK:
      if (!g.lower) return launch<T, CfgDeepK>(g, s);
      return 1;
    default: return 1;
  }
}

template <class T>
int gemm_dispatch(const GemmArgs& g, hipStream_t s) {
  if (g.m <= 0 || g.n <= 0) return THC_OK;
  if (g.lower && g.m != g.n) return THC_ERR_ARG;
  if (g.c_f32 && (sizeof(T) != 8 || g.beta != 0.0 || g.epi_square || g.epi_hadamard || g.epi_addsquare))
    return THC_ERR_ARG;
  if (g.cfg < 0 || g.cfg >= GEMM_CFG_COUNT) return THC_ERR_ARG;
  if (g.epi_digits) return launch<T, CfgDeepK>(g, s);
  const int cfg = g.cfg != GEMM_CFG_AUTO ? g.cfg : tuned_cfg(g);
  if (cfg != GEMM_CFG_AUTO) {
    const int r = launch_cfg<T>(g, cfg, s);
    if (r != 1) return r;
    if (g.cfg != GEMM_CFG_AUTO) return THC_ERR_UNSUPPORTED;  // a forced configuration that does not apply
  }
  const long tiles128 = ((g.m + 127) / 128) * ((g.n + 127) / 128) * g.batch;
  long tiles64 = ((g.m + 63) / 64) * ((g.n + 63) / 64);
  if (g.lower) tiles64 = (tiles64 + (g.m + 63) / 64) / 2;
  tiles64 *= g.batch;
  // The ZC^T GEMM (N x K output, inner dimension M): with one workgroup per CU
  // a CU empties completely whenever a tile finishes, and the latency-bound
  // CC^T Cholesky and inverse on the second stream get CUs between tiles; two
  // workgroups per CU finish out of step and keep every CU occupied.  Measured:
  // ZC^T 3.8 ms with the second stream hidden, against 2.9 ms + 1.6 ms exposed.
  if (g.k <= 256 && g.m >= 65536 && !g.lower) return launch<T, CfgDeepK>(g, s);
  // too few 64 x 64 tiles to fill 256 CUs a few times over: single-wave tiles
  if (tiles64 < 1024) return launch<T, CfgTiny>(g, s);
  if (tiles128 < 512) return launch<T, CfgSmall>(g, s);
  if (sizeof(T) == 8 && dtl_fits(g)) return launch_dtl(g, s);
  return launch<T, CfgLarge>(g, s);
}

constexpr int SY_BM = CfgLarge::BM, SY_BN = CfgLarge::BN, SY_BK = CfgLarge::BK;

}  // namespace

int gemm_tune_flags(const GemmArgs& g) {
  return (g.a_rows || g.b_rows ? 1 : 0) | (g.lower ? 2 : 0) | (g.epi_square ? 4 : 0) | (g.tri ? 8 : 0);
}

int gemm_set_tuning(const int* entries, int n) {
  if (n < 0 || (n > 0 && !entries)) return THC_ERR_ARG;
  std::vector<TuneEntry> t(n);
  for (int i = 0; i < n; ++i) {
    const int* e = entries + 5 * i;
    if (e[4] < GEMM_CFG_AUTO || e[4] >= GEMM_CFG_COUNT) return THC_ERR_ARG;
    t[i] = TuneEntry{e[0], e[1], e[2], e[3], e[4]};
  }
  std::lock_guard<std::mutex> lock(tune_mu);
  tune_tab.swap(t);
  return THC_OK;
}

int gemm_f64(const GemmArgs& g, hipStream_t s) { return gemm_dispatch<double>(g, s); }
int gemm_f32(const GemmArgs& g, hipStream_t s) { return gemm_dispatch<float>(g, s); }

static void syrk_plan(long m, long k, long& ntiles, int& nsplit, long& kchunk) {
  const long nb = (m + SY_BM - 1) / SY_BM;
  ntiles = nb * (nb + 1) / 2;
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
                hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  if (m <= 0) return THC_OK;
  long ntiles, kchunk;
  int nsplit;
  syrk_plan(m, k, ntiles, nsplit, kchunk);
  KArgs p = {};
  p.m = m; p.n = m; p.k = k;
  p.a = a; p.lda = lda; p.a_rows = nullptr;
  p.b = a; p.ldb = lda;
  p.c = work; p.ldc = 0;
  p.alpha = 1.0; p.beta = 0.0;
  p.lower = 1;
  p.kchunk = kchunk;
  dim3 grid((unsigned)ntiles, (unsigned)nsplit, 1);
  const dim3 rgrid((unsigned)(ntiles * (SY_BM / 32) * (SY_BN / 32)));  // 32 x 32 blocks of the tiles
  const int vw = prec == 0 ? 2 : 4;
  const bool vec = aligned16(a) && lda % vw == 0;
  if (e0) THC_HIP(hipEventRecord(e0, s));
  if (prec == 0) {
    // direct-to-LDS tiles when every split covers whole 16-deep k-steps
    static_assert(SY_BM == DTL_BM && SY_BN == DTL_BN && SY_BK == DTL_BK, "same tile as the direct-to-LDS kernel");
    if (vec && k % DTL_BK == 0)
      hipLaunchKernelGGL(gemm_dtl_kernel<true>, grid, dim3(DTL_THREADS), 0, s, p);
    else
      launch_vec<double, CfgLarge, EPI_PARTIAL, true, true, false, false>(p, vec, grid, s);
    THC_HIP(hipGetLastError());
    if (e1) THC_HIP(hipEventRecord(e1, s));
    hipLaunchKernelGGL((syrk_reduce_kernel<double, SY_BM, SY_BN>), rgrid, dim3(256), 0, s,
                       (const double*)work, nsplit, ntiles, m, v, ldv);
  } else {
    launch_vec<float, CfgLarge, EPI_PARTIAL, true, true, false, false>(p, vec, grid, s);
    THC_HIP(hipGetLastError());
    if (e1) THC_HIP(hipEventRecord(e1, s));
    hipLaunchKernelGGL((syrk_reduce_kernel<float, SY_BM, SY_BN>), rgrid, dim3(256), 0, s,
                       (const float*)work, nsplit, ntiles, m, v, ldv);
  }
  THC_HIP(hipGetLastError());
  return THC_OK;
}

}  // namespace thc
