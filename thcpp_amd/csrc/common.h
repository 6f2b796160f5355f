// Internal definitions shared by the HIP translation units of libthcpp_amd.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/thcpp_amd.h"

namespace thc {

void set_last_error(const char* what, const char* file, int line);

#define THC_HIP(expr)                                                   \
  do {                                                                  \
    hipError_t e_ = (expr);                                             \
    if (e_ != hipSuccess) {                                             \
      ::thc::set_last_error(hipGetErrorString(e_), __FILE__, __LINE__); \
      return THC_ERR_HIP;                                               \
    }                                                                   \
  } while (0)

#define THC_TRY(expr)          \
  do {                         \
    int s_ = (expr);           \
    if (s_ != THC_OK && s_ < 100) return s_; \
  } while (0)

// Bump allocator over one device allocation.  reset() forgets every carve;
// reserve() grows the backing store (freeing the old one) when a call needs
// more than it holds.  Carves are 256-byte aligned.
struct Arena {
  char* base = nullptr;
  size_t cap = 0;
  size_t used = 0;
  int reserve(size_t bytes);
  void reset() { used = 0; }
  template <class T>
  T* take(size_t count) {
    size_t off = (used + 255) & ~size_t(255);
    size_t bytes = count * sizeof(T);
    if (off + bytes > cap) return nullptr;
    used = off + bytes;
    return reinterpret_cast<T*>(base + off);
  }
  void release();
};

}  // namespace thc

struct KmGraphCache;

struct thc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  bool use_graphs = false;  // k-means loop as a replayed HIP graph (thc_ctx_set_graphs)
  thc::Arena ws;       // per-call scratch
  thc::Arena kmws;     // k-means state (kept across calls for graph replay)
  KmGraphCache* kmgraph = nullptr;
  hipEvent_t ev[15] = {};  // isdf stage boundaries (0-7), solve GEMM (8, 9), V kernel (10, 11), FFT passes (12-14)
  hipStream_t aux = nullptr;      // second stream: CC^T factorisation beside the ZC^T GEMM
  hipEvent_t fork = nullptr, join = nullptr;
  hipEvent_t kev[3] = {};  // k-means: centroid sums done, loop control done (by iteration parity)
  bool km_overlap = true;  // k-means loop control on 'aux' beside the next iteration (off in multi-lane batches)
  void* pinned = nullptr;  // small pinned host buffer for status read-back
  size_t pinned_bytes = 0;
  std::vector<thc_ctx*> lanes;  // extra stream contexts for batched calls (owned)
  thc_ctx* kaux = nullptr;      // batched calls: the second k-means run (owned)
};

namespace thc {

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// Byte count a carve of 'count' T's takes in an arena (with alignment slack).
template <class T>
inline size_t carve_bytes(size_t count) { return align_up(count * sizeof(T), 256) + 256; }

int ensure_pinned(thc_ctx* ctx, size_t bytes);

// ---- GEMM (gemm.hip) ------------------------------------------------------
// C[m, n] = alpha * sum_k op(A)[m, k] op(B)[k, n] + beta * C[m, n]
//   a_trans = 0: A[m * lda + k]  (row-major m x k)   1: A[k * lda + m]
//   b_trans = 0: B[k * ldb + n]  (row-major k x n)   1: B[n * ldb + k]
//   a_rows: optional row gather for a_trans = 0 (row m of op(A) = row a_rows[m] of A)
//   b_rows: optional row gather for b_trans = 1 (column n of op(B) = row b_rows[n] of B)
//   c_trans: write the result transposed (element (m, n) at c[n * ldc + m])
//   epi_square: C = (alpha * acc)^2 (beta ignored)
//   epi_hadamard: C = (alpha * acc) * C elementwise (beta ignored; gathered
//     operand, both operands k-contiguous)
//   epi_addsquare: C = C + (alpha * acc)^2 (same layouts as epi_hadamard)
//   lower: only tiles on/below the diagonal are computed (m, n square)
//   batch: blockIdx.z, strides in elements
struct GemmArgs {
  long m, n, k;
  const void* a; long lda; int a_trans; const int* a_rows;
  long a_nrows;  // rows of A addressable through a_rows (gathered rows outside [0, a_nrows) read as 0)
  const void* b; long ldb; int b_trans; const int* b_rows;
  long b_nrows;  // rows of B addressable through b_rows
  void* c; long ldc;
  int c_trans;   // store C^T: element (m, n) at c[n * ldc + m]
  double alpha, beta;
  int epi_square;
  int epi_hadamard;
  int epi_addsquare;
  int lower;
  int tri;       // GEMM_TRI_* bits: known zero blocks of A / B that bound each tile's k range
  int c_f32;     // FP64 GEMM storing C as float (plain store, beta = 0)
  int cfg;       // GEMM_CFG_*: force a tile configuration (0: the tuning table, then the built-in rule)
  // digit epilogue of the mixed solve (FP64 only, B rows gathered, both
  // operands k-contiguous, 128 x 128 x 32 tiles): instead of C, write the 7
  // int8 digit planes of x = (acc^2 * dig_inv2[col]) / dig_s[row] into
  // dig_out, k-blocked (byte (row, col) of plane q at q * dig_nr * dig_kp +
  // (col / 32) * dig_nr * 32 + row * 32 + col % 32), zero for rows >= m or
  // cols >= n, rows < dig_nr and cols < dig_kp
  int epi_digits;
  signed char* dig_out;
  const double* dig_s;
  const double* dig_inv2;
  long dig_nr;
  int dig_kp;
  int batch;
  long stride_a, stride_b, stride_c;
};
// Triangular operands: the k range of the output tile at rows m0 .. m0+BM-1,
// columns n0 .. n0+BN-1 is cut to where both operands can be nonzero.
enum {
  GEMM_TRI_A_KGE_M = 1,  // A[i][q] = 0 for q < i  (A = L^T of a lower L)
  GEMM_TRI_B_KGE_N = 2,  // B[q][j] = 0 for q < j  (B lower triangular)
  GEMM_TRI_A_KLE_M = 4,  // A[i][q] = 0 for q > i  (A lower triangular)
};
// Tile configurations of gemm_f64 / gemm_f32 (the tuning table's values).
enum {
  GEMM_CFG_AUTO = 0,
  GEMM_CFG_TINY = 1,   // 32 x 32 x 16 on one wave
  GEMM_CFG_SMALL = 2,  // 64 x 64 x 16 on 4 waves
  GEMM_CFG_LARGE = 3,  // 128 x 128 x 16 on 8 waves, register-staged
  GEMM_CFG_DTL = 4,    // 128 x 128 x 16 on 8 waves, direct-to-LDS (FP64, plain store, k-contiguous operands)
  GEMM_CFG_DEEPK = 5,  // 128 x 128 x 32 on 8 waves, one workgroup per CU
  GEMM_CFG_COUNT = 6
};
// Process-wide tuning table: entries of (ceil log2 m, ceil log2 n, ceil log2 k,
// flags, config), flags = 1 gathered operand | 2 lower | 4 square epilogue |
// 8 triangular operands.  n = 0 clears it.
int gemm_set_tuning(const int* entries, int n);
int gemm_tune_flags(const GemmArgs& g);
int gemm_f64(const GemmArgs& g, hipStream_t s);
int gemm_f32(const GemmArgs& g, hipStream_t s);
// Lower-triangle SYRK with split inner dimension and FP64 reduction:
// V = A A^T for A (m x k row-major, lda), written full symmetric into v (ldv).
// A is double (prec 0) or float (prec 1); partials live in 'work'.
size_t syrk_work_bytes(long m, long k, int prec);
// (e0, e1 non-null: recorded on s around the partial-tile kernel alone)
int syrk_splitk(const void* a, long m, long k, long lda, int prec, double* v, long ldv,
                void* work, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);

// ---- dense FP64 factorisations (linalg.hip) --------------------------------
size_t chol_work_bytes(int k);
// Device address of the no-allocation Cholesky's static scratch, taken once
// per device (that device current); thc_ctx_create calls it.
int noalloc_scratch_init(int device);
// In-place lower Cholesky (row-major, lda); info written to device int.
int cholesky_f64(double* a, int k, long lda, int* d_info, void* work, hipStream_t s);
size_t spd_inverse_work_bytes(int k);
// ainv = A^-1 from its lower Cholesky factor l (k x k, lda).
int spd_inverse_from_chol(const double* l, int k, long lda, double* ainv, void* work, hipStream_t s);
// Householder reduction of a symmetric k x k matrix (k <= 4096) to tridiagonal
// form; the diagonal (k) and off-diagonal (k - 1) are copied to host arrays
// asynchronously on s (valid once s has drained).
size_t tridiag_work_bytes(int k);
int sym_tridiag(const double* a, int k, long lda, double* d_host, double* e_host, void* work, hipStream_t s);
// Rank by the rule of MatrixOperations::rank (src/matrix_operations.h:713-735):
// singular values below k * s0 * 1e-16 count as zero; s0 = largest.  Host code
// on the tridiagonal form (Sturm counts and bisection).
void tridiag_rank(const double* d, const double* e, int k, int* rank, double* smax);
int zero_upper(double* a, int k, long lda, hipStream_t s);
int symmetrize_from_lower(double* a, int k, long lda, hipStream_t s);

constexpr int OZ_DIGITS = 5;  // digits per value of the mixed V (ozaki.hip, FFT pass B)
// Mixed-path output of the FFT: instead of the weighted half spectrum, its
// OZ_DIGITS int8 digit planes for the V contraction.  Each (row mu, y-line)
// of pass B is one chunk: one power-of-two exponent e[mu * n1 + iy] (the
// line's max |value| < 2^e) and columns iy * lp + 2 j + (0 re, 1 im) for
// the j-th complex value of the line (j = kx * n2h + kz), zero up to lp; byte
// (row, col) of plane q at q * k * g2p + row * g2p + col, g2p = n1 * lp.
struct FftDigits {
  signed char* s;
  int* e;
  long g2p;
  int lp;
};

// ---- int8-emulated Gram and solve (ozaki.hip) ---------------------------------
// V = B B^T from the digit planes the FFT's pass B writes (FftDigits, one
// chunk per y-line), by exact int8 digit products, written full symmetric
// into v; oz_lines_layout carves the planes and the partial tiles from 'work'.
int oz_line_len(const int n[3]);
size_t oz_lines_work_bytes(int k, const int n[3]);
FftDigits oz_lines_layout(int k, const int n[3], void* work);
int oz_syrk_lines(const FftDigits& d, int k, const int n[3], double* v, long ldv, void* work, hipStream_t s,
                  hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// X = ainv * Z (k x n) from Zt (n x k row-major, ZC^T transposed) by exact
// int8 digit products of the scaled operands ainv diag(s(idx)) and
// Zt / (s(r) s(idx)), s(r) = sqrt(sum phi_a(r)^2 sum phi_b(r)^2) (phi_b null:
// sum phi_a(r)^2).  X is float (x_f32) or double with row stride ldx.
size_t oz_solve_work_bytes(int k, long n);
int oz_solve(const double* ainv, long lda, const double* zt, int k, long n, const int* idx, const double* phi_a,
             int ma, const double* phi_b, int mb, void* x, int x_f32, long ldx, void* work, hipStream_t s,
             hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, int zt_split = 0);
// With the ZC^T digits written by the ZC^T GEMM's digit epilogue: oz_solve_prepare
// computes s and the point scales into the solve workspace and fills the
// epilogue fields of g (dig_*); oz_solve with zt_split = 1 then skips them and
// the split of Zt (zt unused).
int oz_solve_prepare(int k, long n, const int* idx, const double* phi_a, int ma, void* work, GemmArgs& g,
                     hipStream_t s);

// ---- QR with column pivoting (qrcp.hip) ---------------------------------------
// First k pivots of QRCP of the orbital-product matrix (occ = nullptr: all
// orbitals on both sides), sorted ascending into piv_out (device); rdiag_out
// (host, nullable) gets the k pivot magnitudes |R_jj| in pivot order.
// cplx: phi / occ hold complex orbitals, (re, im) interleaved; m and mo are
// the real row widths (twice the orbital counts).
size_t qrcp_work_bytes(long n, int k, bool cplx = false);
int qrcp_points(const double* phi, const double* occ, long n, int m, int mo, int k, int* piv_out, double* rdiag_out,
                Arena& ws, hipStream_t s, bool cplx = false);

// ---- FFT (fft.hip) ----------------------------------------------------------
size_t fft_work_bytes(int k, const int n[3], int prec);
// x: k x n0 n1 n2, double or (x_f32, prec 1 only) float; arithmetic and out
// (k x g2 weighted half spectrum) in double (prec 0) or float (prec 1)
// pass_ev (optional, 3 events): recorded before the first pass, between the
// passes and after the second
int fft_coulomb(const void* x, int x_f32, int k, const int n[3], const double* coulG, void* out, int prec,
                void* work, hipStream_t s, hipEvent_t* pass_ev = nullptr, const FftDigits* dig = nullptr);
inline long half_spectrum_len(const int n[3]) { return 2L * n[0] * n[1] * (n[2] / 2 + 1); }

}  // namespace thc
