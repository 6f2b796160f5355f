// V = B B^T for the mixed-precision path, evaluated exactly in int8 MFMA
// arithmetic on splits of B (an Ozaki-style emulation of the FP64 Gram).
//
// Split (done by the FFT's pass B as it produces B, fft.hip): the columns of
// a row of B come in chunks, one per y-line of the half spectrum (2 n0
// (n2/2+1) values, zero padded to a multiple of 64); each (row, chunk) gets a
// power-of-two exponent e with max |b| < 2^e, and b 2^-e in (-1, 1) is
// written as OZ_S signed 7-bit digits, b = 2^e sum_p d_p 2^-7(p+1) + r,
// |r| < 2^(e - 7 OZ_S).  Every step (scale by 128, truncate, subtract) is
// exact in double.
// Product: for one chunk, sum_g b_rg b_cg = 2^(e_r + e_c) sum_{p+q<OZ_S}
// 2^-7(p+q+2) sum_g d_p,rg d_q,cg.  The inner sums are int8 x int8 products
// accumulated in int32 by v_mfma_i32_16x16x64_i8, grouped by p + q (products
// with the same p + q share a scale), exact as long as
// OZ_S * chunk * 127^2 < 2^31 (chunks up to 26624 columns: any grid edge up
// to 128).  Digit products with p + q >= OZ_S are dropped; they are below
// 2^-7(OZ_S+1) of |b_r| |b_c| per term.
// Each workgroup takes one 128 x 128 lower tile and one chunk, converts its
// groups to double, scales by 2^(e_r + e_c) and writes a partial tile; the
// partials are summed in chunk order in FP64.
// Slabs (5 digit planes x 128 rows x 64 columns per operand) are double
// buffered in LDS (2 x 80 KB, the whole LDS of a CU) and filled by direct
// global-to-LDS loads (global_load_lds_dwordx4: lane L of a wave writes bytes
// 16 L .. 16 L + 15 of the wave's 1 KB block), so the next step's slabs stream
// in while the current step's MFMAs run.  Rows are 64 bytes with no padding;
// 16-byte segment s of row r is stored at segment (s + r / 4) mod 4.
#include <algorithm>

#include "common.h"

namespace thc {
namespace {

typedef int i4 __attribute__((ext_vector_type(4)));

constexpr int OZ_S = OZ_DIGITS;  // digits per value
constexpr int OZ_T = 128;     // tile edge
constexpr int OZ_KS = 64;     // k per MFMA step

__device__ inline double block_max(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  v = red[0];
  for (int i = 1; i < (int)(blockDim.x >> 6); ++i) v = fmax(v, red[i]);
  return v;
}

// One lower 128 x 128 tile (tm >= tn) and one chunk per workgroup; 8 waves
// in a 2 (rows) x 4 (columns) grid, each wave 64 x 32 = 4 x 2 MFMA blocks.
// Work items w = chunk * ntiles + tile are dealt to the 8 XCDs in contiguous
// runs (workgroup i runs on XCD i % 8): the workgroups resident on one XCD then
// hold tiles of the same chunk and walk its columns together, so the digit
// rows they share are fetched into that XCD's L2 once rather than once per
// tile.
typedef __attribute__((address_space(3))) void lds_void;

__device__ inline int oz_seg(int r, int s) { return (s + (r >> 2)) & 3; }

__global__ __launch_bounds__(512) void oz_syrk(const signed char* __restrict__ S, long g2, long g2p, int k,
                                               const int* __restrict__ E, int nchunk, int ch, int spl,
                                               double* __restrict__ work, long ntiles) {
  // one variable per buffer: the compiler then knows that the direct loads
  // into one buffer do not alias the fragment reads of the other, and does not
  // wait for the loads before reading
  constexpr int SLAB = OZ_S * OZ_T * OZ_KS;
  __shared__ __attribute__((aligned(16))) signed char As0[SLAB], As1[SLAB];
  __shared__ __attribute__((aligned(16))) signed char Bs0[SLAB], Bs1[SLAB];
  const long total = ntiles * nchunk;
  const long per_xcd = (total + 7) / 8;
  const long w = (long)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (w >= total) return;
  const int c = (int)(w / ntiles);
  const long tile_lin = w % ntiles;
  int tm = (int)((sqrt(8.0 * tile_lin + 1.0) - 1.0) * 0.5);
  while ((long)(tm + 1) * (tm + 2) / 2 <= tile_lin) ++tm;
  while ((long)tm * (tm + 1) / 2 > tile_lin) --tm;
  const int tn = (int)(tile_lin - (long)tm * (tm + 1) / 2);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 2, wn = wave & 3;  // 2 x 4 waves
  const int l16 = lane & 15, l4 = lane >> 4;

  i4 acc[OZ_S][4][2];
#pragma unroll
  for (int d = 0; d < OZ_S; ++d)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[d][i][j] = i4{0, 0, 0, 0};

  // A slab is OZ_S * OZ_T rows of 64 bytes = 40 blocks of 16 rows (1 KB);
  // block b = wave + 8 p holds rows 16 wave .. 16 wave + 15 of digit plane p.
  // Lane L fills row 16 wave + L / 4 of every plane, physical segment L % 4,
  // from the logical segment that maps there, so one base pointer per operand
  // and the plane stride address all its loads.  Rows past k read row k - 1
  // (valid memory); they only feed output rows or columns >= k, which are never
  // stored.
  static_assert(OZ_T == 16 * 8, "one 16-row block of every plane per wave");
  const int r_ld = 16 * wave + (lane >> 2);
  const int ls = ((lane & 3) - (r_ld >> 2)) & 3;  // logical segment stored at physical segment L % 4
  const long plane = (long)k * g2p, kbase = (long)c * ch;
  const signed char* sa = S + (long)min(tm * OZ_T + r_ld, k - 1) * g2p + kbase + ls * 16;
  const signed char* sb = S + (long)min(tn * OZ_T + r_ld, k - 1) * g2p + kbase + ls * 16;
  auto issue = [&](int ks, signed char* as, signed char* bs) {
    const long k0 = (long)ks * OZ_KS;
#pragma unroll
    for (int p = 0; p < OZ_S; ++p) {
      const int b = wave + 8 * p;
      __builtin_amdgcn_global_load_lds(sa + p * plane + k0, (lds_void*)(as + b * 1024), 16, 0, 0);
      __builtin_amdgcn_global_load_lds(sb + p * plane + k0, (lds_void*)(bs + b * 1024), 16, 0, 0);
    }
  };
  // k-steps of this chunk that hold columns < g2 (the rest are zero padding)
  const int nks = (int)min((long)(ch / OZ_KS), (g2 - kbase + OZ_KS - 1) / OZ_KS);
  // one k-step: fragments from (as, bs) while the next slabs stream into (na, nb)
  auto step = [&](int ks, const signed char* as, const signed char* bs, signed char* na, signed char* nb) {
    if (ks + 1 < nks) issue(ks + 1, na, nb);
    i4 bq[OZ_S][2];
#pragma unroll
    for (int q = 0; q < OZ_S; ++q)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int r = wn * 32 + j * 16 + l16;
        bq[q][j] = *reinterpret_cast<const i4*>(bs + (q * OZ_T + r) * OZ_KS + oz_seg(r, l4) * 16);
      }
#pragma unroll
    for (int p = 0; p < OZ_S; ++p) {
      i4 ap[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = wm * 64 + i * 16 + l16;
        ap[i] = *reinterpret_cast<const i4*>(as + (p * OZ_T + r) * OZ_KS + oz_seg(r, l4) * 16);
      }
#pragma unroll
      for (int q = 0; q < OZ_S - p; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[p + q][i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ap[i], bq[q][j], acc[p + q][i][j], 0, 0, 0);
    }
    // the next slabs have landed (this wave's loads) and every wave is done
    // reading the current ones before they are refilled
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  };
  issue(0, As0, Bs0);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int ks = 0; ks < nks; ks += 2) {
    step(ks, As0, Bs0, As1, Bs1);
    if (ks + 1 < nks) step(ks + 1, As1, Bs1, As0, Bs0);
  }
  // scale and write the partial tile: D[4 (l/16) + v][l % 16]
  double* W = work + ((long)c * ntiles + tile_lin) * (OZ_T * OZ_T);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int rl = wm * 64 + i * 16 + 4 * l4 + v;
        const int cl = wn * 32 + j * 16 + l16;
        const int r = tm * OZ_T + rl, col = tn * OZ_T + cl;
        double val = 0.0;
        if (r < k && col < k) {
          double sum = 0.0;
#pragma unroll
          for (int d = OZ_S - 1; d >= 0; --d) sum += ldexp((double)acc[d][i][j][v], -7 * (d + 2));
          // exponents are per (row, y-line); a y-line is spl consecutive chunks
          const int nl = nchunk / spl, ln = c / spl;
          val = ldexp(sum, E[(long)r * nl + ln] + E[(long)col * nl + ln]);
        }
        W[rl * OZ_T + cl] = val;
      }
}

// V = sum over chunks (in chunk order) of the partial tiles, mirrored to the
// upper triangle.  One workgroup per 32 x 32 block of a lower tile (16 per
// tile): thread t owns row t / 8, columns 4 (t % 8) .. + 3, four independent
// chunk-ordered sums; the mirror image is written through an LDS transpose so
// both stores are row-contiguous.
constexpr int RB = 32;
__global__ __launch_bounds__(256) void oz_reduce(const double* __restrict__ work, int nchunk, long ntiles, int k,
                                                 double* __restrict__ v, long ldv) {
  __shared__ double S[RB][RB + 1];
  constexpr int SUB = OZ_T / RB;
  const long tile_lin = blockIdx.x / (SUB * SUB);
  const int sb = blockIdx.x % (SUB * SUB), sr = sb / SUB, sc = sb % SUB;
  int tm = (int)((sqrt(8.0 * tile_lin + 1.0) - 1.0) * 0.5);
  while ((long)(tm + 1) * (tm + 2) / 2 <= tile_lin) ++tm;
  while ((long)tm * (tm + 1) / 2 > tile_lin) --tm;
  const int tn = (int)(tile_lin - (long)tm * (tm + 1) / 2);
  const int r = threadIdx.x >> 3, c0 = (threadIdx.x & 7) * 4;
  const int rl = sr * RB + r, cl = sc * RB + c0;  // within the tile
  const double* src = work + tile_lin * (OZ_T * OZ_T) + rl * OZ_T + cl;
  const long cstride = ntiles * (OZ_T * OZ_T);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int ch = 0;
  for (; ch + 4 <= nchunk; ch += 4) {
    double4 q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) q[u] = *reinterpret_cast<const double4*>(src + (long)(ch + u) * cstride);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      s0 += q[u].x;
      s1 += q[u].y;
      s2 += q[u].z;
      s3 += q[u].w;
    }
  }
  for (; ch < nchunk; ++ch) {
    const double4 q = *reinterpret_cast<const double4*>(src + (long)ch * cstride);
    s0 += q.x;
    s1 += q.y;
    s2 += q.z;
    s3 += q.w;
  }
  const long gr = (long)tm * OZ_T + rl, gc = (long)tn * OZ_T + cl;
  const double sv[4] = {s0, s1, s2, s3};
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    S[r][c0 + u] = sv[u];
    if (gr < k && gc + u < k && gc + u <= gr) v[gr * ldv + gc + u] = sv[u];
  }
  __syncthreads();
  // mirror: row gc' = tn T + sc RB + r of the upper triangle, columns gr' = tm T + sr RB + c0 .. + 3
  const long mr = (long)tn * OZ_T + sc * RB + r, mc = (long)tm * OZ_T + sr * RB + c0;
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (mr < k && mc + u < k && mr < mc + u) v[mr * ldv + mc + u] = S[c0 + u][r];
}


}  // namespace


// ---- V from the FFT's digit planes: one chunk per (row, y-line) ----------------

int oz_line_len(const int n[3]) {
  const long len = 2L * n[0] * (n[2] / 2 + 1);
  return (int)((len + OZ_KS - 1) / OZ_KS * OZ_KS);
}

// Chunks per y-line: the fewest that divide the line into whole k-steps of at
// most OZ_CH bytes.  The workgroups of one XCD run the tiles of one chunk at a
// time, and a 128-row slab of a chunk (5 planes) of up to OZ_CH = 4352 bytes,
// 2.8 MB, stays in the XCD's 4 MB L2 for the tiles that share it; a 96-point
// grid's 9408-byte line (6 MB slabs) did not, and V ran at 0.43 of the int8
// peak against 0.51 for 64-point lines.
constexpr int OZ_CH = 4352;
static int oz_chunks_per_line(int lp) {
  const int steps = lp / OZ_KS;
  for (int d = 1; d <= steps; ++d)
    if (steps % d == 0 && lp / d <= OZ_CH) return d;
  return steps;
}

static void oz_lines_plan(int k, const int n[3], int& lp, long& g2p, long& ntiles, int& spl) {
  lp = oz_line_len(n);
  g2p = (long)n[1] * lp;
  spl = oz_chunks_per_line(lp);
  const long nb = (k + OZ_T - 1) / OZ_T;
  ntiles = nb * (nb + 1) / 2;
}

size_t oz_lines_work_bytes(int k, const int n[3]) {
  int lp, spl;
  long g2p, ntiles;
  oz_lines_plan(k, n, lp, g2p, ntiles, spl);
  return align_up((size_t)OZ_S * k * g2p, 256) + align_up(sizeof(int) * (size_t)k * n[1], 256) +
         align_up(sizeof(double) * (size_t)n[1] * spl * ntiles * OZ_T * OZ_T, 256) + 256;
}

FftDigits oz_lines_layout(int k, const int n[3], void* work) {
  int lp, spl;
  long g2p, ntiles;
  oz_lines_plan(k, n, lp, g2p, ntiles, spl);
  char* w = reinterpret_cast<char*>(work);
  FftDigits d;
  d.s = reinterpret_cast<signed char*>(w);
  w += align_up((size_t)OZ_S * k * g2p, 256);
  d.e = reinterpret_cast<int*>(w);
  d.g2p = g2p;
  d.lp = lp;
  return d;
}

int oz_syrk_lines(const FftDigits& d, int k, const int n[3], double* v, long ldv, void* work, hipStream_t s,
                  hipEvent_t e0, hipEvent_t e1) {
  if (k <= 0) return THC_OK;
  static_assert((long)OZ_S * 26624 * 127 * 127 < 2147483647L, "int32 sums of a 128-edge grid's line");
  int lp, spl;
  long g2p, ntiles;
  oz_lines_plan(k, n, lp, g2p, ntiles, spl);
  if (lp != d.lp || g2p != d.g2p || lp > 26624) return THC_ERR_ARG;
  const int nchunk = n[1] * spl, ch = lp / spl;
  double* part = reinterpret_cast<double*>(reinterpret_cast<char*>(work) + align_up((size_t)OZ_S * k * g2p, 256) +
                                           align_up(sizeof(int) * (size_t)k * n[1], 256));
  const long slots = (ntiles * nchunk + 7) / 8 * 8;  // every w < ntiles * nchunk gets one workgroup
  if (slots > 0x7fffffffL) return THC_ERR_ARG;
  if (e0) THC_HIP(hipEventRecord(e0, s));
  // every column of a chunk is data or zero padding: g2 = g2p
  hipLaunchKernelGGL(oz_syrk, dim3((unsigned)slots), dim3(512), 0, s, d.s, g2p, g2p, k, d.e, nchunk, ch, spl, part,
                     ntiles);
  THC_HIP(hipGetLastError());
  if (e1) THC_HIP(hipEventRecord(e1, s));
  hipLaunchKernelGGL(oz_reduce, dim3((unsigned)(ntiles * (OZ_T / RB) * (OZ_T / RB))), dim3(256), 0, s, part, nchunk,
                     ntiles, k, v, ldv);
  THC_HIP(hipGetLastError());
  return THC_OK;
}


}  // namespace thc

namespace thc {
namespace {

// ============================================================================
// Least-squares application X = CC^T^-1 ZC^T in exact int8 digit products.
//
// The product is scaled so that both operands have a known range:
//   A'[mu][nu] = (CC^T)^-1[mu][nu] s[nu]          (s = s(r_nu) at the points)
//   Zs[r][nu]  = ZC^T[nu][r] / (s(r_nu) s(r))     |Zs| <= 1 by Cauchy-Schwarz
//   X[mu][r]   = s(r) sum_nu A'[mu][nu] Zs[r][nu]
// with s(r) = sqrt(rho_a(r) rho_b(r)), rho = sum of squared orbitals of the two
// orbital sets whose pair products form ZC^T (rho itself when they are the
// same set).  Each row of A' gets one power-of-two exponent and SV_D signed
// 7-bit digits; Zs/2 is cut into SV_D digits with a fixed exponent (no row
// maximum is needed).  The digit products with p + q < SV_D are accumulated
// exactly in int32 by v_mfma_i32_32x32x32_i8, grouped by p + q, and combined
// in FP64 in the epilogue; the dropped products are below 2^-7(SV_D+1) of
// |A'| |Zs| per term.  Seven digits hold the interpolating vectors of the
// 64-atom, 2048-point system to 1.4e-7 and V to 5e-8 of the FP64 solve (six
// digits give 5.6e-6 in V).
//
// Digit planes are k-blocked: byte (row, k) of plane p lives at
//   p * rows * kp + (k / 32) * rows * 32 + row * 32 + k % 32,
// so the 32 contraction steps of 32 consecutive rows that one stage of the
// GEMM needs are one contiguous KB.  Row counts are padded (A to 128, Z to 64
// rows) with zero digits, so no load needs a bounds check.
//
// Tile: 128 (mu) x 64 (r) per 512-thread workgroup, 8 waves in a 4 x 2 grid of
// 32 x 32 outputs (one MFMA block, 7 accumulator groups = 112 registers).
// Stages of 32 contraction steps (7 planes x (128 + 64) rows x 32 bytes =
// 42 KB) circulate through three LDS buffers filled by direct global-to-LDS
// loads, two stages ahead of the MFMAs; the barrier between stages waits only
// for this wave's loads of the stage about to be read.
// ============================================================================

constexpr int SV_D = 7;
constexpr int SV_GM = 4, SV_GN = 8;  // an XCD's 32 workgroups: SV_GM mu-tiles x SV_GN r-tiles
constexpr int SV_TM = 128, SV_TN = 64, SV_KS = 32;
constexpr int SV_ABLK = SV_D * SV_TM * SV_KS / 1024;  // 1 KB blocks of 32 rows x 32 bytes
constexpr int SV_BBLK = SV_D * SV_TN * SV_KS / 1024;
constexpr int SV_NBLK = SV_ABLK + SV_BBLK;           // 42 per stage
constexpr int SV_LPW = (SV_NBLK + 7) / 8;            // loads a wave issues per stage (6)
static_assert((long)SV_D * 4096 * 127 * 127 < 2147483647L, "int32 group sums must not overflow for K <= 4096");

typedef int i16v __attribute__((ext_vector_type(16)));

// physical 16-byte segment of logical segment ls in row r of a 32-row block:
// rows 4 apart in the same bank group alternate segments, so an 8- or 16-lane
// phase of a fragment read touches every bank once
__device__ inline int sv_seg(int r, int ls) { return ls ^ ((r >> 2) & 1) ^ ((r >> 3) & 1); }

// byte offset of (row, k) in a k-blocked plane of 'rows' rows
__device__ inline long sv_off(long row, long k, long rows) { return (k >> 5) * rows * 32 + row * 32 + (k & 31); }

// s at an interpolation point; indices outside the grid (refused by the caller
// afterwards) read as 0
__device__ inline double sv_at(const double* s, const int* idx, int j, long n) {
  const int i = idx[j];
  return (i >= 0 && i < n) ? s[i] : 0.0;
}

// A'[mu][nu] = a[mu][nu] s[idx[nu]] -> digits SA (k-blocked, kr rows, kp
// columns, zero padded) and exponent EA[mu].  One workgroup per row mu < kr;
// rows >= k are zero.  Thread t takes columns 4 t .. 4 t + 3 of every 1024.
__global__ __launch_bounds__(256) void sv_split_rows(const double* __restrict__ a, long lda, int k, int kr, int kp,
                                                     const double* __restrict__ s, long n,
                                                     const int* __restrict__ idx, signed char* __restrict__ SA,
                                                     int* __restrict__ EA) {
  __shared__ double red[4];
  const int row = blockIdx.x;
  const bool live = row < k;
  double mx = 0.0;
  if (live)
    for (int j = threadIdx.x; j < k; j += 256) mx = fmax(mx, fabs(a[(long)row * lda + j] * sv_at(s, idx, j, n)));
  mx = block_max(mx, red);
  int ex = 0;
  if (mx > 0.0) frexp(mx, &ex);
  if (threadIdx.x == 0) EA[row] = ex;
  const long plane = (long)kr * kp;
  for (int j0 = 4 * threadIdx.x; j0 < kp; j0 += 1024) {
    double x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = j0 + u;
      x[u] = (live && j < k) ? ldexp(a[(long)row * lda + j] * sv_at(s, idx, j, n), -ex) : 0.0;
    }
#pragma unroll
    for (int p = 0; p < SV_D; ++p) {
      unsigned w = 0;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        x[u] *= 128.0;
        const double d = trunc(x[u]);
        x[u] -= d;
        w |= ((unsigned)(int)d & 0xffu) << (8 * u);
      }
      *reinterpret_cast<unsigned*>(SA + p * plane + sv_off(row, j0, kr)) = w;
    }
  }
}

// inv2[j] = 1 / (2 s[idx[j]]) for j < k (0 where s = 0 or j >= k), j < kp
__global__ __launch_bounds__(256) void sv_point_scale(const double* __restrict__ s, long n,
                                                      const int* __restrict__ idx, int k, int kp,
                                                      double* __restrict__ inv2) {
  for (int j = blockIdx.x * 256 + threadIdx.x; j < kp; j += 256 * gridDim.x) {
    const double v = j < k ? sv_at(s, idx, j, n) : 0.0;
    inv2[j] = v > 0.0 ? 0.5 / v : 0.0;
  }
}

// Zs[r][nu] / 2 = zt[r][nu] inv2[nu] / s[r] -> digits SZ (k-blocked, nr rows).
// Block: 64 rows x one 32-column chunk; thread t takes row t / 4, columns
// 8 (t % 4) .. + 7 of the chunk (two 32-byte reads, one 8-byte store per
// plane), so a wave stores 16 rows x 32 bytes contiguous per plane.  Rows
// with s[r] = 0 (their ZC^T column is zero too) and padding are zero.
__global__ __launch_bounds__(256) void sv_split_zt(const double* __restrict__ zt, long n, int k, long nr, int kp,
                                                   const double* __restrict__ s, const double* __restrict__ inv2,
                                                   signed char* __restrict__ SZ) {
  const long r = (long)blockIdx.x * 64 + (threadIdx.x >> 2);
  const int j0 = blockIdx.y * 32 + 8 * (threadIdx.x & 3);
  const double sr = r < n ? s[r] : 0.0;
  const double isr = sr > 0.0 ? 1.0 / sr : 0.0;
  double x[8];
  if (r < n && j0 + 8 <= k && (k & 1) == 0) {
    const double2* src = reinterpret_cast<const double2*>(zt + r * k + j0);
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const double2 v = src[h];
      x[2 * h] = v.x * inv2[j0 + 2 * h] * isr;
      x[2 * h + 1] = v.y * inv2[j0 + 2 * h + 1] * isr;
    }
  } else {
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = (r < n && j0 + u < k) ? zt[r * k + j0 + u] * inv2[j0 + u] * isr : 0.0;
  }
  const long plane = nr * kp;
  const long off = sv_off(r, j0, nr);
#pragma unroll
  for (int q = 0; q < SV_D; ++q) {
    unsigned w0 = 0, w1 = 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      x[u] *= 128.0;
      const double d = trunc(x[u]);
      x[u] -= d;
      if (u < 4) w0 |= ((unsigned)(int)d & 0xffu) << (8 * u);
      else w1 |= ((unsigned)(int)d & 0xffu) << (8 * (u - 4));
    }
    *reinterpret_cast<uint2*>(SZ + q * plane + off) = make_uint2(w0, w1);
  }
}

// s[r] = sqrt(sum_i a[r][i]^2 * sum_i b[r][i]^2) (b = null: s[r] = sum_i a[r][i]^2).
// One wave per row at a time: the lanes read consecutive entries of the row
// (coalesced) and the sums are combined across the wave.
__device__ inline double sv_row_sq(const double* __restrict__ row, int m, int lane) {
  double v = 0.0;
  for (int i = lane; i < m; i += 64) v = fma(row[i], row[i], v);
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__global__ __launch_bounds__(256) void sv_scale(const double* __restrict__ a, int ma, const double* __restrict__ b,
                                                int mb, long n, double* __restrict__ s) {
  const int lane = threadIdx.x & 63;
  const long waves = (long)gridDim.x * 4;
  for (long r = (long)blockIdx.x * 4 + (threadIdx.x >> 6); r < n; r += waves) {
    const double ra = sv_row_sq(a + r * ma, ma, lane);
    double v = ra;
    if (b) v = sqrt(ra * sv_row_sq(b + r * mb, mb, lane));
    if (lane == 0) s[r] = v;
  }
}

__device__ inline void sv_wait_vm(int ahead) {
  // s_waitcnt vmcnt(ahead * SV_LPW) with expcnt and lgkmcnt left free (gfx9
  // encoding: vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8)
  static_assert(SV_LPW <= 15, "vmcnt fits the low field");
  if (ahead == 0) __builtin_amdgcn_s_waitcnt(0x0F70);
  else __builtin_amdgcn_s_waitcnt(0x0F70 | SV_LPW);
}

// The k loop of sv_gemm.  The three stage buffers are separate LDS objects, so
// the compiler can tell that the direct loads into one never touch the
// fragments read from another and adds no wait for them; the waits on this
// wave's own loads are the explicit ones.
__device__ __forceinline__ void sv_mainloop(signed char* __restrict__ b0, signed char* __restrict__ b1,
                                            signed char* __restrict__ b2, const signed char* const (&src)[SV_LPW],
                                            const long (&inc)[SV_LPW], const int (&dst)[SV_LPW], int nst,
                                            int aoff, int boff, i16v (&acc)[SV_D]) {
  // stages past the last reload the last one into a buffer nobody reads again
  auto issue = [&](int st, signed char* buf) {
    const long sc = min(st, nst - 1);
#pragma unroll
    for (int j = 0; j < SV_LPW; ++j)
      __builtin_amdgcn_global_load_lds(src[j] + sc * inc[j], (lds_void*)(buf + dst[j]), 16, 0, 0);
  };
  auto compute = [&](const signed char* buf) {
    i4 bq[SV_D];
#pragma unroll
    for (int q = 0; q < SV_D; ++q) bq[q] = *reinterpret_cast<const i4*>(buf + boff + q * (SV_TN / 32) * 1024);
#pragma unroll
    for (int p = 0; p < SV_D; ++p) {
      const i4 ap = *reinterpret_cast<const i4*>(buf + aoff + p * (SV_TM / 32) * 1024);
#pragma unroll
      for (int q = 0; q < SV_D - p; ++q)
        acc[p + q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ap, bq[q], acc[p + q], 0, 0, 0);
    }
  };
  // stage st lives in buffer st % 3.  Before reading stage st every wave
  // waits for its own loads of st (the SV_LPW loads of st + 1 may still fly),
  // then the barrier makes all waves' loads visible and frees buffer
  // (st + 2) % 3, whose stage st - 1 everyone has read.
  auto step = [&](int st, const signed char* cur, signed char* ahead) {
    sv_wait_vm(1);
    asm volatile("s_barrier" ::: "memory");
    issue(st + 2, ahead);
    compute(cur);
  };
  issue(0, b0);
  issue(1, b1);
  // whole triples first: a conditional step inside the loop body makes the
  // compiler lose track of which buffer each load in flight fills, and it then
  // waits for all of them before every fragment read
  int st = 0;
  for (; st + 3 <= nst; st += 3) {
    step(st, b0, b2);
    step(st + 1, b1, b0);
    step(st + 2, b2, b1);
  }
  if (st < nst) step(st, b0, b2);
  if (st + 1 < nst) step(st + 1, b1, b0);
  sv_wait_vm(0);
}

template <class TX>
__global__ __launch_bounds__(512) void sv_gemm(const signed char* __restrict__ SA, const int* __restrict__ EA,
                                              const signed char* __restrict__ SZ, const double* __restrict__ s,
                                              int k, int kr, int kp, long n, long nr, TX* __restrict__ x, long ldx,
                                              int mt, long nt, long ngroups) {
  __shared__ __attribute__((aligned(16))) signed char L0[SV_NBLK * 1024], L1[SV_NBLK * 1024],
      L2[SV_NBLK * 1024];
  // XCD-aware order: groups of SV_GM mu-tiles x SV_GN r-tiles (32 workgroups, one per CU
  // of an XCD) are dealt to the XCDs in turn; workgroup i runs on XCD i % 8
  const long b = blockIdx.x;
  const long g = (b / 256) * 8 + (b % 8);  // group
  const int w = (int)((b / 8) % 32);       // member of the group
  if (g >= ngroups) return;
  const long ngm = (mt + SV_GM - 1) / SV_GM;
  const long tm = (g % ngm) * SV_GM + (w % SV_GM);
  const long tn = (g / ngm) * SV_GN + (w / SV_GM);
  if (tm >= mt || tn >= nt) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 3, wn = wave >> 2;
  const int l32 = lane & 31, lh = lane >> 5;

  // loads: block bl = wave + 8 j, one contiguous KB of a k-blocked plane; lane
  // L fills row L / 2, physical segment L % 2 of the LDS block from the
  // logical segment that belongs there.  Every wave issues SV_LPW loads per
  // stage (waves whose last block index is past the stage reload their
  // previous block: the same bytes to the same place), so the count of loads
  // in flight is the same at every barrier.
  const int lrow = lane >> 1;
  const int lseg = (lane & 1) ^ ((lrow >> 2) & 1) ^ ((lrow >> 3) & 1);
  const long planeA = (long)kr * kp, planeZ = nr * kp;
  const signed char* src[SV_LPW];
  long inc[SV_LPW];
  int dst[SV_LPW];
#pragma unroll
  for (int j = 0; j < SV_LPW; ++j) {
    const int bl = wave + 8 * j < SV_NBLK ? wave + 8 * j : wave + 8 * (j - 1);
    dst[j] = bl * 1024;
    if (bl < SV_ABLK) {
      const int p = bl / (SV_TM / 32), rb = bl % (SV_TM / 32);
      src[j] = SA + p * planeA + (tm * SV_TM + rb * 32 + lrow) * 32 + lseg * 16;
      inc[j] = (long)kr * 32;
    } else {
      const int q = (bl - SV_ABLK) / (SV_TN / 32), rb = (bl - SV_ABLK) % (SV_TN / 32);
      src[j] = SZ + q * planeZ + (tn * SV_TN + rb * 32 + lrow) * 32 + lseg * 16;
      inc[j] = nr * 32;
    }
  }
  const int nst = kp / SV_KS;
  i16v acc[SV_D];
#pragma unroll
  for (int d = 0; d < SV_D; ++d)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[d][v] = 0;
  // fragment reads: row l32 of the wave's 32-row block, logical segment lh
  const int aoff = wm * 1024 + l32 * 32 + sv_seg(l32, lh) * 16;
  const int boff = SV_ABLK * 1024 + wn * 1024 + l32 * 32 + sv_seg(l32, lh) * 16;
  sv_mainloop(L0, L1, L2, src, inc, dst, nst, aoff, boff, acc);
  // epilogue: acc[d][v] is row 8 (v / 4) + 4 lh + v % 4, column l32 of the block
  const long c = tn * SV_TN + wn * 32 + l32;
  if (c >= n) return;
  const double sc = s[c];
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const long r = tm * SV_TM + wm * 32 + 8 * (v / 4) + 4 * lh + (v % 4);
    if (r < k) {
      double sum = 0.0;
#pragma unroll
      for (int d = SV_D - 1; d >= 0; --d) sum += ldexp((double)acc[d][v], -7 * (d + 2));
      x[r * ldx + c] = (TX)(ldexp(sum, EA[r] + 1) * sc);
    }
  }
}

struct SvPlan {
  int kp, kr, mt;
  long nr, nt, ngroups;
};

SvPlan sv_plan(int k, long n) {
  SvPlan p;
  p.kp = (k + SV_KS - 1) / SV_KS * SV_KS;
  p.mt = (k + SV_TM - 1) / SV_TM;
  p.kr = p.mt * SV_TM;
  p.nt = (n + SV_TN - 1) / SV_TN;
  p.nr = p.nt * SV_TN;
  p.ngroups = (long)((p.mt + SV_GM - 1) / SV_GM) * ((p.nt + SV_GN - 1) / SV_GN);
  return p;
}

// the solve workspace: digit planes of A' and Zs, A' row exponents, s, inv2
struct SvWork {
  signed char *SA, *SZ;
  int* EA;
  double *sc, *inv2;
};
SvWork sv_carve(const SvPlan& P, long n, void* work) {
  SvWork W;
  char* w = reinterpret_cast<char*>(work);
  W.SA = reinterpret_cast<signed char*>(w);
  w += align_up((size_t)SV_D * P.kr * P.kp, 256);
  W.SZ = reinterpret_cast<signed char*>(w);
  w += align_up((size_t)SV_D * P.nr * P.kp, 256);
  W.EA = reinterpret_cast<int*>(w);
  w += align_up(sizeof(int) * (size_t)P.kr, 256);
  W.sc = reinterpret_cast<double*>(w);
  w += align_up(sizeof(double) * (size_t)n, 256);
  W.inv2 = reinterpret_cast<double*>(w);
  return W;
}

// s(r) and the point scales inv2
int sv_scales(const SvPlan& P, const SvWork& W, int k, long n, const int* idx, const double* phi_a, int ma,
              const double* phi_b, int mb, hipStream_t s) {
  const int gs = (int)std::min<long>(8192, (n + 255) / 256);
  hipLaunchKernelGGL(sv_scale, dim3(gs), dim3(256), 0, s, phi_a, ma, phi_b, mb, n, W.sc);
  hipLaunchKernelGGL(sv_point_scale, dim3((P.kp + 255) / 256), dim3(256), 0, s, W.sc, n, idx, k, P.kp, W.inv2);
  THC_HIP(hipGetLastError());
  return THC_OK;
}

}  // namespace

size_t oz_solve_work_bytes(int k, long n) {
  const SvPlan P = sv_plan(k, n);
  return align_up((size_t)SV_D * P.kr * P.kp, 256) + align_up((size_t)SV_D * P.nr * P.kp, 256) +
         align_up(sizeof(int) * (size_t)P.kr, 256) + align_up(sizeof(double) * (size_t)n, 256) +
         align_up(sizeof(double) * (size_t)P.kp, 256) + 1024;
}

int oz_solve_prepare(int k, long n, const int* idx, const double* phi_a, int ma, void* work, GemmArgs& g,
                     hipStream_t s) {
  if (k <= 0 || n <= 0) return THC_OK;
  const SvPlan P = sv_plan(k, n);
  const SvWork W = sv_carve(P, n, work);
  THC_TRY(sv_scales(P, W, k, n, idx, phi_a, ma, nullptr, 0, s));
  g.epi_digits = 1;
  g.dig_out = W.SZ;
  g.dig_s = W.sc;
  g.dig_inv2 = W.inv2;
  g.dig_nr = P.nr;
  g.dig_kp = P.kp;
  return THC_OK;
}

namespace {
int oz_solve_impl(const double* ainv, long lda, const double* zt, int k, long n, const int* idx,
                  const double* phi_a, int ma, const double* phi_b, int mb, void* x, int x_f32, long ldx,
                  void* work, hipStream_t s, hipEvent_t e0, hipEvent_t e1, int zt_split) {
  const SvPlan P = sv_plan(k, n);
  const SvWork W = sv_carve(P, n, work);
  signed char* SA = W.SA;
  signed char* SZ = W.SZ;
  int* EA = W.EA;
  double* sc = W.sc;
  if (!zt_split) THC_TRY(sv_scales(P, W, k, n, idx, phi_a, ma, phi_b, mb, s));
  hipLaunchKernelGGL(sv_split_rows, dim3(P.kr), dim3(256), 0, s, ainv, lda, k, P.kr, P.kp, sc, n, idx, SA, EA);
  THC_HIP(hipGetLastError());
  if (!zt_split) {
    hipLaunchKernelGGL(sv_split_zt, dim3((unsigned)(P.nr / 64), (unsigned)(P.kp / 32)), dim3(256), 0, s, zt, n, k,
                       P.nr, P.kp, sc, W.inv2, SZ);
    THC_HIP(hipGetLastError());
  }
  const long blocks = (P.ngroups + 7) / 8 * 256;  // every group gets its 32 workgroups on one XCD
  if (blocks > 0x7fffffffL) return THC_ERR_ARG;
  if (e0) THC_HIP(hipEventRecord(e0, s));
  if (x_f32)
    hipLaunchKernelGGL(sv_gemm<float>, dim3((unsigned)blocks), dim3(512), 0, s, SA, EA, SZ, sc, k, P.kr, P.kp, n,
                       P.nr, (float*)x, ldx, P.mt, P.nt, P.ngroups);
  else
    hipLaunchKernelGGL(sv_gemm<double>, dim3((unsigned)blocks), dim3(512), 0, s, SA, EA, SZ, sc, k, P.kr, P.kp, n,
                       P.nr, (double*)x, ldx, P.mt, P.nt, P.ngroups);
  THC_HIP(hipGetLastError());
  if (e1) THC_HIP(hipEventRecord(e1, s));
  return THC_OK;
}
}  // namespace

int oz_solve(const double* ainv, long lda, const double* zt, int k, long n, const int* idx, const double* phi_a,
             int ma, const double* phi_b, int mb, void* x, int x_f32, long ldx, void* work, hipStream_t s,
             hipEvent_t e0, hipEvent_t e1, int zt_split) {
  if (k <= 0 || n <= 0) return THC_OK;
  return oz_solve_impl(ainv, lda, zt, k, n, idx, phi_a, ma, phi_b, mb, x, x_f32, ldx, work, s, e0, e1, zt_split);
}

}  // namespace thc
