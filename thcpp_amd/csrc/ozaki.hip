// V = B B^T for the mixed-precision path, evaluated exactly in int8 MFMA
// arithmetic on splits of B (an Ozaki-style emulation of the FP64 Gram).
//
// Split: every row of B is cut into chunks of OZ_CH columns; each (row, chunk)
// gets a power-of-two exponent e with max |b| < 2^e, and b 2^-e in (-1, 1) is
// written as OZ_S signed 7-bit digits, b = 2^e sum_p d_p 2^-7(p+1) + r,
// |r| < 2^(e - 7 OZ_S).  Every step (scale by 128, truncate, subtract) is
// exact in double.
// Product: for one chunk, sum_g b_rg b_cg = 2^(e_r + e_c) sum_{p+q<OZ_S}
// 2^-7(p+q+2) sum_g d_p,rg d_q,cg.  The inner sums are int8 x int8 products
// accumulated in int32 by v_mfma_i32_16x16x64_i8, grouped by p + q (products
// with the same p + q share a scale), exact as long as
// OZ_S * OZ_CH * 127^2 < 2^31.  Digit products with p + q >= OZ_S are
// dropped; they are below 2^-7(OZ_S+1) of |b_r| |b_c| per term.
// Each workgroup takes one 128 x 128 lower tile and one chunk, converts its
// groups to double, scales by 2^(e_r + e_c) and writes a partial tile; the
// partials are summed in chunk order in FP64.
// Slabs (5 digit planes x 128 rows x 64 columns per operand) are double
// buffered in LDS (2 x 80 KB, the whole LDS of a CU) and filled by direct
// global-to-LDS loads (global_load_lds_dwordx4: lane L of a wave writes bytes
// 16 L .. 16 L + 15 of the wave's 1 KB block), so the next step's slabs stream
// in while the current step's MFMAs run.  Rows are 64 bytes with no padding;
// 16-byte segment s of row r is stored at segment (s + r / 4) mod 4, which
// keeps the 16-row fragment reads free of bank conflicts.
#include "common.h"

namespace thc {
namespace {

typedef int i4 __attribute__((ext_vector_type(4)));

constexpr int OZ_S = 5;       // digits per value
constexpr int OZ_CH = 8192;   // columns per chunk (one exponent per row and chunk)
constexpr int OZ_T = 128;     // tile edge
constexpr int OZ_KS = 64;     // k per MFMA step
static_assert((long)OZ_S * OZ_CH * 127 * 127 < 2147483647L, "int32 group sums must not overflow");

__device__ inline double block_max(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  v = red[0];
  for (int i = 1; i < (int)(blockDim.x >> 6); ++i) v = fmax(v, red[i]);
  return v;
}

// four consecutive values of a row as doubles (vector loads when aligned)
__device__ inline void load4(const float* p, bool vec, long g, long g2, double (&v)[4]) {
  if (vec && g + 4 <= g2) {
    const float4 q = *reinterpret_cast<const float4*>(p + g);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = g + j < g2 ? (double)p[g + j] : 0.0;
  }
}
__device__ inline void load4(const double* p, bool vec, long g, long g2, double (&v)[4]) {
  if (vec && g + 4 <= g2) {
    const double2 q0 = *reinterpret_cast<const double2*>(p + g), q1 = *reinterpret_cast<const double2*>(p + g + 2);
    v[0] = q0.x; v[1] = q0.y; v[2] = q1.x; v[3] = q1.y;
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = g + j < g2 ? p[g + j] : 0.0;
  }
}

// digits S[p][row][g] (g < g2p, zero padded) and exponents E[row][chunk].
// Thread t takes runs of 4 consecutive columns (one vector load), and stores
// each digit plane's 4 bytes as one word.
template <class T>
__global__ __launch_bounds__(256) void oz_split(const T* __restrict__ b, long ld, long g2, long g2p, int k,
                                                int nchunk, signed char* __restrict__ S, int* __restrict__ E) {
  __shared__ double red[4];
  const int c = blockIdx.x, row = blockIdx.y;
  const T* src = b + (long)row * ld;
  const long g0 = (long)c * OZ_CH;
  constexpr int GRP = OZ_CH / (256 * 4);
  const bool vec = ((reinterpret_cast<uintptr_t>(src) & 15) == 0);
  double v[GRP][4];
  double mx = 0.0;
#pragma unroll
  for (int q = 0; q < GRP; ++q) {
    load4(src, vec, g0 + ((long)q * 256 + threadIdx.x) * 4, g2, v[q]);
#pragma unroll
    for (int j = 0; j < 4; ++j) mx = fmax(mx, fabs(v[q][j]));
  }
  mx = block_max(mx, red);
  int ex = 0;
  if (mx > 0.0) frexp(mx, &ex);  // mx < 2^ex
  if (threadIdx.x == 0) E[(long)row * nchunk + c] = ex;
#pragma unroll
  for (int q = 0; q < GRP; ++q) {
    const long g = g0 + ((long)q * 256 + threadIdx.x) * 4;
    double x[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) x[j] = ldexp(v[q][j], -ex);
#pragma unroll
    for (int p = 0; p < OZ_S; ++p) {
      unsigned w = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        x[j] *= 128.0;
        const double d = trunc(x[j]);
        x[j] -= d;
        w |= ((unsigned)(int)d & 0xffu) << (8 * j);
      }
      *reinterpret_cast<unsigned*>(S + ((long)p * k + row) * g2p + g) = w;
    }
  }
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
                                               const int* __restrict__ E, int nchunk, double* __restrict__ work,
                                               long ntiles) {
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
  const long plane = (long)k * g2p, kbase = (long)c * OZ_CH;
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
  const int nks = (int)min((long)(OZ_CH / OZ_KS), (g2 - kbase + OZ_KS - 1) / OZ_KS);
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
          val = ldexp(sum, E[(long)r * nchunk + c] + E[(long)col * nchunk + c]);
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

void oz_plan(int k, long g2, int& nchunk, long& g2p, long& ntiles) {
  nchunk = (int)((g2 + OZ_CH - 1) / OZ_CH);
  g2p = (long)nchunk * OZ_CH;
  const long nb = (k + OZ_T - 1) / OZ_T;
  ntiles = nb * (nb + 1) / 2;
}

}  // namespace

size_t oz_syrk_work_bytes(int k, long g2) {
  int nchunk;
  long g2p, ntiles;
  oz_plan(k, g2, nchunk, g2p, ntiles);
  return align_up((size_t)OZ_S * k * g2p, 256) + align_up(sizeof(int) * (size_t)k * nchunk, 256) +
         align_up(sizeof(double) * (size_t)nchunk * ntiles * OZ_T * OZ_T, 256) + 256;
}

int oz_syrk(const void* b, int k, long g2, long ld, int prec, double* v, long ldv, void* work, hipStream_t s) {
  if (k <= 0) return THC_OK;
  int nchunk;
  long g2p, ntiles;
  oz_plan(k, g2, nchunk, g2p, ntiles);
  char* w = reinterpret_cast<char*>(work);
  signed char* S = reinterpret_cast<signed char*>(w);
  w += align_up((size_t)OZ_S * k * g2p, 256);
  int* E = reinterpret_cast<int*>(w);
  w += align_up(sizeof(int) * (size_t)k * nchunk, 256);
  double* part = reinterpret_cast<double*>(w);
  if (prec == 0)
    hipLaunchKernelGGL(oz_split<double>, dim3(nchunk, k), dim3(256), 0, s, (const double*)b, ld, g2, g2p, k, nchunk,
                       S, E);
  else
    hipLaunchKernelGGL(oz_split<float>, dim3(nchunk, k), dim3(256), 0, s, (const float*)b, ld, g2, g2p, k, nchunk,
                       S, E);
  THC_HIP(hipGetLastError());
  const long slots = (ntiles * nchunk + 7) / 8 * 8;  // every w < ntiles * nchunk gets one workgroup
  if (slots > 0x7fffffffL) return THC_ERR_ARG;
  hipLaunchKernelGGL(oz_syrk, dim3((unsigned)slots), dim3(512), 0, s, S, g2, g2p, k, E, nchunk, part, ntiles);
  THC_HIP(hipGetLastError());
  hipLaunchKernelGGL(oz_reduce, dim3((unsigned)(ntiles * (OZ_T / RB) * (OZ_T / RB))), dim3(256), 0, s, part, nchunk,
                     ntiles, k, v, ldv);
  THC_HIP(hipGetLastError());
  return THC_OK;
}

}  // namespace thc
