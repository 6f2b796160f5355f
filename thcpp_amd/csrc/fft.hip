// Batched 3-D real-to-complex FFT of the interpolating vectors with the
// Coulomb weighting fused into the last pass (replaces the per-column FFTW
// loop and the sqrt(coulG) scaling of the reference,
// src/interpolating_vectors.cpp:371-383 and 258-263).
//
// Convention: unnormalised forward transform, X(k) = sum_r x(r) e^{-2 pi i k.r/n},
// over the row-major grid (n0, n1, n2), z fastest.  Only kz <= n2/2 is kept
// (x is real).  Output row mu holds, for every kept G = (kx, ky, kz) in C
// order, (Re, Im) * sqrt(w coulG(G)) with w = 1 when kz = 0 or 2 kz = n2
// (those planes hold both members of each conjugate pair) and 2 otherwise, so
// sum_G conj(A_mu(G)) coulG(G) A_nu(G) over the full grid equals the real dot
// product of rows mu and nu.
//
// Pass A, one workgroup per (mu, x-plane): two real z-rows are packed as one
// complex row, transformed along z, unpacked into the half spectrum, then
// transformed along y.  Pass B, one workgroup per (mu, y-index): transform along
// x, weight, store -- or, for the mixed path's V, cut the weighted line into
// int8 digits with one exponent per line and store the digit planes
// (FftDigits), so the spectrum itself is never written.  Both passes run a
// Stockham autosort FFT in LDS with radix
// 8, 4, 2, 3 and 5 stages; twiddles come from a table of e^{-2 pi i j/n}.
//
// LDS layout: every transform line is contiguous, lines ls = n rounded up to
// 32 elements apart, and element pos of line l sits at l ls + (pos ^ h(l)).
// The XOR spreads the lines of one wave over the banks: a Stockham read takes
// elements t + q nb of 8 lines (t < 8 across lanes) and a first-stage write
// elements 8 t + q; h puts the line in bits 0-2 and bits 3-4 of the position,
// so both touch every bank pair once per 32 lanes.  The transposes between
// transforms (z-lines to y-lines, the strided x-lines of the global layout)
// are done on the way into LDS, so no stage reads a strided line.
//
// The passes are bound by their VALU work (index arithmetic of the LDS
// stages as much as the butterflies), so cubic grids of the common edges (32,
// 64, 96) run instantiations compiled for their size: every divisor, stride,
// twiddle step and loop bound is a constant (64^3 FP32 pass A: 1.60 -> 1.18 ms).
#include <tuple>
#include <algorithm>
#include <type_traits>

#include "common.h"

namespace thc {
namespace {

template <class T> struct C2 { T x, y; };

template <class T>
__device__ inline C2<T> cmul(C2<T> a, C2<T> b) {
  return C2<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class T>
__device__ inline C2<T> cadd(C2<T> a, C2<T> b) { return C2<T>{a.x + b.x, a.y + b.y}; }
template <class T>
__device__ inline C2<T> csub(C2<T> a, C2<T> b) { return C2<T>{a.x - b.x, a.y - b.y}; }

constexpr int MAXR = 8;  // max stages

// 4-point forward DFT (e^{-2 pi i / 4} = -i)
template <class T>
__device__ inline void dft4(C2<T> x0, C2<T> x1, C2<T> x2, C2<T> x3, C2<T>& y0, C2<T>& y1, C2<T>& y2, C2<T>& y3) {
  const C2<T> a = cadd(x0, x2), b = csub(x0, x2);
  const C2<T> c = cadd(x1, x3), d = csub(x1, x3);
  y0 = cadd(a, c);
  y2 = csub(a, c);
  y1 = C2<T>{b.x + d.y, b.y - d.x};  // b - i d
  y3 = C2<T>{b.x - d.y, b.y + d.x};  // b + i d
}

struct Plan1D {
  int n;
  int nst;
  int rad[MAXR];
};

struct FftParams {
  int k, n0, n1, n2, n2h;
  int lsz, lsy, lsx;  // LDS line strides of the z, y and x transforms
  Plan1D pz, py, px;
  const void* twz; const void* twy; const void* twx;
  int vec_out;  // output aligned to two elements: one vector store per complex value
  signed char* dig_s;  // digit output of pass B (FftDigits): planes, exponents, strides
  int* dig_e;
  long dig_plane, dig_row;
  int dig_lp;
};

// Division by a run-time divisor d >= 1 of dividends e with e * d < 2^32:
// q = mulhi(e, floor((2^32 - 1) / d) + 1) is exact there, one multiply in place
// of a division sequence (the index arithmetic of every butterfly and load).
// Here e < 2^14 and d <= 128.
struct FDiv {
  unsigned d, m;
  __device__ explicit FDiv(int dd) : d((unsigned)dd), m(dd > 1 ? 0xFFFFFFFFu / (unsigned)dd + 1u : 0u) {}
  __device__ int div(int e) const { return d > 1 ? (int)__umulhi((unsigned)e, m) : e; }
};

inline __host__ __device__ int line_stride(int n) { return (n + 31) / 32 * 32; }

// LDS index of element pos of line 'line' (pos < ls, ls a multiple of 32, so
// pos ^ h stays inside the line)
__device__ inline int swz(int line, int pos, int ls) {
  const int h = (line & 7) | ((((line & 3) ^ ((line >> 3) & 3))) << 3);
  return line * ls + (pos ^ h);
}

// Radix-R butterfly of one Stockham stage: twiddles w^(j q tstep) on the
// inputs (ns > 1), then the R-point DFT.
template <class T, int R>
__device__ inline void butterfly(C2<T> (&x)[R], C2<T> (&y)[R], int n, int ns, int j, int tstep,
                                 const C2<T>* __restrict__ tw) {
  if (ns > 1) {
#pragma unroll
    for (int q = 1; q < R; ++q) x[q] = cmul(x[q], tw[j * q * tstep]);
  }
  if (R == 2) {
    y[0] = cadd(x[0], x[1]);
    y[1] = csub(x[0], x[1]);
  } else if (R == 4) {
    dft4(x[0], x[1], x[2], x[3], y[0], y[1], y[2], y[3]);
  } else if (R == 8) {
    // even / odd halves by radix 4, joined with w8^p = e^{-2 pi i p / 8}
    C2<T> e[4], o[4];
    dft4(x[0], x[2], x[4], x[6], e[0], e[1], e[2], e[3]);
    dft4(x[1], x[3], x[5], x[7], o[0], o[1], o[2], o[3]);
    const T h = T(0.70710678118654752440);
    const C2<T> w1 = C2<T>{(o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h};  // w8 o1
    const C2<T> w2 = C2<T>{o[2].y, -o[2].x};                                // -i o2
    const C2<T> w3 = C2<T>{(o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h};  // w8^3 o3
    y[0] = cadd(e[0], o[0]);
    y[4] = csub(e[0], o[0]);
    y[1] = cadd(e[1], w1);
    y[5] = csub(e[1], w1);
    y[2] = cadd(e[2], w2);
    y[6] = csub(e[2], w2);
    y[3] = cadd(e[3], w3);
    y[7] = csub(e[3], w3);
  } else if (R == 3) {
    // w = e^{-2 pi i / 3} = -1/2 - i sqrt(3)/2: y1,2 = x0 - t1/2 -/+ i sqrt(3)/2 t2
    const T s3 = T(0.86602540378443864676);
    const C2<T> t1 = cadd(x[1], x[2]), t2 = csub(x[1], x[2]);
    const C2<T> m = C2<T>{x[0].x - T(0.5) * t1.x, x[0].y - T(0.5) * t1.y};
    y[0] = cadd(x[0], t1);
    y[1] = C2<T>{m.x + s3 * t2.y, m.y - s3 * t2.x};
    y[2] = C2<T>{m.x - s3 * t2.y, m.y + s3 * t2.x};
  } else if (R == 5) {
    // w = e^{-2 pi i / 5} = c1 - i s1, w^2 = c2 - i s2; with a1,2 = x1 + x4, x2 + x3
    // and b1,2 = x1 - x4, x2 - x3: y1,4 = x0 + c1 a1 + c2 a2 -/+ i (s1 b1 + s2 b2),
    // y2,3 = x0 + c2 a1 + c1 a2 -/+ i (s2 b1 - s1 b2)
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
    const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
    const C2<T> a1 = cadd(x[1], x[4]), b1 = csub(x[1], x[4]);
    const C2<T> a2 = cadd(x[2], x[3]), b2 = csub(x[2], x[3]);
    y[0] = cadd(x[0], cadd(a1, a2));
    const C2<T> m1 = C2<T>{x[0].x + c1 * a1.x + c2 * a2.x, x[0].y + c1 * a1.y + c2 * a2.y};
    const C2<T> m2 = C2<T>{x[0].x + c2 * a1.x + c1 * a2.x, x[0].y + c2 * a1.y + c1 * a2.y};
    const C2<T> n1 = C2<T>{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
    const C2<T> n2 = C2<T>{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
    y[1] = C2<T>{m1.x + n1.y, m1.y - n1.x};
    y[4] = C2<T>{m1.x - n1.y, m1.y + n1.x};
    y[2] = C2<T>{m2.x + n2.y, m2.y - n2.x};
    y[3] = C2<T>{m2.x - n2.y, m2.y + n2.x};
  } else {
    const int step = n / R;
#pragma unroll
    for (int p = 0; p < R; ++p) {
      C2<T> acc = x[0];
#pragma unroll
      for (int q = 1; q < R; ++q) acc = cadd(acc, cmul(x[q], tw[((p * q) % R) * step]));
      y[p] = acc;
    }
  }
}

// One Stockham stage over 'nlines' lines of length n, read src write dst.
template <class T, int R>
__device__ inline void stage(const C2<T>* __restrict__ src, C2<T>* __restrict__ dst, int n, int ns, int nlines,
                             int ls, const C2<T>* __restrict__ tw) {
  const int nb = n / R;
  const int total = nlines * nb;
  const int tstep = n / (ns * R);
  const FDiv fnb(nb), fns(ns);
  for (int w = threadIdx.x; w < total; w += blockDim.x) {
    const int line = fnb.div(w), t = w - line * nb;
    const int tq = fns.div(t), j = t - tq * ns;
    C2<T> x[R], y[R];
#pragma unroll
    for (int q = 0; q < R; ++q) x[q] = src[swz(line, t + q * nb, ls)];
    butterfly<T, R>(x, y, n, ns, j, tstep, tw);
    const int d0 = tq * ns * R + j;
#pragma unroll
    for (int q = 0; q < R; ++q) dst[swz(line, d0 + q * ns, ls)] = y[q];
  }
}

// Full 1-D transforms of 'nlines' lines; returns the buffer holding the result.
template <class T>
__device__ C2<T>* fft_lines(C2<T>* a, C2<T>* b, const Plan1D& pl, int nlines, int ls, const C2<T>* tw) {
  int ns = 1;
  for (int s = 0; s < pl.nst; ++s) {
    const int r = pl.rad[s];
    if (r == 8) stage<T, 8>(a, b, pl.n, ns, nlines, ls, tw);
    else if (r == 4) stage<T, 4>(a, b, pl.n, ns, nlines, ls, tw);
    else if (r == 2) stage<T, 2>(a, b, pl.n, ns, nlines, ls, tw);
    else if (r == 3) stage<T, 3>(a, b, pl.n, ns, nlines, ls, tw);
    else stage<T, 5>(a, b, pl.n, ns, nlines, ls, tw);
    __syncthreads();
    ns *= r;
    C2<T>* t = a; a = b; b = t;
  }
  return a;
}

// In-place variant: each thread reads the inputs of all its butterflies (at
// most IPV / R of them) into registers, then after a barrier writes the
// outputs into the same buffer.  One LDS buffer instead of two lets a CU hold
// up to twice the workgroups, which is what these latency-bound passes need;
// the smaller IPV budget (16 complex values) keeps FP32 passes at about 64
// VGPRs and seven workgroups per CU, 24 covers 96-point lines.
constexpr int FT = 256;  // threads per workgroup of both passes

template <class T, int R, int IPV>
__device__ inline void stage_ip(C2<T>* buf, int n, int ns, int nlines, int ls, const C2<T>* __restrict__ tw) {
  constexpr int NB = IPV / R > 0 ? IPV / R : 1;
  const int nb = n / R;
  const int total = nlines * nb;
  const int tstep = n / (ns * R);
  const FDiv fnb(nb), fns(ns);
  C2<T> x[NB][R];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int w = threadIdx.x + i * FT;
    if (w < total) {
      const int line = fnb.div(w), t = w - line * nb;
#pragma unroll
      for (int q = 0; q < R; ++q) x[i][q] = buf[swz(line, t + q * nb, ls)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int w = threadIdx.x + i * FT;
    if (w < total) {
      const int line = fnb.div(w), t = w - line * nb;
      const int tq = fns.div(t), j = t - tq * ns;
      C2<T> y[R];
      butterfly<T, R>(x[i], y, n, ns, j, tstep, tw);
      const int d0 = tq * ns * R + j;
#pragma unroll
      for (int q = 0; q < R; ++q) buf[swz(line, d0 + q * ns, ls)] = y[q];
    }
  }
  __syncthreads();
}

template <class T, int IPV>
__device__ void fft_lines_ip(C2<T>* buf, const Plan1D& pl, int nlines, int ls, const C2<T>* tw) {
  int ns = 1;
  for (int s = 0; s < pl.nst; ++s) {
    const int r = pl.rad[s];
    if (r == 8) stage_ip<T, 8, IPV>(buf, pl.n, ns, nlines, ls, tw);
    else if (r == 4) stage_ip<T, 4, IPV>(buf, pl.n, ns, nlines, ls, tw);
    else if (r == 2) stage_ip<T, 2, IPV>(buf, pl.n, ns, nlines, ls, tw);
    else if (r == 3) stage_ip<T, 3, IPV>(buf, pl.n, ns, nlines, ls, tw);
    else stage_ip<T, 5, IPV>(buf, pl.n, ns, nlines, ls, tw);
    ns *= r;
  }
}

// Whether every stage of a plan over nlines lines keeps to ipv values per thread.
bool fits_ip(const Plan1D& pl, int nlines, int ipv) {
  for (int s = 0; s < pl.nst; ++s) {
    const int r = pl.rad[s];
    const long per = ((long)nlines * (pl.n / r) + FT - 1) / FT;
    if (per > ipv / r) return false;
  }
  return true;
}
constexpr int UPV = 12;  // elements per thread of the in-place unpack

// The plan of a transform length known at compile time (make_plan's rule:
// radix 8, then 4, 2, 3, 5).
constexpr Plan1D plan_of(int n) {
  Plan1D pl{};
  pl.n = n;
  pl.nst = 0;
  int m = n;
  const int rs[5] = {8, 4, 2, 3, 5};
  for (int i = 0; i < 5; ++i)
    while (m % rs[i] == 0 && pl.nst < MAXR) {
      pl.rad[pl.nst++] = rs[i];
      m /= rs[i];
    }
  return pl;
}

// fft_lines_ip for a length N fixed at compile time: the stages unrolled, so
// every index divisor, twiddle stride and loop bound is a constant.
template <class T, int IPV, int N, int S = 0, int NS = 1>
__device__ __forceinline__ void fft_lines_ct(C2<T>* buf, int nlines, int ls, const C2<T>* tw) {
  constexpr Plan1D pl = plan_of(N);
  if constexpr (S < pl.nst) {
    constexpr int R = pl.rad[S];
    stage_ip<T, R, IPV>(buf, N, NS, nlines, ls, tw);
    fft_lines_ct<T, IPV, N, S + 1, NS * R>(buf, nlines, ls, tw);
  }
}

// x: (k, n0, n1, n2) real, FP64 or FP32 (rounded to T as it is read); t: (k,
// n0, n1, n2h) complex.  CN > 0: a cubic grid of edge CN, compiled for that
// size (in-place stages only); 0: sizes and plans from p.
template <class T, int IPV, class X, int CN = 0>  // IPV = 0: ping-pong stages
__global__ __launch_bounds__(FT) void pass_a(const X* __restrict__ x, C2<T>* __restrict__ t, FftParams p) {
  constexpr bool ip = IPV > 0;
  static_assert(CN == 0 || ip, "compile-time sizes use the in-place stages");
  extern __shared__ unsigned char smem_raw[];
  constexpr int CLS = (CN + 31) / 32 * 32;
  const int n1 = CN ? CN : p.n1, n2 = CN ? CN : p.n2, n2h = CN ? CN / 2 + 1 : p.n2h;
  const int lsz = CN ? CLS : p.lsz, lsy = CN ? CLS : p.lsy;
  const int n1p = (n1 + 1) / 2;
  const int cap = max(n1p * lsz, n2h * lsy);
  C2<T>* bufa = reinterpret_cast<C2<T>*>(smem_raw);
  C2<T>* bufb = ip ? bufa : bufa + cap;  // in-place stages use one buffer
  C2<T>* twz = bufb + cap;  // twiddle tables staged in LDS (each stage reads them)
  C2<T>* twy = twz + n2;
  const long mu = blockIdx.y, ix = blockIdx.x;
  for (int j = threadIdx.x; j < n2; j += blockDim.x) twz[j] = reinterpret_cast<const C2<T>*>(p.twz)[j];
  for (int j = threadIdx.x; j < n1; j += blockDim.x) twy[j] = reinterpret_cast<const C2<T>*>(p.twy)[j];
  const int n0 = CN ? CN : p.n0;
  const X* src = x + (mu * n0 + ix) * (long)n1 * n2;
  // pack rows (2r, 2r+1) as re/im; LB loads per thread are issued before the
  // first LDS store so the plane is fetched in one or two memory round trips
  constexpr int LB = 8;
  const int tot = n1p * n2;
  const FDiv fn2(n2), fn2h(n2h);
  for (int e0 = threadIdx.x; e0 < tot; e0 += LB * blockDim.x) {
    X re[LB], im[LB];
#pragma unroll
    for (int b = 0; b < LB; ++b) {
      const int e = e0 + b * (int)blockDim.x;
      const int r = fn2.div(e), z = e - r * n2;
      const bool ok = e < tot, ok1 = ok && 2 * r + 1 < n1;
      re[b] = src[ok ? (long)(2 * r) * n2 + z : 0];
      im[b] = ok1 ? src[(long)(2 * r + 1) * n2 + z] : X(0);
    }
#pragma unroll
    for (int b = 0; b < LB; ++b) {
      const int e = e0 + b * (int)blockDim.x;
      const int r = fn2.div(e), z = e - r * n2;
      if (e < tot) bufa[swz(r, z, lsz)] = C2<T>{(T)re[b], (T)im[b]};
    }
  }
  __syncthreads();
  // unpack: A(k) = (Z(k) + conj Z(-k)) / 2, B(k) = (Z(k) - conj Z(-k)) / (2i),
  // written transposed: y-line kz holds A and B of every row pair
  C2<T>* fin;
  if (ip) {
    if constexpr (CN > 0) fft_lines_ct<T, IPV, CN>(bufa, n1p, lsz, twz);
    else fft_lines_ip<T, IPV>(bufa, p.pz, n1p, lsz, twz);
    C2<T> z1[UPV], z2[UPV];
#pragma unroll
    for (int i = 0; i < UPV; ++i) {
      const int e = threadIdx.x + i * FT;
      if (e < n1p * n2h) {
        const int r = fn2h.div(e), kz = e - r * n2h;
        z1[i] = bufa[swz(r, kz, lsz)];
        z2[i] = bufa[swz(r, kz == 0 ? 0 : n2 - kz, lsz)];
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < UPV; ++i) {
      const int e = threadIdx.x + i * FT;
      if (e < n1p * n2h) {
        const int r = fn2h.div(e), kz = e - r * n2h;
        bufa[swz(kz, 2 * r, lsy)] = C2<T>{(z1[i].x + z2[i].x) * T(0.5), (z1[i].y - z2[i].y) * T(0.5)};
        if (2 * r + 1 < n1)
          bufa[swz(kz, 2 * r + 1, lsy)] = C2<T>{(z1[i].y + z2[i].y) * T(0.5), (z2[i].x - z1[i].x) * T(0.5)};
      }
    }
    __syncthreads();
    if constexpr (CN > 0) fft_lines_ct<T, IPV, CN>(bufa, n2h, lsy, twy);
    else fft_lines_ip<T, IPV>(bufa, p.py, n2h, lsy, twy);
    fin = bufa;
  } else {
    C2<T>* res = fft_lines<T>(bufa, bufb, p.pz, n1p, lsz, twz);
    C2<T>* h = (res == bufa) ? bufb : bufa;
    for (int e = threadIdx.x; e < n1p * n2h; e += blockDim.x) {
      const int r = fn2h.div(e), kz = e - r * n2h;
      const C2<T> z1 = res[swz(r, kz, lsz)];
      const C2<T> z2 = res[swz(r, kz == 0 ? 0 : n2 - kz, lsz)];
      const C2<T> a = C2<T>{(z1.x + z2.x) * T(0.5), (z1.y - z2.y) * T(0.5)};
      const C2<T> bb = C2<T>{(z1.y + z2.y) * T(0.5), (z2.x - z1.x) * T(0.5)};
      h[swz(kz, 2 * r, lsy)] = a;
      if (2 * r + 1 < n1) h[swz(kz, 2 * r + 1, lsy)] = bb;
    }
    __syncthreads();
    C2<T>* other = (h == bufa) ? bufb : bufa;
    fin = fft_lines<T>(h, other, p.py, n2h, lsy, twy);
  }
  C2<T>* dst = t + (mu * n0 + ix) * (long)n1 * n2h;
  for (int e = threadIdx.x; e < n1 * n2h; e += blockDim.x) {
    const int y = fn2h.div(e), kz = e - y * n2h;
    dst[e] = fin[swz(kz, y, lsy)];
  }
}

// Half-spectrum weights sqrt(w(kz) coulG(kx, ky, kz)) rounded to T, w = 1 for
// kz = 0 and kz = n2 / 2, else 2 (the conjugate half the R2C layout omits),
// computed once per call for all k vectors.
template <class T>
__global__ void weight_kernel(const double* __restrict__ coulG, int n1, int n2, int n2h, long total,
                              T* __restrict__ wt) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long row = i / n2h;  // kx * n1 + ky
  const int kz = (int)(i - row * n2h);
  const double w = (kz == 0 || 2 * kz == n2) ? 1.0 : 2.0;
  wt[i] = (T)sqrt(w * coulG[row * n2 + kz]);
}

// t: (k, n0, n1, n2h) complex -> out (k, n0, n1, n2h, 2) weighted by wt, or
// (DIG) its int8 digit planes (FftDigits).
template <class T, int IPV, bool DIG, int CN = 0>  // IPV = 0: ping-pong stages; CN as in pass_a
__global__ __launch_bounds__(FT) void pass_b(const C2<T>* __restrict__ t, const T* __restrict__ wt,
                                             T* __restrict__ out, FftParams p) {
  constexpr bool ip = IPV > 0;
  static_assert(CN == 0 || ip, "compile-time sizes use the in-place stages");
  extern __shared__ unsigned char smem_raw[];
  const int n0 = CN ? CN : p.n0, n1 = CN ? CN : p.n1, n2h = CN ? CN / 2 + 1 : p.n2h;
  const int lsx = CN ? (CN + 31) / 32 * 32 : p.lsx;
  C2<T>* bufa = reinterpret_cast<C2<T>*>(smem_raw);
  C2<T>* bufb = ip ? bufa : bufa + n2h * lsx;
  C2<T>* twx = bufb + n2h * lsx;  // twiddle table staged in LDS
  for (int j = threadIdx.x; j < n0; j += blockDim.x) twx[j] = reinterpret_cast<const C2<T>*>(p.twx)[j];
  const long mu = blockIdx.y;
  const int iy = blockIdx.x;
  const C2<T>* src = t + mu * (long)n0 * n1 * n2h;
  // loads in flight per thread: 8 for run-time sizes; the compile-time sizes
  // unroll the loads of every batch into registers, and 3 keeps those kernels
  // at about 50 VGPRs, eight waves per SIMD (64^3 FP32 with digits: 1.64 ->
  // 1.34 ms against 8 at 112 VGPRs)
  constexpr int LB = CN > 0 ? 3 : 8;
  const int tot = n0 * n2h;
  const FDiv fn2h(n2h);
  for (int e0 = threadIdx.x; e0 < tot; e0 += LB * blockDim.x) {
    C2<T> v[LB];
#pragma unroll
    for (int b = 0; b < LB; ++b) {
      const int e = e0 + b * (int)blockDim.x;
      const int ix = fn2h.div(e), kz = e - ix * n2h;
      v[b] = src[e < tot ? ((long)ix * n1 + iy) * n2h + kz : 0];
    }
#pragma unroll
    for (int b = 0; b < LB; ++b) {
      const int e = e0 + b * (int)blockDim.x;
      const int ix = fn2h.div(e), kz = e - ix * n2h;
      if (e < tot) bufa[swz(kz, ix, lsx)] = v[b];  // x-line kz
    }
  }
  __syncthreads();
  C2<T>* fin = bufa;
  if constexpr (CN > 0) fft_lines_ct<T, IPV, CN>(bufa, n2h, lsx, twx);
  else if (ip) fft_lines_ip<T, IPV>(bufa, p.px, n2h, lsx, twx);
  else fin = fft_lines<T>(bufa, bufb, p.px, n2h, lsx, twx);
  if constexpr (DIG) {
    // the line's weighted values a = v.x f, b = v.y f (rounded to T as the
    // stored spectrum would be); exponent: max |value| < 2^ex; then each value
    // times 2^-ex cut into digits (x *= 128, d = trunc(x), x -= d).  The loop
    // runs in T = FP32 and is exact there: x is a power-of-two multiple of an
    // FP32 value, |x| < 1, and only a value below 2^-126 (rounded or flushed)
    // can lose bits, where every digit is 0 anyway (|x| 128^5 < 1).
    __shared__ double red[FT / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // (the weighted values go back into the thread's own LDS slots for the
    // digit pass)
    double mx = 0.0;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
      const int kx = fn2h.div(e), kz = e - kx * n2h;
      const T f = wt[((long)kx * n1 + iy) * n2h + kz];
      C2<T>& slot = fin[swz(kz, kx, lsx)];
      const C2<T> v = slot;
      const T a = v.x * f, b = v.y * f;
      slot = C2<T>{a, b};
      mx = fmax(mx, fmax(fabs((double)a), fabs((double)b)));
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    if (lane == 0) red[wv] = mx;
    __syncthreads();
    mx = red[0];
    for (int w = 1; w < FT / 64; ++w) mx = fmax(mx, red[w]);
    int ex = 0;
    if (mx > 0.0) frexp(mx, &ex);
    if (threadIdx.x == 0) p.dig_e[mu * n1 + iy] = ex;
    signed char* row = p.dig_s + mu * p.dig_row + (long)iy * p.dig_lp;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
      const int kx = fn2h.div(e), kz = e - kx * n2h;
      const C2<T> v = fin[swz(kz, kx, lsx)];
      T xa = ldexp(v.x, -ex), xb = ldexp(v.y, -ex);
#pragma unroll
      for (int q = 0; q < OZ_DIGITS; ++q) {
        xa *= T(128);
        const T da = trunc(xa);
        xa -= da;
        xb *= T(128);
        const T db = trunc(xb);
        xb -= db;
        *reinterpret_cast<unsigned short*>(row + q * p.dig_plane + 2 * e) =
            (unsigned short)(((unsigned)(int)da & 0xffu) | (((unsigned)(int)db & 0xffu) << 8));
      }
    }
    for (int c = 2 * tot + threadIdx.x; c < p.dig_lp; c += blockDim.x)
#pragma unroll
      for (int q = 0; q < OZ_DIGITS; ++q) row[q * p.dig_plane + c] = 0;
    return;
  }
  typedef T v2 __attribute__((ext_vector_type(2)));
  T* dst = out + mu * (long)n0 * n1 * n2h * 2;
  for (int e = threadIdx.x; e < n0 * n2h; e += blockDim.x) {
    const int kx = fn2h.div(e), kz = e - kx * n2h;
    const long h = ((long)kx * n1 + iy) * n2h + kz;
    const T f = wt[h];
    const C2<T> v = fin[swz(kz, kx, lsx)];
    if (p.vec_out) {
      *reinterpret_cast<v2*>(dst + 2 * h) = v2{v.x * f, v.y * f};
    } else {
      dst[2 * h] = v.x * f;
      dst[2 * h + 1] = v.y * f;
    }
  }
}

bool make_plan(int n, Plan1D& pl) {
  pl.n = n;
  pl.nst = 0;
  int m = n;
  while (m % 8 == 0 && pl.nst < MAXR) { pl.rad[pl.nst++] = 8; m /= 8; }
  while (m % 4 == 0 && pl.nst < MAXR) { pl.rad[pl.nst++] = 4; m /= 4; }
  while (m % 2 == 0 && pl.nst < MAXR) { pl.rad[pl.nst++] = 2; m /= 2; }
  while (m % 3 == 0 && pl.nst < MAXR) { pl.rad[pl.nst++] = 3; m /= 3; }
  while (m % 5 == 0 && pl.nst < MAXR) { pl.rad[pl.nst++] = 5; m /= 5; }
  return m == 1;
}

// e^{-2 pi i j / n}, j < n, from sincospi (about 1 ulp in double).
template <class T>
__global__ void twiddle_kernel(int n, C2<T>* tw) {
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double sn, cs;
    sincospi(-2.0 * (double)j / (double)n, &sn, &cs);
    tw[j] = C2<T>{(T)cs, (T)sn};
  }
}

constexpr int MAXN = 128;

// two ping-pong buffers plus the twiddle tables of the pass
template <class T>
size_t lds_a(int n1, int n2, bool ip) {
  const int n2h = n2 / 2 + 1, n1p = (n1 + 1) / 2;
  const size_t cap = (size_t)std::max(n1p * line_stride(n2), n2h * line_stride(n1));
  return (ip ? 1 : 2) * sizeof(C2<T>) * cap + sizeof(C2<T>) * (size_t)(n1 + n2);
}
template <class T>
size_t lds_b(int n0, int n2, bool ip) {
  return (ip ? 1 : 2) * sizeof(C2<T>) * (size_t)(n2 / 2 + 1) * line_stride(n0) + sizeof(C2<T>) * (size_t)n0;
}

template <class T, int IPV, class X, int CN = 0>
int launch_a(dim3 g, size_t la, const X* x, C2<T>* tmid, const FftParams& p, hipStream_t s) {
  const void* fn = (const void*)pass_a<T, IPV, X, CN>;
  THC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)la));
  hipLaunchKernelGGL((pass_a<T, IPV, X, CN>), g, dim3(FT), la, s, x, tmid, p);
  THC_HIP(hipGetLastError());
  return THC_OK;
}

template <class T, int IPV, int CN = 0>
int launch_b(dim3 g, size_t lb, const C2<T>* tmid, const T* wt, T* out, const FftParams& p, hipStream_t s) {
  if (p.dig_s) {
    if constexpr (std::is_same<T, float>::value) {
      const void* fn = (const void*)pass_b<T, IPV, true, CN>;
      THC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb));
      hipLaunchKernelGGL((pass_b<T, IPV, true, CN>), g, dim3(FT), lb, s, tmid, wt, out, p);
    } else {
      return THC_ERR_ARG;  // digits come from the FP32 transform only
    }
  } else {
    const void* fn = (const void*)pass_b<T, IPV, false, CN>;
    THC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb));
    hipLaunchKernelGGL((pass_b<T, IPV, false, CN>), g, dim3(FT), lb, s, tmid, wt, out, p);
  }
  THC_HIP(hipGetLastError());
  return THC_OK;
}

// Cubic grids whose edges have kernels compiled for their size (values per
// thread, edge): the 32^3, 64^3 and 96^3 grids, and the 24-48 point edges of
// small cells (products of 2, 3 and 5).  A grid of another shape, or whose per-thread budget differs
// from the one listed, runs the kernels with run-time plans.
template <int IPV, int N>
struct CtEdge {
  static constexpr int ipv = IPV, n = N;
};
using CtEdges = std::tuple<CtEdge<16, 64>, CtEdge<24, 96>, CtEdge<16, 32>, CtEdge<16, 48>, CtEdge<16, 45>,
                           CtEdge<16, 40>, CtEdge<16, 36>, CtEdge<16, 30>, CtEdge<16, 27>, CtEdge<16, 25>,
                           CtEdge<16, 24>>;

template <class T, class X, class... E>
bool launch_a_ct(std::tuple<E...>*, int ipa, int cube, dim3 g, size_t la, const X* x, C2<T>* tmid,
                 const FftParams& p, hipStream_t s, int& st) {
  return ((ipa == E::ipv && cube == E::n && ((st = launch_a<T, E::ipv, X, E::n>(g, la, x, tmid, p, s)), true)) ||
          ...);
}
template <class T, class... E>
bool launch_b_ct(std::tuple<E...>*, int ipb, int cube, dim3 g, size_t lb, const C2<T>* tmid, const T* wt, T* out,
                 const FftParams& p, hipStream_t s, int& st) {
  return ((ipb == E::ipv && cube == E::n && ((st = launch_b<T, E::ipv, E::n>(g, lb, tmid, wt, out, p, s)), true)) ||
          ...);
}

template <class T, class X>
int run(const X* x, int k, const int n[3], const double* coulG, void* out, void* work, hipStream_t s,
        hipEvent_t* pass_ev, const FftDigits* dig) {
  FftParams p = {};
  if (dig) {
    p.dig_s = dig->s;
    p.dig_e = dig->e;
    p.dig_row = dig->g2p;
    p.dig_plane = (long)k * dig->g2p;
    p.dig_lp = dig->lp;
    if (dig->lp < 2L * n[0] * (n[2] / 2 + 1) || dig->lp % 2 || dig->g2p != (long)n[1] * dig->lp) return THC_ERR_ARG;
  }
  p.k = k; p.n0 = n[0]; p.n1 = n[1]; p.n2 = n[2]; p.n2h = n[2] / 2 + 1;
  p.lsz = line_stride(n[2]); p.lsy = line_stride(n[1]); p.lsx = line_stride(n[0]);
  if (!make_plan(n[2], p.pz) || !make_plan(n[1], p.py) || !make_plan(n[0], p.px)) return THC_ERR_UNSUPPORTED;
  // in-place stages (one LDS buffer) when every stage and the unpack keep to
  // their per-thread register budgets
  const int n1p = (n[1] + 1) / 2, n2h = n[2] / 2 + 1;
  auto fa = [&](int v) { return fits_ip(p.pz, n1p, v) && fits_ip(p.py, n2h, v) && ((long)n1p * n2h + FT - 1) / FT <= UPV; };
  const int ipa = fa(16) ? 16 : (fa(24) ? 24 : 0);
  const int ipb = fits_ip(p.px, n2h, 16) ? 16 : (fits_ip(p.px, n2h, 24) ? 24 : 0);
  const size_t la = lds_a<T>(n[1], n[2], ipa > 0), lb = lds_b<T>(n[0], n[2], ipb > 0);
  if (la > 160 * 1024 || lb > 160 * 1024) return THC_ERR_UNSUPPORTED;
  // twiddle tables at the head of the workspace, then the intermediate spectrum
  char* w = reinterpret_cast<char*>(work);
  const size_t tb = align_up(sizeof(T) * 2 * MAXN, 256);
  C2<T>* dz = reinterpret_cast<C2<T>*>(w);
  C2<T>* dy = reinterpret_cast<C2<T>*>(w + tb);
  C2<T>* dx = reinterpret_cast<C2<T>*>(w + 2 * tb);
  const long nh = (long)n[0] * n[1] * n2h;  // half-spectrum points
  T* wt = reinterpret_cast<T*>(w + 3 * tb);
  C2<T>* tmid = reinterpret_cast<C2<T>*>(w + 3 * tb + align_up(sizeof(T) * (size_t)nh, 256));
  hipLaunchKernelGGL(weight_kernel<T>, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s, coulG, n[1], n[2], n2h, nh,
                     wt);
  hipLaunchKernelGGL(twiddle_kernel<T>, dim3(1), dim3(128), 0, s, n[2], dz);
  hipLaunchKernelGGL(twiddle_kernel<T>, dim3(1), dim3(128), 0, s, n[1], dy);
  hipLaunchKernelGGL(twiddle_kernel<T>, dim3(1), dim3(128), 0, s, n[0], dx);
  THC_HIP(hipGetLastError());
  p.twz = dz; p.twy = dy; p.twx = dx;
  T* o = reinterpret_cast<T*>(out);
  p.vec_out = (reinterpret_cast<uintptr_t>(out) % (2 * sizeof(T))) == 0;
  if (pass_ev) THC_HIP(hipEventRecord(pass_ev[0], s));
  const int cube = (n[0] == n[1] && n[1] == n[2]) ? n[0] : 0;
  const dim3 ga(n[0], k), gb(n[1], k);
  int st = THC_OK;
  if (!launch_a_ct<T, X>((CtEdges*)nullptr, ipa, cube, ga, la, x, tmid, p, s, st)) {
    if (ipa == 16) st = launch_a<T, 16, X>(ga, la, x, tmid, p, s);
    else if (ipa == 24) st = launch_a<T, 24, X>(ga, la, x, tmid, p, s);
    else st = launch_a<T, 0, X>(ga, la, x, tmid, p, s);
  }
  THC_TRY(st);
  if (pass_ev) THC_HIP(hipEventRecord(pass_ev[1], s));
  if (!launch_b_ct<T>((CtEdges*)nullptr, ipb, cube, gb, lb, tmid, wt, o, p, s, st)) {
    if (ipb == 16) st = launch_b<T, 16>(gb, lb, tmid, wt, o, p, s);
    else if (ipb == 24) st = launch_b<T, 24>(gb, lb, tmid, wt, o, p, s);
    else st = launch_b<T, 0>(gb, lb, tmid, wt, o, p, s);
  }
  THC_TRY(st);
  THC_HIP(hipGetLastError());
  if (pass_ev) THC_HIP(hipEventRecord(pass_ev[2], s));
  return THC_OK;
}

}  // namespace

size_t fft_work_bytes(int k, const int n[3], int prec) {
  const size_t es = prec == 0 ? sizeof(double) : sizeof(float);
  const size_t tb = align_up(es * 2 * MAXN, 256);
  const size_t nh = (size_t)n[0] * n[1] * (n[2] / 2 + 1);
  return 3 * tb + align_up(es * nh, 256) + align_up(2 * es * (size_t)k * nh, 256) + 256;
}

int fft_coulomb(const void* x, int x_f32, int k, const int n[3], const double* coulG, void* out, int prec,
                void* work, hipStream_t s, hipEvent_t* pass_ev, const FftDigits* dig) {
  if (k <= 0) return THC_OK;
  for (int d = 0; d < 3; ++d)
    if (n[d] < 1 || n[d] > MAXN) return THC_ERR_UNSUPPORTED;
  if (prec == 0) {
    if (x_f32) return THC_ERR_ARG;  // FP64 arithmetic keeps FP64 input
    if (dig) return THC_ERR_ARG;
    return run<double>(reinterpret_cast<const double*>(x), k, n, coulG, out, work, s, pass_ev, nullptr);
  }
  if (x_f32) return run<float>(reinterpret_cast<const float*>(x), k, n, coulG, out, work, s, pass_ev, dig);
  return run<float>(reinterpret_cast<const double*>(x), k, n, coulG, out, work, s, pass_ev, dig);
}

}  // namespace thc
