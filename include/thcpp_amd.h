/*
 * thcpp_amd: MI355X (gfx950) ISDF factorisation hot path behind a C ABI.
 *
 * This is the boundary a host (the C++ drop-in classes, Python via ctypes, or
 * any FFI) binds.  Every array argument is a DEVICE pointer on the context's
 * device unless marked "host".  No C++ types, exceptions or allocations cross
 * the boundary; every call returns an int status (THC_OK = 0).  A context owns
 * one HIP stream and a device workspace; it is not internally synchronised
 * (one context per host thread).
 *
 * Ordering: every call runs on the context's stream and returns after that
 * stream has drained, so outputs are complete when a call returns.  Inputs
 * must be complete before the call starts on that stream: either pass the
 * stream the caller produces them on to thc_ctx_create, or create the context
 * with stream = NULL, in which case the library creates a BLOCKING stream
 * (hipStreamDefault), which HIP orders after earlier work on the legacy
 * default (NULL) stream -- the stream torch and most hosts use by default.
 * Inputs produced on any other non-blocking stream must be synchronised by
 * the caller first.
 *
 * Reference interfaces replaced (paths relative to JoonhoLee-Group/thcpp):
 *   KMeans::KMeansSolver::kernel            src/kmeans.cpp:187-264
 *   InterpolatingVectors::IVecs ctor        src/interpolating_vectors.cpp:17-57
 *   InterpolatingVectors::IVecs::kernel     src/interpolating_vectors.cpp:406-455
 * and the per-stage routines they call (cited on each entry point below).
 */
#ifndef THCPP_AMD_H
#define THCPP_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Errors stop the call; warnings (>= 100) mean the outputs were
 * written, mirroring where the reference prints a warning and continues. */
enum thc_status {
  THC_OK = 0,
  THC_ERR_ARG = 1,              /* bad argument (null pointer, size, unsupported shape) */
  THC_ERR_HIP = 2,              /* HIP runtime error; thc_last_error() has the text */
  THC_ERR_EMPTY_CLUSTER = 3,    /* a k-means cluster weight < 1e-8 (kmeans.cpp:132,137-143, the reference exits) */
  THC_ERR_NOT_POSDEF = 4,       /* FP64 Cholesky of CC^T failed (the reference's QR solve would return garbage) */
  THC_ERR_NO_MEMORY = 5,        /* device workspace allocation failed */
  THC_ERR_UNSUPPORTED = 6,      /* an FFT edge with a prime factor > 5 or above THC_MAX_FFT_EDGE, more than
                                   THC_MAX_CLUSTERS k-means clusters, or an option combination listed as
                                   refused on the entry point */
  THC_WARN_DUPLICATE_INDICES = 100, /* map_to_grid found repeated indices (kmeans.cpp:151-155) */
  THC_WARN_LUV_NOT_POSDEF = 101,    /* chol(Muv) failed; Luv still written (interpolating_vectors.cpp:296-304) */
  THC_WARN_RANK_DEFICIENT = 102     /* CC^T has numerical rank < N_mu under the reference's rule
                                       (interpolating_vectors.cpp:192-208, matrix_operations.h:713-735);
                                       outputs still written */
};
/* An FFT grid with prod(fft_grid) != N is refused with THC_ERR_ARG.  The
 * reference prints a warning there and goes on (interpolating_vectors.cpp:
 * 368-370): it transforms the first prod(fft_grid) entries of each column and
 * leaves the rest untransformed when the product is smaller, and reads and
 * writes past the column when it is larger, so it has no result to match. */

/* Size limits. */
#define THC_MAX_FFT_EDGE 128   /* per-axis FFT length (fft_grid[i]) */
#define THC_MAX_CLUSTERS 4096  /* k-means clusters (N_mu) */

/* Arithmetic of the ISDF stages.  ZC^T, CC^T and the Cholesky factor and
 * inverse of CC^T are FP64 in every mode (CC^T is ill-conditioned). */
enum thc_precision {
  THC_PREC_FP64 = 0,  /* FP64 MFMA throughout: solve, FFT, G contraction */
  THC_PREC_FP32 = 1,  /* FP64 solve; FP32 FFT; G contraction on FP32 MFMA with FP32 accumulation
                         over each split-K chunk and an FP64 sum of the chunks */
  THC_PREC_MIXED = 3  /* solve X = (CC^T)^-1 ZC^T from exact int8 MFMA products of 7 signed 7-bit
                         digits of the rho-scaled operands (28 digit products, int32 sums, FP64
                         epilogue); FP32 FFT; V from exact int8 MFMA products of 5 digits per value
                         (per row and y-line exponents; the FFT writes the digits), summed in FP64.
                         Relative V error against the FP64 path about 2e-7 on 64^3 grids with N_mu
                         1024-2048 */
};

typedef struct thc_ctx thc_ctx;

/* Per-run k-means report. */
typedef struct thc_kmeans_stats {
  int iterations;          /* iterations executed */
  int converged;           /* 1 if diff < threshold ended the loop */
  int duplicates;          /* adjacent repeated indices after sorting */
  int interp_zeroed;       /* 1 if the loop ended with diff == threshold (indices left 0, as the reference does) */
  double last_diff;        /* normed difference of the final iteration */
  double *diffs;           /* host, optional: max_it entries, normed difference per iteration */
  double *pcts;            /* host, optional: max_it entries, % of grid points that changed cluster */
  double elapsed_ms;       /* device time of the call */
} thc_kmeans_stats;

typedef struct thc_isdf_opts {
  int precision;           /* enum thc_precision */
  int refine_steps;        /* FP64 iterative-refinement steps of the least-squares solve (0 = none) */
  int want_luv;            /* 1: also factor V = Luv Luv^T */
  int check_rank;          /* 1: numerical rank of CC^T as IVecs::check_rank computes it (thc_cct_rank),
                              beside the ZC^T GEMM; THC_WARN_RANK_DEFICIENT when rank < N_mu */
} thc_isdf_opts;

typedef struct thc_isdf_stats {
  int cct_chol_info;       /* 0, or the 1-based order of the first non-positive leading minor of CC^T */
  int luv_chol_info;       /* same for V */
  int cct_rank;            /* numerical rank of CC^T (opts->check_rank), else -1 */
  double cct_smax;         /* largest singular value of CC^T (opts->check_rank), else 0 */
  double stage_ms[12];     /* device time: 0 ZC^T, 1 the part of CC^T + Cholesky + inverse not hidden behind ZC^T
                              (they run on a second stream beside it), 2 solve, 3 FFT, 4 V, 5 Luv, 6 total,
                              7 refinement, 8 the solve's GEMM kernel alone (inside 2), 9 the V contraction's
                              main kernel alone (inside 4), 10 and 11 the FFT's first (z and y) and second
                              (x, weighting) pass kernels (inside 3) */
} thc_isdf_stats;

/* ---------------------------------------------------------------- context */

/* Create a context on HIP device 'device'.  'stream' may be NULL (the context
 * creates a blocking, default-priority stream of its own; see "Ordering"
 * above) or a hipStream_t to use. */
int thc_ctx_create(int device, void *stream, thc_ctx **out);
int thc_ctx_destroy(thc_ctx *ctx);
/* Capture the single-system k-means iteration loop once as a HIP graph and
 * replay it (default 0).  Measured on MI355X, replaying the graph saved no
 * time and slowed the later work on the context's stream: the 32-cell batch
 * on a context that had run graph-replayed k-means took 67-69 ms, 59-63 ms
 * without graphs. */
int thc_ctx_set_graphs(thc_ctx *ctx, int enable);
const char *thc_status_string(int status);
/* Text of the last HIP error seen on this thread. */
const char *thc_last_error(void);
/* Library version string. */
const char *thc_version(void);

/* ---------------------------------------------------------------- k-means */

/* Initial interpolation indices: iota(N), std::shuffle with std::mt19937(seed),
 * first K sorted ascending.  Host arrays.  Replaces
 * KMeansSolver::guess_initial_centroids (src/kmeans.cpp:159-175). */
int thc_kmeans_initial_indices(long n, int k, unsigned int seed, int *host_idx_out);

/* Diagnostic with no reference counterpart: y = sqrt(x) elementwise (n
 * doubles, device) with the same device square root the k-means distance
 * kernels use.  Bit-identical assignments need it correctly rounded (the
 * reference's std::sqrt is, src/kmeans.cpp:96-104); this lets a host check
 * that on its GPU. */
int thc_device_sqrt(thc_ctx *ctx, const double *x, long n, double *y);

/* Nearest centroid per grid point, ties to the first index, bit-identical to
 * KMeansSolver::classify_grid_points (src/kmeans.cpp:88-106).
 * grid: n x 3 f64 (xyz interleaved), centroids: k x 3 f64, grid_map: n int out. */
int thc_classify_grid_points(thc_ctx *ctx, const double *grid, long n,
                             const double *centroids, int k, int *grid_map);

/* Weighted centroids from an assignment, summed in ascending grid index and
 * divided as KMeansSolver::update_centroids does (src/kmeans.cpp:108-144).
 * Returns THC_ERR_EMPTY_CLUSTER when a weight is < 1e-8 (centroids still written). */
int thc_update_centroids(thc_ctx *ctx, const double *rho, const double *grid, long n,
                         const int *grid_map, int k, double *centroids);

/* Nearest grid point per centroid, sorted ascending (KMeansSolver::map_to_grid,
 * src/kmeans.cpp:146-157).  interp_idx: k int out.  Returns
 * THC_WARN_DUPLICATE_INDICES when indices repeat. */
int thc_map_to_grid(thc_ctx *ctx, const double *grid, long n, const double *centroids,
                    int k, int *interp_idx);

/* Interpolation points by QR with column pivoting of the orbital-product
 * matrix ZT[(i k), a] = phi_i(r_a) phi_k(r_a) (i over the occupied set aoR_half
 * when given): the first k pivot columns in ascending grid order, as
 * QRCPSolver::kernel keeps them (src/qrcp.cpp:100-173, pzgeqpf at
 * src/matrix_operations.h:756-788).  ZT is not formed: the pivots come from
 * pivoted Cholesky of its Gram matrix (P o Pocc, P = Phi Phi^T), which selects
 * the same columns in exact arithmetic.  interp_idx: k ints out (device);
 * rdiag (host, nullable): the k pivot magnitudes |R_jj| in pivot order.
 * Returns THC_ERR_NOT_POSDEF when k exceeds the numerical rank. */
int thc_qrcp_points(thc_ctx *ctx, const double *aoR, const double *aoR_half, long n, int m, int m_occ, int k,
                    int *interp_idx, double *rdiag);

/* thc_qrcp_points for complex orbitals: ZT[(i k), a] = conj(phi_i(r_a))
 * phi_k(r_a) (src/qrcp.cpp:33-66 with complex aoR / aoR_half).  The pivots
 * come from pivoted Cholesky of ZT^H ZT = conj(Pocc) o P (|P|^2 without
 * half rotation; complex Hermitian with it).  aoR_ri: n x m complex,
 * aoR_half_ri (nullable): n x m_occ complex, both (re, im) interleaved. */
int thc_qrcp_points_complex(thc_ctx *ctx, const double *aoR_ri, const double *aoR_half_ri, long n, int m,
                            int m_occ, int k, int *interp_idx, double *rdiag);

/* Full weighted k-means (KMeansSolver::kernel, src/kmeans.cpp:187-264) at one
 * rank: initial centroids from host_init_idx (host, k ints) or, if NULL, from
 * thc_kmeans_initial_indices(n, k, seed); then max_it Lloyd iterations that stop
 * when the normed centroid difference drops below 'threshold'; then map_to_grid.
 * interp_idx: k int out (sorted).  grid_map (n int) and centroids (k x 3 f64)
 * are optional outputs (the final assignment and final centroids). */
int thc_kmeans(thc_ctx *ctx, const double *grid, const double *rho, long n, int k,
               int max_it, double threshold, unsigned int seed, const int *host_init_idx,
               int *interp_idx, int *grid_map, double *centroids, thc_kmeans_stats *stats);

/* ---------------------------------------------------------------- ISDF stages */

/* ZC^T = (Phi_mu Phi^T)^2 elementwise, k x n row-major f64 (IVecs::setup_pseudo_dm
 * + setup_CZt, src/interpolating_vectors.cpp:59-141).  aoR: n x m row-major f64
 * (the orbital file's (N, M) layout), idx: k grid indices. */
int thc_czt(thc_ctx *ctx, const double *aoR, long n, int m, const int *idx, int k, double *czt);

/* CC^T = ZC^T[:, idx], k x k (IVecs::setup_CCt, src/interpolating_vectors.cpp:163-190). */
int thc_cct(thc_ctx *ctx, const double *czt, long n, const int *idx, int k, double *cct);

/* Numerical rank of the symmetric k x k matrix cct (k <= THC_MAX_CLUSTERS):
 * the number of singular values >= k * s0 * 1e-16, s0 the largest, the rule
 * of MatrixOperations::rank (src/matrix_operations.h:713-735) that
 * IVecs::check_rank applies to CC^T (src/interpolating_vectors.cpp:192-208).
 * The singular values of a symmetric matrix are its |eigenvalues|; they are
 * counted on the GPU's Householder tridiagonal form by Sturm sequences, with
 * no eigenvector work.  rank (host) out; smax (host, nullable) gets s0.
 * Returns THC_WARN_RANK_DEFICIENT when the rank is below k. */
int thc_cct_rank(thc_ctx *ctx, const double *cct, int k, int *rank, double *smax);

/* ---------------------------------------------------------------- GEMM tile tuning
 * No reference counterpart (ScaLAPACK chooses its own blocking).  The FP64 /
 * FP32 GEMMs of every stage pick a tile configuration by a built-in rule on
 * the tile count, unless the process-wide tuning table has an entry for the
 * shape class.  The table is written by the project's autotuner
 * (python -m thcpp_amd.autotune), which times every configuration with
 * thc_gemm_time on the shapes of a workload. */
enum thc_gemm_cfg {
  THC_GEMM_AUTO = 0,   /* built-in rule */
  THC_GEMM_TINY = 1,   /* 32 x 32 x 16 tile on one wave */
  THC_GEMM_SMALL = 2,  /* 64 x 64 x 16 on 4 waves */
  THC_GEMM_LARGE = 3,  /* 128 x 128 x 16 on 8 waves, register-staged */
  THC_GEMM_DTL = 4,    /* 128 x 128 x 16 on 8 waves, direct-to-LDS loads (FP64 plain store only) */
  THC_GEMM_DEEPK = 5   /* 128 x 128 x 32 on 8 waves, one workgroup per CU */
};
/* Replace the table: n entries of 5 ints (ceil(log2 m), ceil(log2 n),
 * ceil(log2 k), flags, thc_gemm_cfg) with flags = 1 gathered operand |
 * 2 lower triangle | 4 square epilogue | 8 triangular operands.  n = 0 clears
 * it.  A configuration that does not apply to a call falls back to the rule. */
int thc_gemm_tuning(const int *entries, int n);
/* Average device time (ms, over reps launches after one warm-up) of one FP64
 * GEMM on configuration cfg: kind 0 is C (m x n) = A (m x k) B (n x k)^T, the
 * solve's layout; kind 1 is the ZC^T GEMM, C = (A A[idx]^T)^2 elementwise
 * with n gathered rows.  Operands live in the context's workspace.
 * THC_ERR_UNSUPPORTED when cfg does not apply to the shape. */
int thc_gemm_time(thc_ctx *ctx, long m, long n, long k, int kind, int cfg, int reps, double *ms);

/* In-place lower Cholesky of a k x k row-major SPD matrix (strict upper part
 * zeroed).  *info = 0 or the 1-based failing leading minor (pzpotrf 'L',
 * src/matrix_operations.h:259-267). */
int thc_cholesky(thc_ctx *ctx, double *a, int k, int *info);

/* thc_cholesky without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The Cholesky's scratch
 * (two 64 x 64 FP64 panels and the info word, 65 792 bytes whatever k is) is
 * a static device array each GPU holds for the life of the process, shared
 * by every call on that GPU: calls that overlap on one GPU race.  The status
 * word goes through the context's pinned buffer, made by thc_ctx_create.
 * Same result as thc_cholesky, bit for bit. */
int thc_cholesky_noalloc(thc_ctx *ctx, double *a, int k, int *info);

/* X = CC^T^-1 ZC^T, k x n: FP64 Cholesky of CC^T, explicit FP64 inverse from
 * the factor, one MFMA GEMM, plus 'refine_steps' FP64 residual corrections.
 * Replaces MatrixOperations::least_squares / pzgels (src/matrix_operations.h:184-202,
 * called at src/interpolating_vectors.cpp:416). */
int thc_least_squares(thc_ctx *ctx, const double *cct, const double *czt, long n, int k,
                      int refine_steps, double *x, int *info);

/* Coulomb-weighted half spectrum of the interpolating vectors: for each row mu
 * of x (k x n, real), the unnormalised forward 3-D DFT over fft_grid (C order),
 * kept for kz <= n2/2 and written as out[mu][g] (re, im interleaved) scaled by
 * sqrt(w(G) coulG(G)) with w = 1 on self-conjugate planes and 2 elsewhere, so
 * that V = out out^T.  out is k x (2 n0 n1 (n2/2+1)) f64.  Replaces
 * IVecs::fft_vectors + the sqrt(coulG) scaling of construct_muv
 * (src/interpolating_vectors.cpp:331-404, 258-263). */
int thc_fft_vectors(thc_ctx *ctx, const double *x, int k, const int fft_grid[3],
                    const double *coulG, double *out);

/* V = B B^T (k x k, full symmetric) from the weighted half spectrum B
 * (k x g2).  Replaces the pzherk of IVecs::construct_muv
 * (src/interpolating_vectors.cpp:278). */
int thc_construct_muv(thc_ctx *ctx, const double *b, int k, long g2, double *v);

/* ---------------------------------------------------------------- whole path */

/* IVecs ctor + kernel at one rank: ZC^T, CC^T, least squares, FFT, V and
 * (opts->want_luv) Luv = chol(V) lower with zero upper triangle
 * (src/interpolating_vectors.cpp:17-57, 406-455, 283-315).
 *   aoR n x m f64, idx k ints (sorted interpolation points), coulG n f64.
 *   v k x k out; luv k x k out (nullable); xi k x n out (nullable: the
 *   interpolating vectors X = CC^T^-1 ZC^T, row mu = xi_mu(r)).
 * Returns THC_ERR_NOT_POSDEF if CC^T is not numerically positive definite,
 * THC_WARN_LUV_NOT_POSDEF if V is not (Luv still written). */
int thc_isdf(thc_ctx *ctx, const double *aoR, long n, int m, const int *idx, int k,
             const int fft_grid[3], const double *coulG, const thc_isdf_opts *opts,
             double *v, double *luv, double *xi, thc_isdf_stats *stats);

/* One complete factorisation: thc_kmeans then thc_isdf on its points
 * (src/main.cpp:50-60 at one rank).  interp_idx k ints out. */
int thc_factorize(thc_ctx *ctx, const double *grid, const double *rho, const double *aoR,
                  long n, int m, int thc_cfac, int max_it, double threshold, unsigned int seed,
                  const int fft_grid[3], const double *coulG, const thc_isdf_opts *opts,
                  int *interp_idx, double *v, double *luv,
                  thc_kmeans_stats *kstats, thc_isdf_stats *istats);

/* ---------------------------------------------------------------- complex orbitals */

/* ZC^T for complex orbitals (k-point supercells; the reference's native
 * complex128 aoR): P = Phi_mu^H Phi, ZC^T = P o conj(P) = |P|^2, which is real
 * (IVecs::setup_pseudo_dm + setup_CZt, src/interpolating_vectors.cpp:59-141).
 * aoR_ri: n x m complex orbitals, (re, im) interleaved (n x 2m doubles, the
 * orbital file's compound layout); czt: k x n real out. */
int thc_czt_complex(thc_ctx *ctx, const double *aoR_ri, long n, int m, const int *idx, int k, double *czt);

/* thc_isdf for complex orbitals: ZC^T and CC^T = |P|^2 are real, so the
 * interpolating vectors, the FFT, V and Luv are the real path's
 * (src/interpolating_vectors.cpp:17-57, 406-455).  aoR_ri as in
 * thc_czt_complex. */
int thc_isdf_complex(thc_ctx *ctx, const double *aoR_ri, long n, int m, const int *idx, int k,
                     const int fft_grid[3], const double *coulG, const thc_isdf_opts *opts,
                     double *v, double *luv, double *xi, thc_isdf_stats *stats);

/* ---------------------------------------------------------------- half-rotated THC */

/* ZC^T of the half-rotated factorisation, k x n row-major:
 * (Phi_occ,mu Phi_occ^T) o (Phi_mu Phi^T) elementwise.  Replaces
 * IVecs::setup_CZt_half (src/interpolating_vectors.cpp:143-161), which forms
 * Pocc o conj(P); the orbitals here are real.  aoR_half: n x m_occ row-major
 * (the orbital file's aoR_half). */
int thc_czt_half(thc_ctx *ctx, const double *aoR, const double *aoR_half, long n, int m, int m_occ,
                 const int *idx, int k, double *czt);

/* IVecs with rotate = true at one rank (src/interpolating_vectors.cpp:17-57
 * with setup_CZt_half, then kernel 406-455): the half-rotated ZC^T and CC^T,
 * the least-squares solve, the FFT and
 * Muv = sum_G IVMG(G)^T coulG(G) IVG(G) (construct_muv, :271-273), which for
 * real interpolating vectors (IVMG = conj IVG, :384-391) is V of thc_isdf;
 * written full (k x k), with no Cholesky (HalfTransformedMuv, :317-328).
 * opts->want_luv is ignored. */
int thc_isdf_half(thc_ctx *ctx, const double *aoR, const double *aoR_half, long n, int m, int m_occ,
                  const int *idx, int k, const int fft_grid[3], const double *coulG,
                  const thc_isdf_opts *opts, double *muv, double *xi, thc_isdf_stats *stats);

/* Half-rotated ZC^T for complex orbitals, k x n complex (re, im interleaved):
 * Pocc o conj(P) with P[mu][r] = sum_i conj(phi_i(r_mu)) phi_i(r) over the
 * m orbitals and Pocc the same over the m_occ occupied ones
 * (IVecs::setup_CZt_half, src/interpolating_vectors.cpp:143-161, with complex
 * aoR / aoR_half).  aoR_ri: n x m and aoR_half_ri: n x m_occ complex,
 * interleaved (n x 2m, n x 2 m_occ doubles). */
int thc_czt_half_complex(thc_ctx *ctx, const double *aoR_ri, const double *aoR_half_ri, long n, int m, int m_occ,
                         const int *idx, int k, double *czt_ri);

/* thc_isdf_half for complex orbitals (src/interpolating_vectors.cpp:143-161,
 * 331-404 with both FFT directions, 258-273, 317-328): ZC^T and CC^T are
 * complex (CC^T Hermitian), X = CC^T^-1 ZC^T is complex, and
 * Muv[mu][nu] = sum_G IVMG[G][mu] coulG(G) IVG[G][nu] with IVG / IVMG the
 * forward / backward transforms of conj(X[mu]) -- complex symmetric, not
 * Hermitian.  The solve runs on the real 2k x 2k embedding of CC^T, the FFT
 * and the G contraction on the 2k real rows Re X, Im X (coulG must be even in
 * G, as the Coulomb kernel on an FFT grid is).  muv_ri: k x k complex out
 * (interleaved, full); xi_ri (nullable): X, k x n complex out.  precision as
 * for thc_isdf (MIXED needs 2k <= 4096); opts->want_luv is ignored. */
int thc_isdf_half_complex(thc_ctx *ctx, const double *aoR_ri, const double *aoR_half_ri, long n, int m, int m_occ,
                          const int *idx, int k, const int fft_grid[3], const double *coulG,
                          const thc_isdf_opts *opts, double *muv_ri, double *xi_ri, thc_isdf_stats *stats);

/* ---------------------------------------------------------------- batches */

/* One system of a batch: inputs as for thc_factorize (device pointers), with
 * k = thc_cfac * m interpolation points.  The out fields are written by
 * thc_factorize_batched; kstats.diffs / kstats.pcts may point to host arrays
 * of max_it entries (or be NULL) before the call. */
typedef struct thc_system {
  const double *grid;      /* n x 3 */
  const double *rho;       /* n */
  const double *aoR;       /* n x m */
  const double *coulG;     /* n */
  long n;
  int m;
  int thc_cfac;
  int max_it;
  double threshold;
  unsigned int seed;
  int fft_grid[3];
  int *interp_idx;         /* out, k ints */
  double *v;               /* out, k x k */
  double *luv;             /* out, k x k (nullable) */
  int status;              /* out: thc_factorize's status for this system */
  thc_kmeans_stats kstats; /* out */
  thc_isdf_stats istats;   /* out */
} thc_system;

/* Factorise many independent systems of uneven sizes (the reference runs one
 * system per thcpp.x invocation, src/main.cpp:18-77; this is that loop).
 * The k-means run as two batched runs, every kernel launch covering one step
 * of every system of its run (kstats.elapsed_ms is the time of the system's
 * whole run): the max_streams largest systems on the context's stream, the
 * rest on another stream beside their ISDF stages.  The ISDF stages run on up
 * to max_streams HIP streams (<= 0: 4; at most 16), each with its own
 * workspace and driven by its own host thread, which takes the next system,
 * largest first, whenever its last one is done (a stream reaching the systems
 * of the second k-means run waits for it), so that small systems overlap on
 * the GPU.  Every stream first tries to reserve the largest workspace of the
 * batch (one whose reservation fails reserves per system instead).  Each system's results equal
 * thc_factorize's on the same inputs.  Returns THC_OK when every system
 * returned THC_OK, else the first error status by system order (or, without
 * errors, the first warning); per-system statuses are in systems[i].status. */
int thc_factorize_batched(thc_ctx *ctx, int nsys, thc_system *systems, const thc_isdf_opts *opts,
                          int max_streams);

#ifdef __cplusplus
}
#endif
#endif /* THCPP_AMD_H */
