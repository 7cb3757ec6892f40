/*
 * linqt_amd.h -- C ABI of the MI355X (gfx950) Chebyshev-moment engine.
 *
 * The reference exposes this path as C++ classes backed by MKL/FFTW; each
 * entry point below replaces one of them (cited per function).  Plain
 * pointers and sizes only.  Complex numbers are interleaved (re, im)
 * doubles.  Every function returns LQ_OK (0) or a negative LQ_ERR_* code and
 * never throws; lq_last_error() gives the calling thread's last message.
 *
 * Threading: one lq_ctx owns one HIP stream, its device buffers and scratch.
 * Calls on one context must not overlap; use one context per host thread.
 *
 * Layouts on the device:
 *   - a block of k vectors of length n is stored interleaved, element (i, v)
 *     at index i*k + v ("[n][k]");
 *   - state arrays (phi) are state-major, state r at phi[2*(r*n + i)],
 *     as the reference's `vector` state file holds them
 *     (include/quantum_states.hpp:155-183); the solvers accept phi in host
 *     memory or in device memory of the context's GPU (lq_malloc);
 *   - 2D moment tables are row major, mu(m0, m1) at m0*M1 + m1
 *     (include/chebyshev_moments.hpp:255).
 */
#ifndef LINQT_AMD_H
#define LINQT_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LQ_ABI_VERSION 1

enum lq_status {
  LQ_OK = 0,
  LQ_ERR_ARG = -1,         /* invalid argument / size / handle */
  LQ_ERR_HIP = -2,         /* HIP runtime error */
  LQ_ERR_NOMEM = -3,       /* device allocation failed */
  LQ_ERR_UNSUPPORTED = -4, /* a case whose reference result is undefined */
  LQ_ERR_NODEV = -5        /* no GPU / HIP runtime unavailable */
};

/* SparseMatrixType descriptor fill (include/sparse_matrix.hpp:50-55). */
enum lq_fill {
  LQ_FILL_HERMITIAN_UPPER = 0, /* MKL HERMITIAN + FILL_MODE_UPPER: the reference's only mode */
  LQ_FILL_GENERAL = 1          /* entries used as stored */
};

/* chebyshev::formula (include/Kubo_solver_FFT.hpp:17-21). */
enum lq_formula { LQ_KUBO_GREENWOOD = 0, LQ_KUBO_BASTIN = 1 };

/* 1D moment algorithm. */
enum lq_algo1d {
  LQ_ALGO_DIRECT = 0,  /* mu_m = <phi|T_m>, the reference's loop */
  LQ_ALGO_DOUBLING = 1 /* mu_2n = 2<T_n|T_n> - mu_0, mu_2n+1 = 2<T_n+1|T_n> - mu_1 (OP = identity only) */
};

/* Contraction precision of the 2D table. */
enum lq_precision {
  LQ_PREC_FP64 = 0,   /* FP64 MFMA */
  /* Floating-point splits (the forms the split-precision literature uses):
   * each FP64 operand is cut into lower-precision pieces whose products are
   * summed on MFMA in FP32 accumulators, folded into FP64 every 256
   * K-elements.  The FP32 accumulation limits them to ~1e-7 of the table;
   * they exist to measure that, not for production tables. */
  LQ_PREC_FP32X2 = 2, /* FP32 hi + lo, products hh + hl + lh on FP32 MFMA */
  LQ_PREC_BF16X3 = 3, /* three bf16 pieces, products of levels i + j <= 2 on BF16 MFMA */
  /* Split precision: each operand is cut, per 4096-element group of a row,
   * into S int8 digits under a shared power-of-two exponent; digit products
   * of levels t + u <= S + 1 are summed exactly on INT8 MFMA (int32) and
   * combined in FP64.  Error per product ~ (S + 1) 2^-7S of the group maxima. */
  LQ_PREC_INT8X4 = 4, /* S = 4 */
  LQ_PREC_INT8X5 = 5, /* S = 5 */
  LQ_PREC_INT8X6 = 6, /* S = 6 */
  LQ_PREC_INT8X7 = 7, /* S = 7 */
  LQ_PREC_INT8X8 = 8  /* S = 8 */
};

typedef struct lq_ctx_s *lq_ctx;
typedef struct lq_csr_s *lq_csr;
typedef struct lq_batch_s *lq_batch;

/* Per-context kernel timing (HIP events on the context stream), summed since
 * the last lq_ctx_reset_timing.  Bytes / flops are algorithmic counts. */
typedef struct lq_timing {
  double spmv_ms;        /* recursion SpMV/SpMM and operator applies */
  double contract_ms;    /* moment contraction (Lᴴ R) incl. split-K reduce */
  double post_ms;        /* epilogues, Jackson, energy-grid sums */
  double spmv_bytes;     /* sum of 20*nnz + 4*(n+1) + (48 or 32)*k*n per launch */
  double contract_flops; /* sum of 8*M0*M1*K per contraction */
  long long spmv_launches;
  long long contract_launches;
} lq_timing;

/* ---- runtime ----------------------------------------------------------- */
const char *lq_last_error(void);
int lq_abi_version(void);
int lq_device_count(int *count);
int lq_ctx_create(int device, lq_ctx *out);
int lq_ctx_destroy(lq_ctx ctx);
/* No counterpart in the reference (its buffers live as long as its objects).
 * Waits for the context's work and frees the device scratch buffers, pinned
 * staging and captured graphs it keeps between calls (they grow to the
 * largest solve run on it); operators and the context stay valid.  A server
 * calls it between requests so a context holds no memory while idle. */
int lq_ctx_release(lq_ctx ctx);
int lq_ctx_synchronize(lq_ctx ctx);
int lq_ctx_set_timing(lq_ctx ctx, int enabled);
int lq_ctx_reset_timing(lq_ctx ctx);
int lq_ctx_get_timing(lq_ctx ctx, lq_timing *out);
/* Device memory of the context's device (plain device pointers). */
int lq_malloc(lq_ctx ctx, size_t bytes, void **dptr);
int lq_free(lq_ctx ctx, void *dptr);
int lq_memcpy_h2d(lq_ctx ctx, void *dst, const void *src, size_t bytes);
int lq_memcpy_d2h(lq_ctx ctx, void *dst, const void *src, size_t bytes);
int lq_memcpy_d2d(lq_ctx ctx, void *dst, const void *src, size_t bytes);

/* ---- sparse operators ---------------------------------------------------- */
/* Replaces SparseMatrixType::ConvertFromCSR (src/mkl_sparse_matrix.cpp:39-54)
 * + SparseMatrixBuilder::BuildOPFromCSRFile (include/sparse_matrix.hpp:100-109).
 * Host CSR in (zero based); the operator is Hermitian-completed (for
 * LQ_FILL_HERMITIAN_UPPER) and uploaded in a sliced-ELL layout. */
int lq_csr_create(lq_ctx ctx, int n, int nnz, const int *rowptr, const int *col,
                  const double *val, int fill, lq_csr *out);
int lq_csr_destroy(lq_csr A);
/* Replaces SparseMatrixType::Rescale (src/mkl_sparse_matrix.cpp:73-97):
 * A <- a*A + b*I, as Moments::Rescale2ChebyshevDomain calls it with
 * a = 1/HalfWidth, b = -BandCenter/HalfWidth (include/chebyshev_moments.hpp:95-99). */
int lq_csr_rescale(lq_csr A, double a_re, double a_im, double b_re, double b_im);
/* n, completed nnz, stored (padded) entries. */
int lq_csr_info(lq_csr A, int *n, long long *nnz, long long *nnz_stored);
/* The operator's stored entries (upper triangle with the diagonal for
 * LQ_FILL_HERMITIAN_UPPER, all for LQ_FILL_GENERAL) as zero-based CSR, the
 * arrays a reference CSR file holds; after lq_csr_rescale the rescaled ones.
 * rowptr == NULL: *nnz only; otherwise arrays of n+1, nnz, 2*nnz entries. */
int lq_csr_export(lq_csr A, long long *nnz, int *rowptr, int *col, double *val);
/* Replaces wannier2sparse's hopping_list::wrap_in_supercell +
 * save_hopping_list_as_csr (utilities/wannier2sparse/src/hopping_list.cpp:
 * 156-285) and the Hermitian completion of the CSR file it writes: the
 * operator of a unit-cell hopping list on a dims[0] x dims[1] x dims[2]
 * supercell, generated on the device (LQ_FILL_HERMITIAN_UPPER).  Per hopping:
 * cell[3] (unit-cell vector), edge[2] (zero-based orbitals from, to), val
 * (re, im).  Hoppings are taken in the order of the reference's multimap
 * (tag string "c0 c1 c2 e0 e1 "), those with |v| <= 1e-10 mean|v| and those
 * below the diagonal are dropped, duplicates are summed in that order; the
 * stored entries equal the file's up to its 6-digit text rounding. */
int lq_csr_from_hoppings(lq_ctx ctx, int nwann, int nhop, const int *cell, const int *edge,
                         const double *val, const int *dims, lq_csr *out);

/* Replaces SparseMatrixType::Multiply(a, x, b, y) (src/mkl_sparse_matrix.cpp:56-69)
 * and BatchMultiply (:102-106): y <- a*A*x + b*y for k interleaved vectors.
 * x, y are DEVICE pointers ([n][k]); b == 0 overwrites y. */
int lq_spmv(lq_csr A, double a_re, double a_im, const double *x, double b_re, double b_im,
            double *y, int k);

/* No counterpart in the reference (MKL tunes its own kernels through
 * mkl_sparse_set_mv_hint + mkl_sparse_optimize, src/mkl_sparse_matrix.cpp:48-51):
 * times every schedule of A's recursion step at k vectors (mode 0: the step,
 * 1: the step with <phi|T> fused) -- gather chunk 4/8/12/16, grid of
 * resident blocks or 2048, prefetched row groups on/off -- and keeps the
 * fastest in A.  Results do not change (the schedules do the same
 * arithmetic).  Choices persist in LQ_TUNE_CACHE (default
 * $HOME/.cache/linqt_amd/spmm_tune.txt) keyed by GPU and operator shape;
 * *from_cache is 1 when read from there.  Outputs may be NULL. */
int lq_csr_tune(lq_csr A, int k, int mode, int *unroll, int *gridcap, int *prefetch, int *from_cache);

/* Replaces chebyshev::ComputeMomTable (src/chebyshev_solver.cpp:26-42):
 * sub[m0*ldsub + m1] = sum_k conj(L[m0*ldl + k]) * R[m1*ldr + k] for
 * m0 < ML, m1 < MR, k < K.  All DEVICE pointers; FP64 MFMA. */
int lq_compute_mom_table(lq_ctx ctx, const double *L, long long ldl, int ML, const double *R,
                         long long ldr, int MR, long long K, double *sub, int ldsub);
/* The same table in a chosen lq_precision (LQ_PREC_FP64 equals lq_compute_mom_table). */
int lq_compute_mom_table_prec(lq_ctx ctx, const double *L, long long ldl, int ML, const double *R,
                              long long ldr, int MR, long long K, int precision, double *sub, int ldsub);

/* ---- moment solvers (host phi in, host moments out) ---------------------- */
/* Replaces chebyshev::SpectralMoments (src/chebyshev_solver.cpp:176-200).
 * H already rescaled.  OP == NULL is the identity (SparseMatrixType "1").
 * mu (M complex) is accumulated into, as the reference's Moments1D is. */
int lq_spectral_moments(lq_csr H, lq_csr OP, const double *phi, int R, int M, int algo,
                        double *mu);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP stream/graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): lq_spectral_moments with LQ_ALGO_DIRECT, same
 * groups, kernels and summation order (the same moments, bit for bit), but
 * every buffer beyond phi and mu lives in one fixed 64 MiB device array per
 * GPU that the library holds until the process ends.  phi (R states,
 * state-major) and mu (M complex, accumulated into) must be DEVICE memory of
 * H's GPU.  The call only enqueues work on the context's stream (no
 * allocation, no host synchronisation): mu is ready after
 * lq_ctx_synchronize.  Sizes whose scratch exceeds 64 MiB -- about
 * 16*M + 48*n*g + 256 KiB*kp bytes, g = min(R, 64), kp = g rounded up to a
 * power of two -- return LQ_ERR_ARG with the required size in
 * lq_last_error().  All contexts of one GPU share the array:
 * calls that overlap on one GPU (two host threads, or two contexts whose
 * streams run concurrently) race and give wrong results. */
int lq_spectral_moments_noalloc(lq_csr H, lq_csr OP, const double *phi, int R, int M,
                                double *mu);

/* Replaces chebyshev::CorrelationExpansionMoments (src/chebyshev_solver.cpp:45-172)
 * for nopr right operators sharing one left operator: table o is
 * mu_o(m0,m1) = <phi| OPL T_m0 OPR_o T_m1 |phi>, scaled by 4/R with 1/2 on
 * m0 = 0 and m1 = 0 and made Hermitian as at chebyshev_solver.cpp:149-162.
 * batch <= 0 is the default BATCH_SIZE = M0; other values reproduce the
 * reference's blocked recursion (kpm_noneqop.hpp:27-39, chebyshev_vectors.cpp:4-21).
 * mu: nopr tables of M0*M1 complex, in/out with the reference's semantics
 * (accumulated into unless padding to a multiple of batch happens). */
int lq_correlation_moments(lq_csr H, lq_csr OPL, int nopr, const lq_csr *OPR, const double *phi,
                           int R, int M0, int M1, int batch, int precision, double *mu);

/* Replaces Kubo_solver_FFT::compute (src/Kubo_solver_FFT.cpp:12-136 with
 * Greenwood_FFTs.cpp / Bastin_FFTs.cpp): final_data (nopr x 2*nump complex)
 * receives the mean over the R states of r_data.  Requires M <= nump (the
 * reference overflows its DFT inputs otherwise). */
int lq_kubo_fft(lq_csr H, lq_csr OPL, int nopr, const lq_csr *OPR, const double *phi, int R,
                int M, int num_sections, int nump, int formula, int precision, double *final_data);

/* Replaces Kubo_solver_FFT_postProcess::{Greenwood,Bastin}_postProcess
 * (src/Kubo_solver_FFT_postProcess.cpp:120-397): energies (a*E - Ec,
 * ascending) and sigma as the output file's two columns; integrand (Bastin,
 * may be NULL) as the `_integrand` file's second column. */
int lq_kubo_post(lq_ctx ctx, int formula, const double *final_data, int nump, int n, double W,
                 double Ec, double *E, double *sigma, double *integrand);

/* ---- non-orthogonal basis -------------------------------------------------- */
/* Replaces linalg::orthogonalize (src/mkl_linear_algebra.cpp:100-131):
 * x = S^-1 b by BiCGSTAB as Eigen runs it (diagonal preconditioner, zero
 * initial guess, stop at |r| <= tol |b| or after maxit iterations; tol <= 0
 * and maxit <= 0 select the reference's 1e-4 and 2000).  b, x: distinct
 * DEVICE vectors of n complex.  The reference maps the overlap's CSR arrays
 * as a general matrix: create S with LQ_FILL_GENERAL from the file's arrays
 * to solve with the same matrix.  iterations / error (|r| / |b|) optional. */
int lq_overlap_solve(lq_csr S, const double *b, double *x, double tol, int maxit, int *iterations,
                     double *error);
/* Replaces chebyshev::SpectralMoments_nonOrth (src/chebyshev_solver.cpp:203-229)
 * with Moments1D_nonOrth::SetInitVectors_nonOrthogonal / Iterate_nonOrthogonal
 * (src/chebyshev_moments.cpp:63-153): T0 = [OP] S^-1 phi, T1 = S^-1 H T0,
 * T_{m+1} = 2 S^-1 H T_m - T_{m-1}, every solve as lq_overlap_solve;
 * mu (M complex) += (2/R)(1/2 for m = 0) <phi|T_m>.  H already rescaled;
 * OP == NULL is the identity.  iterations: total BiCGSTAB iterations (optional). */
int lq_spectral_moments_nonorth(lq_csr H, lq_csr OP, lq_csr S, const double *phi, int R, int M, double tol,
                                int maxit, double *mu, long long *iterations);

/* No counterpart in the reference: measurement of the achievable HBM rate
 * that the recursion kernels are compared against.  Streams `bytes` of a
 * buffer of the context (kind 0: read, 1: copy between its halves, 2: write)
 * with hand-written 16-byte non-temporal accesses, reps times; *ms is the
 * mean time of one pass and *moved the bytes it reads plus writes. */
int lq_hbm_probe(lq_ctx ctx, size_t bytes, int kind, int reps, double *ms, double *moved);

/* ---- Jackson kernels and energy-grid post-processors --------------------- */
/* Moments1D::ApplyJacksonKernel (src/chebyshev_moments1D.cpp:68-87);
 * *M_out receives the truncated count, mu compacted in place. */
int lq_jackson_1d(lq_ctx ctx, double *mu, int M, double W, double broad_meV, int *M_out);
/* Moments2D::ApplyJacksonKernel (src/chebyshev_moments2D.cpp:87-117). */
int lq_jackson_2d(lq_ctx ctx, double *mu, int M0, int M1, double W, double b0_meV,
                  double b1_meV, int *M0_out, int *M1_out);
/* spectralFunctionFromChebmom_FFTgrid.cpp:35-73: num_div rows. */
int lq_dos_fftgrid(lq_ctx ctx, const double *mu, int M, int n, double W, double Ec, int num_div,
                   double *E, double *rho);
/* kuboGreenwoodFromChebmom_FFTgrid.cpp:35-76: num_div-1 rows. */
int lq_greenwood_fftgrid(lq_ctx ctx, const double *mu, int M0, int M1, int n, double W, double Ec,
                         int num_div, double *E, double *sigma);
/* kuboBastinFromChebmom_FFTgrid.cpp:57-106: num_div-1 rows (cumulative). */
int lq_bastin_fftgrid(lq_ctx ctx, const double *mu, int M0, int M1, int n, double W, double Ec,
                      int num_div, double *E, double *sigma);

/* ---- time-dependent correlations ------------------------------------------ */
/* Replaces chebyshev::TimeDependentCorrelations (src/chebyshev_solver.cpp:231-275)
 * with MomentsTD::Evolve (src/chebyshev_momentsTD.cpp:4-45): mu (M x nT complex,
 * mu(m, n) at m*nT + n) += (2/R)(1/2)^{delta_m0} <OPL phi(t_n)|T_m(H~) OPR|phi(t_n)>,
 * phi(t) = exp(-iHt/hbar) phi evolved in steps dt (fs) by the Bessel-Chebyshev
 * expansion (terms while |J_k| > 1e-10).  H already rescaled with bandwidth W
 * and centre Ec (eV), which set the expansion's frequencies.  OPL/OPR NULL = 1. */
int lq_time_correlations(lq_csr H, lq_csr OPL, lq_csr OPR, const double *phi, int R, int M, int nT,
                         double dt, double W, double Ec, double *mu);
/* MomentsTD::ApplyJacksonKernel (src/chebyshev_momentsTD.cpp:135-154): the first
 * *M_out rows of the M x nT table are kept and damped in place. */
int lq_jackson_td(lq_ctx ctx, double *mu, int M, int nT, double W, double broad_meV, int *M_out);
/* timeCorrelationsFromChebmom.cpp:35-70: nT rows (n*dt, N/a sum_m delta_m(x_F) Re mu(m, n)),
 * x_F = (EF - Ec)/a. */
int lq_timecorr_post(lq_ctx ctx, const double *mu, int M, int nT, int n, double W, double Ec, double dt,
                     double EF, double *t, double *out);

/* Linear-grid post-processors: the tools' grid is num_div energies
 * x_i = -xb + i*2xb/(num_div-1) (xb = 0.99 for Bastin, 1 otherwise; the tools
 * use num_div = 30 * max moments), E = a*x + Ec.  Rows written: num_div for
 * LQ_LIN_SPECTRAL (mu: M0 moments, M1 ignored), num_div-1 otherwise (mu:
 * M0 x M1).  Jackson-damped moments in, as the tools apply it first.  With
 * xb = 1 the end points divide by sqrt(1 - x^2) = 0 and print NaN, and the
 * running sums of BastinII and BastinSea are then NaN throughout, as in the
 * reference. */
enum lq_lingrid {
  LQ_LIN_SPECTRAL = 0,    /* spectralFunctionFromChebmom.cpp:35-67 */
  LQ_LIN_GREENWOOD = 1,   /* kuboGreenwoodFromChebmom.cpp:35-76 */
  LQ_LIN_BASTIN = 2,      /* kuboBastinFromChebmom.cpp:35-86 */
  LQ_LIN_BASTIN_I = 3,    /* kuboBastinIFromChebmom.cpp:35-88 */
  LQ_LIN_BASTIN_II = 4,   /* kuboBastinIIFromChebmom.cpp:35-93 */
  LQ_LIN_BASTIN_SEA = 5,  /* kuboBastinSeaFromChebmom.cpp:35-93 */
  LQ_LIN_BASTIN_SURF = 6  /* kuboBastinSurfFromChebmom.cpp:35-86 */
};
int lq_linear_grid(lq_ctx ctx, int kind, const double *mu, int M0, int M1, int n, double W, double Ec,
                   int num_div, double *E, double *out);

/* ---- batched spectral-moment jobs ------------------------------------------ */
/* Many SpectralMoments jobs (OP = identity, src/chebyshev_solver.cpp:176-200),
 * each with its own rescaled H, states and M, advanced together: one grouped
 * launch per recursion step per vector-width class.  This is the submit/poll
 * surface a serving loop drives (continuous batching): jobs submitted between
 * lq_batch_run calls join at the next chunk boundary when their device
 * footprint fits the budget.  A job's moments are bitwise identical alone,
 * in any mix, and to lq_spectral_moments on a fresh mu (R <= 64). */
int lq_batch_create(lq_ctx ctx, size_t hbm_budget /* 0: 80% of free */, int chunk /* steps, 0: 32 */,
                    lq_batch *b);
int lq_batch_destroy(lq_batch b);
/* Device bytes a job holds while it runs (the admission footprint). */
int lq_batch_job_bytes(int n, int R, int M, int chunk, size_t *bytes);
/* Queue a job (1 <= R <= 64).  Host phi is copied; a device phi must stay
 * valid until the job is admitted.  H must stay alive until it finishes. */
int lq_batch_submit(lq_batch b, lq_csr H, const double *phi, int R, int M, int *job);
/* Run chunks until nothing is queued or running (max_chunks <= 0) or for
 * max_chunks chunks; reports the jobs still running and queued. */
int lq_batch_run(lq_batch b, int max_chunks, int *running, int *queued);
/* state: 0 queued, 1 running, 2 finished. */
int lq_batch_poll(lq_batch b, int job, int *state);
/* mu (M complex) of a finished job, which is then released. */
int lq_batch_result(lq_batch b, int job, double *mu);
/* Gather unroll the tuner chose for jobs with R states (0: still tuning). */
int lq_batch_unroll(lq_batch b, int R, int *unroll);

/* ---- file ingestion (host side, no GPU needed) ------------------------------ */
/* File formats of the writers.  Every reader accepts the text layout and the
 * binary one (told apart by the binary magic). */
enum lq_file_format {
  LQ_FMT_TEXT = 0,     /* the reference's text layout, 17 significant digits (exact round trip) */
  LQ_FMT_TEXT_REF = 1, /* the reference's text layout at its writers' precision: 6 digits for
                          CSR (ofstream default), 15 for moment files (precision(digits10)) */
  LQ_FMT_BINARY = 2    /* "LQCSR001" / "LQMOM001" binary, little endian */
};

/* Header of a moment file: kind 1 = .chebmom1D (M1 = 1), 2 = .chebmom2D,
 * 3 = .chebmomTD (M0 moments x M1 time steps, dt set). */
typedef struct lq_mom_header {
  int kind;
  int M0, M1;
  long long n;
  double W, Ec, dt;
} lq_mom_header;

/* Replaces Sparse::OPERATOR_FromCSRFile (src/sparse_matrix.cpp:3-50): the
 * CSR file (text or binary) as zero-based host arrays.  Call with rowptr ==
 * NULL for n and nnz only, then with arrays of n+1, nnz and 2*nnz entries.
 * Text is parsed by several threads (LQ_IO_THREADS caps them); every value is
 * bitwise what `ifstream >> double` reads. */
int lq_csr_file_read(const char *path, int *n, long long *nnz, int *rowptr, int *col, double *val);
/* Writes a CSR file; LQ_FMT_TEXT_REF reproduces
 * hopping_list::save_hopping_list_as_csr byte for byte
 * (utilities/wannier2sparse/src/hopping_list.cpp:262-283). */
int lq_csr_file_write(const char *path, int format, int n, long long nnz, const int *rowptr, const int *col,
                      const double *val);
/* Replaces SparseMatrixBuilder::BuildOPFromCSRFile
 * (include/sparse_matrix.hpp:100-109): lq_csr_file_read + lq_csr_create. */
int lq_csr_file_load(lq_ctx ctx, const char *path, int fill, lq_csr *out);
/* Replaces the Moments1D / Moments2D / MomentsTD file constructors
 * (src/chebyshev_moments1D.cpp:19-51, src/chebyshev_moments2D.cpp:16-51,
 * src/chebyshev_momentsTD.cpp:48-79): header, then (mu != NULL) the
 * M0*M1 complex moments, row major. */
int lq_mom_file_read(const char *path, lq_mom_header *h, double *mu);
/* Moments1D/2D/TD::saveIn (src/chebyshev_moments1D.cpp:91-103,
 * src/chebyshev_moments2D.cpp:53-68, src/chebyshev_momentsTD.cpp:81-93). */
int lq_mom_file_write(const char *path, int format, const lq_mom_header *h, const double *mu);

#ifdef __cplusplus
}
#endif
#endif
