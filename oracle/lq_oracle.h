/*
 * lq_oracle.h -- CPU FP64 oracle for the LinQT Chebyshev-moment path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is
 * compared against (tests/, __graft_entry__.smoke(), and the cpu_baseline leg
 * of bench.py).  No product code links or calls it.
 *
 * It restates, in plain C99, the reference's algorithm and loop structure:
 *   - MKL Hermitian/upper CSR mv ........ src/mkl_sparse_matrix.cpp:56-69, include/sparse_matrix.hpp:50-55
 *   - Rescale (C = aH + bI) ............. src/mkl_sparse_matrix.cpp:73-97
 *   - SpectralMoments ................... src/chebyshev_solver.cpp:176-200
 *   - CorrelationExpansionMoments ....... src/chebyshev_solver.cpp:26-172, src/chebyshev_vectors.cpp:4-51
 *   - Jackson kernels ................... src/chebyshev_moments1D.cpp:68-87, src/chebyshev_moments2D.cpp:87-117,
 *                                         include/Kubo_solver_FFT.hpp:99-107
 *   - Kubo_solver_FFT (Greenwood/Bastin). src/Kubo_solver_FFT.cpp, src/Greenwood_FFTs.cpp, src/Bastin_FFTs.cpp,
 *                                         src/Kubo_solver_FFT_postProcess.cpp
 *   - non-orthogonal SpectralMoments .... src/chebyshev_solver.cpp:203-229, src/chebyshev_moments.cpp:63-153,
 *                                         src/mkl_linear_algebra.cpp:100-131 (Eigen BiCGSTAB)
 *   - supercell operators ............... utilities/wannier2sparse/src/hopping_list.cpp:156-285
 *   - FFT-grid post-processors .......... src/spectralFunctionFromChebmom_FFTgrid.cpp:35-73,
 *                                         src/kuboGreenwoodFromChebmom_FFTgrid.cpp:35-76,
 *                                         src/kuboBastinFromChebmom_FFTgrid.cpp:57-106
 *
 * Complex numbers cross this interface as interleaved (re, im) doubles.
 * Every function returns 0 on success and a negative code on bad arguments.
 */
#ifndef LQ_ORACLE_H
#define LQ_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Sparse operator as held by the reference's SparseMatrixType: a copy of the
 * CSR arrays (zero based, rows sorted).  mv uses MKL's HERMITIAN / UPPER
 * descriptor semantics: entries with col < row are ignored, every entry with
 * col > row also acts at (col, row) conjugated. */
typedef struct lqo_op lqo_op;

lqo_op *lqo_op_create(int n, int nnz, const int *rowptr, const int *col, const double *val);
void    lqo_op_destroy(lqo_op *op);
int     lqo_op_nnz(const lqo_op *op);
/* H <- a*H + b*I (MKL mkl_sparse_z_add of H and b*I); inserts missing diagonals. */
int     lqo_op_rescale(lqo_op *op, double a_re, double a_im, double b_re, double b_im);
/* y <- a*Herm(U)*x + b*y.  When b == 0, y is overwritten. */
int     lqo_op_mv(const lqo_op *op, double a_re, double a_im, const double *x,
                  double b_re, double b_im, double *y);
/* Dense Hermitian completion of op (n*n complex, row major) -- tiny n only. */
int     lqo_op_dense(const lqo_op *op, double *dense);

/* chebyshev::SpectralMoments.  OP == NULL means the identity operator.
 * phi: R states of length n, state r at phi[2*(r*n + i)].  mu (M complex) is
 * accumulated into, as the reference's Moments1D is. */
int lqo_spectral_moments(const lqo_op *H, const lqo_op *OP, int n,
                         const double *phi, int R, int M, double *mu);

/* chebyshev::CorrelationExpansionMoments (the generator-driven overload).
 * batch <= 0 selects the default BATCH_SIZE = M0.  mu (M0*M1 complex, row
 * major m0*M1+m1) is in/out exactly as the reference's Moments2D is:
 * accumulated into when no padding happens, replaced when it does. */
int lqo_correlation_moments(const lqo_op *H, const lqo_op *OPL, const lqo_op *OPR,
                            int n, const double *phi, int R, int M0, int M1,
                            int batch, double *mu);

/* Unnormalised backward DFT, out[k] = sum_m in[m] exp(+2 pi i m k / len)
 * (FFTW_BACKWARD definition).  Radix-2 when len is a power of two. */
int lqo_dft_backward(int len, const double *in, double *out);

/* Kubo_solver_FFT::Jackson_kernel(m, M). */
double lqo_kubo_jackson(int m, int M);

/* Kubo_solver_FFT::compute, restated with the per-site DFTs.  formula 0 =
 * KUBO_GREENWOOD, 1 = KUBO_BASTIN.  final_data (2*nump complex) receives the
 * running mean over states; last_r_data (optional, 2*nump complex) the last
 * state's r_data. */
int lqo_kubo_fft(const lqo_op *H, const lqo_op *OPL, const lqo_op *OPR, int n,
                 const double *phi, int R, int M, int num_sections, int nump,
                 int formula, double *final_data, double *last_r_data);

/* Kubo_solver_FFT_postProcess: energies written as a*E - Ec (reference
 * convention), ascending.  sigma receives the file's second column;
 * integrand (Bastin only, may be NULL) the `_integrand` file column. */
int lqo_kubo_post(int formula, const double *final_data, int nump, int n,
                  double W, double Ec, double *E_out, double *sigma_out,
                  double *integrand_out);

/* Moments1D::ApplyJacksonKernel: returns the truncated moment count
 * (mu compacted in place). */
int lqo_jackson_1d(double *mu, int M, double W, double broad_meV);
/* Moments2D::ApplyJacksonKernel: mu compacted in place to M0o x M1o. */
int lqo_jackson_2d(double *mu, int M0, int M1, double W, double b0_meV,
                   double b1_meV, int *M0_out, int *M1_out);

/* FFT-grid post-processors on Jackson-damped moments.  num_div energies
 * E_i = cos(pi(2i+1/2)/num_div) (DOS: cos(pi(i+1/2)/num_div)).
 * DOS writes num_div rows; Greenwood and Bastin write num_div-1 rows. */
int lqo_dos_fftgrid(const double *mu, int M, int n, double W, double Ec,
                    int num_div, double *E_out, double *rho_out);
int lqo_greenwood_fftgrid(const double *mu, int M0, int M1, int n, double W,
                          double Ec, int num_div, double *E_out, double *out);
int lqo_bastin_fftgrid(const double *mu, int M0, int M1, int n, double W,
                       double Ec, int num_div, double *E_out, double *out);

/* chebyshev::TimeDependentCorrelations (src/chebyshev_solver.cpp:231-275):
 * mu (M x nT complex, row m stride nT) += (2/R)(1/2)^{delta_m0}
 * <OPL phi(t_n)|T_m(H~) OPR|phi(t_n)> with phi(t) evolved by MomentsTD::Evolve
 * (Bessel-Chebyshev expansion of exp(-iHt/hbar)), t_n = n dt (fs); H already
 * rescaled with W, Ec (the frequencies of the expansion).  OPL/OPR NULL = 1. */
int lqo_time_correlations(const lqo_op *H, const lqo_op *OPL, const lqo_op *OPR, int n, const double *phi,
                          int R, int M, int nT, double dt, double W, double Ec, double *mu);
/* MomentsTD::ApplyJacksonKernel; *M_out = truncated moment count (rows kept). */
int lqo_jackson_td(double *mu, int M, int nT, double W, double broad_meV, int *M_out);
/* timeCorrelationsFromChebmom.cpp: nT rows (t, value) at Fermi energy EF. */
int lqo_timecorr_post(const double *mu, int M, int nT, int n, double W, double Ec, double dt, double EF,
                      double *t_out, double *out);

/* Linear-grid post-processors (src/...FromChebmom.cpp): kind 0 spectral
 * function (1D mu of M0 moments), 1 Greenwood, 2 Bastin, 3 BastinI,
 * 4 BastinII, 5 BastinSea, 6 BastinSurf (2D mu, M0 x M1).
 * lqo_linear_kernel: the tool's kernel (with its scale factor, before any
 * running sum) at nx normalised energies x.
 * lqo_linear_grid: the whole tool on its grid of num_div energies
 * x_i = -xb + i*2xb/(num_div-1) (xb = 0.99 for Bastin, 1 otherwise); writes
 * the printed rows (num_div for kind 0, num_div-1 otherwise) and returns
 * their count, or -1. */
int lqo_linear_kernel(int kind, const double *mu, int M0, int M1, int n, double W, int nx,
                      const double *x, double *kern);
int lqo_linear_grid(int kind, const double *mu, int M0, int M1, int n, double W, double Ec,
                    int num_div, double *E_out, double *out);

/* ---- non-orthogonal basis ------------------------------------------------ */
/* linalg::orthogonalize (src/mkl_linear_algebra.cpp:100-131): x = S^-1 b by
 * Eigen's BiCGSTAB with its diagonal preconditioner, zero initial guess,
 * stopping when |r|^2 <= tol^2 |b|^2 or after maxit iterations.  S is the CSR
 * (n, rowptr, col, val) read as a general matrix -- the reference maps the
 * stored arrays into an Eigen::SparseMatrix, so a Hermitian-upper file acts
 * by its upper triangle.  Returns the iteration count (>= 0). */
int lqo_overlap_solve(int n, const int *rowptr, const int *col, const double *val, const double *b,
                      double *x, double tol, int maxit, double *error);
/* chebyshev::SpectralMoments_nonOrth (src/chebyshev_solver.cpp:203-229) with
 * Moments1D_nonOrth::SetInitVectors_nonOrthogonal / Iterate_nonOrthogonal
 * (src/chebyshev_moments.cpp:63-153): T0 = [OP] S^-1 phi, T1 = S^-1 H T0,
 * T_{m+1} = 2 S^-1 H T_m - T_{m-1}; mu_m += (2/R)(1/2 if m = 0) <phi|T_m>.
 * S as in lqo_overlap_solve; OP == NULL means the identity.  mu accumulated. */
int lqo_spectral_moments_nonorth(const lqo_op *H, const lqo_op *OP, int n, const int *s_rowptr, const int *s_col,
                                 const double *s_val, const double *phi, int R, int M, double tol, int maxit,
                                 double *mu);

/* ---- supercell operators (wannier2sparse) ------------------------------ */
/* hopping_list::wrap_in_supercell + save_hopping_list_as_csr
 * (utilities/wannier2sparse/src/hopping_list.cpp:156-225, 227-285) on a
 * unit-cell list of constant hoppings: cell (3 ints), edge (from, to; zero
 * based orbitals), val (complex) per hopping; dims = supercell size.  The
 * list is first put in the order of the reference's multimap (the tag
 * string "c0 c1 c2 e0 e1 ", equal tags in input order), every hopping is
 * replicated over the supercell shifts (s2, s1, s0 loops), hoppings with
 * |v| / mean|v| <= 1e-10 or to < from are dropped, and duplicates are summed
 * in that order (Eigen setFromTriplets).  Output: the stored upper CSR.
 * Call with rowptr == NULL for *n and *nnz; the mean is over the unit list
 * (every hopping appears Ncells times in the supercell). */
int lqo_supercell_csr(int nwann, int nhop, const int *cell, const int *edge, const double *val,
                      const int *dims, int *n, long long *nnz, int *rowptr, int *col, double *out_val);

/* Sparse::OPERATOR_FromCSRFile (src/sparse_matrix.cpp:3-50): the text CSR
 * read one token at a time from a stream (here C stdio, as the reference
 * reads with ifstream >>).  rowptr == NULL: *n and *nnz only. */
int lqo_csr_read_stream(const char *path, int *n, long long *nnz, int *rowptr, int *col, double *val);

/* ---- OpenMP CPU baseline (same algorithm, threads over rows/pairs) ----- */
int  lqo_omp_threads(void);
/* One Chebyshev step on a completed operator: y <- 2*H~*x - y for k vectors
 * (vector r at x[2*(r*n+i)]), row-parallel over the completed CSR. */
int  lqo_omp_cheb_step(const lqo_op *H, int k, const double *x, double *y);
/* zdotc table sub(m0,m1) = <L_m0|R_m1> over `count` pairs starting at m0,
 * i.e. ComputeMomTable restricted to rows [m0_begin, m0_end). */
int  lqo_omp_momtable(int n, int M1, const double *L, const double *Rv,
                      int m0_begin, int m0_end, double *sub);
/* y <- op*x for k vectors (vector r at x[2*(r*n+i)]), row-parallel. */
int  lqo_omp_mv(const lqo_op *op, int k, const double *x, double *y);
/* Bastin_FFTs' per-site stage over `size` sites, threads over sites with
 * per-thread accumulators as the reference does (src/Bastin_FFTs.cpp:120-267).
 * L, Rv: M vectors of length size (vector m at L[2*(m*size + l)]).
 * r_data (2*nump complex) is accumulated into. */
int  lqo_omp_bastin_ffts(const double *L, const double *Rv, int M, int nump, int size,
                         double *r_data);
/* Full OpenMP SpectralMoments (row-parallel Hermitian mv + zdotc). */
int  lqo_omp_spectral_moments(const lqo_op *H, int n, const double *phi,
                              int R, int M, double *mu);

#ifdef __cplusplus
}
#endif
#endif
