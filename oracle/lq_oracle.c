/*
 * lq_oracle.c -- CPU FP64 oracle (test infrastructure, see lq_oracle.h).
 *
 * A plain C99 restatement of the reference's Chebyshev-moment path.  It keeps
 * the reference's loop structure (state loop, recursion with vector swaps,
 * BATCH_SIZE blocks, per-site DFTs) so that it can serve both as the parity
 * checker and as the "reference algorithm, CPU restatement" baseline.  Every
 * function cites the reference file:line it follows.
 */
#include "lq_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

struct lqo_op {
  int n;
  int nnz;
  int *rowptr;
  int *col;
  cplx *val;
  /* Hermitian completion (full CSR) used only by the OpenMP baseline. */
  int fnnz;
  int *frow;
  int *fcol;
  cplx *fval;
};

static void op_drop_full(lqo_op *op) {
  free(op->frow);
  free(op->fcol);
  free(op->fval);
  op->frow = op->fcol = NULL;
  op->fval = NULL;
  op->fnnz = 0;
}

/* SparseMatrixType::ConvertFromCSR copies the three arrays
 * (src/mkl_sparse_matrix.cpp:39-54). */
lqo_op *lqo_op_create(int n, int nnz, const int *rowptr, const int *col, const double *val) {
  if (n <= 0 || nnz < 0 || !rowptr || (nnz > 0 && (!col || !val))) return NULL;
  if (rowptr[0] != 0 || rowptr[n] != nnz) return NULL;
  lqo_op *op = (lqo_op *)calloc(1, sizeof(lqo_op));
  op->n = n;
  op->nnz = nnz;
  op->rowptr = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  op->col = (int *)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
  op->val = (cplx *)malloc(sizeof(cplx) * (size_t)(nnz > 0 ? nnz : 1));
  memcpy(op->rowptr, rowptr, sizeof(int) * (size_t)(n + 1));
  if (nnz > 0) {
    memcpy(op->col, col, sizeof(int) * (size_t)nnz);
    memcpy(op->val, val, sizeof(cplx) * (size_t)nnz);
  }
  for (int p = 0; p < nnz; ++p)
    if (op->col[p] < 0 || op->col[p] >= n) {
      lqo_op_destroy(op);
      return NULL;
    }
  return op;
}

void lqo_op_destroy(lqo_op *op) {
  if (!op) return;
  op_drop_full(op);
  free(op->rowptr);
  free(op->col);
  free(op->val);
  free(op);
}

int lqo_op_nnz(const lqo_op *op) { return op ? op->nnz : -1; }

/* SparseMatrixType::Rescale (src/mkl_sparse_matrix.cpp:73-97): C = a*A + B
 * with B = b*I, through mkl_sparse_z_add.  The pattern of C is the union of
 * A's pattern and the diagonal; a row without a stored diagonal gains one. */
int lqo_op_rescale(lqo_op *op, double a_re, double a_im, double b_re, double b_im) {
  if (!op) return -1;
  const cplx a = a_re + I * a_im, b = b_re + I * b_im;
  const int n = op->n;
  int *nrow = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  int *ncol = (int *)malloc(sizeof(int) * (size_t)(op->nnz + n));
  cplx *nval = (cplx *)malloc(sizeof(cplx) * (size_t)(op->nnz + n));
  int q = 0;
  nrow[0] = 0;
  for (int i = 0; i < n; ++i) {
    int has_diag = 0, inserted = 0;
    for (int p = op->rowptr[i]; p < op->rowptr[i + 1]; ++p)
      if (op->col[p] == i) has_diag = 1;
    for (int p = op->rowptr[i]; p < op->rowptr[i + 1]; ++p) {
      const int j = op->col[p];
      if (!has_diag && !inserted && j > i) {
        ncol[q] = i;
        nval[q] = b;
        ++q;
        inserted = 1;
      }
      cplx v = a * op->val[p];
      if (j == i && has_diag == 1) {
        v = v + b;
        has_diag = 2; /* b goes to the first stored diagonal only */
      }
      ncol[q] = j;
      nval[q] = v;
      ++q;
    }
    if (!has_diag && !inserted) {
      ncol[q] = i;
      nval[q] = b;
      ++q;
    }
    nrow[i + 1] = q;
  }
  free(op->rowptr);
  free(op->col);
  free(op->val);
  op->rowptr = nrow;
  op->col = ncol;
  op->val = nval;
  op->nnz = q;
  op_drop_full(op);
  return 0;
}

/* mkl_sparse_z_mv with SPARSE_MATRIX_TYPE_HERMITIAN / SPARSE_FILL_MODE_UPPER
 * (include/sparse_matrix.hpp:50-55, src/mkl_sparse_matrix.cpp:56-69):
 * y = a*(U + U^H - diag U)*x + b*y, where U holds the stored entries with
 * col >= row.  Strictly-lower stored entries are not referenced. */
static void herm_upper_apply(const lqo_op *op, const cplx *x, cplx *t) {
  const int n = op->n;
  for (int i = 0; i < n; ++i) t[i] = 0;
  for (int i = 0; i < n; ++i) {
    cplx acc = 0;
    for (int p = op->rowptr[i]; p < op->rowptr[i + 1]; ++p) {
      const int j = op->col[p];
      if (j < i) continue;
      const cplx v = op->val[p];
      acc += v * x[j];
      if (j != i) t[j] += conj(v) * x[i];
    }
    t[i] += acc;
  }
}

static void op_mv_c(const lqo_op *op, cplx a, const cplx *x, cplx b, cplx *y) {
  const int n = op->n;
  cplx *t = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  herm_upper_apply(op, x, t);
  if (b == 0) {
    for (int i = 0; i < n; ++i) y[i] = a * t[i];
  } else {
    for (int i = 0; i < n; ++i) y[i] = a * t[i] + b * y[i];
  }
  free(t);
}

int lqo_op_mv(const lqo_op *op, double a_re, double a_im, const double *x,
              double b_re, double b_im, double *y) {
  if (!op || !x || !y) return -1;
  op_mv_c(op, a_re + I * a_im, (const cplx *)x, b_re + I * b_im, (cplx *)y);
  return 0;
}

int lqo_op_dense(const lqo_op *op, double *dense) {
  if (!op || !dense) return -1;
  const int n = op->n;
  cplx *d = (cplx *)dense;
  memset(d, 0, sizeof(cplx) * (size_t)n * (size_t)n);
  for (int i = 0; i < n; ++i)
    for (int p = op->rowptr[i]; p < op->rowptr[i + 1]; ++p) {
      const int j = op->col[p];
      if (j < i) continue;
      d[(size_t)i * n + j] += op->val[p];
      if (j != i) d[(size_t)j * n + i] += conj(op->val[p]);
    }
  return 0;
}

/* linalg::vdot = cblas_zdotc_sub: sum conj(x_i) y_i (src/mkl_linear_algebra.cpp:24-30). */
static cplx vdot(int n, const cplx *x, const cplx *y) {
  cplx s = 0;
  for (int i = 0; i < n; ++i) s += conj(x[i]) * y[i];
  return s;
}

/* ---- recursion state: the reference's Moments::Chebyshev0/1 ------------ */
typedef struct {
  const lqo_op *H;
  int n;
  cplx *c0, *c1;
} cheb_state;

/* Moments::SetInitVectors(T0) (src/chebyshev_moments.cpp:45-59):
 * Ch0 = T0, Ch1 = H*Ch0. */
static void set_init(cheb_state *s, const cplx *T0) {
  memcpy(s->c0, T0, sizeof(cplx) * (size_t)s->n);
  op_mv_c(s->H, 1.0, s->c0, 0.0, s->c1);
}

/* Moments::SetInitVectors(OP, T0) (src/chebyshev_moments.cpp:5-22):
 * Ch1 = T0, Ch0 = OP*Ch1, Ch1 = H*Ch0.  OP == NULL acts as the identity. */
static void set_init_op(cheb_state *s, const lqo_op *OP, const cplx *T0) {
  memcpy(s->c1, T0, sizeof(cplx) * (size_t)s->n);
  if (OP)
    op_mv_c(OP, 1.0, s->c1, 0.0, s->c0);
  else
    memcpy(s->c0, s->c1, sizeof(cplx) * (size_t)s->n);
  op_mv_c(s->H, 1.0, s->c0, 0.0, s->c1);
}

/* Moments::Iterate (src/chebyshev_moments.cpp:157-162):
 * Ch0 = 2*H*Ch1 - Ch0, then swap. */
static void iterate(cheb_state *s) {
  op_mv_c(s->H, 2.0, s->c1, -1.0, s->c0);
  cplx *t = s->c0;
  s->c0 = s->c1;
  s->c1 = t;
}

/* chebyshev::SpectralMoments (src/chebyshev_solver.cpp:176-200). */
int lqo_spectral_moments(const lqo_op *H, const lqo_op *OP, int n, const double *phi,
                         int R, int M, double *mu_) {
  if (!H || H->n != n || (OP && OP->n != n) || !phi || R <= 0 || M <= 0 || !mu_) return -1;
  cplx *mu = (cplx *)mu_;
  cheb_state s = {H, n, (cplx *)malloc(sizeof(cplx) * (size_t)n),
                  (cplx *)malloc(sizeof(cplx) * (size_t)n)};
  for (int r = 0; r < R; ++r) {
    const cplx *Phi = (const cplx *)phi + (size_t)r * n;
    if (!OP)
      set_init(&s, Phi);
    else
      set_init_op(&s, OP, Phi);
    for (int m = 0; m < M; ++m) {
      double scal = 2.0 / R;
      if (m == 0) scal *= 0.5;
      mu[m] += scal * vdot(n, Phi, s.c0);
      iterate(&s);
    }
  }
  free(s.c0);
  free(s.c1);
  return 0;
}

/* ---- time-dependent correlations ----------------------------------------- */
static double delta_chebF(double x, double m);
#define LQO_HBAR 0.6582119624 /* eV fs (include/chebyshev_moments.hpp:290) */

/* chebyshev::besselJ (src/special_functions.cpp:3-17): libm's j0, j1, jn. */
static double bessel_j(int n, double x) { return n == 0 ? j0(x) : n == 1 ? j1(x) : jn(n, x); }

/* MomentsTD::Evolve (src/chebyshev_momentsTD.cpp:4-45): Phi <- exp(-i H dt/hbar) Phi
 * as sum_n c_n T_n(H~) Phi, c_0 = J_0(x), c_n = 2 (-i)^n J_n(x), x = dt a/hbar, terms
 * added while |J_n| > 1e-10, then the phase exp(-i Ec dt/hbar). */
static void evolve(cheb_state *s, cplx *Phi, double a, double Ec, double dt) {
  const int n = s->n;
  const double x = dt * (a / LQO_HBAR), tol = 1e-10;
  double momcut = 2.0 * tol;
  memcpy(s->c0, Phi, sizeof(cplx) * (size_t)n);
  op_mv_c(s->H, 1.0, s->c0, 0.0, s->c1);
  int k = 0;
  cplx nIp = 1.0;
  double Jn = 0.5 * bessel_j(k, x);
  for (int i = 0; i < n; ++i) Phi[i] = 0.0;
  while (momcut > tol) {
    const cplx c = nIp * 2.0 * Jn;
    for (int i = 0; i < n; ++i) Phi[i] += c * s->c0[i];
    iterate(s);
    nIp *= -I;
    ++k;
    Jn = bessel_j(k, x);
    momcut = fabs(Jn);
  }
  const cplx e0 = cexp(-I * (Ec / LQO_HBAR) * dt);
  for (int i = 0; i < n; ++i) Phi[i] *= e0;
}

/* chebyshev::TimeDependentCorrelations (src/chebyshev_solver.cpp:231-275). */
int lqo_time_correlations(const lqo_op *H, const lqo_op *OPL, const lqo_op *OPR, int n, const double *phi, int R,
                          int M, int nT, double dt, double W, double Ec, double *mu_) {
  if (!H || H->n != n || (OPL && OPL->n != n) || (OPR && OPR->n != n) || !phi || R <= 0 || M <= 0 || nT <= 0 ||
      !mu_)
    return -1;
  cplx *mu = (cplx *)mu_;
  const double a = W / 2.0;
  cheb_state s = {H, n, (cplx *)malloc(sizeof(cplx) * (size_t)n), (cplx *)malloc(sizeof(cplx) * (size_t)n)};
  cplx *PhiR = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  cplx *PhiL = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  for (int r = 0; r < R; ++r) {
    memcpy(PhiR, (const cplx *)phi + (size_t)r * n, sizeof(cplx) * (size_t)n);
    if (OPL)
      op_mv_c(OPL, 1.0, PhiR, 0.0, PhiL);
    else
      memcpy(PhiL, PhiR, sizeof(cplx) * (size_t)n);
    for (int t = 0; t < nT; ++t) {
      set_init_op(&s, OPR, PhiR);
      for (int m = 0; m < M; ++m) {
        double scal = 2.0 / R;
        if (m == 0) scal *= 0.5;
        mu[(size_t)m * nT + t] += scal * vdot(n, PhiL, s.c0);
        iterate(&s);
      }
      evolve(&s, PhiL, a, Ec, dt);
      evolve(&s, PhiR, a, Ec, dt);
    }
  }
  free(s.c0);
  free(s.c1);
  free(PhiR);
  free(PhiL);
  return 0;
}

/* MomentsTD::ApplyJacksonKernel (src/chebyshev_momentsTD.cpp:135-154): truncate
 * to ceil(pi/eta) moments (eta = 2 b/1000/W) and damp every time column. */
int lqo_jackson_td(double *mu_, int M, int nT, double W, double broad, int *M_out) {
  if (!mu_ || M <= 0 || nT <= 0 || broad <= 0 || !M_out) return -1;
  cplx *mu = (cplx *)mu_;
  const double eta = 2.0 * broad / 1000 / W;
  int maxMom = (int)ceil(M_PI / eta);
  if (maxMom > M) maxMom = M;
  const double phi_J = M_PI / (double)(maxMom + 1.0);
  for (int m = 0; m < maxMom; ++m) {
    const double g = ((maxMom - m + 1) * cos(phi_J * m) + sin(phi_J * m) * cos(phi_J) / sin(phi_J)) * phi_J / M_PI;
    for (int t = 0; t < nT; ++t) mu[(size_t)m * nT + t] *= g;
  }
  *M_out = maxMom;
  return 0;
}

/* timeCorrelationsFromChebmom.cpp:41-70: rows (n dt, N/a sum_m delta_m(x_F) Re mu(m, n))
 * with x_F = (E_F - Ec)/a; mu holds M (truncated) x nT moments, row stride nT. */
int lqo_timecorr_post(const double *mu_, int M, int nT, int n, double W, double Ec, double dt, double EF,
                      double *t_out, double *out) {
  if (!mu_ || M <= 0 || nT <= 0 || !t_out || !out) return -1;
  const cplx *mu = (const cplx *)mu_;
  const double a = W / 2.0, x = (EF - Ec) / a;
  for (int t = 0; t < nT; ++t) {
    double o = 0.0;
    for (int m = 0; m < M; ++m) o += delta_chebF(x, m) * creal(mu[(size_t)m * nT + t]);
    o *= n / a;
    t_out[t] = t * dt;
    out[t] = o;
  }
  return 0;
}

/* ---- stored Chebyshev vectors: Vectors_sliced / Vectors ---------------- */
typedef struct {
  cheb_state st;
  int nvec;     /* NumberOfVectors */
  int ns;       /* num_sections */
  int sec_size; /* SystemSize / ns */
  int last_size;/* SystemSize / ns + SystemSize % ns */
  cplx *list;   /* nvec x last_size (VectorList, zero initialised) */
} vec_store;

static void store_init(vec_store *v, const lqo_op *H, int n, int nvec, int ns) {
  v->st.H = H;
  v->st.n = n;
  v->st.c0 = (cplx *)calloc((size_t)n, sizeof(cplx));
  v->st.c1 = (cplx *)calloc((size_t)n, sizeof(cplx));
  v->nvec = nvec;
  v->ns = ns;
  v->sec_size = n / ns;
  v->last_size = n / ns + n % ns;
  v->list = (cplx *)calloc((size_t)nvec * (size_t)v->last_size, sizeof(cplx));
}

static void store_free(vec_store *v) {
  free(v->st.c0);
  free(v->st.c1);
  free(v->list);
}

static cplx *store_vec(vec_store *v, int m) { return v->list + (size_t)m * v->last_size; }

/* Vectors_sliced::IterateAllSliced (src/chebyshev_vectors.cpp:4-21).  Stores
 * Ch0, then for m >= 1 stores Ch1 and advances; on return Ch0/Ch1 hold
 * T_{k+B-1}/T_{k+B}, so a following call starts one order early. */
static void iterate_all_sliced(vec_store *v, int s) {
  const int seg = (s == v->ns - 1) ? v->last_size : v->sec_size;
  const size_t start = (size_t)s * v->sec_size;
  memcpy(store_vec(v, 0), v->st.c0 + start, sizeof(cplx) * (size_t)seg);
  for (int m = 1; m < v->nvec; ++m) {
    memcpy(store_vec(v, m), v->st.c1 + start, sizeof(cplx) * (size_t)seg);
    op_mv_c(v->st.H, 2.0, v->st.c1, -1.0, v->st.c0);
    cplx *t = v->st.c0;
    v->st.c0 = v->st.c1;
    v->st.c1 = t;
  }
}

/* Vectors_sliced::MultiplySliced (src/chebyshev_vectors.cpp:24-51): each
 * stored section is placed into an otherwise-zero full vector, multiplied by
 * OP, and the same section of the product is stored back. */
static void multiply_sliced(vec_store *v, const lqo_op *OP, int s) {
  const int n = v->st.n;
  const int seg = (s == v->ns - 1) ? v->last_size : v->sec_size;
  const size_t start = (size_t)s * v->sec_size;
  cplx *opv = (cplx *)calloc((size_t)n, sizeof(cplx));
  cplx *tmp2 = (cplx *)calloc((size_t)n, sizeof(cplx));
  for (int m = 0; m < v->nvec; ++m) {
    memcpy(opv + start, store_vec(v, m), sizeof(cplx) * (size_t)seg);
    if (OP)
      op_mv_c(OP, 1.0, opv, 0.0, tmp2);
    else
      memcpy(tmp2, opv, sizeof(cplx) * (size_t)n);
    memcpy(store_vec(v, m), tmp2 + start, sizeof(cplx) * (size_t)seg);
  }
  free(opv);
  free(tmp2);
}

/* chebyshev::CorrelationExpansionMoments, both overloads
 * (src/chebyshev_solver.cpp:26-172).  Square tables follow the reference
 * exactly.  For M0 != M1 the reference indexes past its table
 * (chebyshev_solver.cpp:55-56 use the larger count on both axes); this
 * restatement defines that case as the leading M0 x M1 block of the square
 * max(M0, M1) computation. */
int lqo_correlation_moments(const lqo_op *H, const lqo_op *OPL, const lqo_op *OPR,
                            int n, const double *phi, int R, int M0, int M1,
                            int batch, double *mu_) {
  if (!H || H->n != n || (OPL && OPL->n != n) || (OPR && OPR->n != n) || !phi || R <= 0 ||
      M0 <= 0 || M1 <= 0 || !mu_)
    return -1;
  cplx *mu_io = (cplx *)mu_;
  /* GetBatchSize (include/kpm_noneqop.hpp:27-39): unset/invalid -> M0. */
  const int B = (batch > 0) ? batch : M0;
  const int Msq = (M0 > M1) ? M0 : M1;
  /* padding to a multiple of B (chebyshev_solver.cpp:99-114) */
  int P = Msq;
  int padded = 0;
  if (M0 % B != 0 || M1 % B != 0 || M0 != M1) {
    P = ((Msq + B - 1) / B) * B;
    padded = (M0 % B != 0 || M1 % B != 0);
  }
  cplx *mu = (cplx *)calloc((size_t)P * P, sizeof(cplx));
  if (!padded) /* the caller's table is accumulated into */
    for (int m0 = 0; m0 < M0; ++m0)
      for (int m1 = 0; m1 < M1; ++m1) mu[(size_t)m0 * P + m1] = mu_io[(size_t)m0 * M1 + m1];

  vec_store L, Rv;
  store_init(&L, H, n, B, 1);
  store_init(&Rv, H, n, B, 1);
  cplx *sub = (cplx *)malloc(sizeof(cplx) * (size_t)B * B);
  for (int r = 0; r < R; ++r) {
    const cplx *Phi = (const cplx *)phi + (size_t)r * n;
    /* inner overload (chebyshev_solver.cpp:45-82) */
    set_init_op(&L.st, OPL, Phi);
    for (int mL = 0; mL < P; mL += B) {
      iterate_all_sliced(&L, 0);
      set_init(&Rv.st, Phi);
      for (int mR = 0; mR < P; mR += B) {
        iterate_all_sliced(&Rv, 0);
        multiply_sliced(&Rv, OPR, 0);
        /* ComputeMomTable (chebyshev_solver.cpp:26-42) */
        for (int m0 = 0; m0 < B; ++m0)
          for (int m1 = 0; m1 < B; ++m1)
            sub[(size_t)m0 * B + m1] = vdot(n, store_vec(&L, m0), store_vec(&Rv, m1));
        /* Moments2D::AddSubMatrix (include/chebyshev_moments.hpp:265-270) */
        for (int m0 = 0; m0 < B; ++m0)
          for (int m1 = 0; m1 < B; ++m1) mu[(size_t)(mL + m0) * P + (mR + m1)] += sub[(size_t)m0 * B + m1];
      }
    }
  }
  /* scaling and Hermitian symmetrisation (chebyshev_solver.cpp:149-162) */
  for (int mL = 0; mL < P; ++mL)
    for (int mR = mL; mR < P; ++mR) {
      double scal = 4.0 / R;
      if (mL == 0) scal *= 0.5;
      if (mR == 0) scal *= 0.5;
      const cplx tmp = scal * (mu[(size_t)mL * P + mR] + conj(mu[(size_t)mR * P + mL])) / 2.0;
      mu[(size_t)mL * P + mR] = tmp;
      mu[(size_t)mR * P + mL] = conj(tmp);
    }
  /* truncation back to the requested size (chebyshev_solver.cpp:165-169) */
  for (int m0 = 0; m0 < M0; ++m0)
    for (int m1 = 0; m1 < M1; ++m1) mu_io[(size_t)m0 * M1 + m1] = mu[(size_t)m0 * P + m1];
  free(sub);
  store_free(&L);
  store_free(&Rv);
  free(mu);
  return 0;
}

/* ---- DFT (FFTW_BACKWARD definition, include/fftw_wrapper.hpp:50-96) ----
 * A plan holds the twiddles exp(+2 pi i k / len) and, for powers of two, the
 * bit-reversal permutation of an iterative radix-2 Cooley-Tukey transform;
 * other lengths use the direct O(len^2) sum with the same twiddle table. */
typedef struct {
  int len, pow2;
  int *rev;
  cplx *tw;
} dft_plan;

static dft_plan *plan_create(int len) {
  dft_plan *p = (dft_plan *)calloc(1, sizeof(dft_plan));
  p->len = len;
  p->pow2 = len > 0 && (len & (len - 1)) == 0;
  p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)len);
  for (int k = 0; k < len; ++k) p->tw[k] = cos(2.0 * M_PI * k / len) + I * sin(2.0 * M_PI * k / len);
  if (p->pow2) {
    int bits = 0;
    while ((1 << bits) < len) ++bits;
    p->rev = (int *)malloc(sizeof(int) * (size_t)len);
    for (int i = 0; i < len; ++i) {
      int r = 0;
      for (int b = 0; b < bits; ++b)
        if (i & (1 << b)) r |= 1 << (bits - 1 - b);
      p->rev[i] = r;
    }
  }
  return p;
}

static void plan_free(dft_plan *p) {
  if (!p) return;
  free(p->rev);
  free(p->tw);
  free(p);
}

static void plan_exec(const dft_plan *p, const cplx *in, cplx *out) {
  const int len = p->len;
  if (p->pow2) {
    for (int i = 0; i < len; ++i) out[p->rev[i]] = in[i];
    for (int half = 1; half < len; half <<= 1) {
      const int step = len / (2 * half);
      for (int j = 0; j < half; ++j) {
        const cplx w = p->tw[j * step];
        for (int k = j; k < len; k += 2 * half) {
          const cplx u = out[k], t = w * out[k + half];
          out[k] = u + t;
          out[k + half] = u - t;
        }
      }
    }
  } else {
    for (int k = 0; k < len; ++k) {
      cplx s = 0;
      for (int m = 0; m < len; ++m) s += in[m] * p->tw[(int)(((long long)m * k) % len)];
      out[k] = s;
    }
  }
}

static void dft_backward_c(int len, const cplx *in, cplx *out) {
  dft_plan *p = plan_create(len);
  plan_exec(p, in, out);
  plan_free(p);
}

int lqo_dft_backward(int len, const double *in, double *out) {
  if (len <= 0 || !in || !out || in == out) return -1;
  dft_backward_c(len, (const cplx *)in, (cplx *)out);
  return 0;
}

/* Kubo_solver_FFT::Jackson_kernel (include/Kubo_solver_FFT.hpp:99-107). */
double lqo_kubo_jackson(int m, int M) {
  return 1 / (M + 1.0) *
         ((M - m + 1.0) * cos(M_PI * m / (M + 1.0)) + sin(M_PI * m / (M + 1.0)) / tan(M_PI / (M + 1.0)));
}

/* Kubo_solver_FFT::Greenwood_FFTs (src/Greenwood_FFTs.cpp:6-91), one thread. */
static void greenwood_ffts(vec_store *L, vec_store *Rv, int M, int nump, int size, cplx *r_data) {
  cplx *pre = (cplx *)malloc(sizeof(cplx) * (size_t)M);
  for (int m = 0; m < M; ++m)
    pre[m] = ((2 - (m == 0)) * lqo_kubo_jackson(m, M)) *
             (cos(M_PI * m / (2.0 * nump)) + I * sin(M_PI * m / (2.0 * nump)));
  cplx *in[4], *out[4];
  for (int q = 0; q < 4; ++q) {
    in[q] = (cplx *)calloc((size_t)nump, sizeof(cplx));
    out[q] = (cplx *)malloc(sizeof(cplx) * (size_t)nump);
  }
  double *thread_data = (double *)calloc((size_t)nump, sizeof(double));
  dft_plan *plan = plan_create(nump);
  for (int l = 0; l < size; ++l) {
    for (int m = 0; m < M; ++m) {
      const cplx bl = store_vec(L, m)[l], kr = store_vec(Rv, m)[l];
      in[0][m] = pre[m] * creal(bl);
      in[1][m] = pre[m] * cimag(bl);
      in[2][m] = pre[m] * creal(kr);
      in[3][m] = pre[m] * cimag(kr);
    }
    for (int q = 0; q < 4; ++q) plan_exec(plan, in[q], out[q]);
    for (int k = 0; k < nump; ++k)
      thread_data[k] += creal((creal(out[0][k]) - I * creal(out[1][k])) *
                              (creal(out[2][k]) + I * creal(out[3][k])));
  }
  for (int k = 0; k < nump; ++k) r_data[k] += thread_data[k];
  plan_free(plan);
  free(thread_data);
  for (int q = 0; q < 4; ++q) {
    free(in[q]);
    free(out[q]);
  }
  free(pre);
}

/* Kubo_solver_FFT::Bastin_FFTs (src/Bastin_FFTs.cpp:6-174), one thread. */
static void bastin_ffts(vec_store *L, vec_store *Rv, int M, int nump, int size, cplx *r_data) {
  cplx *pre = (cplx *)malloc(sizeof(cplx) * (size_t)nump);
  for (int m = 0; m < M; ++m)
    pre[m] = ((2 - (m == 0)) * lqo_kubo_jackson(m, M)) *
             (cos(M_PI * m / (2.0 * nump)) + I * sin(M_PI * m / (2.0 * nump)));
  cplx *in[8], *o[8];
  for (int q = 0; q < 8; ++q) {
    in[q] = (cplx *)calloc((size_t)nump, sizeof(cplx));
    o[q] = (cplx *)malloc(sizeof(cplx) * (size_t)nump);
  }
  cplx *p = (cplx *)calloc((size_t)nump, sizeof(cplx));
  cplx *w = (cplx *)calloc((size_t)nump, sizeof(cplx));
  /* o[0]=re_bras o[1]=im_bras o[2]=re_kets o[3]=im_kets, o[4..7] the D_ ones */
  dft_plan *plan = plan_create(nump);
  for (int l = 0; l < size; ++l) {
    for (int m = 0; m < M; ++m) {
      const cplx im_m = (double)m;
      const cplx bra = store_vec(L, m)[l];
      in[0][m] = pre[m] * creal(bra);
      in[1][m] = pre[m] * cimag(bra);
      in[4][m] = im_m * in[0][m];
      in[5][m] = im_m * in[1][m];
      const cplx ket = store_vec(Rv, m)[l];
      in[2][m] = pre[m] * creal(ket);
      in[3][m] = pre[m] * cimag(ket);
      in[6][m] = im_m * in[2][m];
      in[7][m] = im_m * in[3][m];
    }
    for (int q = 0; q < 8; ++q) plan_exec(plan, in[q], o[q]);
    for (int j = 0; j < nump / 2; ++j) {
      p[j] += (creal(o[0][j]) - I * creal(o[1][j])) * (o[2][j] + I * o[3][j]) +
              (conj(o[0][j]) - I * conj(o[1][j])) * (creal(o[2][j]) + I * creal(o[3][j]));
      w[j] += (conj(o[4][j]) - I * conj(o[5][j])) * (creal(o[2][j]) + I * creal(o[3][j])) -
              (creal(o[0][j]) - I * creal(o[1][j])) * (o[6][j] + I * o[7][j]);
    }
    for (int j = nump / 2; j < nump; ++j) {
      p[j] += (creal(o[0][j]) - I * creal(o[1][j])) * (conj(o[2][j]) + I * conj(o[3][j])) +
              (o[0][j] - I * o[1][j]) * (creal(o[2][j]) + I * creal(o[3][j]));
      w[j] += (o[4][j] - I * o[5][j]) * (creal(o[2][j]) + I * creal(o[3][j])) -
              (creal(o[0][j]) - I * creal(o[1][j])) * (conj(o[6][j]) + I * conj(o[7][j]));
    }
  }
  for (int k = 0; k < nump; ++k) {
    r_data[k] += p[k];
    r_data[k + nump] += w[k];
  }
  plan_free(plan);
  free(p);
  free(w);
  for (int q = 0; q < 8; ++q) {
    free(in[q]);
    free(o[q]);
  }
  free(pre);
}

/* Kubo_solver_FFT::compute + polynomial_recursion (src/Kubo_solver_FFT.cpp:12-136).
 * r_data and final_data start at zero (the reference declares them as
 * uninitialised stack arrays, src/Kubo_solver_FFT.cpp:30-32). */
int lqo_kubo_fft(const lqo_op *H, const lqo_op *OPL, const lqo_op *OPR, int n,
                 const double *phi, int R, int M, int num_sections, int nump,
                 int formula, double *final_data_, double *last_r_data_) {
  if (!H || H->n != n || (OPL && OPL->n != n) || (OPR && OPR->n != n) || !phi || R <= 0 ||
      M <= 0 || num_sections <= 0 || num_sections > n || nump <= 0 || M > nump ||
      (formula != 0 && formula != 1) || !final_data_)
    return -1;
  cplx *final_data = (cplx *)final_data_;
  cplx *r_data = (cplx *)calloc((size_t)(2 * nump), sizeof(cplx));
  for (int i = 0; i < 2 * nump; ++i) final_data[i] = 0;
  /* Kubo_solver_FFT::section_size_ (src/Kubo_solver_FFT.cpp:18-19); the
   * dummy Moments' SystemSize() is 0, so every section's FFT loop uses it. */
  const int size = n / num_sections + n % num_sections;
  vec_store L, Rv;
  store_init(&L, H, n, M, num_sections);
  store_init(&Rv, H, n, M, num_sections);
  for (int r = 1; r <= R; ++r) {
    const cplx *Phi = (const cplx *)phi + (size_t)(r - 1) * n;
    for (int s = 0; s < num_sections; ++s) {
      set_init_op(&L.st, OPL, Phi);
      iterate_all_sliced(&L, s);
      set_init(&Rv.st, Phi);
      iterate_all_sliced(&Rv, s);
      multiply_sliced(&Rv, OPR, s);
      if (formula == 0) greenwood_ffts(&L, &Rv, M, nump, size, r_data);
      if (formula == 1) bastin_ffts(&L, &Rv, M, nump, size, r_data);
    }
    /* update_data (include/Kubo_solver_FFT.hpp:93-96) */
    for (int i = 0; i < 2 * nump; ++i)
      final_data[i] = (final_data[i] * (cplx)(r - 1) + r_data[i]) / (cplx)r;
    if (last_r_data_) memcpy(last_r_data_, r_data, sizeof(cplx) * (size_t)(2 * nump));
    for (int i = 0; i < 2 * nump; ++i) r_data[i] = 0;
  }
  store_free(&L);
  store_free(&Rv);
  free(r_data);
  return 0;
}

/* Kubo_solver_FFT_postProcess::rearrange_crescent_order
 * (src/Kubo_solver_FFT_postProcess.cpp:98-116). */
static void rearrange_crescent_order(int nump, double *v) {
  double *orig = (double *)malloc(sizeof(double) * (size_t)nump);
  memcpy(orig, v, sizeof(double) * (size_t)nump);
  for (int k = 0; k < nump / 2; ++k) {
    v[2 * k] = orig[k];
    v[2 * k + 1] = orig[nump - k - 1];
  }
  for (int k = 0; k < nump / 2; ++k) {
    const double t = v[k];
    v[k] = v[nump - k - 1];
    v[nump - k - 1] = t;
  }
  free(orig);
}

/* Kubo_solver_FFT_postProcess ctor, Bastin_postProcess, Greenwood_postProcess,
 * integration_linqt (src/Kubo_solver_FFT_postProcess.cpp:8-16, 36-47, 120-273, 279-397). */
int lqo_kubo_post(int formula, const double *final_data_, int nump, int n, double W,
                  double Ec, double *E_out, double *sigma_out, double *integrand_out) {
  if ((formula != 0 && formula != 1) || !final_data_ || nump <= 1 || n <= 0 || !E_out || !sigma_out)
    return -1;
  const cplx *fd = (const cplx *)final_data_;
  double *Ep = (double *)malloc(sizeof(double) * (size_t)nump);
  double *Er = (double *)malloc(sizeof(double) * (size_t)nump);
  double *res = (double *)calloc((size_t)nump, sizeof(double));
  double *integ = (double *)calloc((size_t)nump, sizeof(double));
  for (int k = 0; k < nump; ++k) Ep[k] = cos(M_PI * (2 * k + 0.5) / nump);
  memcpy(Er, Ep, sizeof(double) * (size_t)nump);
  rearrange_crescent_order(nump, Er);
  const double a = W / 2.0, b = Ec;
  const double omega = n / (a * a);
  if (formula == 1) {
    for (int k = 0; k < nump; ++k) {
      integ[k] = Ep[k] * creal(fd[k]) - (sqrt(1.0 - Ep[k] * Ep[k]) * cimag(fd[k + nump]));
      integ[k] *= 1.0 / pow((1.0 - Ep[k] * Ep[k]), 2.0);
      integ[k] *= omega / (M_PI);
    }
    rearrange_crescent_order(nump, integ);
    double acc = 0;
    for (int i = 0; i < nump - 1; ++i) {
      const double denerg = Er[i + 1] - Er[i];
      acc += integ[i] * denerg;
      res[i] = acc;
    }
  } else {
    for (int k = 0; k < nump; ++k)
      res[k] = 2.0 * omega * creal(fd[k]) / (1.0 - Ep[k] * Ep[k]) / (2 * M_PI);
    rearrange_crescent_order(nump, res);
  }
  for (int e = 0; e < nump; ++e) {
    E_out[e] = a * Er[e] - b;
    sigma_out[e] = res[e];
    if (integrand_out) integrand_out[e] = integ[e];
  }
  free(Ep);
  free(Er);
  free(res);
  free(integ);
  return 0;
}

/* Moments1D::ApplyJacksonKernel (src/chebyshev_moments1D.cpp:68-87). */
int lqo_jackson_1d(double *mu_, int M, double W, double broad) {
  if (!mu_ || M <= 0 || !(broad > 0) || !(W > 0)) return -1;
  cplx *mu = (cplx *)mu_;
  const double eta = 2.0 * broad / 1000 / W;
  int maxMom = (int)ceil(M_PI / eta);
  if (maxMom > M) maxMom = M;
  const double phi_J = M_PI / (double)(maxMom + 1.0);
  for (size_t m = 0; m < (size_t)maxMom; ++m) {
    const double g = ((double)(maxMom - m + 1) * cos(phi_J * m) + sin(phi_J * m) * cos(phi_J) / sin(phi_J)) *
                     phi_J / M_PI;
    mu[m] *= g;
  }
  return maxMom;
}

/* Moments2D::ApplyJacksonKernel (src/chebyshev_moments2D.cpp:87-117). */
int lqo_jackson_2d(double *mu_, int M0, int M1, double W, double b0, double b1, int *M0o, int *M1o) {
  if (!mu_ || M0 <= 0 || M1 <= 0 || !(b0 > 0) || !(b1 > 0) || !(W > 0)) return -1;
  cplx *mu = (cplx *)mu_;
  const double eta0 = 2.0 * b0 / 1000 / W, eta1 = 2.0 * b1 / 1000 / W;
  int mm0 = (int)ceil(M_PI / eta0), mm1 = (int)ceil(M_PI / eta1);
  if (mm0 > M0) mm0 = M0;
  if (mm1 > M1) mm1 = M1;
  /* Moments2D::MomentNumber: compact to the leading mm0 x mm1 block */
  for (int m0 = 0; m0 < mm0; ++m0)
    for (int m1 = 0; m1 < mm1; ++m1) mu[(size_t)m0 * mm1 + m1] = mu[(size_t)m0 * M1 + m1];
  const double phi0 = M_PI / (double)(mm0 + 1.0), phi1 = M_PI / (double)(mm1 + 1.0);
  for (int m0 = 0; m0 < mm0; ++m0) {
    const double g0 = ((mm0 - m0 + 1) * cos(phi0 * m0) + sin(phi0 * m0) * cos(phi0) / sin(phi0)) * phi0 / M_PI;
    for (int m1 = 0; m1 < mm1; ++m1) {
      const double g1 = ((mm1 - m1 + 1) * cos(phi1 * m1) + sin(phi1 * m1) * cos(phi1) / sin(phi1)) * phi1 / M_PI;
      mu[(size_t)m0 * mm1 + m1] *= g0 * g1;
    }
  }
  if (M0o) *M0o = mm0;
  if (M1o) *M1o = mm1;
  return 0;
}

/* include/chebyshev_coefficients.hpp:8-32.  pow(complex, double) follows
 * libstdc++: polar(exp(m*log|z|), m*arg z). */
static double delta_chebF(double x, double m) {
  const double f0 = sqrt(1.0 - x * x);
  const double fm = cos(m * acos(x));
  return fm / f0 / M_PI;
}

static cplx cpow_real(cplx z, double m) {
  const cplx t = clog(z);
  const double rho = exp(m * creal(t)), th = m * cimag(t);
  return rho * cos(th) + I * rho * sin(th);
}

static cplx greenR_chebF(double x, double m) {
  const double f0 = sqrt(1.0 - x * x);
  const cplx fm = cpow_real(x - I * f0, m);
  return -I * fm / f0;
}

static cplx DgreenR_chebF(double x, double m) {
  const double f0 = sqrt(1.0 - x * x);
  const cplx fm = cpow_real(x - I * f0, m) * (x + I * m * f0);
  return -I * fm / f0 / f0 / f0;
}

/* spectralFunctionFromChebmom_FFTgrid.cpp:35-73. */
int lqo_dos_fftgrid(const double *mu_, int M, int n, double W, double Ec, int num_div,
                    double *E_out, double *rho_out) {
  if (!mu_ || M <= 0 || num_div <= 0 || !E_out || !rho_out) return -1;
  const cplx *mu = (const cplx *)mu_;
  const double a = W / 2.0;
  for (int i = 0; i < num_div; ++i) {
    const double energ = cos(M_PI * (i + 0.5) / num_div);
    double out = 0;
    for (int m0 = 0; m0 < M; ++m0) out += delta_chebF(energ, m0) * creal(mu[m0]);
    out *= n / a;
    E_out[i] = energ * a + Ec;
    rho_out[i] = fabs(out);
  }
  return 0;
}

/* kuboGreenwoodFromChebmom_FFTgrid.cpp:35-76 (num_div-1 rows, grid order). */
int lqo_greenwood_fftgrid(const double *mu_, int M0, int M1, int n, double W, double Ec,
                          int num_div, double *E_out, double *out_) {
  if (!mu_ || M0 <= 0 || M1 <= 0 || num_div <= 1 || !E_out || !out_) return -1;
  const cplx *mu = (const cplx *)mu_;
  const double a = W / 2.0;
  for (int i = 0; i < num_div - 1; ++i) {
    const double energ = cos(M_PI * (2 * i + 0.5) / num_div);
    double out = 0.0;
    for (int m0 = 0; m0 < M0; ++m0)
      for (int m1 = 0; m1 < M1; ++m1)
        out += delta_chebF(energ, m0) * cimag(greenR_chebF(energ, m1) * mu[(size_t)m0 * M1 + m1]);
    E_out[i] = energ * a + Ec;
    out_[i] = -1.0 * out * (n / a / a);
  }
  return 0;
}

/* kuboBastinFromChebmom_FFTgrid.cpp:57-106 (ascending grid, cumulative sum). */
int lqo_bastin_fftgrid(const double *mu_, int M0, int M1, int n, double W, double Ec,
                       int num_div, double *E_out, double *out_) {
  if (!mu_ || M0 <= 0 || M1 <= 0 || num_div <= 1 || !E_out || !out_) return -1;
  const cplx *mu = (const cplx *)mu_;
  const double a = W / 2.0;
  double *E = (double *)malloc(sizeof(double) * (size_t)num_div);
  double *kern = (double *)calloc((size_t)num_div, sizeof(double));
  for (int i = 0; i < num_div; ++i) E[i] = cos(M_PI * (2 * i + 0.5) / num_div);
  rearrange_crescent_order(num_div, E);
  for (int i = 0; i < num_div; ++i) {
    for (int m0 = 0; m0 < M0; ++m0)
      for (int m1 = 0; m1 < M1; ++m1)
        kern[i] += delta_chebF(E[i], m0) * cimag(DgreenR_chebF(E[i], m1) * mu[(size_t)m0 * M1 + m1]);
    kern[i] *= -2.0 * n / a / a;
  }
  double acc = 0;
  for (int i = 0; i < num_div - 1; ++i) {
    acc += kern[i] * (E[i + 1] - E[i]);
    E_out[i] = E[i] * a + Ec;
    out_[i] = acc;
  }
  free(E);
  free(kern);
  return 0;
}

/* Kernels of the linear-grid tools, each in its source's loop order:
 *   1 kuboGreenwoodFromChebmom.cpp:64-73, 3 kuboBastinIFromChebmom.cpp:66-74:
 *     -N/a^2 sum delta_m0 Im(gR_m1 mu)
 *   2 kuboBastinFromChebmom.cpp:64-72: -2N/a^2 sum delta_m0 Im(DgR_m1 mu)
 *   4 kuboBastinIIFromChebmom.cpp:66-79:
 *     N/a^2/(2 pi) sum -Re((gR_m0 DgR_m1 - DgR_m0 gR_m1) mu)
 *   5 kuboBastinSeaFromChebmom.cpp:66-79:
 *     N/a^2/(2 pi) sum -Re((Gr - conj Gr)(DGr + DGr) mu), Gr = gR_m0, DGr = Re DgR_m1
 *   6 kuboBastinSurfFromChebmom.cpp:64-72: pi N/a^2 sum delta_m0 delta_m1 Re mu
 *   0 spectralFunctionFromChebmom.cpp:60-67: N/a sum delta_m Re mu_m */
int lqo_linear_kernel(int kind, const double *mu_, int M0, int M1, int n, double W, int nx, const double *x,
                      double *kern) {
  if (!mu_ || M0 <= 0 || (kind != 0 && M1 <= 0) || nx <= 0 || !x || !kern || kind < 0 || kind > 6) return -1;
  const cplx *mu = (const cplx *)mu_;
  const double a = W / 2.0;
  for (int i = 0; i < nx; ++i) {
    const double e = x[i];
    double k = 0.0;
    if (kind == 0) {
      for (int m0 = 0; m0 < M0; ++m0) k += delta_chebF(e, m0) * creal(mu[m0]);
      k *= n / a;
    } else if (kind == 1 || kind == 3) {
      double out = 0.0;
      for (int m0 = 0; m0 < M0; ++m0)
        for (int m1 = 0; m1 < M1; ++m1)
          out += delta_chebF(e, m0) * cimag(greenR_chebF(e, m1) * mu[(size_t)m0 * M1 + m1]);
      k = -1.0 * out * (n / a / a);
    } else if (kind == 2) {
      for (int m0 = 0; m0 < M0; ++m0)
        for (int m1 = 0; m1 < M1; ++m1)
          k += delta_chebF(e, m0) * cimag(DgreenR_chebF(e, m1) * mu[(size_t)m0 * M1 + m1]);
      k *= -2.0 * n / a / a;
    } else if (kind == 4) {
      double out = 0.0;
      for (int m0 = 0; m0 < M0; ++m0)
        for (int m1 = 0; m1 < M1; ++m1) {
          const cplx GrL = greenR_chebF(e, m0), GrR = greenR_chebF(e, m1);
          const cplx DGrL = DgreenR_chebF(e, m0), DGrR = DgreenR_chebF(e, m1);
          out += -creal((GrL * DGrR - DGrL * GrR) * mu[(size_t)m0 * M1 + m1]);
        }
      k = out * n / a / a / 2 / M_PI;
    } else if (kind == 5) {
      for (int m0 = 0; m0 < M0; ++m0)
        for (int m1 = 0; m1 < M1; ++m1) {
          const cplx Gr = greenR_chebF(e, m0), Ga = conj(greenR_chebF(e, m0));
          const double DGr = creal(DgreenR_chebF(e, m1));
          const cplx DGa = DGr - 0.0 * I;
          k += -creal((Gr - Ga) * (DGr + DGa) * mu[(size_t)m0 * M1 + m1]);
        }
      k *= n / a / a / 2 / M_PI;
    } else {
      for (int m0 = 0; m0 < M0; ++m0)
        for (int m1 = 0; m1 < M1; ++m1)
          k += delta_chebF(e, m0) * delta_chebF(e, m1) * creal(mu[(size_t)m0 * M1 + m1]);
      k *= M_PI * n / a / a;
    }
    kern[i] = k;
  }
  return 0;
}

/* The grids and the printed columns of the same tools (lines 35-41 and 76-86
 * of each source; the Bastin tool uses xbound = 0.99). */
int lqo_linear_grid(int kind, const double *mu, int M0, int M1, int n, double W, double Ec, int num_div,
                    double *E_out, double *out) {
  if (num_div <= 1 || !E_out || !out) return -1;
  const double a = W / 2.0, xb = kind == 2 ? 0.99 : 1.0;
  double *x = (double *)malloc(sizeof(double) * (size_t)num_div);
  double *k = (double *)malloc(sizeof(double) * (size_t)num_div);
  for (int i = 0; i < num_div; ++i) x[i] = -xb + i * (2 * xb) / (num_div - 1);
  if (lqo_linear_kernel(kind, mu, M0, M1, n, W, num_div, x, k) != 0) {
    free(x);
    free(k);
    return -1;
  }
  const int rows = kind == 0 ? num_div : num_div - 1;
  double acc = 0;
  for (int i = 0; i < rows; ++i) {
    E_out[i] = x[i] * a + Ec;
    switch (kind) {
      case 0:
      case 1: out[i] = k[i]; break;
      case 2: acc += k[i] * (x[i + 1] - x[i]); out[i] = acc; break;
      case 3: out[i] = 0.5 * (k[i] + k[i + 1]); break;
      case 4: acc += (k[i] + k[i + 1]) * (x[i + 1] - x[i]); out[i] = acc; break;
      case 5: acc += 0.5 * (k[i] + k[i + 1]) * (x[i + 1] - x[i]); out[i] = acc; break;
      default: out[i] = (k[i + 1] + k[i]) / 2.0; break;
    }
  }
  free(x);
  free(k);
  return rows;
}

/* ---- OpenMP CPU baseline --------------------------------------------------- */
int lqo_omp_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Builds the completed (U + U^H - diag U) CSR once, for row-parallel mv. */
static void op_build_full(lqo_op *op) {
  if (op->frow) return;
  const int n = op->n;
  int *cnt = (int *)calloc((size_t)n + 1, sizeof(int));
  for (int i = 0; i < n; ++i)
    for (int p = op->rowptr[i]; p < op->rowptr[i + 1]; ++p) {
      const int j = op->col[p];
      if (j < i) continue;
      cnt[i]++;
      if (j != i) cnt[j]++;
    }
  op->frow = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  op->frow[0] = 0;
  for (int i = 0; i < n; ++i) op->frow[i + 1] = op->frow[i] + cnt[i];
  op->fnnz = op->frow[n];
  op->fcol = (int *)malloc(sizeof(int) * (size_t)(op->fnnz > 0 ? op->fnnz : 1));
  op->fval = (cplx *)malloc(sizeof(cplx) * (size_t)(op->fnnz > 0 ? op->fnnz : 1));
  for (int i = 0; i < n; ++i) cnt[i] = op->frow[i];
  for (int i = 0; i < n; ++i)
    for (int p = op->rowptr[i]; p < op->rowptr[i + 1]; ++p) {
      const int j = op->col[p];
      if (j < i) continue;
      op->fcol[cnt[i]] = j;
      op->fval[cnt[i]++] = op->val[p];
      if (j != i) {
        op->fcol[cnt[j]] = i;
        op->fval[cnt[j]++] = conj(op->val[p]);
      }
    }
  free(cnt);
}

int lqo_omp_cheb_step(const lqo_op *Hc, int k, const double *x_, double *y_) {
  if (!Hc || k <= 0 || !x_ || !y_) return -1;
  lqo_op *H = (lqo_op *)Hc;
  op_build_full(H);
  const int n = H->n;
  const cplx *x = (const cplx *)x_;
  cplx *y = (cplx *)y_;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    for (int v = 0; v < k; ++v) {
      const cplx *xv = x + (size_t)v * n;
      cplx acc = 0;
      for (int p = H->frow[i]; p < H->frow[i + 1]; ++p) acc += H->fval[p] * xv[H->fcol[p]];
      y[(size_t)v * n + i] = 2.0 * acc - y[(size_t)v * n + i];
    }
  }
  return 0;
}

int lqo_omp_mv(const lqo_op *opc, int k, const double *x_, double *y_) {
  if (!opc || k <= 0 || !x_ || !y_) return -1;
  lqo_op *op = (lqo_op *)opc;
  op_build_full(op);
  const int n = op->n;
  const cplx *x = (const cplx *)x_;
  cplx *y = (cplx *)y_;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i)
    for (int v = 0; v < k; ++v) {
      const cplx *xv = x + (size_t)v * n;
      cplx acc = 0;
      for (int p = op->frow[i]; p < op->frow[i + 1]; ++p) acc += op->fval[p] * xv[op->fcol[p]];
      y[(size_t)v * n + i] = acc;
    }
  return 0;
}

int lqo_omp_bastin_ffts(const double *L_, const double *R_, int M, int nump, int size, double *r_data_) {
  if (!L_ || !R_ || M <= 0 || nump < M || size <= 0 || !r_data_) return -1;
  const cplx *Lv = (const cplx *)L_, *Rv = (const cplx *)R_;
  cplx *r_data = (cplx *)r_data_;
  cplx *pre = (cplx *)malloc(sizeof(cplx) * (size_t)nump);
  for (int m = 0; m < M; ++m)
    pre[m] = ((2 - (m == 0)) * lqo_kubo_jackson(m, M)) *
             (cos(M_PI * m / (2.0 * nump)) + I * sin(M_PI * m / (2.0 * nump)));
  dft_plan *plan = plan_create(nump);
#pragma omp parallel
  {
    cplx *in[8], *o[8];
    for (int q = 0; q < 8; ++q) {
      in[q] = (cplx *)calloc((size_t)nump, sizeof(cplx));
      o[q] = (cplx *)malloc(sizeof(cplx) * (size_t)nump);
    }
    cplx *p = (cplx *)calloc((size_t)nump, sizeof(cplx));
    cplx *w = (cplx *)calloc((size_t)nump, sizeof(cplx));
#pragma omp for schedule(static)
    for (int l = 0; l < size; ++l) {
      for (int m = 0; m < M; ++m) {
        const cplx im_m = (double)m;
        const cplx bra = Lv[(size_t)m * size + l], ket = Rv[(size_t)m * size + l];
        in[0][m] = pre[m] * creal(bra);
        in[1][m] = pre[m] * cimag(bra);
        in[4][m] = im_m * in[0][m];
        in[5][m] = im_m * in[1][m];
        in[2][m] = pre[m] * creal(ket);
        in[3][m] = pre[m] * cimag(ket);
        in[6][m] = im_m * in[2][m];
        in[7][m] = im_m * in[3][m];
      }
      for (int q = 0; q < 8; ++q) plan_exec(plan, in[q], o[q]);
      for (int j = 0; j < nump / 2; ++j) {
        p[j] += (creal(o[0][j]) - I * creal(o[1][j])) * (o[2][j] + I * o[3][j]) +
                (conj(o[0][j]) - I * conj(o[1][j])) * (creal(o[2][j]) + I * creal(o[3][j]));
        w[j] += (conj(o[4][j]) - I * conj(o[5][j])) * (creal(o[2][j]) + I * creal(o[3][j])) -
                (creal(o[0][j]) - I * creal(o[1][j])) * (o[6][j] + I * o[7][j]);
      }
      for (int j = nump / 2; j < nump; ++j) {
        p[j] += (creal(o[0][j]) - I * creal(o[1][j])) * (conj(o[2][j]) + I * conj(o[3][j])) +
                (o[0][j] - I * o[1][j]) * (creal(o[2][j]) + I * creal(o[3][j]));
        w[j] += (o[4][j] - I * o[5][j]) * (creal(o[2][j]) + I * creal(o[3][j])) -
                (creal(o[0][j]) - I * creal(o[1][j])) * (conj(o[6][j]) + I * conj(o[7][j]));
      }
    }
#pragma omp critical
    for (int k = 0; k < nump; ++k) {
      r_data[k] += p[k];
      r_data[k + nump] += w[k];
    }
    free(p);
    free(w);
    for (int q = 0; q < 8; ++q) {
      free(in[q]);
      free(o[q]);
    }
  }
  plan_free(plan);
  free(pre);
  return 0;
}

int lqo_omp_momtable(int n, int M1, const double *L_, const double *R_, int m0_begin,
                     int m0_end, double *sub_) {
  if (n <= 0 || M1 <= 0 || !L_ || !R_ || !sub_ || m0_end < m0_begin) return -1;
  const cplx *L = (const cplx *)L_, *Rv = (const cplx *)R_;
  cplx *sub = (cplx *)sub_;
  const long long pairs = (long long)(m0_end - m0_begin) * M1;
#pragma omp parallel for schedule(dynamic, 4)
  for (long long q = 0; q < pairs; ++q) {
    const int m0 = m0_begin + (int)(q / M1), m1 = (int)(q % M1);
    sub[q] = vdot(n, L + (size_t)m0 * n, Rv + (size_t)m1 * n);
  }
  return 0;
}

int lqo_omp_spectral_moments(const lqo_op *Hc, int n, const double *phi, int R, int M,
                             double *mu_) {
  if (!Hc || Hc->n != n || !phi || R <= 0 || M <= 0 || !mu_) return -1;
  lqo_op *H = (lqo_op *)Hc;
  op_build_full(H);
  cplx *mu = (cplx *)mu_;
  cplx *c0 = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  cplx *c1 = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  for (int r = 0; r < R; ++r) {
    const cplx *Phi = (const cplx *)phi + (size_t)r * n;
    memcpy(c0, Phi, sizeof(cplx) * (size_t)n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      cplx acc = 0;
      for (int p = H->frow[i]; p < H->frow[i + 1]; ++p) acc += H->fval[p] * c0[H->fcol[p]];
      c1[i] = acc;
    }
    for (int m = 0; m < M; ++m) {
      double scal = 2.0 / R;
      if (m == 0) scal *= 0.5;
      double sre = 0, sim = 0;
#pragma omp parallel for reduction(+ : sre, sim) schedule(static)
      for (int i = 0; i < n; ++i) {
        const cplx t = conj(Phi[i]) * c0[i];
        sre += creal(t);
        sim += cimag(t);
      }
      mu[m] += scal * (sre + I * sim);
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) {
        cplx acc = 0;
        for (int p = H->frow[i]; p < H->frow[i + 1]; ++p) acc += H->fval[p] * c1[H->fcol[p]];
        c0[i] = 2.0 * acc - c0[i];
      }
      cplx *t = c0;
      c0 = c1;
      c1 = t;
    }
  }
  free(c0);
  free(c1);
  return 0;
}

/* ======================================================================== */
/* Supercell operators: hopping_list::wrap_in_supercell +                   */
/* save_hopping_list_as_csr (utilities/wannier2sparse/src/hopping_list.cpp).*/
/* ======================================================================== */

typedef struct {
  char tag[80];
  int idx;
} sc_tag;

/* std::multimap<string,...> order: by tag string, equal tags in insertion order */
static int sc_tag_cmp(const void *a, const void *b) {
  const sc_tag *x = (const sc_tag *)a, *y = (const sc_tag *)b;
  int c = strcmp(x->tag, y->tag);
  return c ? c : (x->idx > y->idx) - (x->idx < y->idx);
}

typedef struct {
  int row, col;
  long long seq;
  double re, im;
} sc_trip;

static int sc_trip_cmp(const void *a, const void *b) {
  const sc_trip *x = (const sc_trip *)a, *y = (const sc_trip *)b;
  if (x->row != y->row) return x->row < y->row ? -1 : 1;
  if (x->col != y->col) return x->col < y->col ? -1 : 1;
  return (x->seq > y->seq) - (x->seq < y->seq);
}

/* index_aliasing (utilities/wannier2sparse/include/hopping_list_util.hpp:11-15) */
static int sc_alias(const int *c, const int *b) {
  return (((c[2] + b[2]) % b[2]) * b[1] + (c[1] + b[1]) % b[1]) * b[0] + (c[0] + b[0]) % b[0];
}

int lqo_supercell_csr(int nwann, int nhop, const int *cell, const int *edge, const double *val,
                      const int *dims, int *n, long long *nnz, int *rowptr, int *col, double *out_val) {
  if (nwann <= 0 || nhop < 0 || !dims || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0 || !n || !nnz) return -1;
  const long long ncell = (long long)dims[0] * dims[1] * dims[2];
  if (ncell * nwann > 0x7ffffffeLL) return -1;
  *n = (int)(ncell * nwann);
  sc_tag *tags = (sc_tag *)malloc(sizeof(sc_tag) * (nhop > 0 ? nhop : 1));
  double avg = 0.0;
  for (int h = 0; h < nhop; ++h) {
    if (edge[2 * h] < 0 || edge[2 * h] >= nwann || edge[2 * h + 1] < 0 || edge[2 * h + 1] >= nwann) {
      free(tags);
      return -1;
    }
    snprintf(tags[h].tag, sizeof(tags[h].tag), "%d %d %d %d %d ", cell[3 * h], cell[3 * h + 1], cell[3 * h + 2],
             edge[2 * h], edge[2 * h + 1]);
    tags[h].idx = h;
  }
  qsort(tags, (size_t)nhop, sizeof(sc_tag), sc_tag_cmp);
  for (int q = 0; q < nhop; ++q) {
    const int h = tags[q].idx;
    avg += sqrt(val[2 * h] * val[2 * h] + val[2 * h + 1] * val[2 * h + 1]);
  }
  avg /= nhop; /* NaN for an empty list: then nothing passes the filter */
  sc_trip *t = (sc_trip *)malloc(sizeof(sc_trip) * (size_t)(ncell * nhop > 0 ? ncell * nhop : 1));
  long long nt = 0, seq = 0;
  for (int q = 0; q < nhop; ++q) {
    const int h = tags[q].idx;
    const double re = val[2 * h], im = val[2 * h + 1];
    const int keep = sqrt(re * re + im * im) / avg > 1e-10;
    int s[3];
    for (s[2] = 0; s[2] < dims[2]; ++s[2])
      for (s[1] = 0; s[1] < dims[1]; ++s[1])
        for (s[0] = 0; s[0] < dims[0]; ++s[0], ++seq) {
          int c[3];
          for (int d = 0; d < 3; ++d) c[d] = (cell[3 * h + d] + s[d] + dims[d]) % dims[d];
          /* the reference wraps with one + dims: keep the same arithmetic, then fold */
          for (int d = 0; d < 3; ++d) c[d] = ((c[d] % dims[d]) + dims[d]) % dims[d];
          const int row = edge[2 * h] + sc_alias(s, dims) * nwann;
          const int cl = edge[2 * h + 1] + sc_alias(c, dims) * nwann;
          if (!keep || cl < row) continue;
          t[nt].row = row;
          t[nt].col = cl;
          t[nt].seq = seq;
          t[nt].re = re;
          t[nt].im = im;
          ++nt;
        }
  }
  free(tags);
  qsort(t, (size_t)nt, sizeof(sc_trip), sc_trip_cmp);
  long long m = 0;
  for (long long q = 0; q < nt; ++q) {
    if (m > 0 && t[m - 1].row == t[q].row && t[m - 1].col == t[q].col) {
      t[m - 1].re += t[q].re;
      t[m - 1].im += t[q].im;
    } else {
      t[m++] = t[q];
    }
  }
  *nnz = m;
  if (rowptr) {
    for (int i = 0; i <= *n; ++i) rowptr[i] = 0;
    for (long long q = 0; q < m; ++q) {
      rowptr[t[q].row + 1]++;
      col[q] = t[q].col;
      out_val[2 * q] = t[q].re;
      out_val[2 * q + 1] = t[q].im;
    }
    for (int i = 0; i < *n; ++i) rowptr[i + 1] += rowptr[i];
  }
  free(t);
  return 0;
}

/* Sparse::OPERATOR_FromCSRFile (src/sparse_matrix.cpp:3-50): sequential
 * token-by-token extraction from a buffered stream. */
int lqo_csr_read_stream(const char *path, int *n, long long *nnz, int *rowptr, int *col, double *val) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  int dim = 0, nz = 0;
  if (fscanf(f, "%d %d", &dim, &nz) != 2 || dim <= 0 || nz < 0) {
    fclose(f);
    return -1;
  }
  *n = dim;
  *nnz = nz;
  if (!rowptr) {
    fclose(f);
    return 0;
  }
  int bad = 0;
  for (int i = 0; i < nz && !bad; ++i) bad = fscanf(f, "%lf %lf", &val[2 * i], &val[2 * i + 1]) != 2;
  for (int i = 0; i < nz && !bad; ++i) bad = fscanf(f, "%d", &col[i]) != 1;
  for (int i = 0; i <= dim && !bad; ++i) bad = fscanf(f, "%d", &rowptr[i]) != 1;
  fclose(f);
  return bad ? -1 : 0;
}

/* ======================================================================== */
/* Non-orthogonal basis: Eigen::BiCGSTAB (Eigen/src/IterativeLinearSolvers/ */
/* BiCGSTAB.h, 3.3/3.4) with DiagonalPreconditioner, as                     */
/* linalg::orthogonalize (src/mkl_linear_algebra.cpp:100-131) runs it.      */
/* ======================================================================== */

/* y = S x, S general CSR, row sums in stored order */
static void gen_mv(int n, const int *rp, const int *cl, const cplx *v, const cplx *x, cplx *y) {
  for (int i = 0; i < n; ++i) {
    cplx t = 0;
    for (int q = rp[i]; q < rp[i + 1]; ++q) t += v[q] * x[cl[q]];
    y[i] = t;
  }
}

static double sqnorm(int n, const cplx *x) {
  double s = 0;
  for (int i = 0; i < n; ++i) s += creal(x[i]) * creal(x[i]) + cimag(x[i]) * cimag(x[i]);
  return s;
}

int lqo_overlap_solve(int n, const int *rp, const int *cl, const double *val_, const double *b_, double *x_,
                      double tol, int maxit, double *error) {
  if (n <= 0 || !rp || !b_ || !x_) return -1;
  const cplx *v = (const cplx *)val_, *b = (const cplx *)b_;
  cplx *x = (cplx *)x_;
  /* DiagonalPreconditioner::factorize: 1/diag where a nonzero diagonal is stored, else 1 */
  cplx *dinv = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    dinv[i] = 1.0;
    for (int q = rp[i]; q < rp[i + 1]; ++q)
      if (cl[q] == i) {
        if (v[q] != 0) dinv[i] = 1.0 / v[q];
        break;
      }
  }
  cplx *r = (cplx *)calloc((size_t)n, sizeof(cplx)), *r0 = (cplx *)calloc((size_t)n, sizeof(cplx));
  cplx *p = (cplx *)calloc((size_t)n, sizeof(cplx)), *vv = (cplx *)calloc((size_t)n, sizeof(cplx));
  cplx *y = (cplx *)calloc((size_t)n, sizeof(cplx)), *z = (cplx *)calloc((size_t)n, sizeof(cplx));
  cplx *s = (cplx *)calloc((size_t)n, sizeof(cplx)), *t = (cplx *)calloc((size_t)n, sizeof(cplx));
  for (int i = 0; i < n; ++i) x[i] = 0; /* solve() starts from zero */
  const double rhs_sq = sqnorm(n, b);
  int it = 0;
  if (rhs_sq == 0) {
    if (error) *error = 0;
  } else {
    for (int i = 0; i < n; ++i) r[i] = r0[i] = b[i]; /* r = b - S*0 */
    double r0_sq = sqnorm(n, r0);
    cplx rho = 1, alpha = 1, w = 1;
    const double tol2 = tol * tol * rhs_sq;
    const double eps = 2.220446049250313e-16, eps2 = eps * eps;
    int restarts = 0;
    while (sqnorm(n, r) > tol2 && it < maxit) {
      const cplx rho_old = rho;
      rho = vdot(n, r0, r);
      if (cabs(rho) < eps2 * r0_sq) { /* r became too orthogonal to r0: restart */
        gen_mv(n, rp, cl, v, x, t);
        for (int i = 0; i < n; ++i) r[i] = r0[i] = b[i] - t[i];
        rho = r0_sq = sqnorm(n, r);
        if (restarts++ == 0) it = 0;
      }
      const cplx beta = (rho / rho_old) * (alpha / w);
      for (int i = 0; i < n; ++i) {
        p[i] = r[i] + beta * (p[i] - w * vv[i]);
        y[i] = dinv[i] * p[i];
      }
      gen_mv(n, rp, cl, v, y, vv);
      alpha = rho / vdot(n, r0, vv);
      for (int i = 0; i < n; ++i) {
        s[i] = r[i] - alpha * vv[i];
        z[i] = dinv[i] * s[i];
      }
      gen_mv(n, rp, cl, v, z, t);
      const double tmp = sqnorm(n, t);
      w = tmp > 0 ? vdot(n, t, s) / tmp : 0;
      for (int i = 0; i < n; ++i) {
        x[i] += alpha * y[i] + w * z[i];
        r[i] = s[i] - w * t[i];
      }
      ++it;
    }
    if (error) *error = sqrt(sqnorm(n, r) / rhs_sq);
  }
  free(dinv); free(r); free(r0); free(p); free(vv); free(y); free(z); free(s); free(t);
  return it;
}

int lqo_spectral_moments_nonorth(const lqo_op *H, const lqo_op *OP, int n, const int *srp, const int *scl,
                                 const double *sval, const double *phi, int R, int M, double tol, int maxit,
                                 double *mu_) {
  if (!H || H->n != n || (OP && OP->n != n) || !phi || R <= 0 || M <= 0 || !mu_) return -1;
  cplx *mu = (cplx *)mu_;
  cplx *c0 = (cplx *)malloc(sizeof(cplx) * (size_t)n), *c1 = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)n), *tmp2 = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  for (int r = 0; r < R; ++r) {
    const cplx *Phi = (const cplx *)phi + (size_t)r * n;
    /* SetInitVectors_nonOrthogonal: T0 = [OP] S^-1 phi, T1 = S^-1 H T0 */
    if (!OP) {
      lqo_overlap_solve(n, srp, scl, sval, (const double *)Phi, (double *)c0, tol, maxit, NULL);
    } else {
      lqo_overlap_solve(n, srp, scl, sval, (const double *)Phi, (double *)tmp, tol, maxit, NULL);
      op_mv_c(OP, 1.0, tmp, 0.0, c0);
    }
    op_mv_c(H, 1.0, c0, 0.0, tmp);
    lqo_overlap_solve(n, srp, scl, sval, (const double *)tmp, (double *)c1, tol, maxit, NULL);
    for (int m = 0; m < M; ++m) {
      double scal = 2.0 / R;
      if (m == 0) scal *= 0.5;
      mu[m] += scal * vdot(n, Phi, c0);
      if (m + 1 == M) break; /* the reference's last Iterate leaves mu unchanged */
      /* Iterate_nonOrthogonal: T0 <- 2 S^-1 H T1 - T0, swap */
      op_mv_c(H, 1.0, c1, 0.0, tmp);
      lqo_overlap_solve(n, srp, scl, sval, (const double *)tmp, (double *)tmp2, tol, maxit, NULL);
      for (int i = 0; i < n; ++i) c0[i] = 2.0 * tmp2[i] - c0[i];
      cplx *sw = c0;
      c0 = c1;
      c1 = sw;
    }
  }
  free(c0); free(c1); free(tmp); free(tmp2);
  return 0;
}

