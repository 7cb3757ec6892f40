// Non-orthogonal basis: overlap solves and the non-orthogonal spectral moments.
//
// linalg::orthogonalize (src/mkl_linear_algebra.cpp:100-131) solves S x = b
// with Eigen's BiCGSTAB (diagonal preconditioner, zero initial guess, stop
// when |r|^2 <= tol^2 |b|^2 or after maxit iterations; tol 1e-4 and 2000 in
// the reference), where S is the overlap CSR mapped as a general matrix.
// Moments1D_nonOrth (src/chebyshev_moments.cpp:63-153) and
// SpectralMoments_nonOrth (src/chebyshev_solver.cpp:203-229) run the
// Chebyshev recursion with such a solve after every Hamiltonian product:
//   T0 = [OP] S^-1 phi,  T1 = S^-1 H T0,  T_{m+1} = 2 S^-1 H T_m - T_{m-1},
//   mu_m += (2/R)(1/2 if m = 0) <phi|T_m>.
//
// Here every BiCGSTAB vector lives in HBM.  An iteration is two SpMVs with S
// (the sliced-ELL kernel), three fused vector kernels (p/y, s/z, x/r) and
// three reductions; the x/r update also produces |r|^2 and <r0|r> for the
// next iteration, so the host reads three small results per iteration.
// Reductions are two-level with a fixed order (per-block partials, then one
// block), so a run is deterministic.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "lq_internal.h"

namespace lq {
namespace {

constexpr int kNbThreads = 256;
constexpr int kNbBlocks = 512;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 cdot1(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// Block sum of two complex values; thread 0 writes partial[blockIdx.x * 2 + j].
__device__ void block_store2(double2 a, double2 b, double2 *__restrict__ partial) {
  __shared__ double2 sh[2][kNbThreads];
  const int t = threadIdx.x;
  sh[0][t] = a;
  sh[1][t] = b;
  __syncthreads();
  for (int w = kNbThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
      sh[0][t] = cadd(sh[0][t], sh[0][t + w]);
      sh[1][t] = cadd(sh[1][t], sh[1][t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
    partial[2 * blockIdx.x] = sh[0][0];
    partial[2 * blockIdx.x + 1] = sh[1][0];
  }
}

// partial <- per-block sums of conj(x1) y1 and conj(x2) y2 (x2 == nullptr: 0).
__global__ __launch_bounds__(kNbThreads) void k_nb_dot2(int n, const double2 *__restrict__ x1,
                                                         const double2 *__restrict__ y1,
                                                         const double2 *__restrict__ x2,
                                                         const double2 *__restrict__ y2, double2 *__restrict__ partial) {
  double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
  for (int i = blockIdx.x * kNbThreads + threadIdx.x; i < n; i += gridDim.x * kNbThreads) {
    a = cadd(a, cdot1(x1[i], y1[i]));
    if (x2) b = cadd(b, cdot1(x2[i], y2[i]));
  }
  block_store2(a, b, partial);
}

// out[j] = sum over blocks of partial[b*2 + j]; with acc != nullptr also
// acc[0] += scal * out[0] (a moment).
__global__ __launch_bounds__(kNbThreads) void k_nb_final(const double2 *__restrict__ partial, int nblk,
                                                          double2 *__restrict__ out, double scal,
                                                          double2 *__restrict__ acc) {
  __shared__ double2 sh[2][kNbThreads];
  const int t = threadIdx.x;
  double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
  for (int q = t; q < nblk; q += kNbThreads) {
    a = cadd(a, partial[2 * q]);
    b = cadd(b, partial[2 * q + 1]);
  }
  sh[0][t] = a;
  sh[1][t] = b;
  __syncthreads();
  for (int w = kNbThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
      sh[0][t] = cadd(sh[0][t], sh[0][t + w]);
      sh[1][t] = cadd(sh[1][t], sh[1][t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = sh[0][0];
    out[1] = sh[1][0];
    if (acc) acc[0] = make_double2(acc[0].x + scal * sh[0][0].x, acc[0].y + scal * sh[0][0].y);
  }
}

// r = r0 = b, x = p = v = 0; partials of |b|^2.
__global__ __launch_bounds__(kNbThreads) void k_nb_init(int n, const double2 *__restrict__ b, double2 *__restrict__ r,
                                                         double2 *__restrict__ r0, double2 *__restrict__ x,
                                                         double2 *__restrict__ p, double2 *__restrict__ v,
                                                         double2 *__restrict__ partial) {
  double2 a = make_double2(0.0, 0.0);
  const double2 z = make_double2(0.0, 0.0);
  for (int i = blockIdx.x * kNbThreads + threadIdx.x; i < n; i += gridDim.x * kNbThreads) {
    const double2 bi = b[i];
    r[i] = bi;
    r0[i] = bi;
    x[i] = z;
    p[i] = z;
    v[i] = z;
    a = cadd(a, cdot1(bi, bi));
  }
  block_store2(a, z, partial);
}

// p = r + beta (p - w v); y = dinv p
__global__ __launch_bounds__(kNbThreads) void k_nb_p(int n, double2 beta, double2 w, const double2 *__restrict__ r,
                                                      double2 *__restrict__ p, const double2 *__restrict__ v,
                                                      const double2 *__restrict__ dinv, double2 *__restrict__ y) {
  for (int i = blockIdx.x * kNbThreads + threadIdx.x; i < n; i += gridDim.x * kNbThreads) {
    const double2 pi = cadd(r[i], cmul(beta, csub(p[i], cmul(w, v[i]))));
    p[i] = pi;
    y[i] = cmul(dinv[i], pi);
  }
}

// s = r - alpha v; z = dinv s
__global__ __launch_bounds__(kNbThreads) void k_nb_s(int n, double2 alpha, const double2 *__restrict__ r,
                                                      const double2 *__restrict__ v, const double2 *__restrict__ dinv,
                                                      double2 *__restrict__ s, double2 *__restrict__ z) {
  for (int i = blockIdx.x * kNbThreads + threadIdx.x; i < n; i += gridDim.x * kNbThreads) {
    const double2 si = csub(r[i], cmul(alpha, v[i]));
    s[i] = si;
    z[i] = cmul(dinv[i], si);
  }
}

// x += alpha y + w z; r = s - w t; partials of |r|^2 and <r0|r>.
__global__ __launch_bounds__(kNbThreads) void k_nb_xr(int n, double2 alpha, double2 w, double2 *__restrict__ x,
                                                       const double2 *__restrict__ y, const double2 *__restrict__ z,
                                                       const double2 *__restrict__ s, const double2 *__restrict__ t,
                                                       double2 *__restrict__ r, const double2 *__restrict__ r0,
                                                       double2 *__restrict__ partial) {
  double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
  for (int i = blockIdx.x * kNbThreads + threadIdx.x; i < n; i += gridDim.x * kNbThreads) {
    x[i] = cadd(x[i], cadd(cmul(alpha, y[i]), cmul(w, z[i])));
    const double2 ri = csub(s[i], cmul(w, t[i]));
    r[i] = ri;
    a = cadd(a, cdot1(ri, ri));
    b = cadd(b, cdot1(r0[i], ri));
  }
  block_store2(a, b, partial);
}

// restart: r = r0 = b - t; partials of |r|^2
__global__ __launch_bounds__(kNbThreads) void k_nb_restart(int n, const double2 *__restrict__ b,
                                                            const double2 *__restrict__ t, double2 *__restrict__ r,
                                                            double2 *__restrict__ r0, double2 *__restrict__ partial) {
  double2 a = make_double2(0.0, 0.0);
  for (int i = blockIdx.x * kNbThreads + threadIdx.x; i < n; i += gridDim.x * kNbThreads) {
    const double2 ri = csub(b[i], t[i]);
    r[i] = ri;
    r0[i] = ri;
    a = cadd(a, cdot1(ri, ri));
  }
  block_store2(a, make_double2(0.0, 0.0), partial);
}

// c0 = 2 u - c0
__global__ __launch_bounds__(kNbThreads) void k_nb_cheb(int n, const double2 *__restrict__ u, double2 *__restrict__ c0) {
  for (int i = blockIdx.x * kNbThreads + threadIdx.x; i < n; i += gridDim.x * kNbThreads) {
    const double2 ui = u[i], ci = c0[i];
    c0[i] = make_double2(2.0 * ui.x - ci.x, 2.0 * ui.y - ci.y);
  }
}

int nb_grid(int n) { return std::max(1, std::min(kNbBlocks, (n + kNbThreads - 1) / kNbThreads)); }

struct NbBuf {
  std::vector<void *> ptrs;
  ~NbBuf() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  int alloc(size_t count, double2 **out) {
    void *d = nullptr;
    LQ_HIP_CHECK(hipMalloc(&d, sizeof(double2) * std::max<size_t>(count, 1)));
    ptrs.push_back(d);
    *out = (double2 *)d;
    return LQ_OK;
  }
};

// BiCGSTAB state on the device for systems with S (n rows).
struct Bicg {
  lq_ctx ctx = nullptr;
  lq_csr S = nullptr;
  int n = 0, grid = 1;
  double2 *r = nullptr, *r0 = nullptr, *p = nullptr, *v = nullptr, *y = nullptr, *z = nullptr, *s = nullptr,
          *t = nullptr, *dinv = nullptr, *partial = nullptr, *res = nullptr;
  long long iterations = 0;

  int init(lq_ctx c, lq_csr S_, NbBuf &buf) {
    ctx = c;
    S = S_;
    n = S->n;
    grid = nb_grid(n);
    for (double2 **q : {&r, &r0, &p, &v, &y, &z, &s, &t, &dinv}) LQ_TRY(buf.alloc((size_t)n, q));
    LQ_TRY(buf.alloc(2 * (size_t)kNbBlocks, &partial));
    LQ_TRY(buf.alloc(2, &res));
    // DiagonalPreconditioner: 1 / (first stored diagonal entry) when nonzero, else 1
    std::vector<double2> h((size_t)n, make_double2(1.0, 0.0));
    for (int i = 0; i < n; ++i)
      for (int q = S->rowptr[i]; q < S->rowptr[i + 1]; ++q) {
        if (S->kind[q] == 1) continue;  // mirrored entries are not part of the stored matrix
        if (S->col[q] != i) continue;
        const std::complex<double> d = S->val[q];
        if (d != std::complex<double>(0.0)) {
          const std::complex<double> inv = 1.0 / d;
          h[i] = make_double2(inv.real(), inv.imag());
        }
        break;
      }
    LQ_HIP_CHECK(hipMemcpyAsync(dinv, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, ctx->stream));
    LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return LQ_OK;
  }

  // reduce the partials to res and read res[0..1] on the host
  int reduce_read(std::complex<double> *a, std::complex<double> *b) {
    hipLaunchKernelGGL(k_nb_final, dim3(1), dim3(kNbThreads), 0, ctx->stream, partial, grid, res, 0.0, nullptr);
    LQ_HIP_CHECK(hipGetLastError());
    double2 h[2];
    LQ_HIP_CHECK(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (a) *a = {h[0].x, h[0].y};
    if (b) *b = {h[1].x, h[1].y};
    return LQ_OK;
  }

  int apply_S(const double2 *in, double2 *out) {
    return lq_spmv(S, 1.0, 0.0, (const double *)in, 0.0, 0.0, (double *)out, 1);
  }

  // x = S^-1 b (Eigen's BiCGSTAB); b and x are distinct device vectors of n.
  int solve(const double2 *b, double2 *x, double tol, int maxit, int *iters, double *error) {
    typedef std::complex<double> C;
    const dim3 g(grid), blk(kNbThreads);
    hipLaunchKernelGGL(k_nb_init, g, blk, 0, ctx->stream, n, b, r, r0, x, p, v, partial);
    LQ_HIP_CHECK(hipGetLastError());
    C rhs;
    LQ_TRY(reduce_read(&rhs, nullptr));
    const double rhs_sq = rhs.real();
    int it = 0;
    if (rhs_sq == 0.0) {
      if (iters) *iters = 0;
      if (error) *error = 0.0;
      return LQ_OK;
    }
    double r0_sq = rhs_sq, r_sq = rhs_sq;
    C r0r = rhs_sq, rho = 1.0, alpha = 1.0, w = 1.0;
    const double tol2 = tol * tol * rhs_sq;
    const double eps = 2.220446049250313e-16, eps2 = eps * eps;
    int restarts = 0;
    while (r_sq > tol2 && it < maxit) {
      const C rho_old = rho;
      rho = r0r;
      if (std::abs(rho) < eps2 * r0_sq) {  // r too orthogonal to r0: restart from the current x
        LQ_TRY(apply_S(x, t));
        hipLaunchKernelGGL(k_nb_restart, g, blk, 0, ctx->stream, n, b, t, r, r0, partial);
        LQ_HIP_CHECK(hipGetLastError());
        C rr;
        LQ_TRY(reduce_read(&rr, nullptr));
        r0_sq = rr.real();
        rho = r0_sq;
        if (restarts++ == 0) it = 0;
      }
      const C beta = (rho / rho_old) * (alpha / w);
      hipLaunchKernelGGL(k_nb_p, g, blk, 0, ctx->stream, n, make_double2(beta.real(), beta.imag()),
                         make_double2(w.real(), w.imag()), r, p, v, dinv, y);
      LQ_HIP_CHECK(hipGetLastError());
      LQ_TRY(apply_S(y, v));
      hipLaunchKernelGGL(k_nb_dot2, g, blk, 0, ctx->stream, n, r0, v, nullptr, nullptr, partial);
      LQ_HIP_CHECK(hipGetLastError());
      C r0v;
      LQ_TRY(reduce_read(&r0v, nullptr));
      alpha = rho / r0v;
      hipLaunchKernelGGL(k_nb_s, g, blk, 0, ctx->stream, n, make_double2(alpha.real(), alpha.imag()), r, v, dinv, s,
                         z);
      LQ_HIP_CHECK(hipGetLastError());
      LQ_TRY(apply_S(z, t));
      hipLaunchKernelGGL(k_nb_dot2, g, blk, 0, ctx->stream, n, t, t, t, s, partial);
      LQ_HIP_CHECK(hipGetLastError());
      C tt, ts;
      LQ_TRY(reduce_read(&tt, &ts));
      const double tmp = tt.real();
      w = tmp > 0.0 ? ts / tmp : C(0.0);
      hipLaunchKernelGGL(k_nb_xr, g, blk, 0, ctx->stream, n, make_double2(alpha.real(), alpha.imag()),
                         make_double2(w.real(), w.imag()), x, y, z, s, t, r, r0, partial);
      LQ_HIP_CHECK(hipGetLastError());
      C rr;
      LQ_TRY(reduce_read(&rr, &r0r));
      r_sq = rr.real();
      ++it;
    }
    iterations += it;
    if (iters) *iters = it;
    if (error) *error = std::sqrt(r_sq / rhs_sq);
    return LQ_OK;
  }
};

int check_square(lq_csr A, int n, const char *what) {
  if (A && A->n != n) {
    set_error("%s: operator sizes differ (%d vs %d)", what, A->n, n);
    return LQ_ERR_ARG;
  }
  return LQ_OK;
}

}  // namespace
}  // namespace lq

using namespace lq;

extern "C" int lq_overlap_solve(lq_csr S, const double *b, double *x, double tol, int maxit, int *iterations,
                                double *error) try {
  if (!S || !S->ctx || !b || !x || b == x) {
    set_error("lq_overlap_solve: need an operator and distinct device vectors b, x");
    return LQ_ERR_ARG;
  }
  lq_ctx ctx = S->ctx;
  DeviceGuard g(ctx->device);
  NbBuf buf;
  Bicg bc;
  LQ_TRY(bc.init(ctx, S, buf));
  return bc.solve((const double2 *)b, (double2 *)x, tol > 0 ? tol : 1e-4, maxit > 0 ? maxit : 2000, iterations,
                  error);
} LQ_ABI_CATCH

extern "C" int lq_spectral_moments_nonorth(lq_csr H, lq_csr OP, lq_csr S, const double *phi, int R, int M,
                                           double tol, int maxit, double *mu, long long *iterations) try {
  if (!H || !S || !phi || !mu || R <= 0 || M <= 0) {
    set_error("lq_spectral_moments_nonorth: invalid arguments");
    return LQ_ERR_ARG;
  }
  const int n = H->n;
  LQ_TRY(check_square(S, n, "lq_spectral_moments_nonorth"));
  LQ_TRY(check_square(OP, n, "lq_spectral_moments_nonorth"));
  if (S->ctx != H->ctx || (OP && OP->ctx != H->ctx)) {
    set_error("lq_spectral_moments_nonorth: operators belong to different contexts");
    return LQ_ERR_ARG;
  }
  tol = tol > 0 ? tol : 1e-4;
  maxit = maxit > 0 ? maxit : 2000;
  lq_ctx ctx = H->ctx;
  DeviceGuard g(ctx->device);
  NbBuf buf;
  Bicg bc;
  LQ_TRY(bc.init(ctx, S, buf));
  double2 *c0, *c1, *tmp, *tmp2, *ph, *dmu;
  for (double2 **q : {&c0, &c1, &tmp, &tmp2, &ph}) LQ_TRY(buf.alloc((size_t)n, q));
  LQ_TRY(buf.alloc((size_t)M, &dmu));
  LQ_HIP_CHECK(hipMemcpyAsync(dmu, mu, sizeof(double2) * (size_t)M, hipMemcpyHostToDevice, ctx->stream));
  const dim3 gr(bc.grid), blk(kNbThreads);
  for (int r = 0; r < R; ++r) {
    LQ_TRY(upload_states(ctx, phi, n, r, 1, ph));
    // SetInitVectors_nonOrthogonal: T0 = [OP] S^-1 phi, T1 = S^-1 H T0
    if (!OP) {
      LQ_TRY(bc.solve(ph, c0, tol, maxit, nullptr, nullptr));
    } else {
      LQ_TRY(bc.solve(ph, tmp, tol, maxit, nullptr, nullptr));
      LQ_TRY(lq_spmv(OP, 1.0, 0.0, (const double *)tmp, 0.0, 0.0, (double *)c0, 1));
    }
    LQ_TRY(lq_spmv(H, 1.0, 0.0, (const double *)c0, 0.0, 0.0, (double *)tmp, 1));
    LQ_TRY(bc.solve(tmp, c1, tol, maxit, nullptr, nullptr));
    for (int m = 0; m < M; ++m) {
      double scal = 2.0 / R;
      if (m == 0) scal *= 0.5;
      hipLaunchKernelGGL(k_nb_dot2, gr, blk, 0, ctx->stream, n, ph, c0, nullptr, nullptr, bc.partial);
      LQ_HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(k_nb_final, dim3(1), blk, 0, ctx->stream, bc.partial, bc.grid, bc.res, scal, dmu + m);
      LQ_HIP_CHECK(hipGetLastError());
      if (m + 1 == M) break;  // the reference's last Iterate leaves mu unchanged
      // Iterate_nonOrthogonal: T0 <- 2 S^-1 H T1 - T0, swap
      LQ_TRY(lq_spmv(H, 1.0, 0.0, (const double *)c1, 0.0, 0.0, (double *)tmp, 1));
      LQ_TRY(bc.solve(tmp, tmp2, tol, maxit, nullptr, nullptr));
      hipLaunchKernelGGL(k_nb_cheb, gr, blk, 0, ctx->stream, n, tmp2, c0);
      LQ_HIP_CHECK(hipGetLastError());
      std::swap(c0, c1);
    }
  }
  LQ_HIP_CHECK(hipMemcpyAsync(mu, dmu, sizeof(double2) * (size_t)M, hipMemcpyDeviceToHost, ctx->stream));
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (iterations) *iterations = bc.iterations;
  return LQ_OK;
} LQ_ABI_CATCH
