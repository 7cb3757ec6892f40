// lq_post.hip -- moment epilogues, Jackson damping and energy-grid sums.
//
// All FP64.  The single-thread sequential scans (cumulative integrals) keep
// the reference's left-to-right summation order.
#include "lq_internal.h"

namespace lq {

namespace {

constexpr double kPi = 3.14159265358979323846;

// Order of the Chebyshev vector the reference stores at padded index p when
// BATCH_SIZE = B: every IterateAll call after the first restarts at the last
// stored order (src/chebyshev_vectors.cpp:4-21), so block q covers orders
// q*(B-1) .. q*(B-1)+B-1.
__device__ __forceinline__ int batch_order(int p, int B) { return (p / B) * (B - 1) + (p % B); }

// Scaling + Hermitian symmetrisation of CorrelationExpansionMoments
// (src/chebyshev_solver.cpp:149-162) on the padded table, restricted to the
// M0 x M1 block that survives the final truncation (:165-169).  raw holds
// true-order moments summed over states; init (nullable) the caller's table.
__global__ void k_corr_epilogue(const double2 *__restrict__ raw, int ldr, const double2 *__restrict__ init, int M0,
                                int M1, int B, int R, double2 *__restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)M0 * M1) return;
  const int p = (int)(t / M1), q = (int)(t % M1);
  const int op = batch_order(p, B), oq = batch_order(q, B);
  double2 apq = raw[(size_t)op * ldr + oq];
  double2 aqp = raw[(size_t)oq * ldr + op];
  if (init) {
    const double2 i1 = init[(size_t)p * M1 + q];
    apq.x += i1.x;
    apq.y += i1.y;
    if (q < M0 && p < M1) {
      const double2 i2 = init[(size_t)q * M1 + p];
      aqp.x += i2.x;
      aqp.y += i2.y;
    }
  }
  // the reference visits each unordered pair once with mL <= mR
  const int lo = p <= q ? p : q, hi = p <= q ? q : p;
  const double2 first = p <= q ? apq : aqp;   // mu(lo, hi)
  const double2 second = p <= q ? aqp : apq;  // mu(hi, lo)
  double scal = 4.0 / R;
  if (lo == 0) scal *= 0.5;
  if (hi == 0) scal *= 0.5;
  const double tx = scal * (first.x + second.x) / 2.0;
  const double ty = scal * (first.y - second.y) / 2.0;
  out[t] = p <= q ? make_double2(tx, ty) : make_double2(tx, -ty);
}

__global__ void k_scale(double2 *__restrict__ x, long long count, double s) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) {
    x[t].x *= s;
    x[t].y *= s;
  }
}

// Per-energy sums of Kubo_solver_FFT's per-site DFT products
// (src/Greenwood_FFTs.cpp:59-83, src/Bastin_FFTs.cpp:172-257) written through
// the moment table mu(m, n) = <L_m|R_n> (mean over states):
//   theta_k = pi (2k + 1/2) / nump, s = +1 for k < nump/2 else -1,
//   t_m = c_m cos(m theta), e_m = c_m exp(i s m theta), d_m = m e_m,
//   Greenwood: r_k = Re sum t_m t_n mu_mn
//   Bastin:    p_k = sum mu_mn (t_m e_n + conj(e_m) t_n)
//              w_k = sum mu_mn (conj(d_m) t_n - t_m d_n)
// with c_m = (2 - delta_m0) g_m (Jackson, Kubo_solver_FFT.hpp:99-107).
// One block per energy; thread j owns columns n = j, j+256, ...
__global__ __launch_bounds__(256) void k_kubo_pw(const double2 *__restrict__ mu, int ldm, int M, int nump,
                                                 int formula, const double *__restrict__ coef,
                                                 double2 *__restrict__ out) {
  extern __shared__ double tab[];  // [3*M]: t_m, e_m (re, im) -> (c cos, c sin)
  double *tc = tab, *ts = tab + M;
  const int k = blockIdx.x;
  const double theta = kPi * (2 * k + 0.5) / nump;
  const double sgn = (k < nump / 2) ? 1.0 : -1.0;
  for (int m = threadIdx.x; m < M; m += blockDim.x) {
    double sn, cs;
    sincos(m * theta, &sn, &cs);
    tc[m] = coef[m] * cs;
    ts[m] = coef[m] * sgn * sn;
  }
  __syncthreads();
  double pr = 0, pi = 0, wr = 0, wi = 0, rr = 0;
  for (int nn = threadIdx.x; nn < M; nn += blockDim.x) {
    // z^t = sum_m t_m mu_mn ; z^e = sum_m conj(e_m) mu_mn ; z^d = sum_m m conj(e_m) mu_mn
    double ztr = 0, zti = 0, zer = 0, zei = 0, zdr = 0, zdi = 0;
    for (int m = 0; m < M; ++m) {
      const double2 a = mu[(size_t)m * ldm + nn];
      const double t = tc[m], er = tc[m], ei = -ts[m];  // conj(e_m) = c cos - i c s sin
      ztr = fma(t, a.x, ztr);
      zti = fma(t, a.y, zti);
      const double cr = er * a.x - ei * a.y, ci = er * a.y + ei * a.x;
      zer += cr;
      zei += ci;
      zdr = fma((double)m, cr, zdr);
      zdi = fma((double)m, ci, zdi);
    }
    const double tn = tc[nn], enr = tc[nn], eni = ts[nn];
    if (formula == LQ_KUBO_GREENWOOD) {
      rr += tn * ztr;
    } else {
      const double dnr = nn * enr, dni = nn * eni;
      // p += z^t * e_n + z^e * t_n
      pr += ztr * enr - zti * eni + zer * tn;
      pi += ztr * eni + zti * enr + zei * tn;
      // w += z^d * t_n - z^t * d_n
      wr += zdr * tn - (ztr * dnr - zti * dni);
      wi += zdi * tn - (ztr * dni + zti * dnr);
    }
  }
  __shared__ double red[5][256];
  red[0][threadIdx.x] = pr;
  red[1][threadIdx.x] = pi;
  red[2][threadIdx.x] = wr;
  red[3][threadIdx.x] = wi;
  red[4][threadIdx.x] = rr;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int q = 0; q < 5; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (formula == LQ_KUBO_GREENWOOD) {
      out[k] = make_double2(red[4][0], 0.0);
      out[k + nump] = make_double2(0.0, 0.0);
    } else {
      out[k] = make_double2(red[0][0], red[1][0]);
      out[k + nump] = make_double2(red[2][0], red[3][0]);
    }
  }
}

// rearrange_crescent_order (src/Kubo_solver_FFT_postProcess.cpp:98-116) as an
// index map: position j of the rearranged array holds original index src(j).
__device__ __forceinline__ int crescent_src(int j, int nump) {
  // second pass reverses the first nump/2 pairs: j <-> nump-1-j for j < nump/2
  int i = j;
  if (j < nump / 2)
    i = nump - 1 - j;
  else if (j >= nump - nump / 2)
    i = nump - 1 - j;
  // first pass: even 2k <- k, odd 2k+1 <- nump-1-k (k < nump/2); others unchanged
  if (i < 2 * (nump / 2)) return (i % 2 == 0) ? i / 2 : nump - 1 - i / 2;
  return i;
}

// Kubo_solver_FFT_postProcess (src/Kubo_solver_FFT_postProcess.cpp:8-16, 36-47,
// 120-273, 279-397), one block; the cumulative sum runs on one thread.
__global__ void k_kubo_post(int formula, const double2 *__restrict__ fd, int nump, int n, double W, double Ec,
                            double *__restrict__ Eo, double *__restrict__ so, double *__restrict__ integ) {
  const double a = W / 2.0, b = Ec;
  const double omega = n / (a * a);
  for (int j = threadIdx.x; j < nump; j += blockDim.x) {
    const int k = crescent_src(j, nump);
    const double e = cos(kPi * (2 * k + 0.5) / nump);
    Eo[j] = a * e - b;
    if (formula == LQ_KUBO_BASTIN) {
      double v = e * fd[k].x - (sqrt(1.0 - e * e) * fd[k + nump].y);
      v *= 1.0 / pow((1.0 - e * e), 2.0);
      v *= omega / (kPi);
      integ[j] = v;
    } else {
      so[j] = 2.0 * omega * fd[k].x / (1.0 - e * e) / (2 * kPi);
      if (integ) integ[j] = 0.0;
    }
  }
  if (formula != LQ_KUBO_BASTIN) return;
  // ascending energies in LDS, then the left-Riemann cumulative sum in the
  // reference's order on one thread
  extern __shared__ double er[];
  for (int j = threadIdx.x; j < nump; j += blockDim.x) er[j] = cos(kPi * (2 * crescent_src(j, nump) + 0.5) / nump);
  __syncthreads();
  if (threadIdx.x == 0) {
    double acc = 0;
    for (int i = 0; i < nump - 1; ++i) {
      acc += integ[i] * (er[i + 1] - er[i]);
      so[i] = acc;
    }
    so[nump - 1] = 0.0;
  }
}

// Moments1D::ApplyJacksonKernel (src/chebyshev_moments1D.cpp:79-86).
__global__ void k_jackson_1d(double2 *__restrict__ mu, int maxMom) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= maxMom) return;
  const double phi_J = kPi / (double)(maxMom + 1.0);
  const double g = ((double)(maxMom - m + 1) * cos(phi_J * m) + sin(phi_J * m) * cos(phi_J) / sin(phi_J)) * phi_J /
                   kPi;
  mu[m].x *= g;
  mu[m].y *= g;
}

// Moments2D::MomentNumber + ApplyJacksonKernel (src/chebyshev_moments2D.cpp:87-117).
__global__ void k_jackson_2d(const double2 *__restrict__ in, int M1_in, double2 *__restrict__ out, int mm0, int mm1) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)mm0 * mm1) return;
  const int m0 = (int)(t / mm1), m1 = (int)(t % mm1);
  const double phi0 = kPi / (double)(mm0 + 1.0), phi1 = kPi / (double)(mm1 + 1.0);
  const double g0 = ((mm0 - m0 + 1) * cos(phi0 * m0) + sin(phi0 * m0) * cos(phi0) / sin(phi0)) * phi0 / kPi;
  const double g1 = ((mm1 - m1 + 1) * cos(phi1 * m1) + sin(phi1 * m1) * cos(phi1) / sin(phi1)) * phi1 / kPi;
  const double g = g0 * g1;
  const double2 v = in[(size_t)m0 * M1_in + m1];
  out[t] = make_double2(v.x * g, v.y * g);
}

// include/chebyshev_coefficients.hpp:8-32; pow(complex, m) as libstdc++ does
// it: polar(exp(m log|z|), m arg z).
__device__ __forceinline__ double delta_chebF(double x, double m) {
  const double f0 = sqrt(1.0 - x * x);
  return cos(m * acos(x)) / f0 / kPi;
}

__device__ __forceinline__ double2 zpow(double zr, double zi, double m) {
  const double lr = log(hypot(zr, zi)), li = atan2(zi, zr);
  const double rho = exp(m * lr);
  double sn, cs;
  sincos(m * li, &sn, &cs);
  return make_double2(rho * cs, rho * sn);
}

__device__ __forceinline__ double2 greenR_chebF(double x, double m) {
  const double f0 = sqrt(1.0 - x * x);
  const double2 fm = zpow(x, -f0, m);
  // -I * fm / f0
  return make_double2(fm.y / f0, -fm.x / f0);
}

__device__ __forceinline__ double2 DgreenR_chebF(double x, double m) {
  const double f0 = sqrt(1.0 - x * x);
  const double2 p = zpow(x, -f0, m);
  const double2 fm = make_double2(p.x * x - p.y * (m * f0), p.x * (m * f0) + p.y * x);
  // -I * fm / f0 / f0 / f0
  return make_double2(fm.y / f0 / f0 / f0, -fm.x / f0 / f0 / f0);
}

// spectralFunctionFromChebmom_FFTgrid.cpp:60-73.
// One wavefront per energy: lane l sums the moments m = l, l + 64, ..., then
// the 64 lane sums are added in a fixed tree (deterministic).
__global__ __launch_bounds__(64) void k_dos_fftgrid(const double2 *__restrict__ mu, int M, int n, double W, double Ec,
                                                    int num_div, double *__restrict__ E, double *__restrict__ rho) {
  __shared__ double sh[64];
  const int i = blockIdx.x;
  const int l = threadIdx.x;
  const double a = W / 2.0;
  const double energ = cos(kPi * (i + 0.5) / num_div);
  double part = 0;
  for (int m0 = l; m0 < M; m0 += 64) part += delta_chebF(energ, m0) * mu[m0].x;
  sh[l] = part;
  __syncthreads();
  for (int h = 32; h > 0; h >>= 1) {
    if (l < h) sh[l] += sh[l + h];
    __syncthreads();
  }
  if (l == 0) {
    E[i] = energ * a + Ec;
    rho[i] = fabs(sh[0] * (n / a));
  }
}

// kuboGreenwoodFromChebmom_FFTgrid.cpp:57-66 and kuboBastinFromChebmom_FFTgrid.cpp:77-86:
// one block per energy; thread j owns columns m1 = j, j+256, ...:
//   kernel = sum_m1 Im( G(E, m1) * sum_m0 delta(E, m0) mu(m0, m1) ).
template <int DERIV>
__global__ __launch_bounds__(256) void k_grid2d(const double2 *__restrict__ mu, int M0, int M1, const double *__restrict__ energies,
                                                double *__restrict__ kern) {
  extern __shared__ double dl[];  // delta(E, m0)
  const int i = blockIdx.x;
  const double energ = energies[i];
  for (int m0 = threadIdx.x; m0 < M0; m0 += blockDim.x) dl[m0] = delta_chebF(energ, m0);
  __syncthreads();
  double acc = 0;
  for (int m1 = threadIdx.x; m1 < M1; m1 += blockDim.x) {
    double zr = 0, zi = 0;
    for (int m0 = 0; m0 < M0; ++m0) {
      const double2 v = mu[(size_t)m0 * M1 + m1];
      zr = fma(dl[m0], v.x, zr);
      zi = fma(dl[m0], v.y, zi);
    }
    const double2 g = DERIV ? DgreenR_chebF(energ, m1) : greenR_chebF(energ, m1);
    acc += g.x * zi + g.y * zr;  // Im(g * z)
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) kern[i] = red[0];
}

// Linear-grid tools (src/*FromChebmom.cpp), kinds as LQ_LIN_*: one block per
// energy evaluates the tool's double sum in separable form,
//   sum_m1 proj( b_m1(x) * sum_m0 a_m0(x) mu(m0, m1) ),
// with (a, b, proj) = (delta, gR, Im) Greenwood/BastinI, (delta, DgR, Im)
// Bastin, (delta, delta, Re) Surf, (Im gR, Re DgR, 4 Im) Sea; BastinII uses
// the two left vectors gR and DgR: -Re(DgR_m1 z_gR - gR_m1 z_DgR).
__device__ __forceinline__ double lin_xgrid(int kind, int i, int num_div) {
  const double xb = kind == 2 ? 0.99 : 1.0;
  return -xb + i * (2 * xb) / (num_div - 1);
}

template <int KIND>
__global__ __launch_bounds__(256) void k_linear_kernel(const double2 *__restrict__ mu, int M0, int M1, int num_div,
                                                       double scale, double *__restrict__ kern) {
  extern __shared__ double2 lv[];  // left vector(s): [M0] (BastinII: [2][M0])
  const int i = blockIdx.x;
  const double x = lin_xgrid(KIND, i, num_div);
  for (int m0 = threadIdx.x; m0 < M0; m0 += blockDim.x) {
    if (KIND == 4) {
      lv[m0] = greenR_chebF(x, m0);
      lv[M0 + m0] = DgreenR_chebF(x, m0);
    } else if (KIND == 5) {
      lv[m0] = make_double2(greenR_chebF(x, m0).y, 0.0);
    } else {
      lv[m0] = make_double2(delta_chebF(x, m0), 0.0);
    }
  }
  __syncthreads();
  double acc = 0;
  if (KIND == 0) {
    for (int m0 = threadIdx.x; m0 < M0; m0 += blockDim.x) acc += lv[m0].x * mu[m0].x;
  } else {
    for (int m1 = threadIdx.x; m1 < M1; m1 += blockDim.x) {
      double zr = 0, zi = 0, wr = 0, wi = 0;
      for (int m0 = 0; m0 < M0; ++m0) {
        const double2 v = mu[(size_t)m0 * M1 + m1];
        const double2 l = lv[m0];
        if (KIND == 4) {
          zr += l.x * v.x - l.y * v.y;
          zi += l.x * v.y + l.y * v.x;
          const double2 dl = lv[M0 + m0];
          wr += dl.x * v.x - dl.y * v.y;
          wi += dl.x * v.y + dl.y * v.x;
        } else {
          zr = fma(l.x, v.x, zr);
          zi = fma(l.x, v.y, zi);
        }
      }
      if (KIND == 1 || KIND == 3) {
        const double2 g = greenR_chebF(x, m1);
        acc += g.x * zi + g.y * zr;  // Im(g z)
      } else if (KIND == 2) {
        const double2 g = DgreenR_chebF(x, m1);
        acc += g.x * zi + g.y * zr;
      } else if (KIND == 6) {
        acc += delta_chebF(x, m1) * zr;
      } else if (KIND == 5) {
        acc += 4.0 * DgreenR_chebF(x, m1).x * zi;
      } else {  // 4: -Re(DgR z_gR - gR z_DgR)
        const double2 d = DgreenR_chebF(x, m1), g = greenR_chebF(x, m1);
        acc += -((d.x * zr - d.y * zi) - (g.x * wr - g.y * wi));
      }
    }
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) kern[i] = red[0] * scale;
}

// The printed columns, in the sources' left-to-right order (one thread: the
// running sums must keep that order).
__global__ void k_linear_finish(int kind, int num_div, double a, double Ec, const double *__restrict__ kern,
                                double *__restrict__ E, double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int rows = kind == 0 ? num_div : num_div - 1;
  double acc = 0;
  for (int i = 0; i < rows; ++i) {
    const double x = lin_xgrid(kind, i, num_div);
    E[i] = x * a + Ec;
    if (kind == 0 || kind == 1) {
      out[i] = kern[i];
    } else {
      const double dx = lin_xgrid(kind, i + 1, num_div) - x;
      switch (kind) {
        case 2: acc += kern[i] * dx; out[i] = acc; break;
        case 3: out[i] = 0.5 * (kern[i] + kern[i + 1]); break;
        case 4: acc += (kern[i] + kern[i + 1]) * dx; out[i] = acc; break;
        case 5: acc += 0.5 * (kern[i] + kern[i + 1]) * dx; out[i] = acc; break;
        default: out[i] = (kern[i + 1] + kern[i]) / 2.0; break;
      }
    }
  }
}

// y (=|+=) c x over count complex elements (Bessel-Chebyshev time evolution).
__global__ void k_caxpy(double2 c, const double2 *__restrict__ x, double2 *__restrict__ y, size_t count,
                        int overwrite) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const double2 v = x[i];
  double2 r = make_double2(c.x * v.x - c.y * v.y, c.x * v.y + c.y * v.x);
  if (!overwrite) {
    const double2 o = y[i];
    r.x += o.x;
    r.y += o.y;
  }
  y[i] = r;
}

// MomentsTD::ApplyJacksonKernel (src/chebyshev_momentsTD.cpp:135-154): row m of
// the M x nT table (rows kept: maxMom) times g_m.
__global__ void k_jackson_td(double2 *__restrict__ mu, int maxMom, int nT) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)maxMom * nT) return;
  const int m = (int)(t / nT);
  const double phi_J = kPi / (double)(maxMom + 1.0);
  const double g = ((maxMom - m + 1) * cos(phi_J * m) + sin(phi_J * m) * cos(phi_J) / sin(phi_J)) * phi_J / kPi;
  mu[t].x *= g;
  mu[t].y *= g;
}

// timeCorrelationsFromChebmom.cpp:60-70: one thread per time step.
__global__ void k_timecorr_post(const double2 *__restrict__ mu, int M, int nT, int n, double a, double x, double dt,
                                double *__restrict__ tout, double *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nT) return;
  double o = 0.0;
  for (int m = 0; m < M; ++m) o += delta_chebF(x, m) * mu[(size_t)m * nT + t].x;
  tout[t] = t * dt;
  out[t] = o * (n / a);
}

__global__ void k_greenwood_energies(int num_div, double *__restrict__ e) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < num_div) e[i] = cos(kPi * (2 * i + 0.5) / num_div);
}

__global__ void k_bastin_energies(int num_div, double *__restrict__ e) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < num_div) e[j] = cos(kPi * (2 * crescent_src(j, num_div) + 0.5) / num_div);
}

__global__ void k_greenwood_finish(int num_div, int n, double W, double Ec, const double *__restrict__ energies,
                                   const double *__restrict__ kern, double *__restrict__ E, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_div - 1) return;
  const double a = W / 2.0;
  E[i] = energies[i] * a + Ec;
  out[i] = -1.0 * kern[i] * (n / a / a);
}

__global__ void k_bastin_finish(int num_div, int n, double W, double Ec, const double *__restrict__ energies,
                                double *__restrict__ kern, double *__restrict__ E, double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double a = W / 2.0;
  double acc = 0;
  for (int i = 0; i < num_div - 1; ++i) {
    const double k = kern[i] * (-2.0 * n / a / a);
    acc += k * (energies[i + 1] - energies[i]);
    E[i] = energies[i] * a + Ec;
    out[i] = acc;
  }
}

}  // namespace

int launch_linear_grid(hipStream_t st, int kind, const double2 *mu, int M0, int M1, int n, double W, double Ec,
                       int num_div, double *kern, double *E, double *out) {
  const double a = W / 2.0;
  double scale;
  switch (kind) {
    case 0: scale = n / a; break;
    case 1: case 3: scale = -1.0 * (n / a / a); break;
    case 2: scale = -2.0 * n / a / a; break;
    case 4: case 5: scale = n / a / a / 2 / kPi; break;
    case 6: scale = kPi * n / a / a; break;
    default:
      set_error("launch_linear_grid: unknown kind %d", kind);
      return LQ_ERR_ARG;
  }
  const size_t lds = sizeof(double2) * (size_t)M0 * (kind == 4 ? 2 : 1);
  if (lds > 64 * 1024) {
    set_error("launch_linear_grid: %d moments exceed the LDS table (max %d)", M0, kind == 4 ? 2048 : 4096);
    return LQ_ERR_UNSUPPORTED;
  }
  switch (kind) {
#define LQ_LIN(K) \
  case K: hipLaunchKernelGGL((k_linear_kernel<K>), dim3(num_div), dim3(256), lds, st, mu, M0, M1, num_div, scale, kern); break;
    LQ_LIN(0) LQ_LIN(1) LQ_LIN(2) LQ_LIN(3) LQ_LIN(4) LQ_LIN(5) LQ_LIN(6)
#undef LQ_LIN
  }
  hipLaunchKernelGGL(k_linear_finish, dim3(1), dim3(64), 0, st, kind, num_div, a, Ec, kern, E, out);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_caxpy(hipStream_t st, double2 c, const double2 *x, double2 *y, size_t count, int overwrite) {
  hipLaunchKernelGGL(k_caxpy, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, c, x, y, count, overwrite);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_jackson_td(hipStream_t st, double2 *mu, int maxMom, int nT) {
  const long long total = (long long)maxMom * nT;
  if (total <= 0) return LQ_OK;
  hipLaunchKernelGGL(k_jackson_td, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, mu, maxMom, nT);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_timecorr_post(hipStream_t st, const double2 *mu, int M, int nT, int n, double W, double Ec, double dt,
                         double EF, double *t, double *out) {
  const double a = W / 2.0;
  hipLaunchKernelGGL(k_timecorr_post, dim3((nT + 127) / 128), dim3(128), 0, st, mu, M, nT, n, a, (EF - Ec) / a, dt, t,
                     out);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_corr_epilogue(hipStream_t st, const double2 *raw, int ldr, const double2 *init, int M0, int M1, int B,
                         int R, double2 *out) {
  const long long total = (long long)M0 * M1;
  hipLaunchKernelGGL(k_corr_epilogue, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, raw, ldr, init, M0,
                     M1, B, R, out);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_scale(hipStream_t st, double2 *x, size_t count, double s) {
  if (count == 0) return LQ_OK;
  hipLaunchKernelGGL(k_scale, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, x, (long long)count, s);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_kubo_pw(hipStream_t st, const double2 *mu, int ldm, int M, int nump, int formula, const double *coef,
                   double2 *out) {
  const size_t lds = sizeof(double) * 2 * (size_t)M;
  if (lds > 64 * 1024) {
    set_error("kubo_pw: M = %d exceeds the LDS table limit (4096)", M);
    return LQ_ERR_ARG;
  }
  hipLaunchKernelGGL(k_kubo_pw, dim3(nump), dim3(256), lds, st, mu, ldm, M, nump, formula, coef, out);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_kubo_post(hipStream_t st, int formula, const double2 *final_data, int nump, int n, double W, double Ec,
                     double *E, double *sigma, double *integrand) {
  if ((size_t)nump * sizeof(double) > 64 * 1024) {
    set_error("kubo_post: nump = %d exceeds 8192", nump);
    return LQ_ERR_ARG;
  }
  hipLaunchKernelGGL(k_kubo_post, dim3(1), dim3(256), (size_t)nump * sizeof(double), st, formula, final_data, nump, n,
                     W, Ec, E, sigma, integrand);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_jackson_1d(hipStream_t st, double2 *mu, int maxMom) {
  hipLaunchKernelGGL(k_jackson_1d, dim3((maxMom + 255) / 256), dim3(256), 0, st, mu, maxMom);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_jackson_2d(hipStream_t st, const double2 *mu_in, int M1_in, double2 *mu_out, int mm0, int mm1) {
  const long long total = (long long)mm0 * mm1;
  hipLaunchKernelGGL(k_jackson_2d, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, mu_in, M1_in, mu_out, mm0,
                     mm1);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_dos_fftgrid(hipStream_t st, const double2 *mu, int M, int n, double W, double Ec, int num_div, double *E,
                       double *rho) {
  hipLaunchKernelGGL(k_dos_fftgrid, dim3(num_div), dim3(64), 0, st, mu, M, n, W, Ec, num_div, E, rho);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_greenwood_fftgrid(hipStream_t st, const double2 *mu, int M0, int M1, int n, double W, double Ec,
                             int num_div, double *scratch, double *E, double *out) {
  double *energies = scratch, *kern = scratch + num_div;
  const size_t lds = sizeof(double) * (size_t)M0;
  if (lds > 64 * 1024) {
    set_error("greenwood_fftgrid: M0 = %d exceeds 8192", M0);
    return LQ_ERR_ARG;
  }
  hipLaunchKernelGGL(k_greenwood_energies, dim3((num_div + 255) / 256), dim3(256), 0, st, num_div, energies);
  hipLaunchKernelGGL((k_grid2d<0>), dim3(num_div), dim3(256), lds, st, mu, M0, M1, energies, kern);
  hipLaunchKernelGGL(k_greenwood_finish, dim3((num_div + 255) / 256), dim3(256), 0, st, num_div, n, W, Ec, energies,
                     kern, E, out);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_bastin_fftgrid(hipStream_t st, const double2 *mu, int M0, int M1, int n, double W, double Ec, int num_div,
                          double *scratch, double *E, double *out) {
  double *energies = scratch, *kern = scratch + num_div;
  const size_t lds = sizeof(double) * (size_t)M0;
  if (lds > 64 * 1024) {
    set_error("bastin_fftgrid: M0 = %d exceeds 8192", M0);
    return LQ_ERR_ARG;
  }
  hipLaunchKernelGGL(k_bastin_energies, dim3((num_div + 255) / 256), dim3(256), 0, st, num_div, energies);
  hipLaunchKernelGGL((k_grid2d<1>), dim3(num_div), dim3(256), lds, st, mu, M0, M1, energies, kern);
  hipLaunchKernelGGL(k_bastin_finish, dim3(1), dim3(64), 0, st, num_div, n, W, Ec, energies, kern, E, out);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

}  // namespace lq
