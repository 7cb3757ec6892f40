// lq_contract.hip -- the moment contraction mu(m0, m1) = sum_k conj(L_m0[k]) R_m1[k].
//
// Replaces chebyshev::ComputeMomTable (src/chebyshev_solver.cpp:26-42), which
// runs one cblas_zdotc per (m0, m1) pair.  Here the stored vectors of all
// states of a group form two K-contiguous matrices (K = n * states), and the
// table is one complex GEMM  C = A^H B  on FP64 MFMA (v_mfma_f64_16x16x4_f64):
//   Re C = Ar Br + Ai Bi,   Im C = Ar Bi - Ai Br.
//
// Block tile 64 x 64 (four waves of 32 x 32 = 2 x 2 MFMA tiles).  K is split
// into nsplit chunks; chunk = blockIdx % nsplit, so with nsplit a multiple of
// 8 every block of one K-chunk lands on the same XCD and the chunk's rows of
// A and B are shared in that XCD's L2.  A deterministic second pass sums the
// chunks.
#include "lq_internal.h"

namespace lq {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// One K step of 16 complex elements: lane (r16, ks) feeds the k-slot ks of
// four consecutive MFMAs with elements kk + s, s = 0..3.  A and B use the same
// element-to-slot map, so every product pairs equal K indices.
__device__ __forceinline__ void mma_step(const double2 (&a)[2][4], const double2 (&b)[2][4], v4d (&cr)[2][2],
                                         v4d (&ci)[2][2]) {
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) {
        cr[ti][tj] = mfma(a[ti][s].x, b[tj][s].x, cr[ti][tj]);
        cr[ti][tj] = mfma(a[ti][s].y, b[tj][s].y, cr[ti][tj]);
        ci[ti][tj] = mfma(a[ti][s].x, b[tj][s].y, ci[ti][tj]);
        ci[ti][tj] = mfma(-a[ti][s].y, b[tj][s].x, ci[ti][tj]);
      }
}

__global__ __launch_bounds__(256) void k_contract_f64(const double2 *__restrict__ A, long long lda, int M0,
                                                      const double2 *__restrict__ B, long long ldb, int M1,
                                                      long long K, long long kchunk, int nsplit, int tiles_n,
                                                      double2 *__restrict__ work) {
  const int bid = blockIdx.x;
  const int chunk = bid % nsplit;
  const int tile = bid / nsplit;
  const int tm = tile / tiles_n, tn = tile % tiles_n;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r16 = lane & 15, ks = lane >> 4;
  const int m_base = tm * 64 + (wave >> 1) * 32;
  const int n_base = tn * 64 + (wave & 1) * 32;

  const double2 *pa[2], *pb[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    int m = m_base + t * 16 + r16;
    if (m >= M0) m = M0 - 1;  // clamped rows feed outputs that are never stored
    pa[t] = A + (size_t)m * lda;
    int nn = n_base + t * 16 + r16;
    if (nn >= M1) nn = M1 - 1;
    pb[t] = B + (size_t)nn * ldb;
  }
  const long long k0 = (long long)chunk * kchunk;
  const long long k1 = (k0 + kchunk < K) ? k0 + kchunk : K;

  v4d cr[2][2], ci[2][2];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
      cr[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};
      ci[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};
    }

  long long kb = k0;
  for (; kb + 16 <= k1; kb += 16) {
    const long long kk = kb + ks * 4;
    double2 a[2][4], b[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        a[t][s] = pa[t][kk + s];
        b[t][s] = pb[t][kk + s];
      }
    mma_step(a, b, cr, ci);
  }
  if (kb < k1) {
    const long long kk = kb + ks * 4;
    double2 a[2][4], b[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const bool ok = kk + s < k1;
        a[t][s] = ok ? pa[t][kk + s] : make_double2(0.0, 0.0);
        b[t][s] = ok ? pb[t][kk + s] : make_double2(0.0, 0.0);
      }
    mma_step(a, b, cr, ci);
  }

  // D layout of v_mfma_f64_16x16x4: element r of lane l is D[4*(l/16) + r][l % 16].
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m_base + ti * 16 + ks * 4 + r;
        const int nn = n_base + tj * 16 + r16;
        if (m < M0 && nn < M1) work[((size_t)chunk * M0 + m) * M1 + nn] = make_double2(cr[ti][tj][r], ci[ti][tj][r]);
      }
}

__global__ void k_contract_reduce(const double2 *__restrict__ work, int nsplit, int M0, int M1,
                                  double2 *__restrict__ out, int ldo, int accumulate) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)M0 * M1) return;
  const int m = (int)(t / M1), nn = (int)(t % M1);
  double sx = 0.0, sy = 0.0;
  for (int c = 0; c < nsplit; ++c) {
    const double2 w = work[((size_t)c * M0 + m) * M1 + nn];
    sx += w.x;
    sy += w.y;
  }
  double2 *o = out + (size_t)m * ldo + nn;
  if (accumulate) {
    sx += o->x;
    sy += o->y;
  }
  *o = make_double2(sx, sy);
}

int choose_nsplit(int M0, int M1, long long K) {
  const long long tiles = (long long)((M0 + 63) / 64) * ((M1 + 63) / 64);
  long long ns = (2048 + tiles - 1) / tiles;  // aim at >= 2048 blocks
  ns = ((ns + 7) / 8) * 8;                     // whole XCD groups
  if (ns > 256) ns = 256;
  long long by_k = K / 4096;  // keep every chunk >= 4096 elements long
  if (by_k < 1) by_k = 1;
  if (ns > by_k) ns = by_k;
  return (int)ns;
}

}  // namespace

int contract_workspace_bytes(int M0, int M1, long long K, size_t *bytes, int *nsplit) {
  const int ns = choose_nsplit(M0, M1, K);
  if (nsplit) *nsplit = ns;
  if (bytes) *bytes = (size_t)ns * M0 * M1 * sizeof(double2);
  return LQ_OK;
}

int launch_contract_f64(hipStream_t st, const double2 *A, long long lda, int M0, const double2 *B, long long ldb,
                        int M1, long long K, double2 *work, int nsplit) {
  if (M0 <= 0 || M1 <= 0 || K <= 0 || nsplit <= 0) {
    set_error("launch_contract_f64: bad sizes M0=%d M1=%d K=%lld nsplit=%d", M0, M1, K, nsplit);
    return LQ_ERR_ARG;
  }
  long long kchunk = (K + nsplit - 1) / nsplit;
  kchunk = ((kchunk + 15) / 16) * 16;
  const int tiles_m = (M0 + 63) / 64, tiles_n = (M1 + 63) / 64;
  const long long blocks = (long long)tiles_m * tiles_n * nsplit;
  if (blocks > 0x7fffffffLL) {
    set_error("launch_contract_f64: grid too large");
    return LQ_ERR_ARG;
  }
  hipLaunchKernelGGL(k_contract_f64, dim3((unsigned)blocks), dim3(256), 0, st, A, lda, M0, B, ldb, M1, K, kchunk,
                     nsplit, tiles_n, work);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_contract_reduce(hipStream_t st, const double2 *work, int nsplit, int M0, int M1, double2 *out, int ldo,
                           int accumulate) {
  const long long total = (long long)M0 * M1;
  hipLaunchKernelGGL(k_contract_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, work, nsplit, M0,
                     M1, out, ldo, accumulate);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

}  // namespace lq
