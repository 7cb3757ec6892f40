// lq_split.hip -- floating-point split-precision forms of the moment
// contraction mu(m0, m1) = sum_k conj(A[m0][k]) B[m1][k]
// (chebyshev::ComputeMomTable, src/chebyshev_solver.cpp:26-42).
//
// Both forms compute the three real products of the three-multiplication
// complex product, P1 = sum Ar Br, P2 = sum Ai Bi, P3 = sum (Ar - Ai)(Br + Bi)
// (Re = P1 + P2, Im = P3 - P1 + P2), from FP64 operands split into pieces of
// lower precision:
//   FP32 hi/lo:  x = h + l, h = fp32(x), l = fp32(x - h); products h h', h l',
//                l h' on v_mfma_f32_16x16x4_f32;
//   3 x BF16:    x = x0 + x1 + x2, x0 = bf16(x), x1 = bf16(x - x0),
//                x2 = bf16(x - x0 - x1); products x_i x'_j for i + j <= 2 on
//                v_mfma_f32_16x16x32_bf16 (products of bf16 pieces are exact
//                in FP32).
// The products accumulate in FP32 inside the MFMA accumulators for kFold
// K-elements, are then added into FP64 registers and restarted; blocks of a
// K chunk write FP64 partial tables that the deterministic split-K reduce of
// the FP64 path sums.  The error is that of the FP32 accumulation (about
// 2^-24 of the chunk sums) plus the dropped piece products.
#include "lq_internal.h"

namespace lq {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f4;

constexpr int kST = 64;     // block tile: rows of A and rows of B
constexpr int kSK = 32;     // K elements per LDS stage
constexpr int kFold = 256;  // K elements accumulated in FP32 before the FP64 fold

__device__ __forceinline__ unsigned short bf16_bits(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite values)
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ double bf16_value(unsigned short b) { return (double)__uint_as_float((unsigned)b << 16); }

// Three bf16 pieces of x (each from the exact remainder in FP64).
__device__ __forceinline__ void split3(double x, unsigned short &p0, unsigned short &p1, unsigned short &p2) {
  p0 = bf16_bits((float)x);
  const double r1 = x - bf16_value(p0);
  p1 = bf16_bits((float)r1);
  const double r2 = r1 - bf16_value(p1);
  p2 = bf16_bits((float)r2);
}

union Frag {
  unsigned short s[8];
  bf16x8 v;
};

// Operand value q of product p from a complex element: p = 0: re,
// p = 1: im, p = 2: re - im (left) or re + im (right).
__device__ __forceinline__ double operand(const double2 x, int p, bool right) {
  return p == 0 ? x.x : p == 1 ? x.y : (right ? x.x + x.y : x.x - x.y);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_contract_split(const double2 *__restrict__ A, long long lda, int M0,
                                                        const double2 *__restrict__ B, long long ldb, int M1,
                                                        long long K, long long kchunk, int nsplit, int tiles_n,
                                                        double2 *__restrict__ work) {
  __shared__ double2 sA[kST][kSK + 1];
  __shared__ double2 sB[kST][kSK + 1];
  const int tile = blockIdx.x / nsplit, chunk = blockIdx.x % nsplit;
  const int row0 = (tile / tiles_n) * kST, col0 = (tile % tiles_n) * kST;
  const long long k0 = (long long)chunk * kchunk;
  const long long k1 = k0 + kchunk < K ? k0 + kchunk : K;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lr = lane & 15, lg = lane >> 4;
  f4 acc[3][2][2];
  double dacc[3][2][2][4];
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        acc[p][i][j] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < 4; ++r) dacc[p][i][j][r] = 0.0;
      }
  int since_fold = 0;
  for (long long kb = k0; kb < k1; kb += kSK) {
    for (int e = threadIdx.x; e < kST * kSK; e += 256) {
      const int r = e / kSK, kk = e % kSK;
      const long long kg = kb + kk;
      double2 va = make_double2(0.0, 0.0), vb = make_double2(0.0, 0.0);
      if (kg < k1) {
        if (row0 + r < M0) va = A[(size_t)(row0 + r) * lda + kg];
        if (col0 + r < M1) vb = B[(size_t)(col0 + r) * ldb + kg];
      }
      sA[r][kk] = va;
      sB[r][kk] = vb;
    }
    __syncthreads();
    if (MODE == LQ_PREC_BF16X3) {
      // one 16x16x32 product covers the stage; lane group lg holds k = lg*8 .. lg*8+7
      // of its row (A) or column (B), the same k set on both sides
#pragma unroll 1
      for (int p = 0; p < 3; ++p) {
        Frag a[3][2], b[3][2];
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
          for (int t = 0; t < 8; ++t) {
            split3(operand(sA[wr + ti * 16 + lr][lg * 8 + t], p, false), a[0][ti].s[t], a[1][ti].s[t], a[2][ti].s[t]);
            split3(operand(sB[wc + ti * 16 + lr][lg * 8 + t], p, true), b[0][ti].s[t], b[1][ti].s[t], b[2][ti].s[t]);
          }
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
          for (int tj = 0; tj < 2; ++tj) {
            f4 c = acc[p][ti][tj];
            // smallest piece products first
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][ti].v, b[0][tj].v, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][ti].v, b[1][tj].v, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][ti].v, b[2][tj].v, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][ti].v, b[0][tj].v, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][ti].v, b[1][tj].v, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][ti].v, b[0][tj].v, c, 0, 0, 0);
            acc[p][ti][tj] = c;
          }
      }
    } else {
      // 16x16x4 products: step s covers k = s*4 .. s*4+3, lane group lg holds k = s*4 + lg
#pragma unroll 1
      for (int s = 0; s < kSK / 4; ++s)
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          float ah[2], al[2], bh[2], bl[2];
#pragma unroll
          for (int ti = 0; ti < 2; ++ti) {
            const double xa = operand(sA[wr + ti * 16 + lr][s * 4 + lg], p, false);
            const double xb = operand(sB[wc + ti * 16 + lr][s * 4 + lg], p, true);
            ah[ti] = (float)xa;
            al[ti] = (float)(xa - (double)ah[ti]);
            bh[ti] = (float)xb;
            bl[ti] = (float)(xb - (double)bh[ti]);
          }
#pragma unroll
          for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) {
              f4 c = acc[p][ti][tj];
              c = __builtin_amdgcn_mfma_f32_16x16x4f32(al[ti], bh[tj], c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[ti], bl[tj], c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[ti], bh[tj], c, 0, 0, 0);
              acc[p][ti][tj] = c;
            }
        }
    }
    __syncthreads();
    since_fold += kSK;
    if (since_fold >= kFold || kb + kSK >= k1) {
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) dacc[p][i][j][r] += (double)acc[p][i][j][r];
            acc[p][i][j] = f4{0.f, 0.f, 0.f, 0.f};
          }
      since_fold = 0;
    }
  }
  // 16x16 result layout: lane holds column lr, rows 4*lg + r
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = row0 + wr + ti * 16 + 4 * lg + r, nn = col0 + wc + tj * 16 + lr;
        if (m < M0 && nn < M1) {
          const double p1 = dacc[0][ti][tj][r], p2 = dacc[1][ti][tj][r], p3 = dacc[2][ti][tj][r];
          work[((size_t)chunk * M0 + m) * M1 + nn] = make_double2(p1 + p2, p3 - p1 + p2);
        }
      }
}

}  // namespace

int split_nsplit(int M0, int M1, long long K) {
  const long long tiles = (long long)((M0 + kST - 1) / kST) * ((M1 + kST - 1) / kST);
  long long ns = (1024 + tiles - 1) / tiles;
  ns = ((ns + 7) / 8) * 8;
  if (ns > 256) ns = 256;
  long long by_k = K / 1024;
  if (by_k < 1) by_k = 1;
  if (ns > by_k) ns = by_k;
  return (int)ns;
}

int launch_contract_split(hipStream_t st, int precision, const double2 *A, long long lda, int M0, const double2 *B,
                          long long ldb, int M1, long long K, double2 *work, int nsplit) {
  if (M0 <= 0 || M1 <= 0 || K <= 0 || nsplit <= 0) {
    set_error("launch_contract_split: bad sizes M0=%d M1=%d K=%lld nsplit=%d", M0, M1, K, nsplit);
    return LQ_ERR_ARG;
  }
  long long kchunk = (K + nsplit - 1) / nsplit;
  kchunk = ((kchunk + kSK - 1) / kSK) * kSK;
  const int tiles_m = (M0 + kST - 1) / kST, tiles_n = (M1 + kST - 1) / kST;
  const long long blocks = (long long)tiles_m * tiles_n * nsplit;
  if (blocks > 0x7fffffffLL) {
    set_error("launch_contract_split: grid too large");
    return LQ_ERR_ARG;
  }
  if (precision == LQ_PREC_BF16X3)
    hipLaunchKernelGGL(k_contract_split<LQ_PREC_BF16X3>, dim3((unsigned)blocks), dim3(256), 0, st, A, lda, M0, B,
                       ldb, M1, K, kchunk, nsplit, tiles_n, work);
  else if (precision == LQ_PREC_FP32X2)
    hipLaunchKernelGGL(k_contract_split<LQ_PREC_FP32X2>, dim3((unsigned)blocks), dim3(256), 0, st, A, lda, M0, B,
                       ldb, M1, K, kchunk, nsplit, tiles_n, work);
  else {
    set_error("launch_contract_split: precision %d is not a floating-point split", precision);
    return LQ_ERR_ARG;
  }
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

}  // namespace lq
