// lq_ozaki.hip -- split-precision moment contraction on INT8 MFMA.
//
// mu(m0, m1) = sum_k conj(A[m0][k]) B[m1][k] with FP64 complex operands,
// computed as exact integer products (Ozaki-style slicing):
//
//  1. k_slice: every row is cut into groups of kGroup K-elements.  For each
//     group, e = 1 + the exponent with max(|re|, |im|) < 2^(e-1), and
//     x / 2^e (exact, |.| < 1/2) is written as S int8 digits rounded to nearest,
//        a_0 = x / 2^e,  d_t = rint(128 a_{t-1}),  a_t = 128 a_{t-1} - d_t,
//     so x = 2^e (sum_t d_t 2^{-7t} + a_S 2^{-7S}), |d_t| <= 64, |a_S| <= 1/2;
//     every step is exact in FP64.  Rounding (not truncation) keeps the
//     remainders' signs independent of x, so rows that are correlated (as the
//     left and right vectors of sigma_xx are) do not add them coherently.
//  2. k_contract_i8: per group, the digit products of level l = t + u
//     (t, u >= 1, l <= S + 1) are summed exactly in int32 on
//     v_mfma_i32_16x16x64_i8 (|sum| <= 2 S 64^2 kGroup < 2^31), then added in
//     FP64 with weight 2^(eA + eB - 7 l).  Dropping the levels above S + 1 and
//     the rounding remainders bounds the error of a complex component by
//     (S + 1) 2^{-7S} / 2 relative to 2^(eA + eB).
//  Re C = Ar Br + Ai Bi and Im C = Ar Bi + Ai (-Br): -Br is formed in
//  registers with a byte-wise negate, so each level needs two accumulators.
//  Digit planes are stored step-major (64 K-elements per step), so the rows
//  a block needs for one step are one contiguous range.
//
// Block tile 64 x 64 (S <= 5) or 32 x 64 (S >= 6, register budget), 8 waves of
// (16 TI) x 16.  A K-step is 64 elements (one MFMA depth); its digits of A and
// B (rows x S x 2 x 64 bytes) are staged in LDS, the next step's prefetched
// into registers.  Split-K chunks are whole groups; each block folds finished
// groups into its own partial table, which the deterministic split-K reduce of
// the FP64 path sums.
#include "lq_internal.h"

namespace lq {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kGroup = 4096;  // K-elements per exponent group
constexpr int kStep = 64;     // K-elements per MFMA
constexpr int kOBT = 64;      // block tile
constexpr int kSliceThreads = 256;

// ---- slicing ----------------------------------------------------------------
// Digit planes, step-major: Q[step][row][plane][64] int8 with step = k / 64,
// plane = 2 t + c (c = 0 real, 1 imag) and rows padded to a multiple of 64
// (zero digits), so one K-step of any block's rows is a single contiguous,
// in-bounds range.  Exponents E[group][padded row] int32.  Elements k >= K get
// digit 0.  One block per (padded row, group).
template <int S>
__global__ __launch_bounds__(kSliceThreads) void k_slice(const double2 *__restrict__ X, long long ldx, long long K,
                                                         int rows, int rows_pad, signed char *__restrict__ Q,
                                                         int *__restrict__ E) {
  const int row = blockIdx.y, grp = blockIdx.x;
  const bool real_row = row < rows;
  const long long g0 = (long long)grp * kGroup;
  constexpr int per = kGroup / kSliceThreads;  // 16 elements per thread
  const long long k0 = g0 + (long long)threadIdx.x * per;
  const double2 *xr = X + (size_t)row * ldx;
  double2 v[per];
  double mx = 0.0;
#pragma unroll
  for (int j = 0; j < per; ++j) {
    const long long k = k0 + j;
    v[j] = (real_row && k < K) ? xr[k] : make_double2(0.0, 0.0);
    mx = fmax(mx, fmax(fabs(v[j].x), fabs(v[j].y)));
  }
  __shared__ double red[kSliceThreads];
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = kSliceThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  int e = 0;
  if (red[0] > 0.0) {
    frexp(red[0], &e);  // red[0] < 2^e
    e += 1;             // |x / 2^e| < 1/2: rounded digits stay within [-64, 64]
  }
  if (threadIdx.x == 0) E[(size_t)grp * rows_pad + row] = e;
  // digits packed four to a register: pr[t][w] holds elements 4w..4w+3
  unsigned pr[S][per / 4], pq[S][per / 4];
#pragma unroll
  for (int t = 0; t < S; ++t)
#pragma unroll
    for (int w = 0; w < per / 4; ++w) pr[t][w] = pq[t][w] = 0u;
#pragma unroll
  for (int j = 0; j < per; ++j) {
    double ar = ldexp(v[j].x, -e), ai = ldexp(v[j].y, -e);
#pragma unroll
    for (int t = 0; t < S; ++t) {
      const double yr = ar * 128.0, yi = ai * 128.0;
      const double tr = rint(yr), tq = rint(yi);
      pr[t][j / 4] |= ((unsigned)(int)tr & 0xffu) << (8 * (j % 4));
      pq[t][j / 4] |= ((unsigned)(int)tq & 0xffu) << (8 * (j % 4));
      ar = yr - tr;
      ai = yi - tq;
    }
  }
  const long long step = k0 / kStep;
  signed char *q = Q + ((size_t)step * rows_pad + row) * (2 * S * kStep) + (k0 % kStep);
#pragma unroll
  for (int t = 0; t < S; ++t) {
    *reinterpret_cast<uint4 *>(q + (2 * t + 0) * kStep) = make_uint4(pr[t][0], pr[t][1], pr[t][2], pr[t][3]);
    *reinterpret_cast<uint4 *>(q + (2 * t + 1) * kStep) = make_uint4(pq[t][0], pq[t][1], pq[t][2], pq[t][3]);
  }
}

// ---- the INT8 GEMM -----------------------------------------------------------
__device__ __forceinline__ v4i mfma_i8(v4i a, v4i b, v4i c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// Byte-wise negation of four int8 digits in [-127, 127]: -x = ~x + 1, with
// the +1 applied to the low seven bits of each byte so no carry crosses bytes.
__device__ __forceinline__ int neg4(int x) {
  const unsigned nx = ~(unsigned)x;
  return (int)(((nx & 0x7f7f7f7fu) + 0x01010101u) ^ (nx & 0x80808080u));
}

__device__ __forceinline__ v4i neg(v4i x) { return v4i{neg4(x.x), neg4(x.y), neg4(x.z), neg4(x.w)}; }

// TI 16-row MFMA tiles per wave: block tile (32 TI) x 64, 8 waves as 2 (m) x 4 (n).
template <int S, int TI>
struct I8Layout {
  static constexpr int kBM = 32 * TI;
  static constexpr int kRB = 2 * S * kStep;   // bytes of one row in one K-step
  static constexpr int kRowBytes = kRB + 16;  // padded LDS row (16 rows hit distinct banks)
  static_assert(kRB % 16 == 0, "rows are whole 16-byte chunks");
  static constexpr int kTileA = kBM * kRowBytes, kTileB = kOBT * kRowBytes;
  static constexpr int kChunksA = kBM * kRB / 16, kChunksB = kOBT * kRB / 16;  // 16-byte chunks
  static constexpr int kPerA = (kChunksA + 511) / 512, kPerB = (kChunksB + 511) / 512;
};

template <int S>
constexpr int tiles_per_wave() {
  return S <= 5 ? 2 : 1;
}

// grid: tiles_m * tiles_n * nsplit blocks of 512 threads; block b handles
// K-chunk b % nsplit (whole groups).  Per K-step: barrier, prefetched digits
// (and, at a group's first step, its exponents) go to LDS, barrier, the next
// step's loads are issued, then this step's MFMAs run.  The only global
// accesses inside the loop are those prefetches, so their latency overlaps the
// MFMAs of a whole step.
template <int S, int TI>
__global__ __launch_bounds__(512) void k_contract_i8(const signed char *__restrict__ QA, const int *__restrict__ EA,
                                                     int M0, const signed char *__restrict__ QB,
                                                     const int *__restrict__ EB, int M1, int ngroups,
                                                     int groups_per_chunk, int nsplit, int tiles_n,
                                                     double2 *__restrict__ work) {
  using Lay = I8Layout<S, TI>;
  __shared__ __attribute__((aligned(16))) unsigned char lds[Lay::kTileA + Lay::kTileB];
  __shared__ int sE[2][Lay::kBM + kOBT];  // exponents of the current / next group (by group parity)
  unsigned char *sA = lds, *sB = lds + Lay::kTileA;
  const int bid = blockIdx.x;
  const int chunk = bid % nsplit;
  const int tile = bid / nsplit;
  const int tm = tile / tiles_n, tn = tile % tiles_n;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int r16 = lane & 15, kq = lane >> 4;
  const int wm = (wave >> 2) * 16 * TI, wn = (wave & 3) * 16;

  const int grp0 = chunk * groups_per_chunk;
  int grp1 = grp0 + groups_per_chunk;
  if (grp1 > ngroups) grp1 = ngroups;
  constexpr int steps_per_group = kGroup / kStep;
  const int st0 = grp0 * steps_per_group;
  const int nsteps = grp1 > grp0 ? (grp1 - grp0) * steps_per_group : 0;

  // Chunk c = tid + 512 j of a step's tile: 16 bytes at c * 16 of the
  // contiguous tile, LDS row c / (kRB / 16) (rows padded by 16 bytes).
  const int M0p = (M0 + kOBT - 1) / kOBT * kOBT, M1p = (M1 + kOBT - 1) / kOBT * kOBT;
  const unsigned voff = (unsigned)tid * 16u;
  v4i pa[Lay::kPerA], pb[Lay::kPerB];
  int pe = 0;
  auto gload = [&](int st) {
    const signed char *ba = QA + ((size_t)st * M0p + (size_t)tm * Lay::kBM) * Lay::kRB;
    const signed char *bb = QB + ((size_t)st * M1p + (size_t)tn * kOBT) * Lay::kRB;
#pragma unroll
    for (int j = 0; j < Lay::kPerA; ++j)
      if (tid + j * 512 < Lay::kChunksA) pa[j] = *reinterpret_cast<const v4i *>(ba + (size_t)j * 8192 + voff);
#pragma unroll
    for (int j = 0; j < Lay::kPerB; ++j)
      if (tid + j * 512 < Lay::kChunksB) pb[j] = *reinterpret_cast<const v4i *>(bb + (size_t)j * 8192 + voff);
    if (st % steps_per_group == 0) {
      const int g = st / steps_per_group;
      if (tid < Lay::kBM) pe = EA[(size_t)g * M0p + tm * Lay::kBM + tid];
      else if (tid < Lay::kBM + kOBT) pe = EB[(size_t)g * M1p + tn * kOBT + tid - Lay::kBM];
    }
  };
  auto lstore = [&](int st) {
#pragma unroll
    for (int j = 0; j < Lay::kPerA; ++j) {
      const int c = tid + j * 512;
      if (c < Lay::kChunksA) *reinterpret_cast<v4i *>(sA + 16 * (c + c / (Lay::kRB / 16))) = pa[j];
    }
#pragma unroll
    for (int j = 0; j < Lay::kPerB; ++j) {
      const int c = tid + j * 512;
      if (c < Lay::kChunksB) *reinterpret_cast<v4i *>(sB + 16 * (c + c / (Lay::kRB / 16))) = pb[j];
    }
    if (st % steps_per_group == 0 && tid < Lay::kBM + kOBT) sE[(st / steps_per_group) & 1][tid] = pe;
  };

  v4i accR[TI][S], accI[TI][S];
  double fR[TI][4], fI[TI][4];
#pragma unroll
  for (int ti = 0; ti < TI; ++ti) {
#pragma unroll
    for (int l = 0; l < S; ++l) accR[ti][l] = accI[ti][l] = v4i{0, 0, 0, 0};
#pragma unroll
    for (int r = 0; r < 4; ++r) fR[ti][r] = fI[ti][r] = 0.0;
  }
  // output element r of lane l in a 16x16 tile: row 4*(l/16) + r, column l % 16

  if (nsteps > 0) gload(st0);
  for (int i = 0; i < nsteps; ++i) {
    const int st = st0 + i;
    __syncthreads();  // previous step's LDS reads are done
    lstore(st);
    __syncthreads();
    if (i + 1 < nsteps) gload(st + 1);
    // digits of B for this wave's column tile (all slices), -Br formed here
    v4i br[S], bi[S], nbr[S];
#pragma unroll
    for (int u = 0; u < S; ++u) {
      const int off = (wn + r16) * Lay::kRowBytes + (u * 2) * kStep + kq * 16;
      br[u] = *reinterpret_cast<const v4i *>(sB + off);
      bi[u] = *reinterpret_cast<const v4i *>(sB + off + kStep);
      nbr[u] = neg(br[u]);
    }
#pragma unroll
    for (int ti = 0; ti < TI; ++ti) {
#pragma unroll
      for (int t = 0; t < S; ++t) {
        const int off = (wm + ti * 16 + r16) * Lay::kRowBytes + (t * 2) * kStep + kq * 16;
        const v4i ar = *reinterpret_cast<const v4i *>(sA + off);
        const v4i ai = *reinterpret_cast<const v4i *>(sA + off + kStep);
#pragma unroll
        for (int u = 0; u + t < S; ++u) {
          accR[ti][t + u] = mfma_i8(ar, br[u], accR[ti][t + u]);
          accI[ti][t + u] = mfma_i8(ar, bi[u], accI[ti][t + u]);
          accR[ti][t + u] = mfma_i8(ai, bi[u], accR[ti][t + u]);
          accI[ti][t + u] = mfma_i8(ai, nbr[u], accI[ti][t + u]);
        }
      }
    }
    // end of an exponent group: fold the exact level sums into FP64
    if ((i + 1) % steps_per_group == 0) {
      const int* se = sE[(st / steps_per_group) & 1];
      const int eb = se[Lay::kBM + wn + r16];
#pragma unroll
      for (int ti = 0; ti < TI; ++ti)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int ea = se[wm + ti * 16 + kq * 4 + r];
          double sr = 0.0, si = 0.0;
#pragma unroll
          for (int l = S - 1; l >= 0; --l) {  // smallest weights first
            sr += ldexp((double)accR[ti][l][r], -7 * (l + 2));
            si += ldexp((double)accI[ti][l][r], -7 * (l + 2));
          }
          fR[ti][r] += ldexp(sr, ea + eb);
          fI[ti][r] += ldexp(si, ea + eb);
        }
#pragma unroll
      for (int ti = 0; ti < TI; ++ti)
#pragma unroll
        for (int l = 0; l < S; ++l) accR[ti][l] = accI[ti][l] = v4i{0, 0, 0, 0};
    }
  }
  const int col = tn * kOBT + wn + r16;
#pragma unroll
  for (int ti = 0; ti < TI; ++ti)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = tm * Lay::kBM + wm + ti * 16 + kq * 4 + r;
      if (row < M0 && col < M1) work[((size_t)chunk * M0 + row) * M1 + col] = make_double2(fR[ti][r], fI[ti][r]);
    }
}

}  // namespace

long long ozaki_kp(long long K) { return ((K + kGroup - 1) / kGroup) * kGroup; }

int pad_rows(int rows) { return (rows + kOBT - 1) / kOBT * kOBT; }

int ozaki_slice_bytes(int rows, long long K, int S, size_t *qbytes, size_t *ebytes) {
  const long long Kp = ozaki_kp(K);
  const long long ng = Kp / kGroup;
  if (qbytes) *qbytes = (size_t)pad_rows(rows) * S * 2 * Kp;
  if (ebytes) *ebytes = (size_t)pad_rows(rows) * ng * sizeof(int);
  return LQ_OK;
}

int launch_slice(hipStream_t st, const double2 *X, long long ldx, int rows, long long K, int S, signed char *Q,
                 int *E) {
  const long long Kp = ozaki_kp(K);
  const int ng = (int)(Kp / kGroup);
  dim3 grid(ng, pad_rows(rows));
  switch (S) {
#define LQ_SL(SV)                                                                                         \
  case SV:                                                                                               \
    hipLaunchKernelGGL((k_slice<SV>), grid, dim3(kSliceThreads), 0, st, X, ldx, K, rows, pad_rows(rows), Q, E); \
    break;
    LQ_SL(4)
    LQ_SL(5)
    LQ_SL(6)
    LQ_SL(7)
    LQ_SL(8)
#undef LQ_SL
    default:
      set_error("launch_slice: %d slices unsupported (4..8)", S);
      return LQ_ERR_ARG;
  }
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int ozaki_nsplit(int M0, int M1, long long K) {
  const long long ng = ozaki_kp(K) / kGroup;
  const long long tiles = (long long)((M0 + 31) / 32) * ((M1 + kOBT - 1) / kOBT);
  long long ns = (512 + tiles - 1) / tiles;
  ns = ((ns + 7) / 8) * 8;
  if (ns > ng) ns = ng;
  if (ns < 1) ns = 1;
  return (int)ns;
}

int launch_contract_i8(hipStream_t st, const signed char *QA, const int *EA, int M0, const signed char *QB,
                       const int *EB, int M1, long long K, int S, double2 *work, int nsplit) {
  const long long Kp = ozaki_kp(K);
  const int ng = (int)(Kp / kGroup);
  const int gpc = (ng + nsplit - 1) / nsplit;
  const int tiles_n = (M1 + kOBT - 1) / kOBT;
  switch (S) {
#define LQ_CI(SV)                                                                                        \
  case SV: {                                                                                            \
    using Lay = I8Layout<SV, tiles_per_wave<SV>()>;                                                      \
    const long long blocks = (long long)((M0 + Lay::kBM - 1) / Lay::kBM) * tiles_n * nsplit;            \
    hipLaunchKernelGGL((k_contract_i8<SV, tiles_per_wave<SV>()>), dim3((unsigned)blocks), dim3(512), 0, st, QA, \
                       EA, M0, QB, EB, M1, ng, gpc, nsplit, tiles_n, work);                                     \
    break;                                                                                              \
  }
    LQ_CI(4)
    LQ_CI(5)
    LQ_CI(6)
    LQ_CI(7)
    LQ_CI(8)
#undef LQ_CI
    default:
      set_error("launch_contract_i8: %d slices unsupported (4..8)", S);
      return LQ_ERR_ARG;
  }
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

}  // namespace lq
