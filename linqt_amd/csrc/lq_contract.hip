// lq_contract.hip -- the moment contraction mu(m0, m1) = sum_k conj(L_m0[k]) R_m1[k].
//
// Replaces chebyshev::ComputeMomTable (src/chebyshev_solver.cpp:26-42), which
// runs one cblas_zdotc per (m0, m1) pair.  Here the stored vectors of all
// states of a group form two K-contiguous matrices (K = n * states), and the
// table is one complex GEMM  C = A^H B  on FP64 MFMA (v_mfma_f64_16x16x4_f64):
//   Re C = Ar Br + Ai Bi,   Im C = Ar Bi - Ai Br.
//
// Kernels: lq_contract_f64_3m_t128 (the default for tables of at least
// 128 x 128: three real products per complex product, 128 x 128 block tile,
// operands copied from global memory straight into LDS buffers; built as a
// code object of its own from modules/lq_contract_t128.hip and loaded here),
// k_contract_f64_3m_dl (the same with a 128 x 64 tile, LQ_CONTRACT=dl),
// k_contract_f64_3m_w1 (its one-wave-per-SIMD form, LQ_CONTRACT=w1),
// k_contract_f64_3m (the 128 x 64 products staged through registers,
// LQ_CONTRACT_DL=0), k_contract_f64_lds (the four-product form with a
// 128 x 128 tile, LQ_CONTRACT=4m) and k_contract_f64 (64 x 64 tile, operands
// straight from global memory, for small tables).  The table is a
// GEMM with a tiny output and a huge K: the LDS kernels read each operand tile
// once per block, K is split into nsplit chunks, and the blocks of one chunk
// run together on one XCD (block_chunk_tile).  A deterministic second pass
// sums the chunks.
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

#include "lq_contract_dev.h"
#include "lq_internal.h"

namespace lq {

namespace {


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
                                                      long long K, long long kchunk, int nsplit, int tiles_n, int xmap,
                                                      double2 *__restrict__ work) {
  int chunk, tile;
  block_chunk_tile(blockIdx.x, nsplit, xmap, chunk, tile);
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

  // Main loop over pairs of 16-element steps with the next step's operands
  // loaded while the current step's MFMAs run (two register sets).
  auto load = [&](double2(&a)[2][4], double2(&b)[2][4], long long kbase) {
    const long long kk = kbase + ks * 4;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        a[t][s] = pa[t][kk + s];
        b[t][s] = pb[t][kk + s];
      }
  };
  long long kb = k0;
  const long long kmain = k0 + ((k1 - k0) / 32) * 32;
  if (kb < kmain) {
    double2 a0[2][4], b0[2][4], a1[2][4], b1[2][4];
    load(a0, b0, kb);
    for (; kb < kmain; kb += 32) {
      load(a1, b1, kb + 16);
      mma_step(a0, b0, cr, ci);
      if (kb + 32 < kmain) load(a0, b0, kb + 32);
      mma_step(a1, b1, cr, ci);
    }
  }
  for (; kb + 16 <= k1; kb += 16) {
    double2 a[2][4], b[2][4];
    load(a, b, kb);
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

  // D layout of v_mfma_f64_16x16x4: element r of lane l is D[l/16 + 4*r][l % 16].
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m_base + ti * 16 + ks + 4 * r;
        const int nn = n_base + tj * 16 + r16;
        if (m < M0 && nn < M1) work[((size_t)chunk * M0 + m) * M1 + nn] = make_double2(cr[ti][tj][r], ci[ti][tj][r]);
      }
}

// LDS-tiled variant: 512 threads, block tile 128 x 128, waves 2 (m) x 4 (n)
// of 64 x 32 (4 x 2 MFMA tiles).  K advances in stages of 16 complex
// elements; a stage of A and B (128 rows each) is loaded once from global
// memory by the whole block into one of two LDS buffers while the other
// buffer feeds the MFMAs.  Rows are padded to 17 elements so the 16 rows read
// by one MFMA fragment land in distinct LDS banks.
constexpr int kBT = 128;       // block tile (rows of A and of B)
constexpr int kLdsRow = kKC + 1;

__global__ __launch_bounds__(512) void k_contract_f64_lds(const double2 *__restrict__ A, long long lda, int M0,
                                                          const double2 *__restrict__ B, long long ldb, int M1,
                                                          long long K, long long kchunk, int nsplit, int tiles_n, int xmap,
                                                          double2 *__restrict__ work) {
  __shared__ double2 sA[2][kBT * kLdsRow];
  __shared__ double2 sB[2][kBT * kLdsRow];
  int chunk, tile;
  block_chunk_tile(blockIdx.x, nsplit, xmap, chunk, tile);
  const int tm = tile / tiles_n, tn = tile % tiles_n;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int r16 = lane & 15, ks = lane >> 4;
  const int wm = (wave >> 2) * 64, wn = (wave & 3) * 32;

  // global -> LDS assignment: thread t moves row t/4, elements 4*(t%4)..+3
  const int lrow = tid >> 2, le = (tid & 3) * 4;
  int ga = tm * kBT + lrow;
  if (ga >= M0) ga = M0 - 1;  // clamped rows feed outputs that are never stored
  int gb = tn * kBT + lrow;
  if (gb >= M1) gb = M1 - 1;
  const double2 *pa = A + (size_t)ga * lda;
  const double2 *pb = B + (size_t)gb * ldb;

  const long long k0 = (long long)chunk * kchunk;
  const long long k1 = (k0 + kchunk < K) ? k0 + kchunk : K;
  const int nstage = k1 > k0 ? (int)((k1 - k0 + kKC - 1) / kKC) : 0;

  double2 ra[4], rb[4];
  auto gload = [&](long long kb) {
    if (kb + kKC <= k1) {  // whole stage in range (uniform over the block)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        ra[s] = pa[kb + le + s];
        rb[s] = pb[kb + le + s];
      }
      return;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const long long kk = kb + le + s;
      const bool ok = kk < k1;
      ra[s] = ok ? pa[kk] : make_double2(0.0, 0.0);
      rb[s] = ok ? pb[kk] : make_double2(0.0, 0.0);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      sA[buf][lrow * kLdsRow + le + s] = ra[s];
      sB[buf][lrow * kLdsRow + le + s] = rb[s];
    }
  };

  v4d cr[4][2], ci[4][2];
#pragma unroll
  for (int ti = 0; ti < 4; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
      cr[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};
      ci[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};
    }

  if (nstage > 0) {
    gload(k0);
    lstore(0);
  }
  __syncthreads();
  for (int st = 0; st < nstage; ++st) {
    const int cur = st & 1;
    const bool more = st + 1 < nstage;
    if (more) gload(k0 + (long long)(st + 1) * kKC);
    // lane (r16, ks) feeds k-slot ks of four MFMAs with elements 4*ks + s
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      double2 fa[4], fb[2];
#pragma unroll
      for (int ti = 0; ti < 4; ++ti) fa[ti] = sA[cur][(wm + ti * 16 + r16) * kLdsRow + ks * 4 + s];
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) fb[tj] = sB[cur][(wn + tj * 16 + r16) * kLdsRow + ks * 4 + s];
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) {
          cr[ti][tj] = mfma(fa[ti].x, fb[tj].x, cr[ti][tj]);
          cr[ti][tj] = mfma(fa[ti].y, fb[tj].y, cr[ti][tj]);
          ci[ti][tj] = mfma(fa[ti].x, fb[tj].y, ci[ti][tj]);
          ci[ti][tj] = mfma(-fa[ti].y, fb[tj].x, ci[ti][tj]);
        }
    }
    if (more) lstore(cur ^ 1);
    __syncthreads();
  }

  // D layout of v_mfma_f64_16x16x4: element r of lane l is D[l/16 + 4*r][l % 16].
#pragma unroll
  for (int ti = 0; ti < 4; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = tm * kBT + wm + ti * 16 + ks + 4 * r;
        const int nn = tn * kBT + wn + tj * 16 + r16;
        if (m < M0 && nn < M1) work[((size_t)chunk * M0 + m) * M1 + nn] = make_double2(cr[ti][tj][r], ci[ti][tj][r]);
      }
}

// Three-multiplication variant (the default for tables of at least 128 x 128):
// with P1 = sum Ar Br, P2 = sum Ai Bi and P3 = sum (Ar - Ai)(Br + Bi),
//   Re C = P1 + P2,   Im C = P3 - P1 + P2,
// three real MFMAs per complex product instead of four.  512 threads (two
// waves per SIMD), block tile 128 (A rows) x 64 (B rows), waves 4 (m) x 2 (n)
// of 32 x 32 (2 x 2 MFMA tiles, 3 x 4 accumulators = 96 registers), K in
// stages of 16 complex elements double-buffered in LDS (rows padded to 17).
// The sums Ar - Ai and Br + Bi are formed when a fragment is read.  Rounding:
// Im carries the error of P3 (operands rounded once more) on top of the
// four-multiplication form, a few units in the last place of sum |A||B|.
constexpr int kBM3 = 128, kBN3 = 64;

__device__ __forceinline__ void mma3_stage(const double2 *__restrict__ sA, const double2 *__restrict__ sB, int arow0,
                                           int brow0, int ks, v4d (&p1)[2][2], v4d (&p2)[2][2], v4d (&p3)[2][2]) {
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    double ar[2], ai[2], ad[2], br[2], bi[2], bs[2];
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
      const double2 x = sA[(arow0 + ti * 16) * kLdsRow + ks * 4 + s];
      ar[ti] = x.x;
      ai[ti] = x.y;
      ad[ti] = x.x - x.y;
    }
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
      const double2 y = sB[(brow0 + tj * 16) * kLdsRow + ks * 4 + s];
      br[tj] = y.x;
      bi[tj] = y.y;
      bs[tj] = y.x + y.y;
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) {
        p1[ti][tj] = mfma(ar[ti], br[tj], p1[ti][tj]);
        p2[ti][tj] = mfma(ai[ti], bi[tj], p2[ti][tj]);
        p3[ti][tj] = mfma(ad[ti], bs[tj], p3[ti][tj]);
      }
  }
}

__global__ __launch_bounds__(512) void k_contract_f64_3m(const double2 *__restrict__ A, long long lda, int M0,
                                                         const double2 *__restrict__ B, long long ldb, int M1,
                                                         long long K, long long kchunk, int nsplit, int tiles_n, int xmap,
                                                         double2 *__restrict__ work) {
  __shared__ double2 sA[2][kBM3 * kLdsRow];
  __shared__ double2 sB[2][kBN3 * kLdsRow];
  int chunk, tile;
  block_chunk_tile(blockIdx.x, nsplit, xmap, chunk, tile);
  const int tm = tile / tiles_n, tn = tile % tiles_n;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int r16 = lane & 15, ks = lane >> 4;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;

  // global -> LDS: A row t/4 elements 4(t%4)..+3, B row t/8 elements 2(t%8)..+1
  const int arow = tid >> 2, ae = (tid & 3) * 4;
  const int brow = tid >> 3, be = (tid & 7) * 2;
  int ga = tm * kBM3 + arow;
  if (ga >= M0) ga = M0 - 1;  // clamped rows feed outputs that are never stored
  int gb = tn * kBN3 + brow;
  if (gb >= M1) gb = M1 - 1;
  const double2 *pa = A + (size_t)ga * lda;
  const double2 *pb = B + (size_t)gb * ldb;

  const long long k0 = (long long)chunk * kchunk;
  const long long k1 = (k0 + kchunk < K) ? k0 + kchunk : K;
  const int nstage = k1 > k0 ? (int)((k1 - k0 + kKC - 1) / kKC) : 0;

  double2 ra[4], rb[2];
  auto gload = [&](long long kb) {
    if (kb + kKC <= k1) {  // whole stage in range (uniform over the block)
#pragma unroll
      for (int s = 0; s < 4; ++s) ra[s] = pa[kb + ae + s];
#pragma unroll
      for (int s = 0; s < 2; ++s) rb[s] = pb[kb + be + s];
      return;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const long long kk = kb + ae + s;
      ra[s] = kk < k1 ? pa[kk] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const long long kk = kb + be + s;
      rb[s] = kk < k1 ? pb[kk] : make_double2(0.0, 0.0);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int s = 0; s < 4; ++s) sA[buf][arow * kLdsRow + ae + s] = ra[s];
#pragma unroll
    for (int s = 0; s < 2; ++s) sB[buf][brow * kLdsRow + be + s] = rb[s];
  };

  v4d p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) p1[ti][tj] = p2[ti][tj] = p3[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};

  if (nstage > 0) {
    gload(k0);
    lstore(0);
  }
  __syncthreads();
  for (int st = 0; st < nstage; ++st) {
    const int cur = st & 1;
    const bool more = st + 1 < nstage;
    if (more) gload(k0 + (long long)(st + 1) * kKC);
    mma3_stage(sA[cur], sB[cur], wm + r16, wn + r16, ks, p1, p2, p3);
    if (more) lstore(cur ^ 1);
    __syncthreads();
  }

  // D layout of v_mfma_f64_16x16x4: element r of lane l is D[l/16 + 4*r][l % 16].
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = tm * kBM3 + wm + ti * 16 + ks + 4 * r;
        const int nn = tn * kBN3 + wn + tj * 16 + r16;
        if (m < M0 && nn < M1) {
          const double a1 = p1[ti][tj][r], a2 = p2[ti][tj][r];
          work[((size_t)chunk * M0 + m) * M1 + nn] = make_double2(a1 + a2, p3[ti][tj][r] - a1 + a2);
        }
      }
}

// Direct-load variant of the three-multiplication kernel (same tiles, waves
// and MFMA sequence): operands go from global memory straight into LDS with
// global_load_lds_dwordx4 (lane l of a load writes 16 bytes at base + 16 l),
// so no wave stages them through registers or stores them itself.  NBUF LDS
// buffers (2 by default): the loads of stage st + NBUF - 1 are issued when
// stage st starts, and one counter wait plus one barrier per stage make stage
// st + 1 visible.  These loads are fetched from beyond L2 about twice as often
// as the register-staged kernel's (measured; the cache-policy bits did not
// change that), yet the kernel is faster.  At 81-84 % of the FP64 peak it
// loses most of the rest to that per-stage wait and barrier (measured by
// removing each in turn), which the 128 x 128 tile of lq_contract_f64_3m_t128
// halves per flop.  LDS rows are 16 unpadded slots; element k of row R sits in slot
// k ^ (R & 15), so the 16 rows read by one MFMA operand fetch hit 16
// different slots and a 16-lane load group still covers one row's 256
// contiguous bytes.  A partial last stage is loaded through registers with
// zeros past K into the same slots.
constexpr int kDlA = kBM3 * kKC, kDlB = kBN3 * kKC;  // slots per buffer


__device__ __forceinline__ void mma3_stage_dl(const double2 *sA, const double2 *sB, int arow0, int brow0, int ks,
                                              v4d (&p1)[2][2], v4d (&p2)[2][2], v4d (&p3)[2][2]) {
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    double ar[2], ai[2], ad[2], br[2], bi[2], bs[2];
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
      const double2 x = sA[dl_slot(arow0 + ti * 16, ks * 4 + s)];
      ar[ti] = x.x;
      ai[ti] = x.y;
      ad[ti] = x.x - x.y;
    }
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
      const double2 y = sB[dl_slot(brow0 + tj * 16, ks * 4 + s)];
      br[tj] = y.x;
      bi[tj] = y.y;
      bs[tj] = y.x + y.y;
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) {
        p1[ti][tj] = mfma(ar[ti], br[tj], p1[ti][tj]);
        p2[ti][tj] = mfma(ai[ti], bi[tj], p2[ti][tj]);
        p3[ti][tj] = mfma(ad[ti], bs[tj], p3[ti][tj]);
      }
  }
}


// One global_load_lds_dwordx4: lane l copies 16 bytes from g (its own
// address) to LDS byte address lds + 16 l (lds uniform over the wave).
// Issued as inline assembly so the compiler does not add a wait for it before
// unrelated LDS reads; its completion is awaited explicitly (s_waitcnt vmcnt).
__device__ __forceinline__ void lds_dma16(const double2 *g, double2 *lds) {
  const unsigned a = (unsigned)(size_t)(lds_void_t *)lds;
  asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off"
               :
               : "s"(__builtin_amdgcn_readfirstlane(a)), "v"(g)
               : "memory", "m0");
}

template <int NBUF>
__global__ __launch_bounds__(512) void k_contract_f64_3m_dl(const double2 *__restrict__ A, long long lda, int M0,
                                                            const double2 *__restrict__ B, long long ldb, int M1,
                                                            long long K, long long kchunk, int nsplit, int tiles_n,
                                                            int xmap, double2 *__restrict__ work) {
  // three separate arrays (not one indexed by stage % 3) so the compiler can
  // tell the buffer being read from the ones the loads in flight write
  static_assert(NBUF == 2 || NBUF == 3, "two or three LDS buffers");
  __shared__ double2 sA0[kDlA], sA1[kDlA], sA2[NBUF == 3 ? kDlA : 1];
  __shared__ double2 sB0[kDlB], sB1[kDlB], sB2[NBUF == 3 ? kDlB : 1];
  int chunk, tile;
  block_chunk_tile(blockIdx.x, nsplit, xmap, chunk, tile);
  const int tm = tile / tiles_n, tn = tile % tiles_n;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int r16 = lane & 15, ks = lane >> 4;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;

  // Load q of this wave covers slots 64 c .. 64 c + 63 of a buffer (c = 4 wave + q
  // for A, 2 wave + q for B): lane l fills slot 64 c + l = row R = 4 c + l / 16,
  // element (l % 16) ^ (R & 15).
  const int lrow = lane >> 4, lk = lane & 15;
  const double2 *pa[4];
  const double2 *pb[2];
  int ka[4], kb2[2];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int R = 4 * (4 * wave + q) + lrow;
    int g = tm * kBM3 + R;
    if (g >= M0) g = M0 - 1;  // clamped rows feed outputs that are never stored
    ka[q] = lk ^ (R & 15);
    pa[q] = A + (size_t)g * lda + ka[q];
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int R = 4 * (2 * wave + q) + lrow;
    int g = tn * kBN3 + R;
    if (g >= M1) g = M1 - 1;
    kb2[q] = lk ^ (R & 15);
    pb[q] = B + (size_t)g * ldb + kb2[q];
  }

  const long long k0 = (long long)chunk * kchunk;
  const long long k1 = (k0 + kchunk < K) ? k0 + kchunk : K;
  const int nstage = k1 > k0 ? (int)((k1 - k0 + kKC - 1) / kKC) : 0;

  // Issue the loads of stage j into (sA, sB).  Whole stage in range (uniform
  // over the block): six direct loads per lane.  Partial stage: through
  // registers, zeros past k1.
  auto issue = [&](int j, double2 *sA, double2 *sB) {
    const long long kb = k0 + (long long)j * kKC;
    if (kb + kKC <= k1) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        lds_dma16(pa[q] + kb, sA + 64 * (4 * wave + q));
#pragma unroll
      for (int q = 0; q < 2; ++q)
        lds_dma16(pb[q] + kb, sB + 64 * (2 * wave + q));
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        sA[64 * (4 * wave + q) + lane] = kb + ka[q] < k1 ? pa[q][kb] : make_double2(0.0, 0.0);
#pragma unroll
      for (int q = 0; q < 2; ++q)
        sB[64 * (2 * wave + q) + lane] = kb + kb2[q] < k1 ? pb[q][kb] : make_double2(0.0, 0.0);
    }
  };
  v4d p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) p1[ti][tj] = p2[ti][tj] = p3[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};

  // After computing stage j: this wave's loads of stage j + 1 complete (with
  // three buffers the six of stage j + 2 may still be in flight), then a
  // barrier without a memory fence (a fence would also wait for stage j + 2).
  // Every LDS read of stage j has returned (lgkmcnt) before the barrier, so its
  // buffer is free again.
  auto settle = [&](int j) {
    if (NBUF == 3 && j + 2 < nstage)
      asm volatile("s_waitcnt vmcnt(6) lgkmcnt(0)\n\ts_barrier" ::: "memory");
    else
      asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
  };
  // Stage j from (cA, cB); first the loads of stage j + NBUF - 1 into (nA, nB).
  auto step = [&](int j, const double2 *cA, const double2 *cB, double2 *nA, double2 *nB) {
    if (j + NBUF - 1 < nstage) issue(j + NBUF - 1, nA, nB);
    mma3_stage_dl(cA, cB, wm + r16, wn + r16, ks, p1, p2, p3);
    settle(j);
  };

  if (nstage > 0) issue(0, sA0, sB0);
  if (NBUF == 3 && nstage > 1) issue(1, sA1, sB1);
  settle(-1);
  if (NBUF == 3) {
    for (int st = 0; st < nstage; st += 3) {
      step(st, sA0, sB0, sA2, sB2);
      if (st + 1 < nstage) step(st + 1, sA1, sB1, sA0, sB0);
      if (st + 2 < nstage) step(st + 2, sA2, sB2, sA1, sB1);
    }
  } else {
    for (int st = 0; st < nstage; st += 2) {
      step(st, sA0, sB0, sA1, sB1);
      if (st + 1 < nstage) step(st + 1, sA1, sB1, sA0, sB0);
    }
  }

  // D layout of v_mfma_f64_16x16x4: element r of lane l is D[l/16 + 4*r][l % 16].
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = tm * kBM3 + wm + ti * 16 + ks + 4 * r;
        const int nn = tn * kBN3 + wn + tj * 16 + r16;
        if (m < M0 && nn < M1) {
          const double a1 = p1[ti][tj][r], a2 = p2[ti][tj][r];
          work[((size_t)chunk * M0 + m) * M1 + nn] = make_double2(a1 + a2, p3[ti][tj][r] - a1 + a2);
        }
      }
}

// One wave per SIMD: the same block tile (128 x 64), LDS layout, direct loads
// and MFMA sequence per output element as k_contract_f64_3m_dl, with four
// waves of 64 x 32 (ti = 0..3, tj = 0..1) instead of eight of 32 x 32.  Its
// 192 accumulator registers live in AGPRs (amdgpu_num_vgpr(192) caps both
// register files at 192), so a block holds 3/4 of a SIMD lane's registers and
// leaves room for a recursion wave of up to 128 registers on every SIMD.
// Each output element sees the same products in the same K order as in
// k_contract_f64_3m_dl: the tables are bitwise equal.  Measured 72 % of the
// FP64 peak alone (one wave cannot hide its own LDS latency), and a recursion
// beside it advances at about a quarter of its speed, so it is a schedule
// option (LQ_CONTRACT=w1), not the default.
__global__ __launch_bounds__(256) __attribute__((amdgpu_num_vgpr(192))) void k_contract_f64_3m_w1(const double2 *__restrict__ A, long long lda, int M0,
                                                            const double2 *__restrict__ B, long long ldb, int M1,
                                                            long long K, long long kchunk, int nsplit, int tiles_n,
                                                            int xmap, double2 *__restrict__ work) {
  // buffer b: A slots at sbuf + b (kDlA + kDlB), then the B slots
  __shared__ double2 sbuf[2 * (kDlA + kDlB)];
  int chunk, tile;
  block_chunk_tile(blockIdx.x, nsplit, xmap, chunk, tile);
  const int tm = tile / tiles_n, tn = tile % tiles_n;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int r16 = lane & 15, ks = lane >> 4;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 32;

  // Load q of this wave covers slots 64 c .. 64 c + 63 of a buffer (c = 8 wave + q
  // for A, 4 wave + q for B): lane l fills slot 64 c + l = row R = 4 c + l / 16,
  // element (l % 16) ^ (R & 15).
  const int lrow = lane >> 4, lk = lane & 15;
  const double2 *pa[8];
  const double2 *pb[4];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int R = 4 * (8 * wave + q) + lrow;
    int g = tm * kBM3 + R;
    if (g >= M0) g = M0 - 1;  // clamped rows feed outputs that are never stored
    pa[q] = A + (size_t)g * lda + (lk ^ (R & 15));
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int R = 4 * (4 * wave + q) + lrow;
    int g = tn * kBN3 + R;
    if (g >= M1) g = M1 - 1;
    pb[q] = B + (size_t)g * ldb + (lk ^ (R & 15));
  }

  const long long k0 = (long long)chunk * kchunk;
  const long long len = (k0 + kchunk < K ? k0 + kchunk : K) - k0;
  const int nfull = len > 0 ? (int)(len / kKC) : 0;              // whole stages
  const int nstage = len > 0 ? (int)((len + kKC - 1) / kKC) : 0;  // nfull or nfull + 1
  auto abuf = [&](int j) { return sbuf + (j & 1) * (kDlA + kDlB); };

  // A whole stage: twelve direct loads per lane.
  auto issue_full = [&](int j) {
    const long long kb = k0 + (long long)j * kKC;
    double2 *sA = abuf(j), *sB = sA + kDlA;
#pragma unroll
    for (int q = 0; q < 8; ++q) lds_dma16(pa[q] + kb, sA + 64 * (8 * wave + q));
#pragma unroll
    for (int q = 0; q < 4; ++q) lds_dma16(pb[q] + kb, sB + 64 * (4 * wave + q));
  };
  // The partial last stage: through registers, zeros past the chunk end (the
  // element of a slot is its pointer's offset from the row start, lk ^ (R & 15)).
  auto issue_part = [&](int j) {
    const long long kb = (long long)j * kKC;
    double2 *sA = abuf(j), *sB = sA + kDlA;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int R = 4 * (8 * wave + q) + lrow;
      sA[64 * (8 * wave + q) + lane] = kb + (lk ^ (R & 15)) < len ? pa[q][k0 + kb] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int R = 4 * (4 * wave + q) + lrow;
      sB[64 * (4 * wave + q) + lane] = kb + (lk ^ (R & 15)) < len ? pb[q][k0 + kb] : make_double2(0.0, 0.0);
    }
  };
  v4d p1[4][2], p2[4][2], p3[4][2];
#pragma unroll
  for (int ti = 0; ti < 4; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
      p1[ti][tj] = p2[ti][tj] = p3[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};
      asm("" : "+a"(p1[ti][tj]), "+a"(p2[ti][tj]), "+a"(p3[ti][tj]));
    }

  // This wave's loads of the next stage complete and every LDS read of the
  // current one has returned, then a barrier without a memory fence.
  auto settle = [&]() { asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory"); };
  // The empty asm statements keep each accumulator in AGPRs from one stage to
  // the next (otherwise the compiler carries them in VGPRs and copies all 192
  // registers in and out of AGPRs every stage).
  auto compute = [&](int j) {
    const double2 *sA = abuf(j);
    mma3_stage_64x32(sA, sA + kDlA, wm + r16, wn + r16, ks, p1, p2, p3);
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) asm("" : "+a"(p1[ti][tj]), "+a"(p2[ti][tj]), "+a"(p3[ti][tj]));
  };
  if (nfull > 0)
    issue_full(0);
  else if (nstage > 0)
    issue_part(0);
  settle();
  // The main loop body is one basic block (the next stage is whole): the
  // accumulators stay in their AGPRs across iterations, no copies at joins.
  int st = 0;
  for (; st + 1 < nfull; ++st) {
    issue_full(st + 1);
    compute(st);
    settle();
  }
  if (st < nstage) {  // the last whole stage, then the partial one if any
    if (st + 1 < nstage) issue_part(st + 1);
    compute(st);
    settle();
    if (st + 1 < nstage) {
      compute(st + 1);
      settle();
    }
  }

#pragma unroll
  for (int ti = 0; ti < 4; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = tm * kBM3 + wm + ti * 16 + ks + 4 * r;
        const int nn = tn * kBN3 + wn + tj * 16 + r16;
        if (m < M0 && nn < M1) {
          const double a1 = p1[ti][tj][r], a2 = p2[ti][tj][r];
          work[((size_t)chunk * M0 + m) * M1 + nn] = make_double2(a1 + a2, p3[ti][tj][r] - a1 + a2);
        }
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

// The same sum for a table whose columns hold nout operators' blocks of bc
// columns side by side: column o*bc + s goes to out_o[m*ldo + s].
__global__ void k_contract_reduce_cols(const double2 *__restrict__ work, int nsplit, int M0, int M1, int bc,
                                       double2 *__restrict__ out0, double2 *__restrict__ out1, int ldo,
                                       int accumulate) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)M0 * M1) return;
  const int m = (int)(t / M1), nn = (int)(t % M1);
  double sx = 0.0, sy = 0.0;
  for (int c = 0; c < nsplit; ++c) {
    const double2 w = work[((size_t)c * M0 + m) * M1 + nn];
    sx += w.x;
    sy += w.y;
  }
  const int o = nn / bc, sc = nn % bc;
  double2 *o_ = (o == 0 ? out0 : out1) + (size_t)m * ldo + sc;
  if (accumulate) {
    sx += o_->x;
    sy += o_->y;
  }
  *o_ = make_double2(sx, sy);
}

// Kernel choice for tables of at least 128 x 128: the three-multiplication
// kernel (LQ_CONTRACT=4m selects the four-multiplication LDS kernel); smaller
// tables, or LQ_CONTRACT=wave, use the register-only 64 x 64 kernel.
enum ContractKind { kWave = 0, kLds4m = 1, kLds3m = 2, kLds3m128 = 3 };

// Kernel choice for tables of at least 128 x 128: the 128 x 128 three-product
// kernel (module lq_contract_f64_3m_t128).  LQ_CONTRACT=dl selects the
// 128 x 64 direct-load kernel (also selected by LQ_CONTRACT_DL=0, which takes
// its register-staged variant, and by LQ_CONTRACT_DLBUF=3), =w1 its one-wave-
// per-SIMD variant, =4m the four-product kernel; =wave, or a smaller table,
// the register-only 64 x 64 kernel.
int contract_kind(int M0, int M1) {
  static const int forced = [] {
    const char *e = getenv("LQ_CONTRACT");
    if (e && strcmp(e, "wave") == 0) return (int)kWave;
    if (e && strcmp(e, "4m") == 0) return (int)kLds4m;
    if (e && (strcmp(e, "dl") == 0 || strcmp(e, "w1") == 0)) return (int)kLds3m;
    if (e && strcmp(e, "t128") == 0) return (int)kLds3m128;
    if (getenv("LQ_CONTRACT_DL") || getenv("LQ_CONTRACT_DLBUF")) return (int)kLds3m;
    return -1;
  }();
  if (M0 < kBT || M1 < kBT) return kWave;
  return forced >= 0 ? forced : kLds3m128;
}

bool use_lds(int M0, int M1) { return contract_kind(M0, M1) != kWave; }

// The 128 x 64 three-multiplication kernels: direct global -> LDS loads
// (k_contract_f64_3m_dl) unless LQ_CONTRACT_DL=0 selects the register-staged
// k_contract_f64_3m.  Both compute the same products in the same order.
bool contract_dl() {
  static const bool on = [] {
    const char *e = getenv("LQ_CONTRACT_DL");
    return !(e && e[0] == '0');
  }();
  return on;
}

// LDS buffers of the direct-load kernel: 2 (loads one stage ahead, 96 KB) or
// 3 (LQ_CONTRACT_DLBUF=3: two stages ahead, 144 KB; measured 1-2 % slower).
int contract_dlbuf() {
  static const int nb = [] {
    const char *e = getenv("LQ_CONTRACT_DLBUF");
    return e && e[0] == '3' ? 3 : 2;
  }();
  return nb;
}

// One-wave-per-SIMD three-product kernel (LQ_CONTRACT=w1): the same tables
// as k_contract_f64_3m_dl, in 384 registers per SIMD lane, so a recursion
// wave of up to 128 registers fits beside it on every SIMD.
bool contract_w1() {
  static const bool on = [] {
    const char *e = getenv("LQ_CONTRACT");
    return e && strcmp(e, "w1") == 0;
  }();
  return on;
}

// XCD-major block order (block_chunk_tile); LQ_CONTRACT_XMAP=0 restores
// chunk = block % nsplit.
bool contract_xmap() {
  static const bool on = [] {
    const char *e = getenv("LQ_CONTRACT_XMAP");
    return !(e && e[0] == '0');
  }();
  return on;
}

// The 128 x 128 kernel is a code object of its own (its register split is
// set in its LLVM IR at build time), embedded in this library and loaded once
// per device.
extern "C" const unsigned char lq_module_contract_t128[];

hipFunction_t t128_function(int device) {
  static std::mutex mu;
  static std::map<int, hipFunction_t> fns;
  std::lock_guard<std::mutex> lock(mu);
  auto it = fns.find(device);
  if (it != fns.end()) return it->second;
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  if (hipModuleLoadData(&mod, lq_module_contract_t128) != hipSuccess ||
      hipModuleGetFunction(&fn, mod, "lq_contract_f64_3m_t128") != hipSuccess)
    return nullptr;
  fns[device] = fn;
  return fn;
}

// Its direct loads address each lane's row as an unsigned 32-bit byte offset
// of at most three rows plus 240 bytes from a scalar base, so rows longer than
// lim elements take the 128 x 64 kernel (same tables).  LQ_CONTRACT_T128_MAXLD
// lowers the limit (to test that path at small sizes).
bool t128_fits(long long lda, long long ldb) {
  static const long long lim = [] {
    long long l = ((1LL << 32) - 256) / (3 * (long long)sizeof(double2));
    const char *e = getenv("LQ_CONTRACT_T128_MAXLD");
    if (e && atoll(e) > 0 && atoll(e) < l) l = atoll(e);
    return l;
  }();
  return lda <= lim && ldb <= lim;
}

// Split-K chunks.  The 128 x 128 kernel takes the chunks of the 128 x 64
// kernels (its time does not depend on them: 46.2-46.3 ms at 16, 32 or 64
// chunks for 512 x 512 x 2^21), so every three-product kernel sums the same
// chunks in the same order and gives the same tables.
int choose_nsplit(int M0, int M1, long long K) {
  const int kind = contract_kind(M0, M1);
  const int bm = kind == kWave ? 64 : kBT, bn = (kind == kLds3m || kind == kLds3m128) ? kBN3 : bm;
  const int target = kind == kWave ? 2048 : 512;  // blocks in flight
  const long long tiles = (long long)((M0 + bm - 1) / bm) * ((M1 + bn - 1) / bn);
  long long ns = (target + tiles - 1) / tiles;
  ns = ((ns + 7) / 8) * 8;  // whole XCD groups
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
  const int kind = contract_kind(M0, M1);
  const int bm = kind == kWave ? 64 : kBT, bn = kind == kLds3m ? kBN3 : bm;
  long long kchunk = (K + nsplit - 1) / nsplit;
  kchunk = ((kchunk + 15) / 16) * 16;
  const int tiles_m = (M0 + bm - 1) / bm, tiles_n = (M1 + bn - 1) / bn;
  const long long blocks = (long long)tiles_m * tiles_n * nsplit;
  if (blocks > 0x7fffffffLL) {
    set_error("launch_contract_f64: grid too large");
    return LQ_ERR_ARG;
  }
  const int xmap = (nsplit % 8 == 0 && contract_xmap()) ? 1 : 0;
  if (kind == kLds3m128 && t128_fits(lda, ldb)) {
    int dev = 0;
    LQ_HIP_CHECK(hipGetDevice(&dev));
    hipFunction_t fn = t128_function(dev);
    if (!fn) {
      set_error("launch_contract_f64: cannot load the lq_contract_f64_3m_t128 code object on device %d", dev);
      return LQ_ERR_HIP;
    }
    int m0 = M0, m1 = M1, ns = nsplit, tn = tiles_n, xm = xmap;
    long long la = lda, lb = ldb, kk = K, kc = kchunk;
    const double2 *a = A, *b = B;
    double2 *w = work;
    void *args[] = {&a, &la, &m0, &b, &lb, &m1, &kk, &kc, &ns, &tn, &xm, &w};
    LQ_HIP_CHECK(hipModuleLaunchKernel(fn, (unsigned)blocks, 1, 1, 512, 1, 1, 0, st, args, nullptr));
    return LQ_OK;
  }
  if (kind == kLds3m128) {  // rows too long for 32-bit lane offsets: the 128 x 64 kernel, same tables
    const int tiles_n64 = (M1 + kBN3 - 1) / kBN3;
    const long long blocks64 = (long long)tiles_m * tiles_n64 * nsplit;
    hipLaunchKernelGGL(k_contract_f64_3m_dl<2>, dim3((unsigned)blocks64), dim3(512), 0, st, A, lda, M0, B, ldb, M1,
                       K, kchunk, nsplit, tiles_n64, xmap, work);
  } else if (kind == kLds3m && contract_w1())
    hipLaunchKernelGGL(k_contract_f64_3m_w1, dim3((unsigned)blocks), dim3(256), 0, st, A, lda, M0, B, ldb, M1, K,
                       kchunk, nsplit, tiles_n, xmap, work);
  else if (kind == kLds3m && contract_dl() && contract_dlbuf() == 2)
    hipLaunchKernelGGL(k_contract_f64_3m_dl<2>, dim3((unsigned)blocks), dim3(512), 0, st, A, lda, M0, B, ldb, M1, K,
                       kchunk, nsplit, tiles_n, xmap, work);
  else if (kind == kLds3m && contract_dl())
    hipLaunchKernelGGL(k_contract_f64_3m_dl<3>, dim3((unsigned)blocks), dim3(512), 0, st, A, lda, M0, B, ldb, M1, K,
                       kchunk, nsplit, tiles_n, xmap, work);
  else if (kind == kLds3m)
    hipLaunchKernelGGL(k_contract_f64_3m, dim3((unsigned)blocks), dim3(512), 0, st, A, lda, M0, B, ldb, M1, K,
                       kchunk, nsplit, tiles_n, xmap, work);
  else if (kind == kLds4m)
    hipLaunchKernelGGL(k_contract_f64_lds, dim3((unsigned)blocks), dim3(512), 0, st, A, lda, M0, B, ldb, M1, K,
                       kchunk, nsplit, tiles_n, xmap, work);
  else
    hipLaunchKernelGGL(k_contract_f64, dim3((unsigned)blocks), dim3(256), 0, st, A, lda, M0, B, ldb, M1, K, kchunk,
                       nsplit, tiles_n, xmap, work);
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

int launch_contract_reduce_cols(hipStream_t st, const double2 *work, int nsplit, int M0, int M1, int bc,
                                double2 *const *outs, int ldo, int accumulate) {
  if (bc <= 0 || M1 % bc != 0 || M1 / bc > 2) {
    set_error("launch_contract_reduce_cols: %d columns are not 1 or 2 blocks of %d", M1, bc);
    return LQ_ERR_ARG;
  }
  const long long total = (long long)M0 * M1;
  hipLaunchKernelGGL(k_contract_reduce_cols, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, work, nsplit,
                     M0, M1, bc, outs[0], M1 / bc > 1 ? outs[1] : outs[0], ldo, accumulate);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

}  // namespace lq
