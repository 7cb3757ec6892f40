// lq_contract_t128.hip -- the default FP64 contraction kernel, built as a
// code object of its own and loaded by lq_contract.hip as a module.
//
// Its accumulators need 192 of the 256 registers a wave has at two waves per
// SIMD.  The compiler splits a wave's budget evenly between VGPRs and AGPRs
// unless the kernel's "amdgpu-agpr-alloc" attribute says otherwise, and no
// source attribute sets that; linqt_amd/build.py therefore compiles this file
// to LLVM IR, sets the attribute on the kernel named in the line below, and
// links the code object that lq_contract.hip embeds and loads per device.
//
// lq-module-kernel: lq_contract_f64_3m_t128 agpr-alloc=192
#include "../lq_contract_dev.h"

namespace lq {

// 128 x 128 block tile, two waves per SIMD: eight waves of 64 x 32 (ti =
// 0..3, tj = 0..1), each with its 3 x 8 accumulators (192 registers) pinned in
// AGPRs, so that fragments and addressing fit in the remaining 64 VGPRs of a
// wave.  Against the 128 x 64 tile of k_contract_f64_3m_dl (lq_contract.hip) a stage holds twice
// the MFMA work per barrier and per wait for the next stage's loads, and a
// third fewer operand bytes per flop.  The direct loads use a scalar base per
// (row group, stage) and a fixed 32-bit byte offset per lane (row within the
// group of four, element slot), so they need no per-load address arithmetic
// in vector registers.  Same MFMA sequence per output element as the other
// three-product kernels: bitwise equal tables.
constexpr int kT2 = 128;               // rows of A and of B per block
constexpr int kT2Slots = kT2 * kKC;    // slots of one operand in one buffer

// One global_load_lds_dwordx4 with a scalar base: lane l copies 16 bytes from
// base + off_l to LDS byte address lds + 16 l (lds uniform over the wave).
__device__ __forceinline__ void lds_dma16_s(const void *base, unsigned off, unsigned lds) {
  asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, %2"
               :
               : "s"(lds), "v"(off), "s"(base)
               : "memory", "m0");
}

extern "C" __global__ __launch_bounds__(512) void lq_contract_f64_3m_t128(const double2 *__restrict__ A, long long lda, int M0,
                                                              const double2 *__restrict__ B, long long ldb, int M1,
                                                              long long K, long long kchunk, int nsplit, int tiles_n,
                                                              int xmap, double2 *__restrict__ work) {
  // buffer b: A slots at sbuf + 2 b kT2Slots, then the B slots (128 KB for two)
  __shared__ double2 sbuf[4 * kT2Slots];
  int chunk, tile;
  block_chunk_tile(blockIdx.x, nsplit, xmap, chunk, tile);
  const int tm = tile / tiles_n, tn = tile % tiles_n;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r16 = lane & 15, ks = lane >> 4;
  const int wm = (wave >> 2) * 64, wn = (wave & 3) * 32;
  const unsigned lds0 = __builtin_amdgcn_readfirstlane((unsigned)(size_t)(lds_void_t *)sbuf);

  // Load q (0..3) of this wave covers slots 64 c .. 64 c + 63, c = 4 wave + q, of
  // both operands: lane l fills row R = 4 c + l / 16 = 16 wave + 4 q + l / 16,
  // element (l % 16) ^ (R & 15).  Rows past M0 / M1 are clamped (they feed
  // outputs that are never stored); the group's base row is clamped the same
  // way, so every lane offset is at most three rows plus 240 bytes.
  const int lrow = lane >> 4, lk = lane & 15;
  const double2 *pa[4], *pb[4];
  unsigned offa[4], offb[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r0 = 16 * wave + 4 * q;
    const int ga0 = min(tm * kT2 + r0, M0 - 1), gb0 = min(tn * kT2 + r0, M1 - 1);
    const int ga = min(tm * kT2 + r0 + lrow, M0 - 1), gb = min(tn * kT2 + r0 + lrow, M1 - 1);
    const int x = lk ^ ((4 * q + lrow) & 15);
    pa[q] = A + (size_t)ga0 * lda;
    pb[q] = B + (size_t)gb0 * ldb;
    offa[q] = (unsigned)(((long long)(ga - ga0) * lda + x) * (long long)sizeof(double2));
    offb[q] = (unsigned)(((long long)(gb - gb0) * ldb + x) * (long long)sizeof(double2));
  }

  const long long k0 = (long long)chunk * kchunk;
  const long long len = (k0 + kchunk < K ? k0 + kchunk : K) - k0;
  const int nfull = len > 0 ? (int)(len / kKC) : 0;
  const int nstage = len > 0 ? (int)((len + kKC - 1) / kKC) : 0;

  auto issue_full = [&](int j) {
    const long long kb = k0 + (long long)j * kKC;
    const unsigned la = lds0 + (unsigned)(((j & 1) * 2 * kT2Slots + 64 * 4 * wave) * sizeof(double2));
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      lds_dma16_s(pa[q] + kb, offa[q], la + q * 64 * sizeof(double2));
      lds_dma16_s(pb[q] + kb, offb[q], la + (kT2Slots + q * 64) * sizeof(double2));
    }
  };
  auto issue_part = [&](int j) {
    const long long kb = (long long)j * kKC;
    double2 *sA = sbuf + (j & 1) * 2 * kT2Slots, *sB = sA + kT2Slots;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int x = lk ^ ((4 * q + lrow) & 15);
      const double2 *ga = (const double2 *)((const char *)(pa[q] + k0 + kb) + offa[q]);
      const double2 *gb = (const double2 *)((const char *)(pb[q] + k0 + kb) + offb[q]);
      sA[64 * (4 * wave + q) + lane] = kb + x < len ? *ga : make_double2(0.0, 0.0);
      sB[64 * (4 * wave + q) + lane] = kb + x < len ? *gb : make_double2(0.0, 0.0);
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
  auto settle = [&]() { asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory"); };
  auto compute = [&](int j) {
    const double2 *sA = sbuf + (j & 1) * 2 * kT2Slots;
    mma3_stage_64x32(sA, sA + kT2Slots, wm + r16, wn + r16, ks, p1, p2, p3);
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
  int st = 0;
  for (; st + 1 < nfull; ++st) {
    issue_full(st + 1);
    compute(st);
    settle();
  }
  if (st < nstage) {
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
        const int m = tm * kT2 + wm + ti * 16 + ks + 4 * r;
        const int nn = tn * kT2 + wn + tj * 16 + r16;
        if (m < M0 && nn < M1) {
          const double a1 = p1[ti][tj][r], a2 = p2[ti][tj][r];
          work[((size_t)chunk * M0 + m) * M1 + nn] = make_double2(a1 + a2, p3[ti][tj][r] - a1 + a2);
        }
      }
}

}  // namespace lq
