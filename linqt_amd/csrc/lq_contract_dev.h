// lq_contract_dev.h -- device helpers shared by the contraction kernels of
// lq_contract.hip and by the code-object module lq_contract_t128.hip.
#pragma once
#include <hip/hip_runtime.h>

namespace lq {

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Block -> (K chunk, output tile).  Blocks are dispatched round-robin over
// the 8 XCDs (block b on XCD b % 8), each with its own L2.  xmap (nsplit a
// multiple of 8): XCD x takes chunks x, x + 8, ... one after another, all
// tiles of a chunk together, so every row of A and B in the chunk is read
// from HBM once and then shared through that XCD's L2.  Otherwise
// chunk = b % nsplit.  Either way each (chunk, tile) is one block and writes
// the same partial table, so the result does not depend on the mapping.
__device__ __forceinline__ void block_chunk_tile(int bid, int nsplit, int xmap, int &chunk, int &tile) {
  if (xmap) {
    const int tiles = (int)gridDim.x / nsplit;
    const int j = bid >> 3;
    chunk = (bid & 7) + 8 * (j / tiles);
    tile = j % tiles;
  } else {
    chunk = bid % nsplit;
    tile = bid / nsplit;
  }
}

constexpr int kKC = 16;  // complex elements per stage

// LDS slot of element k of row `row` in the unpadded direct-load layout: rows
// of 16 slots, element k in slot k ^ (row & 15), so the 16 rows read by one
// MFMA operand fetch hit 16 different slots.
__device__ __forceinline__ int dl_slot(int row, int k) { return row * kKC + (k ^ (row & 15)); }

typedef __attribute__((address_space(3))) void lds_void_t;

// One stage (16 K elements) of a 64 x 32 wave tile in the three-product form:
// ti = 0..3 fragments of A rows, tj = 0..1 of B rows, the same MFMA order per
// output element as the 32 x 32 tiles of the other three-product kernels.
__device__ __forceinline__ void mma3_stage_64x32(const double2 *sA, const double2 *sB, int arow0, int brow0, int ks,
                                              v4d (&p1)[4][2], v4d (&p2)[4][2], v4d (&p3)[4][2]) {
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    double ar[4], ai[4], ad[4], br[2], bi[2], bs[2];
#pragma unroll
    for (int ti = 0; ti < 4; ++ti) {
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
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) {
        p1[ti][tj] = mfma(ar[ti], br[tj], p1[ti][tj]);
        p2[ti][tj] = mfma(ai[ti], bi[tj], p2[ti][tj]);
        p3[ti][tj] = mfma(ad[ti], bs[tj], p3[ti][tj]);
      }
  }
}

}  // namespace lq
