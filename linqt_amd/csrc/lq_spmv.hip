// lq_spmv.hip -- complex128 sliced-ELL SpMM for the Chebyshev recursion.
//
// One launch computes, for up to two vector sets sharing the matrix stream,
//     y = alpha * A * x + beta * z          (k interleaved vectors per set)
// which covers SparseMatrixType::Multiply(a, x, b, y) (src/mkl_sparse_matrix.cpp:56-69)
// and the recursion step T_{n+1} = 2 H~ T_n - T_{n-1} (src/chebyshev_moments.cpp:157-162,
// src/chebyshev_vectors.cpp:17) for the left and right recursions of
// CorrelationExpansionMoments at once.
//
// Mapping: a 256-thread block owns 256/KP consecutive rows; the KP lanes of a
// row own one vector each.  A wavefront therefore covers 64/KP rows of one
// 64-row slice, so the column/value loads of one ELL column j are contiguous
// across the wave and the x gathers / y stores are 16-byte per lane.
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>

#include "lq_internal.h"

namespace lq {

namespace {

struct SetArgs {
  const double2 *x[2];
  const double2 *z[2];
  double2 *y[2];
  int k[2];
};

__device__ __forceinline__ double2 conj_mul(double2 a, double2 b) {
  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

template <int KP, int NDOT>
__device__ __forceinline__ void block_reduce_store(double2 (&d)[2], double2 *partial, int slot) {
  __shared__ double2 sh[2][256];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < NDOT; ++q) sh[q][tid] = d[q];
  __syncthreads();
#pragma unroll
  for (int stride = 128; stride >= KP; stride >>= 1) {
    if (tid < stride) {
#pragma unroll
      for (int q = 0; q < NDOT; ++q) {
        sh[q][tid].x += sh[q][tid + stride].x;
        sh[q][tid].y += sh[q][tid + stride].y;
      }
    }
    __syncthreads();
  }
  if (tid < KP) {
#pragma unroll
    for (int q = 0; q < NDOT; ++q) partial[((size_t)slot * NDOT + q) * KP + tid] = sh[q][tid];
  }
}

// MI355X has 8 XCDs and dispatches consecutive blocks round-robin over them.
// Remapping block b to row group (b % 8) * q + b / 8 gives every XCD one
// contiguous range of rows, so the neighbour gathers of a lattice ordering
// hit that XCD's own L2.  Logical groups >= ngroups (padding to 8q) exit.
constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_group(int b, int q) { return (b % kXcds) * q + b / kXcds; }

// Entry accessors of the two operator forms, in two phases so that all
// loads of a chunk's stored words are in flight before any value that depends
// on them is fetched: load(p, w, a) reads stored slot p (plain: column and
// value; dictionary: the packed word), column(w) gives the column and
// value(w, a) the value (dictionary: looked up in the table).  diag() is the
// separate diagonal of the dictionary form.
struct PlainMat {
  const int *__restrict__ sptr;
  const int *__restrict__ col;
  const double2 *__restrict__ val;
  __device__ __forceinline__ void load(size_t p, unsigned &w, double2 &a) const {
    w = (unsigned)col[p];
    a = val[p];
  }
  __device__ __forceinline__ int column(unsigned w) const { return (int)w; }
  __device__ __forceinline__ double2 value(unsigned, double2 a) const { return a; }
  __device__ __forceinline__ bool has_diag() const { return false; }
  __device__ __forceinline__ double2 diag(int) const { return make_double2(0.0, 0.0); }
};

struct DictMat {
  const int *__restrict__ sptr;
  const unsigned *__restrict__ wcol;
  const double2 *dict;  // value table (LDS copy, or global in the batch kernel)
  const double2 *__restrict__ dg;
  __device__ __forceinline__ void load(size_t p, unsigned &w, double2 &) const { w = wcol[p]; }
  __device__ __forceinline__ int column(unsigned w) const { return (int)(w & 0xffffffu); }
  // the table is in LDS: read each value where it is used, not held in registers
  __device__ __forceinline__ double2 value(unsigned w, double2) const { return dict[w >> 24]; }
  __device__ __forceinline__ bool has_diag() const { return dg != nullptr; }
  __device__ __forceinline__ double2 diag(int row) const { return dg[row]; }
};

// Value tables live in LDS for the dictionary-form kernels: a lookup per lane
// per entry from LDS instead of a 16-byte-per-lane load through the vector
// memory path that also carries the gathers.  The grids of these kernels are
// capped (grid-stride) so the one-off staging per block is amortised.
// By default a dictionary-form launch without dot fusion gets one full set of
// resident blocks (CUs x blocks per CU at the kernel's occupancy, see
// resident_blocks): on the fused recursion step that measured 2 % faster than
// twice as many, and a grid that leaves the last round part-full (1.5 sets)
// 13 % slower.  LQ_SPMM_GRIDCAP=n (n >= 256) fixes the cap instead.
constexpr int kDictGridCap = 1024;  // used when the occupancy query fails

// Deterministic sum of the per-block dot partials of one slot, p[b * width + w]
// over b < grid for each w < width (width a power of two <= 256), by a
// 256-thread block: thread t takes w = t % width and the blocks
// b = t / width, t / width + S, ... (S = 256 / width streams), coalesced; the S
// stream sums of each w are then added pairwise in a fixed tree.
__device__ void reduce_slot(const double2 *__restrict__ p, int grid, int width, double2 *__restrict__ out) {
  __shared__ double2 sh[256];
  const int t = threadIdx.x;
  const int streams = 256 / width;
  double sx = 0.0, sy = 0.0;
  if (t < streams * width) {
    const int w = t % width;
    for (int b = t / width; b < grid; b += streams) {
      const double2 v = p[(size_t)b * width + w];
      sx += v.x;
      sy += v.y;
    }
  }
  sh[t] = make_double2(sx, sy);
  __syncthreads();
  for (int h = streams / 2; h > 0; h >>= 1) {
    if (t < h * width) sh[t] = make_double2(sh[t].x + sh[t + h * width].x, sh[t].y + sh[t + h * width].y);
    __syncthreads();
  }
  if (t < width) out[t] = sh[t];
}

// Schedule knobs: LQ_SPMM_GRIDCAP fixes the grid cap of dictionary-form
// launches (default: one set of resident blocks) and LQ_SPMM_PF=0 turns the
// prefetched schedule off (both read when an operator is uploaded),
// LQ_SPMM_UNROLL (4, 8, 12, 16; read once) sets the gather chunk.  None of
// them changes the arithmetic of a row.
int env_int(const char *name, int dflt) {
  const char *e = getenv(name);
  return e && *e ? atoi(e) : dflt;
}

// Non-temporal 16-byte load and store (streamed past the caches).
typedef double dvec2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 nt_load2(const double2 *p) {
  const dvec2 v = __builtin_nontemporal_load((const dvec2 *)p);
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void nt_store2(double2 *p, double2 a) {
  const dvec2 v = {a.x, a.y};
  __builtin_nontemporal_store(v, (dvec2 *)p);
}

__device__ __forceinline__ void stage(double2 *dst, const double2 *__restrict__ src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// Rows of row groups g0, g0 + gstep, ... (RPB = 256 / KP rows each) of one
// launch; dot contributions are accumulated into d.
// NV > 0: the lanes of vector set fa.set also form yv[o] = V_o x (o < NV)
// from the same gathers, V_o's values indexed per stored slot by fa.vidx.
// PF (dictionary form, grid-stride launches): the block's slice offsets are
// staged in LDS (sgeo[i] = {first stored slot, width} of its i-th row group);
// the packed words of the next chunk (of this row, or the next row group's
// first) are loaded while the current chunk's gathers are in flight, and the
// diagonal, the row's own x and z are loaded together after the last chunk, so
// a row group of two chunks waits on three memory round trips instead of six.
// The arithmetic is unchanged.
constexpr int kPfGroups = 64;  // row groups per block covered by the staged offsets

template <int KP, int DOT, int kUnroll, class Mat, int NV = 0, bool PF = false>
__device__ __forceinline__ void spmm_body(int n, const Mat &mat, double2 alpha, double2 beta, const SetArgs &sa,
                                          int nrowgroups, const double2 *__restrict__ phi, int g0, int gstep,
                                          double2 (&d)[2], const FusedArgs &fa = FusedArgs{},
                                          const int2 *sgeo = nullptr) {
  constexpr int RPB = 256 / KP;
  static_assert(!PF || RPB <= 64, "staged offsets need row groups within one slice");
  const int tid = threadIdx.x;
  const int v = tid % KP;
  const int set = (v < sa.k[0]) ? 0 : 1;
  const int vv = set ? v - sa.k[0] : v;
  const int kk = sa.k[set];
  const bool lane_on = v < sa.k[0] + sa.k[1];
  const double2 *__restrict__ x = sa.x[set];
  const double2 *z = sa.z[set];
  double2 *y = sa.y[set];
  const bool vlane = NV > 0 && set == fa.set;
  // PF: slice offsets of the block's i-th row group from LDS; pw holds the
  // packed words of the next chunk (of this row, or the next row group's first)
  unsigned pw[PF ? kUnroll : 1];
  auto fetch_chunk = [&](size_t p, int cnt) {  // words of slots p, p + kSlice, ...
#pragma unroll
    for (int u = 0; u < (PF ? kUnroll : 1); ++u) {
      pw[u] = 0u;
      double2 dummy;
      if (u < cnt) mat.load(p + (size_t)u * kSlice, pw[u], dummy);
    }
  };
  auto geo = [&](int i, int gg, int &base_o, int &w_o) {
    if (PF && i < kPfGroups) {
      const int2 q = sgeo[i];
      base_o = q.x;
      w_o = q.y;
    } else {
      const int ss = (gg * RPB) >> 6;
      base_o = mat.sptr[ss];
      w_o = (mat.sptr[ss + 1] - base_o) >> 6;
    }
  };
  auto fetch = [&](int i, int gg) {  // first chunk of row group gg (the block's i-th)
    if (!PF) return;
    const int rr = gg * RPB + tid / KP;
    int bb = 0, ww = 0;
    if (gg < nrowgroups && rr < n && lane_on) geo(i, gg, bb, ww);
    fetch_chunk((size_t)bb + (rr & 63), ww);
  };
  fetch(0, g0);
  int it = 0;
  for (int g = g0; g < nrowgroups; g += gstep, ++it) {
    const int row = g * RPB + tid / KP;
    if (row < n && lane_on) {
      const int r = row & 63;
      int base, w;
      if (PF) {
        geo(it, g, base, w);
      } else {
        const int s = row >> 6;
        base = mat.sptr[s];
        w = (mat.sptr[s + 1] - base) >> 6;
      }
      const size_t p0 = (size_t)base + r;
      const size_t o = (size_t)row * kk + vv;
      // The separate diagonal (dictionary form) and the row's own x: without
      // PF loaded with the first entries; with PF loaded after the last chunk
      // together with z (one round trip, and no registers held through the
      // chunks).  Either way they enter the sums after the off-diagonal terms.
      double2 dv = make_double2(0.0, 0.0), xo = make_double2(0.0, 0.0);
      double2 vdv[NV > 0 ? NV : 1];
      double2 zv = make_double2(0.0, 0.0);
      auto own_loads = [&]() {
        bool vdg = false;
#pragma unroll
        for (int q = 0; q < NV; ++q) {
          vdv[q] = make_double2(0.0, 0.0);
          if (vlane && fa.vdiag[q]) {
            vdv[q] = fa.vdiag[q][row];
            vdg = true;
          }
        }
        if (mat.has_diag()) dv = mat.diag(row);
        if (mat.has_diag() || vdg) xo = x[o];
      };
      if (!PF) own_loads();
      double ar = 0.0, ai = 0.0;
      double vr[NV > 0 ? NV : 1], vi[NV > 0 ? NV : 1];
#pragma unroll
      for (int q = 0; q < NV; ++q) vr[q] = vi[q] = 0.0;
      // Chunks of kUnroll entries: all column/value loads of a chunk are in
      // flight together, then all of its gathers, so a lattice row costs two
      // memory latencies instead of one per entry.
      for (int j0 = 0; j0 < w; j0 += kUnroll) {
        unsigned wd[kUnroll];
        unsigned vw[NV > 0 ? kUnroll : 1];
        double2 a[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          wd[u] = 0u;
          a[u] = make_double2(0.0, 0.0);
          if (PF)
            wd[u] = pw[PF ? u : 0];
          else if (j0 + u < w)
            mat.load(p0 + (size_t)(j0 + u) * kSlice, wd[u], a[u]);
          if (NV > 0 && !PF) vw[u] = (vlane && j0 + u < w) ? (unsigned)fa.vidx[p0 + (size_t)(j0 + u) * kSlice] : 0u;
        }
        double2 xv[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          if (j0 + u < w) {
            xv[u] = x[(size_t)mat.column(wd[u]) * kk + vv];
          } else {
            xv[u] = make_double2(0.0, 0.0);
          }
        }
        if (PF) {
          // behind this chunk's gathers: its value indices (fused operators)
          // and the next chunk's words
#pragma unroll
          for (int u = 0; u < (NV > 0 ? kUnroll : 0); ++u)
            vw[u] = (vlane && j0 + u < w) ? (unsigned)fa.vidx[p0 + (size_t)(j0 + u) * kSlice] : 0u;
          if (j0 + kUnroll < w)
            fetch_chunk(p0 + (size_t)(j0 + kUnroll) * kSlice, w - j0 - kUnroll);
          else
            fetch(it + 1, g + gstep);
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          // padding slots: xv = 0, and a word of 0 reads table entry 0 (finite)
          const double2 av = mat.value(wd[u], a[u]);
          ar = fma(av.x, xv[u].x, ar);
          ar = fma(-av.y, xv[u].y, ar);
          ai = fma(av.x, xv[u].y, ai);
          ai = fma(av.y, xv[u].x, ai);
        }
        if (vlane) {
#pragma unroll
          for (int u = 0; u < kUnroll; ++u) {
            if (j0 + u < w) {
#pragma unroll
              for (int q = 0; q < NV; ++q) {
                const double2 b = fa.vdict[q][(vw[u] >> (8 * q)) & 0xffu];
                vr[q] = fma(b.x, xv[u].x, vr[q]);
                vr[q] = fma(-b.y, xv[u].y, vr[q]);
                vi[q] = fma(b.x, xv[u].y, vi[q]);
                vi[q] = fma(b.y, xv[u].x, vi[q]);
              }
            }
          }
        }
      }
      if (PF && w <= 0) fetch(it + 1, g + gstep);
      if (PF) {
        own_loads();
        if (z) zv = fa.nt ? nt_load2(z + o) : z[o];
      }
      if (vlane) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
          vr[q] = fma(vdv[q].x, xo.x, vr[q]);
          vr[q] = fma(-vdv[q].y, xo.y, vr[q]);
          vi[q] = fma(vdv[q].x, xo.y, vi[q]);
          vi[q] = fma(vdv[q].y, xo.x, vi[q]);
          if (fa.nt)
            nt_store2(fa.yv[q] + o, make_double2(vr[q], vi[q]));
          else
            fa.yv[q][o] = make_double2(vr[q], vi[q]);
        }
      }
      if (mat.has_diag()) {
        ar = fma(dv.x, xo.x, ar);
        ar = fma(-dv.y, xo.y, ar);
        ai = fma(dv.x, xo.y, ai);
        ai = fma(dv.y, xo.x, ai);
      }
      double2 res = make_double2(alpha.x * ar - alpha.y * ai, alpha.x * ai + alpha.y * ar);
      if (z) {
        if (!PF) zv = fa.nt ? nt_load2(z + o) : z[o];
        res.x += beta.x * zv.x - beta.y * zv.y;
        res.y += beta.x * zv.y + beta.y * zv.x;
      }
      y[o] = res;
      if (DOT == kDotPhi) {
        const double2 t = conj_mul(phi[o], res);
        d[0].x += t.x;
        d[0].y += t.y;
      } else if (DOT == kDotDoubling) {
        const double2 xi = x[o];
        const double2 t0 = conj_mul(xi, xi);
        const double2 t1 = conj_mul(xi, res);
        d[0].x += t0.x;
        d[0].y += t0.y;
        d[1].x += t1.x;
        d[1].y += t1.y;
      }
    }
  }
}

// The operator as seen by one launch: plain or dictionary form.
struct MatArgs {
  const int *sptr;
  const int *col;
  const double2 *val;
  const unsigned *wcol;
  const double2 *dict;
  const double2 *diag;
  int ndict;  // > 0: dictionary form
};

template <int KP, int DOT, int kUnroll, bool DICT, int NV = 0, bool PF = false>
__global__ __launch_bounds__(256) void k_spmm(int n, MatArgs m, double2 alpha, double2 beta, SetArgs sa,
                                              int nrowgroups, const double2 *__restrict__ phi,
                                              double2 *__restrict__ partial, int q, FusedArgs fa) {
  constexpr int NDOT = (DOT == kDotDoubling) ? 2 : 1;
  double2 d[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  // q > 0: one row group per block, XCD-remapped; q == 0: grid-stride loop
  const int g0 = q > 0 ? xcd_group(blockIdx.x, q) : blockIdx.x;
  const int gstep = q > 0 ? nrowgroups : gridDim.x;
  if (DICT) {
    __shared__ double2 sdict[256];
    __shared__ double2 svdict[NV > 0 ? NV : 1][NV > 0 ? 256 : 1];
    stage(sdict, m.dict, m.ndict);
    __shared__ int2 sgeo[PF ? kPfGroups : 1];
    if (PF) {
      constexpr int RPB = 256 / KP;
      for (int i = threadIdx.x; i < kPfGroups; i += blockDim.x) {
        const int gg = g0 + i * gstep;
        if (gg < nrowgroups) {
          const int ss = (gg * RPB) >> 6;
          const int b = m.sptr[ss];
          sgeo[i] = make_int2(b, (m.sptr[ss + 1] - b) >> 6);
        }
      }
    }
    FusedArgs fl = fa;
#pragma unroll
    for (int o = 0; o < NV; ++o) {
      stage(svdict[o], fa.vdict[o], fa.nvdict[o]);
      fl.vdict[o] = svdict[o];
    }
    __syncthreads();
    const DictMat mat{m.sptr, m.wcol, sdict, m.diag};
    spmm_body<KP, DOT, kUnroll, DictMat, NV, PF>(n, mat, alpha, beta, sa, nrowgroups, phi, g0, gstep, d, fl, sgeo);
  } else {
    const PlainMat mat{m.sptr, m.col, m.val};
    spmm_body<KP, DOT, kUnroll>(n, mat, alpha, beta, sa, nrowgroups, phi, g0, gstep, d);
  }
  if (DOT != kDotNone) block_reduce_store<KP, NDOT>(d, partial, blockIdx.x);
}

// Batched recursion step over many independent jobs (one vector-width class):
// block b belongs to the job j with blk0_j <= b < blk0_j + nblk_j and runs that
// job's step s = t - t0_j with the same row schedule, dot fusion and partial
// layout as a single-job launch with nblk_j blocks, so every job's moments are
// independent of what else is in the batch.
//   s == 0:  T_1 = H~ T_0                     (x = buf0, y = buf1)
//   s >= 1:  T_{s+1} = 2 H~ T_s - T_{s-1}     (in place over T_{s-1})
template <int KP, int kUnroll, bool PF = false>
__global__ __launch_bounds__(256) void k_spmm_batch(const BatchJobDev *__restrict__ jobs, int njobs,
                                                    const int *__restrict__ tc, int slot, int ntz) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].blk0 <= (int)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const BatchJobDev &J = jobs[lo];
  const int s = tc[0] + slot - J.t0;
  if (s < 0 || s >= J.steps) return;  // whole block leaves together
  const int lb = blockIdx.x - J.blk0;
  SetArgs sa{};
  const bool odd = s & 1;
  sa.x[0] = sa.x[1] = odd ? J.buf1 : J.buf0;
  double2 *y = odd ? J.buf0 : J.buf1;
  sa.y[0] = sa.y[1] = y;
  sa.z[0] = sa.z[1] = s == 0 ? nullptr : y;
  sa.k[0] = J.k;
  sa.k[1] = 0;
  const double2 alpha = s == 0 ? make_double2(1.0, 0.0) : make_double2(2.0, 0.0);
  const double2 beta = s == 0 ? make_double2(0.0, 0.0) : make_double2(-1.0, 0.0);
  constexpr int RPB = 256 / KP;
  const int groups = (J.n + RPB - 1) / RPB;
  double2 d[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  if (J.ndict > 0) {
    __shared__ double2 sdict[256];
    stage(sdict, J.dict, J.ndict);  // J is the same for the whole block
    __shared__ int2 sgeo[PF ? kPfGroups : 1];
    if (PF) {
      for (int i = threadIdx.x; i < kPfGroups; i += blockDim.x) {
        const int gg = lb + i * J.nblk;
        if (gg < groups) {
          const int ss = (gg * RPB) >> 6;
          const int b = J.dsptr[ss];
          sgeo[i] = make_int2(b, (J.dsptr[ss + 1] - b) >> 6);
        }
      }
    }
    __syncthreads();
    const DictMat mat{J.dsptr, J.wcol, sdict, J.diag};
    FusedArgs bf;
    bf.nt = ntz;
    spmm_body<KP, kDotPhi, kUnroll, DictMat, 0, PF>(J.n, mat, alpha, beta, sa, groups, J.phi, lb, J.nblk, d, bf,
                                                    sgeo);
  } else {
    const PlainMat mat{J.sptr, J.col, J.val};
    FusedArgs bf;
    bf.nt = ntz;
    spmm_body<KP, kDotPhi, kUnroll>(J.n, mat, alpha, beta, sa, groups, J.phi, lb, J.nblk, d, bf);
  }
  block_reduce_store<KP, 1>(d, J.partial + (size_t)slot * J.nblk * KP, lb);
}

// mom[(s + 1) * KP + v] = sum_b partial[slot][b][v] for every step s a job ran
// in this chunk (the same summation order as k_reduce_partials).
__global__ __launch_bounds__(256) void k_batch_reduce(const BatchJobDev *__restrict__ jobs,
                                                      const int *__restrict__ tc, int nslots, int kp) {
  const BatchJobDev &J = jobs[blockIdx.y];
  const int slot = blockIdx.x;
  const int s = tc[0] + slot - J.t0;
  if (s < 0 || s >= J.steps) return;  // uniform over the block
  reduce_slot(J.partial + (size_t)slot * J.nblk * kp, J.nblk, kp, J.mom + (size_t)(s + 1) * kp);
}

__global__ void k_set_int(int *p, int v) { *p = v; }

// Lane-per-row variant: one lane owns one row and all KS vectors of each of
// NSETS sets (registers), so each (col, value) pair is loaded once and the
// x gathers of one row are KS*16 contiguous bytes.
template <int KS, int NSETS>
__global__ __launch_bounds__(256) void k_spmm_rows(int n, const int *__restrict__ sptr, const int *__restrict__ col,
                                                   const double2 *__restrict__ val, double2 alpha, double2 beta,
                                                   SetArgs sa, int ngroups, int q) {
  const int g = q > 0 ? xcd_group(blockIdx.x, q) : blockIdx.x;
  if (g >= ngroups) return;
  const int row = g * 256 + threadIdx.x;
  if (row >= n) return;
  const int s = row >> 6, r = row & 63;
  const int base = sptr[s];
  const int w = (sptr[s + 1] - base) >> 6;
  const int *__restrict__ cp = col + base + r;
  const double2 *__restrict__ vp = val + base + r;
  double ar[NSETS][KS], ai[NSETS][KS];
#pragma unroll
  for (int st = 0; st < NSETS; ++st)
#pragma unroll
    for (int v = 0; v < KS; ++v) ar[st][v] = ai[st][v] = 0.0;
  for (int j = 0; j < w; ++j) {
    const int c = cp[j * kSlice];
    const double2 a = vp[j * kSlice];
#pragma unroll
    for (int st = 0; st < NSETS; ++st) {
      const double2 *__restrict__ xr = sa.x[st] + (size_t)c * KS;
#pragma unroll
      for (int v = 0; v < KS; ++v) {
        const double2 xv = xr[v];
        ar[st][v] = fma(a.x, xv.x, ar[st][v]);
        ar[st][v] = fma(-a.y, xv.y, ar[st][v]);
        ai[st][v] = fma(a.x, xv.y, ai[st][v]);
        ai[st][v] = fma(a.y, xv.x, ai[st][v]);
      }
    }
  }
#pragma unroll
  for (int st = 0; st < NSETS; ++st) {
    const double2 *z = sa.z[st];
    double2 *y = sa.y[st] + (size_t)row * KS;
#pragma unroll
    for (int v = 0; v < KS; ++v) {
      double2 res = make_double2(alpha.x * ar[st][v] - alpha.y * ai[st][v], alpha.x * ai[st][v] + alpha.y * ar[st][v]);
      if (z) {
        const double2 zv = z[(size_t)row * KS + v];
        res.x += beta.x * zv.x - beta.y * zv.y;
        res.y += beta.x * zv.y + beta.y * zv.x;
      }
      y[v] = res;
    }
  }
}

template <int KP>
__global__ __launch_bounds__(256) void k_dot(int n, int k, const double2 *__restrict__ a,
                                             const double2 *__restrict__ b, int nrowgroups,
                                             double2 *__restrict__ partial) {
  constexpr int RPB = 256 / KP;
  const int tid = threadIdx.x;
  const int v = tid % KP;
  double2 d[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  for (int g = blockIdx.x; g < nrowgroups; g += gridDim.x) {
    const int row = g * RPB + tid / KP;
    if (row < n && v < k) {
      const size_t o = (size_t)row * k + v;
      const double2 t = conj_mul(a[o], b[o]);
      d[0].x += t.x;
      d[0].y += t.y;
    }
  }
  block_reduce_store<KP, 1>(d, partial, blockIdx.x);
}

// one block per slot: out[slot * width + w] = sum over b < grid of
// partial[(slot * grid + b) * width + w], in the order of reduce_slot
__global__ __launch_bounds__(256) void k_reduce_partials(const double2 *__restrict__ partial, int nslots, int grid,
                                                         int width, double2 *__restrict__ out) {
  const int slot = blockIdx.x;
  reduce_slot(partial + (size_t)slot * grid * width, grid, width, out + (size_t)slot * width);
}

// dst[i*g + v] = src[(r0 + v)*n + i]: state-major -> interleaved block.
__global__ void k_interleave(const double2 *__restrict__ src, int n, int r0, int g, double2 *__restrict__ dst) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * g) return;
  const int i = (int)(t / g), v = (int)(t % g);
  dst[t] = src[(size_t)(r0 + v) * n + i];
}

int pad_pow2(int k) {
  int p = 1;
  while (p < k) p <<= 1;
  return p;
}

constexpr int kDotGridCap = 1024;

// Schedule knobs, read once: LQ_SPMM_MAPPING=rows selects the lane-per-row
// mapping, LQ_SPMM_XCD=1 the XCD remap (both measured slower on the
// lattice operators; kept as schedule options).
struct SpmmKnobs {
  bool rows = false;
  bool xcd = false;
  SpmmKnobs() {
    const char *m = getenv("LQ_SPMM_MAPPING");
    if (m && strcmp(m, "rows") == 0) rows = true;
    const char *x = getenv("LQ_SPMM_XCD");
    if (x && strcmp(x, "1") == 0) xcd = true;
  }
};

const SpmmKnobs &knobs() {
  static SpmmKnobs k;
  return k;
}

int xcd_q(int groups) { return knobs().xcd ? (groups + kXcds - 1) / kXcds : 0; }

bool use_dict(const SellDev &A) { return A.ndict > 0; }

}  // namespace

void dict_schedule(SellDev &d) {
  const int c = env_int("LQ_SPMM_GRIDCAP", 0);
  d.gridcap = c >= 256 ? c : 0;  // 0: one set of resident blocks
  d.pf = env_int("LQ_SPMM_PF", 1) != 0;
}

namespace {

MatArgs mat_args(const SellDev &A) {
  if (use_dict(A)) return MatArgs{A.dsptr, nullptr, nullptr, A.wcol, A.dict, A.diag, A.ndict};
  return MatArgs{A.sptr, A.col, A.val, nullptr, nullptr, nullptr, 0};
}

// Prefetched schedule (dictionary form, up to 64 rows per group, no fused
// operators): on unless LQ_SPMM_PF=0.  The fused kernel fits it in 122
// registers too, but measured 1 % slower with it: with its fourteen read and
// write streams it is bound by the HBM traffic, not by the round trips.
bool spmm_pf() { return env_int("LQ_SPMM_PF", 1) != 0; }

// Blocks of 256 threads of `kernel` resident on the current device at once
// (CUs x blocks per CU at its occupancy), queried once per kernel and device.
int resident_blocks(const void *kernel) {
  static std::mutex mu;
  static std::map<std::pair<const void *, int>, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return kDictGridCap;
  }
  std::lock_guard<std::mutex> lock(mu);
  const auto key = std::make_pair(kernel, dev);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int per_cu = 0, cus = 0;
  int r = kDictGridCap;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && per_cu > 0 &&
      cus > 0)
    r = per_cu * cus;
  else
    (void)hipGetLastError();
  cache.emplace(key, r);
  return r;
}

// grid < 0: one set of resident blocks of the kernel that runs, at most groups.
template <int KP, int DOT, int U, bool DICT, int NV>
void launch_k_spmm(hipStream_t st, int grid, int n, const MatArgs &m, double2 alpha, double2 beta,
                   const SetArgs &sa, int groups, const double2 *phi, double2 *partial, int q, const FusedArgs &fa,
                   bool pf) {
  auto go = [&](auto kernel) {
    int gr = grid;
    if (gr < 0) {
      gr = resident_blocks((const void *)kernel);
      if (gr > groups) gr = groups;
    }
    hipLaunchKernelGGL(kernel, dim3(gr), dim3(256), 0, st, n, m, alpha, beta, sa, groups, phi, partial, q, fa);
  };
  if constexpr (DICT && NV == 0 && 256 / KP <= 64) {
    if (q == 0 && pf) {
      go(k_spmm<KP, DOT, U, DICT, NV, true>);
      return;
    }
  }
  go(k_spmm<KP, DOT, U, DICT, NV>);
}

template <int KP, int U, bool DICT>
int launch_spmm_kpu(hipStream_t st, const SellDev &A, double2 alpha, double2 beta, const SetArgs &sa,
                    int dot_mode, const double2 *phi, double2 *partial, int *grid_used, const FusedArgs &fa) {
  constexpr int RPB = 256 / KP;
  const int groups = (A.n + RPB - 1) / RPB;
  int grid = groups, q = 0;
  if (dot_mode != kDotNone) {
    grid = groups < kDotGridCap ? groups : kDotGridCap;
  } else if (DICT) {
    grid = A.gridcap > 0 ? (groups < A.gridcap ? groups : A.gridcap) : -1;  // -1: resident set
  } else {
    q = xcd_q(groups);
    if (q > 0) grid = q * kXcds;
  }
  // the block count matters to callers of dot-fused launches (partial sums);
  // launches that size themselves at launch time report 0
  if (grid_used) *grid_used = grid > 0 ? grid : 0;
  const MatArgs m = mat_args(A);
  if (fa.nv > 0) {
    if constexpr (DICT) {
      if (dot_mode != kDotNone) {
        set_error("launch_spmm: fused operators take no dot fusion");
        return LQ_ERR_ARG;
      }
      if (fa.nv == 1)
        launch_k_spmm<KP, kDotNone, U, true, 1>(st, grid, A.n, m, alpha, beta, sa, groups, phi, partial, q, fa, A.pf);
      else
        launch_k_spmm<KP, kDotNone, U, true, 2>(st, grid, A.n, m, alpha, beta, sa, groups, phi, partial, q, fa, A.pf);
      LQ_HIP_CHECK(hipGetLastError());
      return LQ_OK;
    } else {
      set_error("launch_spmm: fused operators need the dictionary form");
      return LQ_ERR_ARG;
    }
  }
  switch (dot_mode) {
    case kDotNone:
      launch_k_spmm<KP, kDotNone, U, DICT, 0>(st, grid, A.n, m, alpha, beta, sa, groups, phi, partial, q, fa, A.pf);
      break;
    case kDotPhi:
      launch_k_spmm<KP, kDotPhi, U, DICT, 0>(st, grid, A.n, m, alpha, beta, sa, groups, phi, partial, 0, fa, A.pf);
      break;
    case kDotDoubling:
      launch_k_spmm<KP, kDotDoubling, U, DICT, 0>(st, grid, A.n, m, alpha, beta, sa, groups, phi, partial, 0, fa, A.pf);
      break;
    default:
      set_error("launch_spmm: bad dot mode %d", dot_mode);
      return LQ_ERR_ARG;
  }
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

// Unroll width: the smallest of 4/8/12/16 covering the widest slice (wider
// operators loop over chunks of 16).
int unroll_for(int wmax, int tuned) {
  static const int forced = env_int("LQ_SPMM_UNROLL", 0);
  if (forced == 4 || forced == 8 || forced == 12 || forced == 16) return forced;
  if (tuned == 4 || tuned == 8 || tuned == 12 || tuned == 16) return tuned;
  return wmax <= 4 ? 4 : wmax <= 8 ? 8 : wmax <= 12 ? 12 : 16;
}

template <int KP>
int launch_spmm_kp(hipStream_t st, const SellDev &A, double2 alpha, double2 beta, const SetArgs &sa,
                   int dot_mode, const double2 *phi, double2 *partial, int *grid_used, const FusedArgs &fa) {
#define LQ_KPU(UV, DV) \
  return launch_spmm_kpu<KP, UV, DV>(st, A, alpha, beta, sa, dot_mode, phi, partial, grid_used, fa)
  if (use_dict(A)) {
    // with fused right-operator applies a 12-entry chunk needs 164 VGPRs (3 waves
    // per SIMD) and 8-entry chunks 128 (4 waves); with the grid sized to the
    // resident blocks of each, one 12-entry chunk per row measured faster on
    // rows of 9-12 entries than two 8-entry chunks
    const int u = unroll_for(A.dwmax, A.unroll);
    switch (u) {
      case 4: LQ_KPU(4, true);
      case 8: LQ_KPU(8, true);
      case 12: LQ_KPU(12, true);
      default: LQ_KPU(16, true);
    }
  }
  switch (unroll_for(A.wmax, A.unroll)) {
    case 4: LQ_KPU(4, false);
    case 8: LQ_KPU(8, false);
    case 12: LQ_KPU(12, false);
    default: LQ_KPU(16, false);
  }
#undef LQ_KPU
}

}  // namespace

int spmm_dot_grid(int n, int kpad) {
  const int RPB = 256 / kpad;
  const int groups = (n + RPB - 1) / RPB;
  return groups < kDotGridCap ? groups : kDotGridCap;
}

int launch_spmm(hipStream_t st, const SellDev &A, double2 alpha, double2 beta, const VecSet *sets, int nsets,
                int dot_mode, const double2 *phi, double2 *partial, int *grid_used, const FusedArgs *fused) {
  // Fused steps store the operator products (read again only by the
  // contraction) and load T_{n-1} (read once) non-temporally, which leaves the
  // caches to the vectors the next step gathers: 5 % faster on the Bastin step.
  // LQ_SPMM_NT=0 turns it off, LQ_SPMM_NT=2 also streams T_{n-1} in launches
  // without fused operators (2 % slower on the cache-resident DOS case).
  // T_{n+1} itself is gathered by the next step and measured 4 % slower when
  // stored the same way.
  FusedArgs fa = fused ? *fused : FusedArgs{};
  static const int nt = env_int("LQ_SPMM_NT", 1);
  fa.nt = fa.nv > 0 ? (nt > 0) : (nt >= 2);
  if (fa.nv < 0 || fa.nv > 2 || (fa.nv > 0 && (fa.set < 0 || fa.set >= nsets || !fa.vidx))) {
    set_error("launch_spmm: bad fused operator arguments");
    return LQ_ERR_ARG;
  }
  if (nsets < 1 || nsets > 2) {
    set_error("launch_spmm: nsets must be 1 or 2");
    return LQ_ERR_ARG;
  }
  SetArgs sa{};
  int ktot = 0;
  for (int s = 0; s < nsets; ++s) {
    if (!sets[s].x || !sets[s].y || sets[s].k <= 0) {
      set_error("launch_spmm: incomplete vector set %d", s);
      return LQ_ERR_ARG;
    }
    sa.x[s] = sets[s].x;
    sa.z[s] = sets[s].z;
    sa.y[s] = sets[s].y;
    sa.k[s] = sets[s].k;
    ktot += sets[s].k;
  }
  if (nsets == 1) {
    sa.x[1] = sa.x[0];
    sa.z[1] = sa.z[0];
    sa.y[1] = sa.y[0];
    sa.k[1] = 0;
  }
  if (dot_mode != kDotNone && nsets != 1) {
    set_error("launch_spmm: dot fusion needs a single vector set");
    return LQ_ERR_ARG;
  }
  // lane-per-row schedule: equal power-of-two widths per set, no dot fusion
  const int ks = sets[0].k;
  const bool rows_ok = knobs().rows && fa.nv == 0 && dot_mode == kDotNone && (nsets == 1 || sets[1].k == ks) &&
                       (ks == 1 || ks == 2 || ks == 4 || ks == 8 || (ks == 16 && nsets == 1));
  if (rows_ok) {
    const int groups = (A.n + 255) / 256;
    const int q = xcd_q(groups);
    const int grid = q > 0 ? q * kXcds : groups;
    if (grid_used) *grid_used = grid;
#define LQ_ROWS(KSV, NS)                                                                                         \
  hipLaunchKernelGGL((k_spmm_rows<KSV, NS>), dim3(grid), dim3(256), 0, st, A.n, A.sptr, A.col, A.val, alpha, beta, \
                     sa, groups, q)
    if (nsets == 1) {
      switch (ks) {
        case 1: LQ_ROWS(1, 1); break;
        case 2: LQ_ROWS(2, 1); break;
        case 4: LQ_ROWS(4, 1); break;
        case 8: LQ_ROWS(8, 1); break;
        default: LQ_ROWS(16, 1); break;
      }
    } else {
      switch (ks) {
        case 1: LQ_ROWS(1, 2); break;
        case 2: LQ_ROWS(2, 2); break;
        case 4: LQ_ROWS(4, 2); break;
        default: LQ_ROWS(8, 2); break;
      }
    }
#undef LQ_ROWS
    LQ_HIP_CHECK(hipGetLastError());
    return LQ_OK;
  }
  const int kp = pad_pow2(ktot);
  switch (kp) {
    case 1: return launch_spmm_kp<1>(st, A, alpha, beta, sa, dot_mode, phi, partial, grid_used, fa);
    case 2: return launch_spmm_kp<2>(st, A, alpha, beta, sa, dot_mode, phi, partial, grid_used, fa);
    case 4: return launch_spmm_kp<4>(st, A, alpha, beta, sa, dot_mode, phi, partial, grid_used, fa);
    case 8: return launch_spmm_kp<8>(st, A, alpha, beta, sa, dot_mode, phi, partial, grid_used, fa);
    case 16: return launch_spmm_kp<16>(st, A, alpha, beta, sa, dot_mode, phi, partial, grid_used, fa);
    case 32: return launch_spmm_kp<32>(st, A, alpha, beta, sa, dot_mode, phi, partial, grid_used, fa);
    case 64: return launch_spmm_kp<64>(st, A, alpha, beta, sa, dot_mode, phi, partial, grid_used, fa);
    default:
      set_error("launch_spmm: %d vectors per row exceed 64", ktot);
      return LQ_ERR_ARG;
  }
}

int launch_spmm_batch(hipStream_t st, int kp, int unroll, const BatchJobDev *jobs, int njobs, int grid,
                      const int *tc, int slot) {
  if (njobs <= 0 || grid <= 0) return LQ_OK;
  // T_{n-1} loaded non-temporally: the batch mix holds jobs whose vector
  // blocks exceed the MALL (0.8 % faster over the mix); LQ_SPMM_NT=0 turns it off
  static const int ntz = env_int("LQ_SPMM_NT", 1) >= 1 ? 1 : 0;
#define LQ_BU(KPV, UV)                                                                                     \
  if (kp == KPV && unroll == UV) {                                                                       \
    if (256 / KPV <= 64 && spmm_pf())                                                                    \
      hipLaunchKernelGGL((k_spmm_batch<KPV, UV, (256 / KPV <= 64)>), dim3(grid), dim3(256), 0, st, jobs,  \
                         njobs, tc, slot, ntz);                                                          \
    else                                                                                                 \
      hipLaunchKernelGGL((k_spmm_batch<KPV, UV>), dim3(grid), dim3(256), 0, st, jobs, njobs, tc, slot,   \
                         ntz);                                                                           \
    LQ_HIP_CHECK(hipGetLastError());                                                                     \
    return LQ_OK;                                                                                        \
  }
#define LQ_BK(KPV) LQ_BU(KPV, 4) LQ_BU(KPV, 8) LQ_BU(KPV, 12) LQ_BU(KPV, 16)
  LQ_BK(1)
  LQ_BK(2)
  LQ_BK(4)
  LQ_BK(8)
  LQ_BK(16)
  LQ_BK(32)
  LQ_BK(64)
#undef LQ_BK
#undef LQ_BU
  set_error("launch_spmm_batch: unsupported width %d / unroll %d", kp, unroll);
  return LQ_ERR_ARG;
}

int launch_batch_reduce(hipStream_t st, const BatchJobDev *jobs, int njobs, const int *tc, int nslots, int kp) {
  if (njobs <= 0 || nslots <= 0) return LQ_OK;
  if (kp <= 0 || kp > 256 || (kp & (kp - 1))) {
    set_error("launch_batch_reduce: width %d is not a power of two <= 256", kp);
    return LQ_ERR_ARG;
  }
  hipLaunchKernelGGL(k_batch_reduce, dim3(nslots, njobs), dim3(256), 0, st, jobs, tc, nslots, kp);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_set_int(hipStream_t st, int *p, int v) {
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, p, v);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_dot(hipStream_t st, int n, int k, const double2 *a, const double2 *b, double2 *partial, int *grid_used) {
  const int kp = pad_pow2(k);
  const int RPB = 256 / kp;
  const int groups = (n + RPB - 1) / RPB;
  const int grid = groups < kDotGridCap ? groups : kDotGridCap;
  if (grid_used) *grid_used = grid;
  switch (kp) {
#define LQ_DOT_CASE(KPV) \
  case KPV: hipLaunchKernelGGL((k_dot<KPV>), dim3(grid), dim3(256), 0, st, n, k, a, b, groups, partial); break;
    LQ_DOT_CASE(1)
    LQ_DOT_CASE(2)
    LQ_DOT_CASE(4)
    LQ_DOT_CASE(8)
    LQ_DOT_CASE(16)
    LQ_DOT_CASE(32)
    LQ_DOT_CASE(64)
#undef LQ_DOT_CASE
    default:
      set_error("launch_dot: k = %d exceeds 64", k);
      return LQ_ERR_ARG;
  }
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_reduce_partials(hipStream_t st, const double2 *partial, int nslots, int grid, int width, double2 *out) {
  if (nslots <= 0 || width <= 0) return LQ_OK;
  if (width > 256 || (width & (width - 1))) {
    set_error("launch_reduce_partials: width %d is not a power of two <= 256", width);
    return LQ_ERR_ARG;
  }
  hipLaunchKernelGGL(k_reduce_partials, dim3(nslots), dim3(256), 0, st, partial, nslots, grid, width, out);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_interleave(hipStream_t st, const double2 *src, int n, int r0, int g, double2 *dst) {
  const size_t total = (size_t)n * g;
  hipLaunchKernelGGL(k_interleave, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, src, n, r0, g, dst);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
}

int launch_copy(hipStream_t st, double2 *dst, const double2 *src, size_t count) {
  LQ_HIP_CHECK(hipMemcpyAsync(dst, src, count * sizeof(double2), hipMemcpyDeviceToDevice, st));
  return LQ_OK;
}

}  // namespace lq
