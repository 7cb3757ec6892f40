// Supercell operators from a unit-cell hopping list, built on the device.
//
// Replaces wannier2sparse's hopping_list::wrap_in_supercell and
// save_hopping_list_as_csr (utilities/wannier2sparse/src/hopping_list.cpp:
// 156-225, 227-285) followed by the solver's Hermitian completion
// (include/sparse_matrix.hpp:50-55; SparseMatrixType::ConvertFromCSR).  The
// reference replicates every hopping (cell c, orbitals e0 -> e1, value v)
// over the supercell shifts s, at row e0 + W idx(s) and column
// e1 + W idx((c + s) mod dims) (index_aliasing: idx = (x2 b1 + x1) b0 + x0),
// keeps |v| / mean|v| > 1e-10 and column >= row, sums duplicates in the order
// of its tag-string multimap (Eigen setFromTriplets: ((v1 + v2) + v3) ...),
// and writes the upper CSR; the solver then mirrors it.
//
// Here the host does the small unit-cell work once: hoppings are put in the
// multimap order, filtered, and merged into groups -- the hoppings of one
// orbital that land on the same column in every row, i.e. equal (c mod dims,
// e1) -- summed in that order, which is the order the reference's duplicates
// are summed in.  The same is done for the mirrored side: at orbital a the
// hoppings e0 -> a grouped by ((-c) mod dims, e0) give, conjugated, the lower
// triangle of the completed operator.  One GPU thread per supercell row then
// emits its row: for a row in cell t the column of a group is
// e + W idx((r + t) mod dims), and sorting a row by column is sorting by
// ((r2 + t2) mod b2, (r1 + t1) mod b1, (r0 + t0) mod b0, e).  The host sorts
// each orbital's groups by (r2, r1, r0, e) into nested blocks of equal r2,
// r1, r0; adding t rotates each level, so a thread walks the blocks of each
// level starting at the first with r >= b - t, cyclically, and the entries
// come out in column order without a sort.  Rows are the completed rows of
// the reference's operator: mirrored entries (row' < i, conjugated, kind 1),
// then stored ones (col >= i; kind 2 on the diagonal, 0 above).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "lq_internal.h"

namespace lq {
namespace {

// One side's group tree on the device.  Orbital a's level-2 blocks are
// l2[a_off[a] .. a_off[a+1]); a block is {r, first child, end child, 0}.
struct SideDev {
  const int *a_off;
  const int4 *l2, *l1, *l0;
  const int *g_e;
  const double2 *g_v;
};

struct GenArgs {
  int W;       // orbitals per cell
  int b0, b1, b2;
  int n;       // rows
  SideDev up, dn;
};

__device__ __forceinline__ int first_rotated(const int4 *blk, int begin, int end, int thresh) {
  for (int j = begin; j < end; ++j)
    if (blk[j].x >= thresh) return j - begin;
  return 0;
}

// Visits the groups of orbital a for a row in cell (t0, t1, t2) in order of
// their column (UP) or row (!UP) index; f(index, group) returns false to stop.
template <class F>
__device__ __forceinline__ void walk(const SideDev &S, const GenArgs &g, int a, int t0, int t1, int t2, F f) {
  const int b2s = S.a_off[a], b2e = S.a_off[a + 1];
  const int n2 = b2e - b2s;
  if (n2 == 0) return;
  const int s2 = first_rotated(S.l2, b2s, b2e, g.b2 - t2);
  for (int j2 = 0; j2 < n2; ++j2) {
    const int4 B2 = S.l2[b2s + (s2 + j2) % n2];
    const int w2 = B2.x + t2 >= g.b2 ? B2.x + t2 - g.b2 : B2.x + t2;
    const int n1 = B2.z - B2.y;
    const int s1 = first_rotated(S.l1, B2.y, B2.z, g.b1 - t1);
    for (int j1 = 0; j1 < n1; ++j1) {
      const int4 B1 = S.l1[B2.y + (s1 + j1) % n1];
      const int w1 = B1.x + t1 >= g.b1 ? B1.x + t1 - g.b1 : B1.x + t1;
      const int n0 = B1.z - B1.y;
      const int s0 = first_rotated(S.l0, B1.y, B1.z, g.b0 - t0);
      for (int j0 = 0; j0 < n0; ++j0) {
        const int4 B0 = S.l0[B1.y + (s0 + j0) % n0];
        const int w0 = B0.x + t0 >= g.b0 ? B0.x + t0 - g.b0 : B0.x + t0;
        const long long base = (long long)g.W * (((long long)w2 * g.b1 + w1) * g.b0 + w0);
        for (int q = B0.y; q < B0.z; ++q)
          if (!f((int)(base + S.g_e[q]), q)) return;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_sc_count(GenArgs g, int *__restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  const int a = i % g.W, tc = i / g.W;
  const int t0 = tc % g.b0, t1 = (tc / g.b0) % g.b1, t2 = tc / (g.b0 * g.b1);
  int c = 0;
  walk(g.dn, g, a, t0, t1, t2, [&](int r, int) {
    if (r >= i) return false;  // sorted: the rest lie on or above the diagonal
    ++c;
    return true;
  });
  walk(g.up, g, a, t0, t1, t2, [&](int col, int) {
    c += col >= i;
    return true;
  });
  count[i] = c;
}

__global__ __launch_bounds__(256) void k_sc_fill(GenArgs g, const int *__restrict__ rowptr, int *__restrict__ col,
                                                  double2 *__restrict__ val, signed char *__restrict__ kind) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  const int a = i % g.W, tc = i / g.W;
  const int t0 = tc % g.b0, t1 = (tc / g.b0) % g.b1, t2 = tc / (g.b0 * g.b1);
  int p = rowptr[i];
  walk(g.dn, g, a, t0, t1, t2, [&](int r, int q) {
    if (r >= i) return false;
    const double2 v = g.dn.g_v[q];
    col[p] = r;
    val[p] = make_double2(v.x, -v.y);
    kind[p] = 1;
    ++p;
    return true;
  });
  walk(g.up, g, a, t0, t1, t2, [&](int c, int q) {
    if (c < i) return true;
    col[p] = c;
    val[p] = g.up.g_v[q];
    kind[p] = c == i ? 2 : 0;
    ++p;
    return true;
  });
}

struct Group {
  int r2, r1, r0, e;
  double re, im;
};

struct SideHost {
  std::vector<int> a_off;
  std::vector<int4> l2, l1, l0;
  std::vector<int> g_e;
  std::vector<double2> g_v;
};

// Groups (already summed) of every orbital -> nested blocks.
void build_side(int W, std::vector<std::vector<Group>> &per, SideHost &S) {
  S.a_off.assign(W + 1, 0);
  for (int a = 0; a < W; ++a) {
    auto &gs = per[a];
    std::sort(gs.begin(), gs.end(), [](const Group &x, const Group &y) {
      return std::tie(x.r2, x.r1, x.r0, x.e) < std::tie(y.r2, y.r1, y.r0, y.e);
    });
    size_t q = 0;
    while (q < gs.size()) {
      const int r2 = gs[q].r2;
      int4 B2 = make_int4(r2, (int)S.l1.size(), 0, 0);
      while (q < gs.size() && gs[q].r2 == r2) {
        const int r1 = gs[q].r1;
        int4 B1 = make_int4(r1, (int)S.l0.size(), 0, 0);
        while (q < gs.size() && gs[q].r2 == r2 && gs[q].r1 == r1) {
          const int r0 = gs[q].r0;
          int4 B0 = make_int4(r0, (int)S.g_e.size(), 0, 0);
          while (q < gs.size() && gs[q].r2 == r2 && gs[q].r1 == r1 && gs[q].r0 == r0) {
            S.g_e.push_back(gs[q].e);
            S.g_v.push_back(make_double2(gs[q].re, gs[q].im));
            ++q;
          }
          B0.z = (int)S.g_e.size();
          S.l0.push_back(B0);
        }
        B1.z = (int)S.l0.size();
        S.l1.push_back(B1);
      }
      B2.z = (int)S.l1.size();
      S.l2.push_back(B2);
    }
    S.a_off[a + 1] = (int)S.l2.size();
  }
}

struct DevBuf {
  std::vector<void *> ptrs;
  ~DevBuf() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <class T>
  int put(const std::vector<T> &h, const T **out) {
    void *d = nullptr;
    const size_t bytes = sizeof(T) * std::max<size_t>(h.size(), 1);
    LQ_HIP_CHECK(hipMalloc(&d, bytes));
    ptrs.push_back(d);
    if (!h.empty()) LQ_HIP_CHECK(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    *out = (const T *)d;
    return LQ_OK;
  }
  int alloc(size_t bytes, void **out) {
    LQ_HIP_CHECK(hipMalloc(out, std::max<size_t>(bytes, 1)));
    ptrs.push_back(*out);
    return LQ_OK;
  }
};

int upload_side(DevBuf &buf, const SideHost &h, SideDev &d) {
  LQ_TRY(buf.put(h.a_off, &d.a_off));
  LQ_TRY(buf.put(h.l2, &d.l2));
  LQ_TRY(buf.put(h.l1, &d.l1));
  LQ_TRY(buf.put(h.l0, &d.l0));
  LQ_TRY(buf.put(h.g_e, &d.g_e));
  LQ_TRY(buf.put(h.g_v, &d.g_v));
  return LQ_OK;
}

inline int wrap(long long x, int b) { return (int)(((x % b) + b) % b); }

}  // namespace
}  // namespace lq

using namespace lq;

extern "C" int lq_csr_from_hoppings(lq_ctx ctx, int nwann, int nhop, const int *cell, const int *edge,
                                    const double *val, const int *dims, lq_csr *out) try {
  if (!ctx || !out || nwann <= 0 || nhop < 0 || !dims || (nhop > 0 && (!cell || !edge || !val))) {
    set_error("lq_csr_from_hoppings: invalid arguments");
    return LQ_ERR_ARG;
  }
  *out = nullptr;
  const int b0 = dims[0], b1 = dims[1], b2 = dims[2];
  if (b0 <= 0 || b1 <= 0 || b2 <= 0) {
    set_error("lq_csr_from_hoppings: supercell dimensions must be positive (%d %d %d)", b0, b1, b2);
    return LQ_ERR_ARG;
  }
  const long long ncell = (long long)b0 * b1 * b2;
  if (ncell * nwann > 0x7ffffffeLL) {
    set_error("lq_csr_from_hoppings: %lld rows exceed the 32-bit index range", ncell * nwann);
    return LQ_ERR_ARG;
  }
  const int W = nwann, n = (int)(ncell * nwann);
  for (int h = 0; h < nhop; ++h)
    if (edge[2 * h] < 0 || edge[2 * h] >= W || edge[2 * h + 1] < 0 || edge[2 * h + 1] >= W) {
      set_error("lq_csr_from_hoppings: hopping %d joins orbitals %d -> %d outside [0, %d)", h, edge[2 * h],
                edge[2 * h + 1], W);
      return LQ_ERR_ARG;
    }
  // the reference's multimap order: tag "c0 c1 c2 e0 e1 ", equal tags in input order
  std::vector<std::pair<std::string, int>> tags(nhop);
  char buf[96];
  for (int h = 0; h < nhop; ++h) {
    snprintf(buf, sizeof buf, "%d %d %d %d %d ", cell[3 * h], cell[3 * h + 1], cell[3 * h + 2], edge[2 * h],
             edge[2 * h + 1]);
    tags[h] = {buf, h};
  }
  std::stable_sort(tags.begin(), tags.end(),
                   [](const std::pair<std::string, int> &x, const std::pair<std::string, int> &y) {
                     return x.first < y.first;
                   });
  // mean |v| over the list (each hopping appears once per supercell shift)
  double avg = 0.0;
  for (const auto &t : tags) {
    const double re = val[2 * t.second], im = val[2 * t.second + 1];
    avg += std::sqrt(re * re + im * im);
  }
  avg /= nhop;  // NaN for an empty list: nothing passes the filter below
  std::vector<std::vector<Group>> up(W), dn(W);
  std::vector<std::map<std::tuple<int, int, int, int>, int>> up_ix(W), dn_ix(W);
  auto add = [](std::vector<Group> &gs, std::map<std::tuple<int, int, int, int>, int> &ix, int r0, int r1, int r2,
                int e, double re, double im) {
    auto key = std::make_tuple(r2, r1, r0, e);
    auto it = ix.find(key);
    if (it == ix.end()) {
      ix.emplace(key, (int)gs.size());
      gs.push_back({r2, r1, r0, e, re, im});
    } else {  // setFromTriplets' duplicate sum, in list order
      gs[it->second].re += re;
      gs[it->second].im += im;
    }
  };
  for (const auto &t : tags) {
    const int h = t.second;
    const double re = val[2 * h], im = val[2 * h + 1];
    if (!(std::sqrt(re * re + im * im) / avg > 1e-10)) continue;
    const int c0 = cell[3 * h], c1 = cell[3 * h + 1], c2 = cell[3 * h + 2];
    const int e0 = edge[2 * h], e1 = edge[2 * h + 1];
    add(up[e0], up_ix[e0], wrap(c0, b0), wrap(c1, b1), wrap(c2, b2), e1, re, im);
    add(dn[e1], dn_ix[e1], wrap(-(long long)c0, b0), wrap(-(long long)c1, b1), wrap(-(long long)c2, b2), e0, re, im);
  }
  SideHost hu, hd;
  build_side(W, up, hu);
  build_side(W, dn, hd);

  DeviceGuard guard(ctx->device);
  DevBuf dbuf;
  GenArgs g;
  g.W = W;
  g.b0 = b0;
  g.b1 = b1;
  g.b2 = b2;
  g.n = n;
  LQ_TRY(upload_side(dbuf, hu, g.up));
  LQ_TRY(upload_side(dbuf, hd, g.dn));
  int *d_count = nullptr;
  LQ_TRY(dbuf.alloc(sizeof(int) * ((size_t)n + 1), (void **)&d_count));
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipEvent_t ev_count;
  LQ_TRY(time_begin(ctx, kCatPost, &ev_count));
  hipLaunchKernelGGL(k_sc_count, dim3(blocks), dim3(256), 0, ctx->stream, g, d_count);
  LQ_HIP_CHECK(hipGetLastError());
  LQ_TRY(time_end(ctx, kCatPost, ev_count, 4.0 * n));
  lq_csr A = new lq_csr_s();
  A->n = n;
  A->fill = LQ_FILL_HERMITIAN_UPPER;
  A->rowptr.assign((size_t)n + 1, 0);
  int rc = LQ_OK;
  DevCsr dsrc;  // the completed CSR on the device, for the sliced-ELL layout
  do {
    if (hipMemcpyAsync(A->rowptr.data() + 1, d_count, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
      set_error("lq_csr_from_hoppings: copying row counts failed");
      rc = LQ_ERR_HIP;
      break;
    }
    long long total = 0;
    for (int i = 0; i < n; ++i) {
      total += A->rowptr[i + 1];
      if (total > 0x7fffffffLL) break;
      A->rowptr[i + 1] = (int)total;
    }
    if (total > 0x7fffffffLL) {
      set_error("lq_csr_from_hoppings: more than 2^31 - 1 completed entries");
      rc = LQ_ERR_ARG;
      break;
    }
    int *d_rowptr = d_count;  // reused: counts are on the host now
    int *d_col = nullptr;
    double2 *d_val = nullptr;
    signed char *d_kind = nullptr;
    if ((rc = dbuf.alloc(sizeof(int) * (size_t)total, (void **)&d_col)) != LQ_OK) break;
    if ((rc = dbuf.alloc(sizeof(double2) * (size_t)total, (void **)&d_val)) != LQ_OK) break;
    if ((rc = dbuf.alloc((size_t)total, (void **)&d_kind)) != LQ_OK) break;
    if (hipMemcpyAsync(d_rowptr, A->rowptr.data(), sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice,
                       ctx->stream) != hipSuccess) {
      set_error("lq_csr_from_hoppings: copying row pointers failed");
      rc = LQ_ERR_HIP;
      break;
    }
    hipEvent_t ev_fill;
    if ((rc = time_begin(ctx, kCatPost, &ev_fill)) != LQ_OK) break;
    hipLaunchKernelGGL(k_sc_fill, dim3(blocks), dim3(256), 0, ctx->stream, g, d_rowptr, d_col, d_val, d_kind);
    if (hipGetLastError() != hipSuccess) {
      set_error("lq_csr_from_hoppings: fill launch failed");
      rc = LQ_ERR_HIP;
      break;
    }
    // bytes written: column, value and kind per completed entry, plus the row pointers read
    if ((rc = time_end(ctx, kCatPost, ev_fill, 21.0 * (double)total + 4.0 * n)) != LQ_OK) break;
    A->col.resize((size_t)total);
    A->val.resize((size_t)total);
    A->kind.resize((size_t)total);
    // through the context's pinned staging buffer (two halves of kStageHalf
    // bytes whatever the operator's size): the copy of one chunk from the
    // device overlaps host threads moving the previous chunk into the
    // operator's arrays
    const size_t bc = sizeof(int) * (size_t)total, bv = sizeof(double2) * (size_t)total, bk = (size_t)total;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
      set_error("lq_csr_from_hoppings: generator kernels failed");
      rc = LQ_ERR_HIP;
      break;
    }
    if (total > 0) {
      rc = staged_d2h(ctx, {{d_col, A->col.data(), bc}, {d_val, A->val.data(), bv}, {d_kind, A->kind.data(), bk}});
      if (rc != LQ_OK) break;
    }
    dsrc.rowptr = d_rowptr;
    dsrc.col = d_col;
    dsrc.val = d_val;
  } while (false);
  if (rc != LQ_OK) {
    delete A;
    return rc;
  }
  return csr_adopt(ctx, A, out, dsrc.rowptr ? &dsrc : nullptr);
} LQ_ABI_CATCH
