// lq_api.hip -- the C ABI (include/linqt_amd.h): contexts, operators and the
// moment pipelines built from the kernels in lq_spmv / lq_contract / lq_post.
#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <numeric>
#include <thread>
#include <vector>

#include "lq_internal.h"

namespace lq {

namespace {
thread_local std::string g_last_error;

constexpr double kPi = 3.14159265358979323846;

inline double2 d2(double x, double y) { return make_double2(x, y); }

}  // namespace

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int abi_fail(const char *what, int code) {
  set_error("%s", what);
  return code;
}

int ctx_pinned(lq_ctx ctx, size_t bytes, void **p) {
  auto &b = ctx->pinned;
  if (b.bytes < bytes) {
    if (b.p) (void)hipHostFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    LQ_HIP_CHECK(hipHostMalloc(&b.p, bytes));
    b.bytes = bytes;
  }
  *p = b.p;
  return LQ_OK;
}

int staged_d2h(lq_ctx ctx, const std::vector<D2H> &parts) {
  char *stage = nullptr;
  LQ_TRY(ctx_pinned(ctx, 2 * kStageHalf, (void **)&stage));
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (auto &e : ev) LQ_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  int rc = LQ_OK;
  struct Chunk {
    const char *src = nullptr;
    char *dst = nullptr;
    size_t n = 0;
    int half = 0;
  } pend;
  auto drain = [&]() -> int {
    if (!pend.n) return LQ_OK;
    if (hipEventSynchronize(ev[pend.half]) != hipSuccess) {
      set_error("staged copy to the host failed");
      return LQ_ERR_HIP;
    }
    host_copy(pend.dst, pend.src, pend.n);
    pend.n = 0;
    return LQ_OK;
  };
  int h = 0;
  for (const auto &p : parts) {
    for (size_t off = 0; off < p.bytes && rc == LQ_OK; off += kStageHalf) {
      const size_t n = std::min(kStageHalf, p.bytes - off);
      char *buf = stage + (size_t)h * kStageHalf;
      if (hipMemcpyAsync(buf, (const char *)p.dev + off, n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
          hipEventRecord(ev[h], ctx->stream) != hipSuccess) {
        set_error("staged copy to the host failed");
        rc = LQ_ERR_HIP;
        break;
      }
      rc = drain();  // the other half's chunk
      pend = {buf, (char *)p.host + off, n, h};
      h ^= 1;
    }
  }
  if (rc == LQ_OK) rc = drain();
  for (auto e : ev) (void)hipEventDestroy(e);
  return rc;
}

void host_copy(void *dst, const void *src, size_t bytes) {
  const size_t chunk = (size_t)8 << 20;
  const int T = (int)std::min<size_t>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())),
                                      (bytes + chunk - 1) / chunk);
  auto part = [&](int t) {
    const size_t a = bytes * t / T, b = bytes * (t + 1) / T;
    std::memcpy((char *)dst + a, (const char *)src + a, b - a);
  };
  if (T <= 1) {
    if (bytes) std::memcpy(dst, src, bytes);
    return;
  }
  std::vector<std::thread> th;  // parts whose thread cannot be started run here
  int t = 1;
  try {
    th.reserve(T);
    for (; t < T; ++t) th.emplace_back(part, t);
  } catch (...) {
  }
  for (int u = t; u < T; ++u) part(u);
  part(0);
  for (auto &x : th) x.join();
}

int ctx_buf(lq_ctx ctx, const char *name, size_t bytes, void **p) {
  auto &b = ctx->bufs[name];
  if (bytes == 0) bytes = 16;
  if (b.bytes < bytes) {
    if (b.p) {
      LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
      LQ_HIP_CHECK(hipFree(b.p));
      b.p = nullptr;
      b.bytes = 0;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      b.p = nullptr;
      set_error("device allocation of %zu bytes for '%s' failed: %s", bytes, name, hipGetErrorString(e));
      return LQ_ERR_NOMEM;
    }
    b.bytes = bytes;
  }
  *p = b.p;
  return LQ_OK;
}

int time_begin_on(lq_ctx ctx, hipStream_t st, int cat, hipEvent_t *ev) {
  (void)cat;
  *ev = nullptr;
  if (!ctx->timing) return LQ_OK;
  if (ctx->event_pool.empty()) {
    hipEvent_t e;
    LQ_HIP_CHECK(hipEventCreate(&e));
    ctx->event_pool.push_back(e);
  }
  *ev = ctx->event_pool.back();
  ctx->event_pool.pop_back();
  LQ_HIP_CHECK(hipEventRecord(*ev, st));
  return LQ_OK;
}

int time_end_on(lq_ctx ctx, hipStream_t st, int cat, hipEvent_t ev, double amount, int launches) {
  if (!ctx->timing || !ev) return LQ_OK;
  if (ctx->event_pool.empty()) {
    hipEvent_t e;
    LQ_HIP_CHECK(hipEventCreate(&e));
    ctx->event_pool.push_back(e);
  }
  hipEvent_t b = ctx->event_pool.back();
  ctx->event_pool.pop_back();
  LQ_HIP_CHECK(hipEventRecord(b, st));
  ctx->pending.push_back({cat, ev, b, amount, launches});
  return LQ_OK;
}

int time_begin(lq_ctx ctx, int cat, hipEvent_t *ev) { return time_begin_on(ctx, ctx->stream, cat, ev); }

int time_end(lq_ctx ctx, int cat, hipEvent_t ev, double amount, int launches) {
  return time_end_on(ctx, ctx->stream, cat, ev, amount, launches);
}

int ctx_stream2(lq_ctx ctx, hipStream_t *st) {
  if (!ctx->stream2) LQ_HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
  *st = ctx->stream2;
  return LQ_OK;
}

int ctx_sync_events(lq_ctx ctx, int n, hipEvent_t **ev) {
  while ((int)ctx->sync_events.size() < n) {
    hipEvent_t e;
    LQ_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ctx->sync_events.push_back(e);
  }
  *ev = ctx->sync_events.data();
  return LQ_OK;
}

int spmv_loop_begin(lq_ctx ctx) {
  ctx->loop_bytes = 0.0;
  ctx->loop_launches = 0;
  LQ_TRY(time_begin(ctx, kCatSpmv, &ctx->loop_ev));
  ctx->spmv_loop = 1;
  return LQ_OK;
}

int spmv_loop_end(lq_ctx ctx) {
  ctx->spmv_loop = 0;
  return time_end(ctx, kCatSpmv, ctx->loop_ev, ctx->loop_bytes, ctx->loop_launches);
}

int collect_timing(lq_ctx ctx) {
  if (ctx->pending.empty()) return LQ_OK;
  for (auto &p : ctx->pending) LQ_HIP_CHECK(hipEventSynchronize(p.b));  // events of both streams
  for (auto &p : ctx->pending) {
    float ms = 0.f;
    LQ_HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
    switch (p.cat) {
      case kCatSpmv:
        ctx->tsum.spmv_ms += ms;
        ctx->tsum.spmv_bytes += p.bytes_or_flops;
        ctx->tsum.spmv_launches += p.launches;
        break;
      case kCatContract:
        ctx->tsum.contract_ms += ms;
        ctx->tsum.contract_flops += p.bytes_or_flops;
        ctx->tsum.contract_launches += p.launches;
        break;
      default:
        ctx->tsum.post_ms += ms;
        break;
    }
    ctx->event_pool.push_back(p.a);
    ctx->event_pool.push_back(p.b);
  }
  ctx->pending.clear();
  return LQ_OK;
}

namespace {

// ---- operators ------------------------------------------------------------

void free_fused(FusedDev &f) {
  if (f.vidx) (void)hipFree(f.vidx);
  for (int o = 0; o < 2; ++o) {
    if (f.vdict[o]) (void)hipFree(f.vdict[o]);
    if (f.vdiag[o]) (void)hipFree(f.vdiag[o]);
  }
  f = FusedDev{};
}

// Value table of the dictionary form: index of a complex value in first-seen
// order, at most 256 values.  Open addressing over the bit patterns of
// (re, im) with +0 and -0 taken as equal (values compare as numbers); the
// table keeps the first-seen value.
struct DictIndex {
  static constexpr int kSlots = 1024;  // load factor <= 1/4
  unsigned long long kre[kSlots], kim[kSlots];
  short id[kSlots];
  std::vector<double2> dict;
  DictIndex() { std::fill(id, id + kSlots, (short)-1); }
  static unsigned long long bits(double v) {
    if (v == 0.0) v = 0.0;  // -0 -> +0
    unsigned long long b;
    std::memcpy(&b, &v, sizeof b);
    return b;
  }
  // index of (re, im), added if new; -1 when a 257th value would be needed
  int find_or_add(double re, double im) {
    const unsigned long long a = bits(re), b = bits(im);
    unsigned h = (unsigned)((a * 0x9E3779B97F4A7C15ull ^ b * 0xC2B2AE3D27D4EB4Full) >> 54);
    for (;; h = (h + 1) & (kSlots - 1)) {
      if (id[h] < 0) {
        if (dict.size() == 256) return -1;
        kre[h] = a;
        kim[h] = b;
        id[h] = (short)dict.size();
        dict.push_back(d2(re, im));
        return id[h];
      }
      if (kre[h] == a && kim[h] == b) return id[h];
    }
  }
};

// Aligns up to two right operators to the stored slots of H's dictionary
// form (see FusedDev).  *ok = false (nothing built) when H has no dictionary
// form, an operator has an off-diagonal entry outside H's pattern, or more
// than 256 distinct per-slot values.  Cached on H by operator id and version.
int fused_prepare(lq_csr H, const lq_csr *V, int nv, bool *ok) {
  *ok = false;
  const char *env = getenv("LQ_FUSE_OPR");
  if (env && env[0] == '0') return LQ_OK;
  if (nv < 1 || nv > 2 || H->dev.ndict == 0) return LQ_OK;
  for (int o = 0; o < nv; ++o)
    if (!V[o] || V[o]->n != H->n || V[o]->ctx != H->ctx) return LQ_OK;
  FusedDev &f = H->fused;
  bool hit = f.nv == nv && f.hver == H->version;
  for (int o = 0; o < nv && hit; ++o) hit = f.key_uid[o] == V[o]->uid && f.key_ver[o] == V[o]->version;
  if (hit) {
    *ok = true;
    return LQ_OK;
  }
  const int n = H->n;
  const int nslices = (n + kSlice - 1) / kSlice;
  std::vector<int> offd(n, 0);
  for (int i = 0; i < n; ++i)
    for (int q = H->rowptr[i]; q < H->rowptr[i + 1]; ++q) offd[i] += H->col[q] != i;
  std::vector<long long> sptr(nslices + 1, 0);
  for (int sl = 0; sl < nslices; ++sl) {
    int w = 0;
    for (int r = 0; r < kSlice; ++r) {
      const int i = sl * kSlice + r;
      if (i < n) w = std::max(w, offd[i]);
    }
    sptr[sl + 1] = sptr[sl] + (long long)w * kSlice;
  }
  if (sptr[nslices] != H->dev.dstored) return LQ_OK;
  const size_t stored = (size_t)sptr[nslices];
  std::vector<unsigned short> vidx(stored > 0 ? stored : 1, 0);
  std::vector<std::vector<double2>> dicts(nv), diags(nv);
  std::vector<bool> has_diag(nv, false);
  std::vector<std::pair<int, int>> slot_of;  // (column, slot position in row) for one row
  for (int o = 0; o < nv; ++o) {
    const lq_csr A = V[o];
    std::vector<std::complex<double>> sv(stored > 0 ? stored : 1, 0.0);
    std::vector<double2> dg(n, d2(0.0, 0.0));
    for (int i = 0; i < n; ++i) {
      slot_of.clear();
      int j = 0;
      for (int q = H->rowptr[i]; q < H->rowptr[i + 1]; ++q)
        if (H->col[q] != i) slot_of.emplace_back(H->col[q], j++);
      std::sort(slot_of.begin(), slot_of.end());
      const int sl = i / kSlice, r = i % kSlice;
      for (int q = A->rowptr[i]; q < A->rowptr[i + 1]; ++q) {
        const int c = A->col[q];
        if (c == i) {
          dg[i].x += A->val[q].real();
          dg[i].y += A->val[q].imag();
          has_diag[o] = true;
          continue;
        }
        auto it = std::lower_bound(slot_of.begin(), slot_of.end(), std::make_pair(c, -1));
        if (it == slot_of.end() || it->first != c) return LQ_OK;  // outside H's pattern
        sv[(size_t)sptr[sl] + (size_t)it->second * kSlice + r] += A->val[q];
      }
    }
    DictIndex index;
    for (size_t p = 0; p < stored; ++p) {
      const int id = index.find_or_add(sv[p].real(), sv[p].imag());
      if (id < 0) return LQ_OK;
      vidx[p] |= (unsigned short)(id << (8 * o));
    }
    dicts[o] = std::move(index.dict);
    if (dicts[o].empty()) dicts[o].push_back(d2(0.0, 0.0));
    diags[o] = std::move(dg);
  }
  free_fused(f);
  LQ_HIP_CHECK(hipMalloc(&f.vidx, sizeof(unsigned short) * vidx.size()));
  LQ_HIP_CHECK(hipMemcpy(f.vidx, vidx.data(), sizeof(unsigned short) * vidx.size(), hipMemcpyHostToDevice));
  for (int o = 0; o < nv; ++o) {
    f.nvdict[o] = (int)dicts[o].size();
    LQ_HIP_CHECK(hipMalloc(&f.vdict[o], sizeof(double2) * dicts[o].size()));
    LQ_HIP_CHECK(hipMemcpy(f.vdict[o], dicts[o].data(), sizeof(double2) * dicts[o].size(), hipMemcpyHostToDevice));
    if (has_diag[o]) {
      LQ_HIP_CHECK(hipMalloc(&f.vdiag[o], sizeof(double2) * n));
      LQ_HIP_CHECK(hipMemcpy(f.vdiag[o], diags[o].data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    }
    f.key_uid[o] = V[o]->uid;
    f.key_ver[o] = V[o]->version;
  }
  f.nv = nv;
  f.hver = H->version;
  *ok = true;
  return LQ_OK;
}

void free_dev(SellDev &d) {
  if (d.sptr) (void)hipFree(d.sptr);
  if (d.col) (void)hipFree(d.col);
  if (d.val) (void)hipFree(d.val);
  if (d.dsptr) (void)hipFree(d.dsptr);
  if (d.wcol) (void)hipFree(d.wcol);
  if (d.dict) (void)hipFree(d.dict);
  if (d.diag) (void)hipFree(d.diag);
  d = SellDev{};
}

// Dictionary form of the completed CSR (see SellDev), when the operator has
// n < 2^24 rows and at most 256 distinct off-diagonal values (a zero value
// counts when padding is needed).  Leaves d.ndict = 0 otherwise.
// Dictionary-form fill from a device CSR: one wavefront per 64-row slice,
// lane r = row r.  Off-diagonal entries in CSR order become packed words at
// dsptr[s] + 64 j + r (padding: own row, table entry zero); the row's diagonal
// entries are summed in CSR order into diag (when diag is not null).
__global__ __launch_bounds__(64) void k_dict_fill(int n, const int *__restrict__ rowptr, const int *__restrict__ col,
                                                  const double2 *__restrict__ val,
                                                  const unsigned char *__restrict__ eid, const int *__restrict__ dsptr,
                                                  unsigned zero, unsigned *__restrict__ wcol,
                                                  double2 *__restrict__ diag) {
  const int s = blockIdx.x, r = threadIdx.x, i = s * kSlice + r;
  const int base = dsptr[s], w = (dsptr[s + 1] - base) / kSlice;
  int j = 0;
  if (i < n) {
    double dx = 0.0, dy = 0.0;
    for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) {
      const int c = col[q];
      if (c == i) {
        dx += val[q].x;
        dy += val[q].y;
        continue;
      }
      wcol[(size_t)base + (size_t)j * kSlice + r] = (unsigned)c | ((unsigned)eid[q] << 24);
      ++j;
    }
    if (diag) diag[i] = make_double2(dx, dy);
  }
  for (; j < w; ++j) wcol[(size_t)base + (size_t)j * kSlice + r] = (unsigned)(i < n ? i : 0) | (zero << 24);
}

int build_dict_form(const lq_csr A, SellDev &d, const DevCsr &csr) {
  const int n = A->n;
  if (n >= (1 << 24)) return LQ_OK;
  const char *env = getenv("LQ_SPMM_DICT");
  if (env && env[0] == '0') return LQ_OK;
  const int nslices = (n + kSlice - 1) / kSlice;
  std::vector<unsigned char> eid(A->col.size() > 0 ? A->col.size() : 1);  // table index of each off-diagonal entry
  std::vector<int> offd(n, 0);
  // Value indices in first-seen order, over contiguous row ranges on several
  // threads: each range indexes its own values, then the range tables are
  // merged in range order (which is the first-seen order over all entries)
  // and the entries renumbered.
  const int T = A->col.size() < ((size_t)1 << 20)
                    ? 1
                    : (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  struct Part {
    DictIndex idx;
    bool diag = false, full = false;
    unsigned char map[256];
  };
  std::vector<Part> parts(T);
  auto rows_of = [&](int t, int &i0, int &i1) {
    i0 = (int)((long long)n * t / T);
    i1 = (int)((long long)n * (t + 1) / T);
  };
  auto index_part = [&](int t) {
    Part &P = parts[t];
    int i0, i1;
    rows_of(t, i0, i1);
    double lre = 0.0, lim = 0.0;  // last value looked up (runs of equal values are common)
    int lid = -1;
    for (int i = i0; i < i1; ++i)
      for (int q = A->rowptr[i]; q < A->rowptr[i + 1]; ++q) {
        if (A->col[q] == i) {
          P.diag = true;
          continue;
        }
        ++offd[i];
        const double re = A->val[q].real(), im = A->val[q].imag();
        if (lid < 0 || !(re == lre && im == lim)) {
          lid = P.idx.find_or_add(re, im);
          if (lid < 0) {
            P.full = true;
            return;
          }
          lre = re;
          lim = im;
        }
        eid[q] = (unsigned char)lid;
      }
  };
  auto run = [&](auto &&f) {  // ranges whose thread cannot be started run here
    std::vector<std::thread> th;
    int t = 1;
    try {
      th.reserve(T);
      for (; t < T; ++t) th.emplace_back(f, t);
    } catch (...) {
    }
    for (int u = t; u < T; ++u) f(u);
    f(0);
    for (auto &x : th) x.join();
  };
  run(index_part);
  DictIndex index;
  bool has_diag = false, renumber = false;
  for (int t = 0; t < T; ++t) {
    if (parts[t].full) return LQ_OK;
    has_diag = has_diag || parts[t].diag;
    const std::vector<double2> &loc = parts[t].idx.dict;
    for (size_t k = 0; k < loc.size(); ++k) {
      const int g = index.find_or_add(loc[k].x, loc[k].y);
      if (g < 0) return LQ_OK;
      parts[t].map[k] = (unsigned char)g;
      renumber = renumber || g != (int)k;
    }
  }
  if (renumber)
    run([&](int t) {
      int i0, i1;
      rows_of(t, i0, i1);
      const unsigned char *m = parts[t].map;
      for (int i = i0; i < i1; ++i)
        for (int q = A->rowptr[i]; q < A->rowptr[i + 1]; ++q)
          if (A->col[q] != i) eid[q] = m[eid[q]];
    });
  std::vector<int> sptr(nslices + 1, 0);
  int wmax = 0;
  bool padded = false;
  for (int s = 0; s < nslices; ++s) {
    int w = 0, wmin = 1 << 30;
    for (int r = 0; r < kSlice; ++r) {
      const int i = s * kSlice + r;
      const int len = i < n ? offd[i] : 0;
      w = std::max(w, len);
      wmin = std::min(wmin, len);
    }
    if (wmin < w) padded = true;
    wmax = std::max(wmax, w);
    const long long next = (long long)sptr[s] + (long long)w * kSlice;
    if (next > 0x7fffffffLL) return LQ_OK;
    sptr[s + 1] = (int)next;
  }
  int zero = -1;
  if (padded) {
    zero = index.find_or_add(0.0, 0.0);
    if (zero < 0) return LQ_OK;
  }
  std::vector<double2> &dict = index.dict;
  const size_t stored = (size_t)sptr[nslices];
  d.ndict = (int)std::max<size_t>(dict.size(), 1);
  if (dict.empty()) dict.push_back(d2(0.0, 0.0));
  d.dstored = (long long)stored;
  d.dwmax = wmax;
  LQ_HIP_CHECK(hipMalloc(&d.dsptr, sizeof(int) * sptr.size()));
  LQ_HIP_CHECK(hipMalloc(&d.wcol, sizeof(unsigned) * (stored > 0 ? stored : 1)));
  LQ_HIP_CHECK(hipMalloc(&d.dict, sizeof(double2) * dict.size()));
  LQ_HIP_CHECK(hipMemcpy(d.dsptr, sptr.data(), sizeof(int) * sptr.size(), hipMemcpyHostToDevice));
  LQ_HIP_CHECK(hipMemcpy(d.dict, dict.data(), sizeof(double2) * dict.size(), hipMemcpyHostToDevice));
  if (has_diag) LQ_HIP_CHECK(hipMalloc(&d.diag, sizeof(double2) * n));
  if (n > 0) {
    unsigned char *t_eid = nullptr;
    LQ_HIP_CHECK(hipMalloc(&t_eid, eid.size()));
    hipError_t e = hipMemcpy(t_eid, eid.data(), eid.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_dict_fill, dim3(nslices), dim3(kSlice), 0, 0, n, csr.rowptr, csr.col, csr.val, t_eid,
                         d.dsptr, (unsigned)zero, d.wcol, d.diag);
      e = hipGetLastError();
    }
    const hipError_t e2 = hipDeviceSynchronize();
    (void)hipFree(t_eid);
    LQ_HIP_CHECK(e);
    LQ_HIP_CHECK(e2);
  }
  return LQ_OK;
}

// Sliced-ELL fill from a device CSR: one wavefront per 64-row slice, lane r
// = row r of the slice; entry j of the row goes to sptr[s] + 64 j + r, the
// slots past the row's length get its own column and value 0.
__global__ __launch_bounds__(64) void k_sell_fill(int n, const int *__restrict__ rowptr, const int *__restrict__ col,
                                                  const double2 *__restrict__ val, const int *__restrict__ sptr,
                                                  int *__restrict__ ocol, double2 *__restrict__ oval) {
  const int s = blockIdx.x, r = threadIdx.x, i = s * kSlice + r;
  const int base = sptr[s], w = (sptr[s + 1] - base) / kSlice;
  const int q0 = i < n ? rowptr[i] : 0, len = i < n ? rowptr[i + 1] - q0 : 0;
  for (int j = 0; j < w; ++j) {
    const size_t p = (size_t)base + (size_t)j * kSlice + r;
    if (j < len) {
      ocol[p] = col[q0 + j];
      oval[p] = val[q0 + j];
    } else {
      ocol[p] = i < n ? i : 0;
      oval[p] = make_double2(0.0, 0.0);
    }
  }
}

// Lays the completed CSR out as sliced ELL (64-row slices, column-major
// inside a slice) on the device: slice offsets from the host row pointers,
// the entries by k_sell_fill from src (or from an upload of A's host CSR).
int upload_sell(lq_csr A, const DevCsr *src = nullptr) {
  const int n = A->n;
  const int nslices = (n + kSlice - 1) / kSlice;
  std::vector<int> sptr(nslices + 1, 0);
  int wmax = 0;
  for (int s = 0; s < nslices; ++s) {
    int w = 0;
    for (int r = 0; r < kSlice; ++r) {
      const int i = s * kSlice + r;
      if (i < n) w = std::max(w, A->rowptr[i + 1] - A->rowptr[i]);
    }
    wmax = std::max(wmax, w);
    const long long next = (long long)sptr[s] + (long long)w * kSlice;
    if (next > 0x7fffffffLL) {
      set_error("operator too large for 32-bit sliced-ELL offsets");
      return LQ_ERR_ARG;
    }
    sptr[s + 1] = (int)next;
  }
  const size_t stored = (size_t)sptr[nslices];
  const size_t nnz = A->col.size();
  free_dev(A->dev);
  SellDev d;
  d.n = n;
  d.nslices = nslices;
  d.stored = (long long)stored;
  d.wmax = wmax;
  dict_schedule(d);
  // the CSR on the device for both fills (A's host arrays uploaded when no
  // device copy was handed in); on an error everything allocated here is freed
  DevCsr up;
  int *t_rowptr = nullptr, *t_col = nullptr;
  double2 *t_val = nullptr;
  auto free_tmp = [&]() {
    if (t_rowptr) (void)hipFree(t_rowptr);
    if (t_col) (void)hipFree(t_col);
    if (t_val) (void)hipFree(t_val);
    t_rowptr = t_col = nullptr;
    t_val = nullptr;
  };
  hipError_t e = hipSuccess;
  auto ok = [&](hipError_t r) {
    if (e == hipSuccess && r != hipSuccess) e = r;
    return e == hipSuccess;
  };
  if (ok(hipMalloc(&d.sptr, sizeof(int) * sptr.size())) &&
      ok(hipMalloc(&d.col, sizeof(int) * (stored > 0 ? stored : 1))) &&
      ok(hipMalloc(&d.val, sizeof(double2) * (stored > 0 ? stored : 1))) &&
      ok(hipMemcpy(d.sptr, sptr.data(), sizeof(int) * sptr.size(), hipMemcpyHostToDevice)) && !src) {
    if (ok(hipMalloc(&t_rowptr, sizeof(int) * ((size_t)n + 1))) &&
        ok(hipMalloc(&t_col, sizeof(int) * (nnz > 0 ? nnz : 1))) &&
        ok(hipMalloc(&t_val, sizeof(double2) * (nnz > 0 ? nnz : 1))) &&
        ok(hipMemcpy(t_rowptr, A->rowptr.data(), sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice)) &&
        nnz > 0 && ok(hipMemcpy(t_col, A->col.data(), sizeof(int) * nnz, hipMemcpyHostToDevice)))
      ok(hipMemcpy(t_val, A->val.data(), sizeof(double2) * nnz, hipMemcpyHostToDevice));
    up.rowptr = t_rowptr;
    up.col = t_col;
    up.val = t_val;
    src = &up;
  }
  if (e == hipSuccess && stored > 0) {
    hipLaunchKernelGGL(k_sell_fill, dim3(nslices), dim3(kSlice), 0, 0, n, src->rowptr, src->col, src->val, d.sptr,
                       d.col, d.val);
    if (ok(hipGetLastError())) ok(hipDeviceSynchronize());
  }
  if (e != hipSuccess) {
    free_tmp();
    free_dev(d);
    LQ_HIP_CHECK(e);
  }
  int rc = build_dict_form(A, d, *src);
  free_tmp();
  A->dev = d;
  return rc;
}

struct Entry {
  int row, col;
  std::complex<double> v;
  signed char kind;
};

// Sorts entries by (row, col) keeping input order among equal keys, and
// fills the host CSR of A.
void set_entries(lq_csr A, std::vector<Entry> &e) {
  std::stable_sort(e.begin(), e.end(), [](const Entry &a, const Entry &b) {
    return a.row != b.row ? a.row < b.row : a.col < b.col;
  });
  A->rowptr.assign(A->n + 1, 0);
  A->col.resize(e.size());
  A->val.resize(e.size());
  A->kind.resize(e.size());
  for (size_t q = 0; q < e.size(); ++q) {
    A->rowptr[e[q].row + 1]++;
    A->col[q] = e[q].col;
    A->val[q] = e[q].v;
    A->kind[q] = e[q].kind;
  }
  for (int i = 0; i < A->n; ++i) A->rowptr[i + 1] += A->rowptr[i];
}

std::vector<Entry> get_entries(const lq_csr A) {
  std::vector<Entry> e(A->col.size());
  for (int i = 0; i < A->n; ++i)
    for (int q = A->rowptr[i]; q < A->rowptr[i + 1]; ++q) e[q] = {i, A->col[q], A->val[q], A->kind[q]};
  return e;
}

int check_op(lq_csr A, const char *what) {
  if (!A || !A->ctx) {
    set_error("%s: null operator", what);
    return LQ_ERR_ARG;
  }
  return LQ_OK;
}

long long spmv_bytes(const lq_csr A, int k, bool with_z) {
  return 20LL * (long long)A->col.size() + 4LL * (A->n + 1) + (with_z ? 48LL : 32LL) * k * (long long)A->n;
}

// Closes an SpMM timing loop left open by an early error return.
struct SpmvLoopGuard {
  lq_ctx ctx;
  ~SpmvLoopGuard() { ctx->spmv_loop = 0; }
};

// One SpMM launch, timed on its own or as part of an open SpMM loop.
// extra_bytes: algorithmic bytes of fused right-operator applies (counted as
// the separate applies they replace).
int spmm(lq_ctx ctx, lq_csr A, double2 alpha, double2 beta, const VecSet *sets, int nsets, int dot_mode = kDotNone,
         const double2 *phi = nullptr, double2 *partial = nullptr, const FusedArgs *fused = nullptr,
         double extra_bytes = 0.0) {
  int k = 0;
  for (int s = 0; s < nsets; ++s) k += sets[s].k;
  const double bytes = (double)spmv_bytes(A, k, sets[0].z != nullptr) + extra_bytes;
  if (ctx->spmv_loop) {
    ctx->loop_bytes += bytes;
    ctx->loop_launches++;
    return launch_spmm(ctx->stream, A->dev, alpha, beta, sets, nsets, dot_mode, phi, partial, nullptr, fused);
  }
  hipEvent_t ev;
  LQ_TRY(time_begin(ctx, kCatSpmv, &ev));
  LQ_TRY(launch_spmm(ctx->stream, A->dev, alpha, beta, sets, nsets, dot_mode, phi, partial, nullptr, fused));
  return time_end(ctx, kCatSpmv, ev, bytes);
}

}  // namespace

// A new operator around a host CSR already filled in (rowptr / col / val /
// kind, rows sorted, columns ascending): uploads its device forms.
int csr_adopt(lq_ctx ctx, lq_csr A, lq_csr *out, const DevCsr *src) {
  static std::atomic<unsigned long long> next_uid{1};
  A->ctx = ctx;
  A->uid = next_uid++;
  DeviceGuard g(ctx->device);
  const int rc = upload_sell(A, src);
  if (rc != LQ_OK) {
    free_dev(A->dev);
    delete A;
    return rc;
  }
  *out = A;
  return LQ_OK;
}

bool is_device_ptr(const void *p, int device) {
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice && attr.device == device;
}

// State-major phi (host, or device memory of ctx's GPU) -> device [n][g]
// interleaved block of states r0..r0+g-1.
int upload_states(lq_ctx ctx, const double *phi, int n, int r0, int g, double2 *dst) {
  if (is_device_ptr(phi, ctx->device)) return launch_interleave(ctx->stream, (const double2 *)phi, n, r0, g, dst);
  std::vector<double2> h((size_t)n * g);
  for (int v = 0; v < g; ++v) {
    const double *src = phi + 2 * ((size_t)(r0 + v) * n);
    for (int i = 0; i < n; ++i) h[(size_t)i * g + v] = d2(src[2 * i], src[2 * i + 1]);
  }
  LQ_HIP_CHECK(hipMemcpyAsync(dst, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, ctx->stream));
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return LQ_OK;
}

namespace {

size_t free_device_bytes() {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return fr;
}

// ---- captured recursion chunks ----------------------------------------------
// A run of recursion steps captured as a HIP graph (kernel nodes only) and
// replayed: one graph launch in place of a launch per step.  By default only
// the 64-step ring chunks are captured (one capture, replayed M/64 times in
// the same solve); LQ_GRAPH=1 also captures a whole recursion loop when the
// same solve comes again (measured at N = 65536, M = 1024, R = 8: replay
// 7.48 ms against 7.52 ms launched, for a 2.7 ms capture), LQ_GRAPH=0 none.
int graph_mode() {
  static const int mode = [] {
    const char *e = getenv("LQ_GRAPH");
    if (e && e[0] == '0') return 0;
    if (e && e[0] == '1') return 2;
    return 1;
  }();
  return mode;
}
bool graphs_enabled() { return graph_mode() > 0; }

// Captures the launches `body` makes on the context's stream into a graph,
// or returns the one cached under the same key (operator and its version,
// buffers, sizes).  The context keeps the most recent graph.
template <class F>
int recursion_chunk_graph(lq_ctx ctx, const GraphKey &key, F &&body, hipGraphExec_t *out) {
  if (ctx->graph.exec && ctx->graph.key == key) {
    *out = ctx->graph.exec;
    return LQ_OK;
  }
  if (ctx->graph.exec) {
    LQ_HIP_CHECK(hipGraphExecDestroy(ctx->graph.exec));
    ctx->graph.exec = nullptr;
  }
  const int saved_loop = ctx->spmv_loop;
  ctx->spmv_loop = 1;  // no timing events inside the capture
  const double bytes = ctx->loop_bytes;
  const int launches = ctx->loop_launches;
  LQ_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc = body();
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
  ctx->spmv_loop = saved_loop;
  ctx->loop_bytes = bytes;
  ctx->loop_launches = launches;
  if (rc != LQ_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) {
    set_error("recursion graph capture: %s", hipGetErrorString(e));
    return LQ_ERR_HIP;
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    set_error("recursion graph instantiate: %s", hipGetErrorString(ei));
    return LQ_ERR_HIP;
  }
  ctx->graph.exec = exec;
  ctx->graph.key = key;
  *out = exec;
  return LQ_OK;
}

// ---- the 2D pipeline --------------------------------------------------------
// For one group of g states (interleaved on the device) it stores
//   L_m = T_m(H~) OPL phi,  m < ML      (left recursion, src/chebyshev_moments.cpp:5-22)
//   R^o_n = OPR_o T_n(H~) phi, n < MR   (right recursion + Multiply, src/chebyshev_vectors.cpp:24-51)
// and accumulates raw_o += L^H R^o (K = n*g).
struct Pipe2D {
  lq_ctx ctx;
  lq_csr H, OPL;
  int nopr;
  const lq_csr *OPR;
  int n, ML, MR;
  double2 *raw;  // nopr tables of ML x MR (ld MR)
  int precision = LQ_PREC_FP64;
};

// out (+)= A^H B for A: MA rows (ld lda), B: MB rows (ld ldb), K columns, in
// the given precision.  For the split-precision path A's digit planes are
// reused when slice_a is false (same A as the previous call on this context).
int contract_table(lq_ctx ctx, int precision, const double2 *A, long long lda, int MA, const double2 *B,
                   long long ldb, int MB, long long K, double2 *out, int ldo, bool accumulate, bool slice_a) {
  hipEvent_t ev;
  double2 *work;
  size_t wbytes = 0;
  int nsplit = 1;
  if (precision == LQ_PREC_FP64) {
    LQ_TRY(contract_workspace_bytes(MA, MB, K, &wbytes, &nsplit));
    LQ_TRY(ctx_buf(ctx, "cwork", wbytes, (void **)&work));
    LQ_TRY(time_begin(ctx, kCatContract, &ev));
    LQ_TRY(launch_contract_f64(ctx->stream, A, lda, MA, B, ldb, MB, K, work, nsplit));
  } else if (!precision_int8(precision)) {
    nsplit = split_nsplit(MA, MB, K);
    LQ_TRY(ctx_buf(ctx, "cwork", (size_t)nsplit * MA * MB * sizeof(double2), (void **)&work));
    LQ_TRY(time_begin(ctx, kCatContract, &ev));
    LQ_TRY(launch_contract_split(ctx->stream, precision, A, lda, MA, B, ldb, MB, K, work, nsplit));
  } else {
    const int S = precision;
    size_t qa, ea, qb, eb;
    LQ_TRY(ozaki_slice_bytes(MA, K, S, &qa, &ea));
    LQ_TRY(ozaki_slice_bytes(MB, K, S, &qb, &eb));
    signed char *QA, *QB;
    int *EA, *EB;
    LQ_TRY(ctx_buf(ctx, "ozQA", qa, (void **)&QA));
    LQ_TRY(ctx_buf(ctx, "ozEA", ea, (void **)&EA));
    LQ_TRY(ctx_buf(ctx, "ozQB", qb, (void **)&QB));
    LQ_TRY(ctx_buf(ctx, "ozEB", eb, (void **)&EB));
    nsplit = ozaki_nsplit(MA, MB, K);
    wbytes = (size_t)nsplit * MA * MB * sizeof(double2);
    LQ_TRY(ctx_buf(ctx, "cwork", wbytes, (void **)&work));
    LQ_TRY(time_begin(ctx, kCatContract, &ev));
    if (slice_a) LQ_TRY(launch_slice(ctx->stream, A, lda, MA, K, S, QA, EA));
    LQ_TRY(launch_slice(ctx->stream, B, ldb, MB, K, S, QB, EB));
    LQ_TRY(launch_contract_i8(ctx->stream, QA, EA, MA, QB, EB, MB, K, S, work, nsplit));
  }
  LQ_TRY(time_end(ctx, kCatContract, ev, 8.0 * MA * (double)MB * (double)K));
  LQ_TRY(time_begin(ctx, kCatPost, &ev));
  LQ_TRY(launch_contract_reduce(ctx->stream, work, nsplit, MA, MB, out, ldo, accumulate ? 1 : 0));
  return time_end(ctx, kCatPost, ev, 0);
}

int pipe2d_group(Pipe2D &P, const double *phi, int r0, int g, bool accumulate) {
  lq_ctx ctx = P.ctx;
  const int n = P.n;
  const size_t vec = (size_t)n * g;
  double2 *phi_d, *Ls, *Rw, *Rs;
  LQ_TRY(ctx_buf(ctx, "phi", vec * sizeof(double2), (void **)&phi_d));
  LQ_TRY(ctx_buf(ctx, "Lstore", vec * (size_t)P.ML * sizeof(double2), (void **)&Ls));
  LQ_TRY(ctx_buf(ctx, "Rwork", 2 * vec * sizeof(double2), (void **)&Rw));
  LQ_TRY(ctx_buf(ctx, "Rstore", vec * (size_t)P.MR * P.nopr * sizeof(double2), (void **)&Rs));
  LQ_TRY(upload_states(ctx, phi, n, r0, g, phi_d));

  const double2 one = d2(1, 0), zero = d2(0, 0), two = d2(2, 0), mone = d2(-1, 0);
  double2 *Rcur = Rw, *Rprev = Rw + vec;  // Rcur = T_n, Rprev = T_{n-1}
  auto L = [&](int m) { return Ls + (size_t)m * vec; };
  auto RS = [&](int o, int m) { return Rs + ((size_t)o * P.MR + m) * vec; };

  SpmvLoopGuard guard{ctx};
  LQ_TRY(spmv_loop_begin(ctx));
  // T0 of the left recursion: OPL * phi
  if (P.OPL) {
    VecSet s{phi_d, nullptr, L(0), g};
    LQ_TRY(spmm(ctx, P.OPL, one, zero, &s, 1));
  } else {
    LQ_TRY(launch_copy(ctx->stream, L(0), phi_d, vec));
  }
  LQ_TRY(launch_copy(ctx->stream, Rprev, phi_d, vec));  // right T0 = phi
  auto apply_opr = [&](const double2 *Tn, int m) -> int {
    for (int o = 0; o < P.nopr; ++o) {
      if (P.OPR[o]) {
        VecSet s{Tn, nullptr, RS(o, m), g};
        LQ_TRY(spmm(ctx, P.OPR[o], one, zero, &s, 1));
      } else {
        LQ_TRY(launch_copy(ctx->stream, RS(o, m), Tn, vec));
      }
    }
    return LQ_OK;
  };
  const int Mmax = std::max(P.ML, P.MR);
  // Right-operator applies fused into the H step when every OPR's pattern lies
  // inside H's: the step that reads T_m as x also writes OPR_o T_m.
  bool fuse = false;
  if (P.nopr <= 2 && P.MR > 1) LQ_TRY(fused_prepare(P.H, P.OPR, P.nopr, &fuse));
  if (fuse) {
    FusedArgs fa;
    fa.nv = P.nopr;
    fa.vidx = P.H->fused.vidx;
    double vbytes = 0.0;
    for (int o = 0; o < P.nopr; ++o) {
      fa.vdict[o] = P.H->fused.vdict[o];
      fa.nvdict[o] = P.H->fused.nvdict[o];
      fa.vdiag[o] = P.H->fused.vdiag[o];
      vbytes += (double)spmv_bytes(P.OPR[o], g, false);
    }
    for (int m = 0; m + 1 < Mmax; ++m) {
      VecSet sets[2];
      int ns = 0;
      if (m + 1 < P.ML) sets[ns++] = m == 0 ? VecSet{L(0), nullptr, L(1), g} : VecSet{L(m), L(m - 1), L(m + 1), g};
      const bool rstep = m + 1 < P.MR;
      if (rstep) {
        // x = T_m (Rcur, or T_0 in Rprev at m = 0); V_o T_m -> RS(o, m)
        sets[ns++] = m == 0 ? VecSet{Rprev, nullptr, Rcur, g} : VecSet{Rcur, Rprev, Rprev, g};
        fa.set = ns - 1;
        for (int o = 0; o < P.nopr; ++o) fa.yv[o] = RS(o, m);
      }
      if (m == 0)
        LQ_TRY(spmm(ctx, P.H, one, zero, sets, ns, kDotNone, nullptr, nullptr, rstep ? &fa : nullptr,
                    rstep ? vbytes : 0.0));
      else
        LQ_TRY(spmm(ctx, P.H, two, mone, sets, ns, kDotNone, nullptr, nullptr, rstep ? &fa : nullptr,
                    rstep ? vbytes : 0.0));
      if (rstep && m > 0) std::swap(Rcur, Rprev);
    }
    // the last right slot, T_{MR-1}, is not the x of any step
    LQ_TRY(apply_opr(P.MR > 1 ? Rcur : Rprev, P.MR - 1));
    LQ_TRY(spmv_loop_end(ctx));
    for (int o = 0; o < P.nopr; ++o)
      LQ_TRY(contract_table(ctx, P.precision, Ls, (long long)vec, P.ML, RS(o, 0), (long long)vec, P.MR,
                            (long long)vec, P.raw + (size_t)o * P.ML * P.MR, P.MR, accumulate, o == 0));
    return LQ_OK;
  }
  if (P.MR > 0) LQ_TRY(apply_opr(Rprev, 0));
  // T1 = H~ T0 on both sides
  if (Mmax > 1) {
    VecSet sets[2];
    int ns = 0;
    if (P.ML > 1) sets[ns++] = VecSet{L(0), nullptr, L(1), g};
    if (P.MR > 1) sets[ns++] = VecSet{Rprev, nullptr, Rcur, g};
    LQ_TRY(spmm(ctx, P.H, one, zero, sets, ns));
    if (P.MR > 1) LQ_TRY(apply_opr(Rcur, 1));
  }
  // T_{m+1} = 2 H~ T_m - T_{m-1}
  for (int m = 1; m + 1 < Mmax; ++m) {
    VecSet sets[2];
    int ns = 0;
    if (m + 1 < P.ML) sets[ns++] = VecSet{L(m), L(m - 1), L(m + 1), g};
    const bool rstep = m + 1 < P.MR;
    if (rstep) sets[ns++] = VecSet{Rcur, Rprev, Rprev, g};
    LQ_TRY(spmm(ctx, P.H, two, mone, sets, ns));
    if (rstep) {
      std::swap(Rcur, Rprev);
      LQ_TRY(apply_opr(Rcur, m + 1));
    }
  }
  LQ_TRY(spmv_loop_end(ctx));
  // raw_o (+)= L^H R^o
  for (int o = 0; o < P.nopr; ++o)
    LQ_TRY(contract_table(ctx, P.precision, Ls, (long long)vec, P.ML, RS(o, 0), (long long)vec, P.MR, (long long)vec,
                          P.raw + (size_t)o * P.ML * P.MR, P.MR, accumulate, o == 0));
  return LQ_OK;
}

// ---- the streamed 2D pipeline ------------------------------------------------
// The same table with bounded memory: the left vectors L_m are stored for a
// pass of BL orders at a time, and the right products R^o_n = OPR_o T_n phi are
// written in blocks of BR orders into nbuf block buffers.  Each right block is
// contracted against the pass's left block (one GEMM per pair of operators,
// reduced into the operators' tables) on the context's second stream as soon
// as both are complete, while the recursion on the first stream goes on into
// the next buffer.  With BL < ML the right recursion restarts for every pass,
// as the reference's BATCH_SIZE loop recomputes its right vectors for every
// left block (src/chebyshev_solver.cpp:45-82); each table entry is still one
// sum over K = n*g of conj(L_m) R_n.
struct StreamPlan {
  int BL = 0, BR = 128, nbuf = 2;
};

// Device buffers of one contraction of a streamed block (allocated before any
// work of the group is queued: a buffer must not be freed while the second
// stream may still read it).
struct CWork {
  double2 *work = nullptr;
  signed char *QA = nullptr, *QB = nullptr;
  int *EA = nullptr, *EB = nullptr;
};

int cwork_alloc(lq_ctx ctx, int precision, int MAmax, int MBmax, long long K, const std::vector<std::pair<int, int>> &shapes,
                CWork *w) {
  size_t wbytes = 0;
  for (auto &sh : shapes) {
    size_t b = 0;
    int ns = 1;
    if (precision == LQ_PREC_FP64) {
      LQ_TRY(contract_workspace_bytes(sh.first, sh.second, K, &b, &ns));
    } else if (!precision_int8(precision)) {
      ns = split_nsplit(sh.first, sh.second, K);
      b = (size_t)ns * sh.first * sh.second * sizeof(double2);
    } else {
      ns = ozaki_nsplit(sh.first, sh.second, K);
      b = (size_t)ns * sh.first * sh.second * sizeof(double2);
    }
    wbytes = std::max(wbytes, b);
  }
  LQ_TRY(ctx_buf(ctx, "cwork", wbytes, (void **)&w->work));
  if (precision_int8(precision)) {
    size_t qa, ea, qb, eb;
    LQ_TRY(ozaki_slice_bytes(MAmax, K, precision, &qa, &ea));
    LQ_TRY(ozaki_slice_bytes(MBmax, K, precision, &qb, &eb));
    LQ_TRY(ctx_buf(ctx, "ozQA", qa, (void **)&w->QA));
    LQ_TRY(ctx_buf(ctx, "ozEA", ea, (void **)&w->EA));
    LQ_TRY(ctx_buf(ctx, "ozQB", qb, (void **)&w->QB));
    LQ_TRY(ctx_buf(ctx, "ozEB", eb, (void **)&w->EB));
  }
  return LQ_OK;
}

// outs[o][m*ldo + s] (+)= sum_k conj(A[m][k]) B[o*bc + s][k] for m < MA,
// s < bc, o < MB/bc (<= 2), on stream st.  slice_a: cut A into digit planes
// (split precision; otherwise the planes of the previous call are reused).
int contract_cols(lq_ctx ctx, hipStream_t st, int precision, const CWork &w, const double2 *A, long long lda, int MA,
                  const double2 *B, long long ldb, int MB, int bc, long long K, double2 *const *outs, int ldo,
                  bool accumulate, bool slice_a) {
  hipEvent_t ev;
  int nsplit = 1;
  if (precision == LQ_PREC_FP64) {
    LQ_TRY(contract_workspace_bytes(MA, MB, K, nullptr, &nsplit));
    LQ_TRY(time_begin_on(ctx, st, kCatContract, &ev));
    LQ_TRY(launch_contract_f64(st, A, lda, MA, B, ldb, MB, K, w.work, nsplit));
  } else if (!precision_int8(precision)) {
    nsplit = split_nsplit(MA, MB, K);
    LQ_TRY(time_begin_on(ctx, st, kCatContract, &ev));
    LQ_TRY(launch_contract_split(st, precision, A, lda, MA, B, ldb, MB, K, w.work, nsplit));
  } else {
    nsplit = ozaki_nsplit(MA, MB, K);
    LQ_TRY(time_begin_on(ctx, st, kCatContract, &ev));
    if (slice_a) LQ_TRY(launch_slice(st, A, lda, MA, K, precision, w.QA, w.EA));
    LQ_TRY(launch_slice(st, B, ldb, MB, K, precision, w.QB, w.EB));
    LQ_TRY(launch_contract_i8(st, w.QA, w.EA, MA, w.QB, w.EB, MB, K, precision, w.work, nsplit));
  }
  LQ_TRY(time_end_on(ctx, st, kCatContract, ev, 8.0 * MA * (double)MB * (double)K));
  LQ_TRY(time_begin_on(ctx, st, kCatPost, &ev));
  LQ_TRY(launch_contract_reduce_cols(st, w.work, nsplit, MA, MB, bc, outs, ldo, accumulate ? 1 : 0));
  return time_end_on(ctx, st, kCatPost, ev, 0);
}

int pipe2d_streamed(Pipe2D &P, const double *phi, int r0, int g, bool accumulate, const StreamPlan &pl) {
  lq_ctx ctx = P.ctx;
  const int n = P.n, ML = P.ML, MR = P.MR, nopr = P.nopr;
  const size_t vec = (size_t)n * g;
  const int BL = std::min(pl.BL > 0 ? pl.BL : ML, ML), BR = std::min(pl.BR, MR);
  const int nbR = (MR + BR - 1) / BR;
  const int nbuf = std::max(1, std::min(pl.nbuf, nbR));
  const int npass = (ML + BL - 1) / BL;
  if (npass > 1 && BL < 2) {
    set_error("pipe2d: a left block of %d orders cannot carry the recursion", BL);
    return LQ_ERR_ARG;
  }
  auto bsize = [&](int j) { return std::min(BR, MR - j * BR); };
  double2 *phi_d, *Lb, *Lc = nullptr, *Rw, *Rb;
  LQ_TRY(ctx_buf(ctx, "phi", vec * sizeof(double2), (void **)&phi_d));
  LQ_TRY(ctx_buf(ctx, "Lstore", vec * (size_t)BL * sizeof(double2), (void **)&Lb));
  if (npass > 1) LQ_TRY(ctx_buf(ctx, "Lcarry", 2 * vec * sizeof(double2), (void **)&Lc));
  LQ_TRY(ctx_buf(ctx, "Rwork", 2 * vec * sizeof(double2), (void **)&Rw));
  LQ_TRY(ctx_buf(ctx, "Rstore", vec * (size_t)nbuf * nopr * BR * sizeof(double2), (void **)&Rb));
  // every shape a contraction of this group can take
  std::vector<std::pair<int, int>> shapes;
  for (int ma : {BL, ML - (npass - 1) * BL})
    for (int bc : {BR, bsize(nbR - 1)})
      for (int no : {1, 2})
        if (no <= nopr) shapes.push_back({ma, no * bc});
  CWork cw;
  LQ_TRY(cwork_alloc(ctx, P.precision, BL, std::min(nopr, 2) * BR, (long long)vec, shapes, &cw));
  hipStream_t st1 = ctx->stream, st2;
  LQ_TRY(ctx_stream2(ctx, &st2));
  hipEvent_t *ev;  // [0, nbuf): right buffer free; nbuf: operands ready; nbuf + 1: pass contracted
  LQ_TRY(ctx_sync_events(ctx, nbuf + 2, &ev));
  hipEvent_t &ev_ready = ev[nbuf], &ev_pass = ev[nbuf + 1];
  LQ_TRY(upload_states(ctx, phi, n, r0, g, phi_d));

  const double2 one = d2(1, 0), zero = d2(0, 0), two = d2(2, 0), mone = d2(-1, 0);
  bool fuse = false;
  if (nopr <= 2 && MR > 1) LQ_TRY(fused_prepare(P.H, P.OPR, nopr, &fuse));
  FusedArgs fa;
  double vbytes = 0.0;
  if (fuse) {
    fa.nv = nopr;
    fa.vidx = P.H->fused.vidx;
    for (int o = 0; o < nopr; ++o) {
      fa.vdict[o] = P.H->fused.vdict[o];
      fa.nvdict[o] = P.H->fused.nvdict[o];
      fa.vdiag[o] = P.H->fused.vdiag[o];
      vbytes += (double)spmv_bytes(P.OPR[o], g, false);
    }
  }
  // slot s of the right products (operator o) in its block buffer
  auto Rslot = [&](int sl, int o) {
    const int j = sl / BR;
    return Rb + ((size_t)(j % nbuf) * nopr * BR + (size_t)o * bsize(j) + sl % BR) * vec;
  };
  SpmvLoopGuard guard{ctx};
  for (int b = 0; b < npass; ++b) {
    const int mL0 = b * BL, mL1 = std::min(ML, mL0 + BL), nl = mL1 - mL0;
    auto Lslot = [&](int m) { return m >= mL0 ? Lb + (size_t)(m - mL0) * vec : Lc + (size_t)(m - (mL0 - 2)) * vec; };
    // the previous pass's contractions read Lb and the right buffers
    if (b > 0) LQ_HIP_CHECK(hipStreamWaitEvent(st1, ev_pass, 0));
    bool ldone = false, sliced = false;
    std::vector<int> pending;
    auto contract_block = [&](int j) -> int {
      const int bc = bsize(j);
      LQ_HIP_CHECK(hipEventRecord(ev_ready, st1));
      LQ_HIP_CHECK(hipStreamWaitEvent(st2, ev_ready, 0));
      const double2 *base = Rb + (size_t)(j % nbuf) * nopr * BR * vec;
      for (int o0 = 0; o0 < nopr; o0 += 2) {
        const int no = std::min(2, nopr - o0);
        double2 *outs[2];
        for (int o = 0; o < no; ++o) outs[o] = P.raw + (size_t)(o0 + o) * ML * MR + (size_t)mL0 * MR + (size_t)j * BR;
        LQ_TRY(contract_cols(ctx, st2, P.precision, cw, Lb, (long long)vec, nl, base + (size_t)o0 * bc * vec,
                             (long long)vec, no * bc, bc, (long long)vec, outs, MR, accumulate, !sliced));
        sliced = true;
      }
      LQ_HIP_CHECK(hipEventRecord(ev[j % nbuf], st2));
      return LQ_OK;
    };
    auto slot_begin = [&](int sl) -> int {  // before the first write of slot sl
      if (sl % BR == 0 && sl / BR >= nbuf) LQ_HIP_CHECK(hipStreamWaitEvent(st1, ev[(sl / BR) % nbuf], 0));
      return LQ_OK;
    };
    auto slot_done = [&](int sl) -> int {
      const int j = sl / BR;
      if (sl != j * BR + bsize(j) - 1) return LQ_OK;
      if (ldone) return contract_block(j);
      pending.push_back(j);
      return LQ_OK;
    };
    auto apply_opr = [&](const double2 *Tn, int sl) -> int {
      LQ_TRY(slot_begin(sl));
      for (int o = 0; o < nopr; ++o) {
        if (P.OPR[o]) {
          VecSet sv{Tn, nullptr, Rslot(sl, o), g};
          LQ_TRY(spmm(ctx, P.OPR[o], one, zero, &sv, 1));
        } else {
          LQ_TRY(launch_copy(st1, Rslot(sl, o), Tn, vec));
        }
      }
      return slot_done(sl);
    };
    auto left_finished = [&]() -> int {
      if (b + 1 < npass) {  // carry L(mL1 - 2), L(mL1 - 1) into the next pass
        LQ_TRY(launch_copy(st1, Lc, Lslot(mL1 - 2), vec));
        LQ_TRY(launch_copy(st1, Lc + vec, Lslot(mL1 - 1), vec));
      }
      ldone = true;
      for (int j : pending) LQ_TRY(contract_block(j));
      pending.clear();
      return LQ_OK;
    };
    LQ_TRY(spmv_loop_begin(ctx));
    int lm = mL0;  // next left order to produce
    if (b == 0) {
      if (P.OPL) {
        VecSet sv{phi_d, nullptr, Lb, g};
        LQ_TRY(spmm(ctx, P.OPL, one, zero, &sv, 1));
      } else {
        LQ_TRY(launch_copy(st1, Lb, phi_d, vec));
      }
      lm = 1;
    }
    double2 *Rcur = Rw, *Rprev = Rw + vec;  // Rcur = T_rn, Rprev = T_{rn-1} (T_0 in Rprev at rn = 0)
    LQ_TRY(launch_copy(st1, Rprev, phi_d, vec));
    int rn = 0;  // next right step: x = T_rn
    if (!fuse) LQ_TRY(apply_opr(Rprev, 0));
    if (lm >= mL1) LQ_TRY(left_finished());
    while (lm < mL1 || rn < MR - 1) {
      bool lstep = lm < mL1, rstep = rn < MR - 1;
      const int rsl = fuse ? rn : rn + 1;  // right slot this step writes
      // while the left block is unfinished, the right side only fills buffers that need no wait
      if (lstep && rstep && rsl / BR >= nbuf) rstep = false;
      const bool lfirst = lstep && lm == 1, rfirst = rstep && rn == 0;
      if (lstep && rstep && lfirst != rfirst) lstep = false;  // one alpha/beta per launch: the right side's first step alone
      VecSet sets[2];
      int ns = 0;
      if (lstep)
        sets[ns++] = lfirst ? VecSet{Lslot(0), nullptr, Lslot(1), g} : VecSet{Lslot(lm - 1), Lslot(lm - 2), Lslot(lm), g};
      const FusedArgs *fp = nullptr;
      if (rstep) {
        sets[ns++] = rfirst ? VecSet{Rprev, nullptr, Rcur, g} : VecSet{Rcur, Rprev, Rprev, g};
        if (fuse) {
          LQ_TRY(slot_begin(rn));
          fa.set = ns - 1;
          for (int o = 0; o < nopr; ++o) fa.yv[o] = Rslot(rn, o);
          fp = &fa;
        }
      }
      const bool first = lstep ? lfirst : rfirst;
      LQ_TRY(spmm(ctx, P.H, first ? one : two, first ? zero : mone, sets, ns, kDotNone, nullptr, nullptr, fp,
                  fp ? vbytes : 0.0));
      if (lstep && ++lm == mL1) LQ_TRY(left_finished());
      if (rstep) {
        if (rn > 0) std::swap(Rcur, Rprev);
        if (fuse)
          LQ_TRY(slot_done(rn));
        else
          LQ_TRY(apply_opr(Rcur, rn + 1));  // after the swap Rcur holds T_{rn+1}
        ++rn;
      }
    }
    // fused steps write V T_rn for x = T_rn: the last slot, T_{MR-1}, is applied on its own
    if (fuse) LQ_TRY(apply_opr(MR > 1 ? Rcur : Rprev, MR - 1));
    LQ_TRY(spmv_loop_end(ctx));
    LQ_HIP_CHECK(hipEventRecord(ev_pass, st2));
  }
  // the caller's later work on the first stream (epilogue, next group) follows the contractions
  LQ_HIP_CHECK(hipStreamWaitEvent(st1, ev_pass, 0));
  return LQ_OK;
}

// States per group: as many as fit (vectors per SpMM row <= 64 in total).
int choose_group(lq_ctx ctx, int R, int n, int ML, int MR, int nopr, int precision) {
  size_t per_state = ((size_t)ML + (size_t)MR * nopr + 3) * (size_t)n * sizeof(double2);
  if (precision_int8(precision)) per_state += ((size_t)ML + MR) * 2 * precision * (size_t)n;  // digit planes
  // the pipeline's own buffers from earlier calls are reused (or freed before
  // they grow), so they count as available: the group size, and with it the
  // order of the sums, then does not depend on what ran before
  size_t budget = free_device_bytes();
  for (const char *name : {"phi", "Lstore", "Rwork", "Rstore", "cwork", "ozQA", "ozEA", "ozQB", "ozEB"}) {
    auto it = ctx->bufs.find(name);
    if (it != ctx->bufs.end()) budget += it->second.bytes;
  }
  budget = budget > ((size_t)4 << 30) ? budget - ((size_t)4 << 30) : budget / 2;  // keep head room
  int g = (int)std::min<size_t>((size_t)R, per_state ? budget / per_state : (size_t)R);
  g = std::max(1, std::min(g, 32));
  const int ngroups = (R + g - 1) / g;
  return (R + ngroups - 1) / ngroups;
}

// Device bytes the pipelines may use: free memory plus the pipeline buffers
// of earlier calls (reused, or freed before they grow), less 4 GiB of head
// room.  Counting the pipeline's own buffers keeps the group size, and with it
// the order of the sums, independent of what ran before.
size_t pipe_budget(lq_ctx ctx) {
  size_t budget = free_device_bytes();
  for (const char *name : {"phi", "Lstore", "Lcarry", "Rwork", "Rstore", "cwork", "ozQA", "ozEA", "ozQB", "ozEB"}) {
    auto it = ctx->bufs.find(name);
    if (it != ctx->bufs.end()) budget += it->second.bytes;
  }
  return budget > ((size_t)4 << 30) ? budget - ((size_t)4 << 30) : budget / 2;
}

// The streamed pipeline's plan: all left orders of a pass of g states when
// that fits (g as large as fits, at most 32), else one state per group and as
// many left orders per pass as fit.  Returns LQ_ERR_NOMEM when not even two
// left orders fit.
int stream_plan(lq_ctx ctx, int R, int n, int ML, int MR, int nopr, int precision, int *g_out, StreamPlan *pl) {
  static const int br_env = [] {
    const char *e = getenv("LQ_PIPE_BR");
    return e ? std::max(16, atoi(e)) : 128;
  }();
  pl->BR = std::min(br_env, MR);
  pl->nbuf = 2;
  const size_t vb = (size_t)n * sizeof(double2);                       // one vector of one state
  const size_t rbuf = (size_t)pl->nbuf * nopr * pl->BR + 5;             // right buffers, phi, ring, carry
  const size_t planes = precision_int8(precision) ? (size_t)2 * precision * n : 0;  // digit bytes per row
  const size_t work = (size_t)512 * 128 * 128 * sizeof(double2);       // split-K partial tables (about)
  const size_t budget = pipe_budget(ctx);
  auto per_state = [&](size_t BL) { return (BL + rbuf) * vb + (BL + 2 * pl->BR) * planes; };
  size_t g = budget > work ? (budget - work) / per_state(ML) : 0;
  if (g >= 1) {
    g = std::min<size_t>({g, (size_t)R, 32});
    const int ngroups = (int)((R + g - 1) / g);
    *g_out = (R + ngroups - 1) / ngroups;
    pl->BL = ML;
    return LQ_OK;
  }
  const size_t avail = budget > work + per_state(0) ? budget - work - per_state(0) : 0;
  const size_t BL = avail / (vb + planes);
  if (BL < 2) {
    set_error("pipe2d: n = %d needs %zu bytes for two left orders and the right blocks; %zu available", n,
              per_state(2) + work, budget);
    return LQ_ERR_NOMEM;
  }
  *g_out = 1;
  pl->BL = (int)std::min<size_t>(BL, (size_t)ML);
  return LQ_OK;
}

// Resident pipeline (every stored vector of a group on the device, one
// contraction per operator after the recursion) when a group of at least one
// state fits; the streamed pipeline otherwise.  LQ_PIPE=stream selects the
// streamed pipeline, LQ_PIPE_BL caps its left orders per pass (to exercise
// the multi-pass form), LQ_PIPE_BR sets its right block (default 128).
int run_pipe2d(Pipe2D &P, const double *phi, int R) {
  static const int mode = [] {
    const char *e = getenv("LQ_PIPE");
    return e && strcmp(e, "stream") == 0 ? 1 : 0;
  }();
  static const int bl_env = [] {
    const char *e = getenv("LQ_PIPE_BL");
    return e ? std::max(2, atoi(e)) : 0;
  }();
  const size_t per_state = ((size_t)P.ML + (size_t)P.MR * P.nopr + 3) * (size_t)P.n * sizeof(double2) +
                           (precision_int8(P.precision) ? ((size_t)P.ML + P.MR) * 2 * P.precision * (size_t)P.n : 0);
  if (mode == 0 && bl_env == 0 && per_state <= pipe_budget(P.ctx)) {
    const int g = choose_group(P.ctx, R, P.n, P.ML, P.MR, P.nopr, P.precision);
    for (int r0 = 0; r0 < R; r0 += g) {
      const int gg = std::min(g, R - r0);
      LQ_TRY(pipe2d_group(P, phi, r0, gg, r0 > 0));
    }
    return LQ_OK;
  }
  int g = 1;
  StreamPlan pl;
  LQ_TRY(stream_plan(P.ctx, R, P.n, P.ML, P.MR, P.nopr, P.precision, &g, &pl));
  if (bl_env > 0) pl.BL = std::min(pl.BL, bl_env);
  for (int r0 = 0; r0 < R; r0 += g) {
    const int gg = std::min(g, R - r0);
    LQ_TRY(pipe2d_streamed(P, phi, r0, gg, r0 > 0, pl));
  }
  return LQ_OK;
}

int ord_of(int p, int B) { return (p / B) * (B - 1) + (p % B); }

// Section-masked copy of an operator: entries whose row and column fall in
// different sections are dropped, which is what MultiplySliced computes with
// num_sections > 1 (src/chebyshev_vectors.cpp:24-51).
int masked_copy(lq_csr A, int ns, lq_csr *out) {
  const int n = A->n, ss = n / ns;
  auto sec = [&](int i) { return std::min(i / ss, ns - 1); };
  std::vector<int> rp(n + 1, 0), cl;
  std::vector<double> vl;
  for (int i = 0; i < n; ++i) {
    for (int q = A->rowptr[i]; q < A->rowptr[i + 1]; ++q)
      if (sec(A->col[q]) == sec(i)) {
        cl.push_back(A->col[q]);
        vl.push_back(A->val[q].real());
        vl.push_back(A->val[q].imag());
      }
    rp[i + 1] = (int)cl.size();
  }
  return lq_csr_create(A->ctx, n, (int)cl.size(), rp.data(), cl.data(), vl.data(), LQ_FILL_GENERAL, out);
}

}  // namespace
}  // namespace lq

using namespace lq;

// ============================================================================
extern "C" {

const char *lq_last_error(void) { return g_last_error.c_str(); }

int lq_abi_version(void) { return LQ_ABI_VERSION; }

int lq_device_count(int *count) try {
  if (!count) return LQ_ERR_ARG;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *count = 0;
    set_error("hipGetDeviceCount: %s", hipGetErrorString(e));
    return LQ_ERR_NODEV;
  }
  return LQ_OK;
} LQ_ABI_CATCH

int lq_ctx_create(int device, lq_ctx *out) try {
  if (!out) {
    set_error("lq_ctx_create: null output");
    return LQ_ERR_ARG;
  }
  *out = nullptr;
  int count = 0;
  LQ_TRY(lq_device_count(&count));
  if (device < 0 || device >= count) {
    set_error("lq_ctx_create: device %d not present (%d GPUs)", device, count);
    return LQ_ERR_NODEV;
  }
  DeviceGuard g(device);
  LQ_TRY(noalloc_init(device));
  lq_ctx c = new lq_ctx_s();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    set_error("hipStreamCreate: %s", hipGetErrorString(e));
    return LQ_ERR_HIP;
  }
  *out = c;
  return LQ_OK;
} LQ_ABI_CATCH

int lq_ctx_destroy(lq_ctx ctx) try {
  if (!ctx) return LQ_OK;
  DeviceGuard g(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &kv : ctx->bufs)
    if (kv.second.p) (void)hipFree(kv.second.p);
  for (auto &p : ctx->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  for (auto e : ctx->sync_events) (void)hipEventDestroy(e);
  if (ctx->graph.exec) (void)hipGraphExecDestroy(ctx->graph.exec);
  if (ctx->pinned.p) (void)hipHostFree(ctx->pinned.p);
  if (ctx->stream2) {
    (void)hipStreamSynchronize(ctx->stream2);
    (void)hipStreamDestroy(ctx->stream2);
  }
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return LQ_OK;
} LQ_ABI_CATCH

int lq_ctx_release(lq_ctx ctx) try {
  if (!ctx) return LQ_ERR_ARG;
  DeviceGuard g(ctx->device);
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (ctx->stream2) LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream2));
  LQ_TRY(collect_timing(ctx));
  for (auto &kv : ctx->bufs)
    if (kv.second.p) LQ_HIP_CHECK(hipFree(kv.second.p));
  ctx->bufs.clear();
  if (ctx->pinned.p) LQ_HIP_CHECK(hipHostFree(ctx->pinned.p));
  ctx->pinned = {};
  if (ctx->graph.exec) LQ_HIP_CHECK(hipGraphExecDestroy(ctx->graph.exec));
  ctx->graph.exec = nullptr;
  ctx->graph.key = GraphKey{};
  ctx->graph_seen = GraphKey{};
  return LQ_OK;
} LQ_ABI_CATCH

int lq_ctx_synchronize(lq_ctx ctx) try {
  if (!ctx) return LQ_ERR_ARG;
  DeviceGuard g(ctx->device);
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (ctx->stream2) LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream2));
  return collect_timing(ctx);
} LQ_ABI_CATCH

int lq_ctx_set_timing(lq_ctx ctx, int enabled) try {
  if (!ctx) return LQ_ERR_ARG;
  ctx->timing = enabled != 0;
  return LQ_OK;
} LQ_ABI_CATCH

int lq_ctx_reset_timing(lq_ctx ctx) try {
  if (!ctx) return LQ_ERR_ARG;
  DeviceGuard g(ctx->device);
  LQ_TRY(collect_timing(ctx));
  ctx->tsum = lq_timing{};
  return LQ_OK;
} LQ_ABI_CATCH

int lq_ctx_get_timing(lq_ctx ctx, lq_timing *out) try {
  if (!ctx || !out) return LQ_ERR_ARG;
  DeviceGuard g(ctx->device);
  LQ_TRY(collect_timing(ctx));
  *out = ctx->tsum;
  return LQ_OK;
} LQ_ABI_CATCH

int lq_malloc(lq_ctx ctx, size_t bytes, void **dptr) try {
  if (!ctx || !dptr) return LQ_ERR_ARG;
  DeviceGuard g(ctx->device);
  hipError_t e = hipMalloc(dptr, bytes ? bytes : 16);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *dptr = nullptr;
    set_error("lq_malloc(%zu): %s", bytes, hipGetErrorString(e));
    return LQ_ERR_NOMEM;
  }
  return LQ_OK;
} LQ_ABI_CATCH

int lq_free(lq_ctx ctx, void *dptr) try {
  if (!ctx) return LQ_ERR_ARG;
  if (!dptr) return LQ_OK;
  DeviceGuard g(ctx->device);
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  LQ_HIP_CHECK(hipFree(dptr));
  return LQ_OK;
} LQ_ABI_CATCH

int lq_memcpy_h2d(lq_ctx ctx, void *dst, const void *src, size_t bytes) try {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return LQ_ERR_ARG;
  DeviceGuard g(ctx->device);
  LQ_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return LQ_OK;
} LQ_ABI_CATCH

int lq_memcpy_d2h(lq_ctx ctx, void *dst, const void *src, size_t bytes) try {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return LQ_ERR_ARG;
  DeviceGuard g(ctx->device);
  LQ_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return LQ_OK;
} LQ_ABI_CATCH

int lq_memcpy_d2d(lq_ctx ctx, void *dst, const void *src, size_t bytes) try {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return LQ_ERR_ARG;
  DeviceGuard g(ctx->device);
  LQ_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return LQ_OK;
} LQ_ABI_CATCH

// ---- operators ---------------------------------------------------------------
int lq_csr_create(lq_ctx ctx, int n, int nnz, const int *rowptr, const int *col, const double *val, int fill,
                  lq_csr *out) try {
  if (!out || n <= 0 || nnz < 0 || !rowptr || (nnz > 0 && (!col || !val))) {
    set_error("lq_csr_create: invalid arguments (n=%d nnz=%d)", n, nnz);
    return LQ_ERR_ARG;
  }
  if (fill != LQ_FILL_HERMITIAN_UPPER && fill != LQ_FILL_GENERAL) {
    set_error("lq_csr_create: unknown fill mode %d", fill);
    return LQ_ERR_ARG;
  }
  *out = nullptr;
  if (rowptr[0] != 0 || rowptr[n] != nnz) {
    set_error("lq_csr_create: rowptr[0]=%d rowptr[n]=%d, expected 0 and nnz=%d", rowptr[0], rowptr[n], nnz);
    return LQ_ERR_ARG;
  }
  // Every row range inside [0, nnz] before col/val are touched: a row pointer
  // that jumps past nnz and comes back would otherwise send the scans below
  // beyond the arrays.
  for (int i = 0; i < n; ++i)
    if (rowptr[i + 1] < rowptr[i] || rowptr[i + 1] > nnz) {
      set_error("lq_csr_create: rowptr decreases or exceeds nnz=%d at row %d (%d -> %d)", nnz, i, rowptr[i],
                rowptr[i + 1]);
      return LQ_ERR_ARG;
    }
  // The arrays are checked first, so a malformed operator is reported as such
  // even without a device context.
  if (!ctx) {
    set_error("lq_csr_create: no context");
    return LQ_ERR_ARG;
  }
  // Rows with non-decreasing columns (every file the reference's writers
  // produce): the completed CSR is built by counting, in O(nnz).  A row i's
  // completed entries are the mirrors of the stored (j, i), j < i, in
  // increasing j, then its own entries with col >= i in input order -- the
  // order the stable sort below produces, ties included.
  bool sorted = true;
  for (int i = 0; i < n && sorted; ++i) {
    if (rowptr[i + 1] < rowptr[i]) break;  // reported below
    for (int q = rowptr[i] + 1; q < rowptr[i + 1]; ++q)
      if (col[q] < col[q - 1]) {
        sorted = false;
        break;
      }
  }
  if (sorted) {
    std::vector<int> cnt((size_t)n + 1, 0);
    for (int i = 0; i < n; ++i) {
      if (rowptr[i + 1] < rowptr[i]) {
        set_error("lq_csr_create: rowptr decreases at row %d", i);
        return LQ_ERR_ARG;
      }
      for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) {
        const int j = col[q];
        if (j < 0 || j >= n) {
          set_error("lq_csr_create: column %d out of range in row %d", j, i);
          return LQ_ERR_ARG;
        }
        if (fill == LQ_FILL_GENERAL) {
          cnt[i + 1]++;
        } else if (j >= i) {
          cnt[i + 1]++;
          if (j != i) cnt[j + 1]++;
        }
      }
    }
    for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    lq_csr A = new lq_csr_s();
    A->n = n;
    A->fill = fill;
    A->rowptr = cnt;
    const size_t total = (size_t)cnt[n];
    A->col.resize(total);
    A->val.resize(total);
    A->kind.resize(total);
    std::vector<int> cur(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < n; ++i)
      for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) {
        const int j = col[q];
        if (fill != LQ_FILL_GENERAL && j < i) continue;  // MKL's upper-fill descriptor ignores the lower triangle
        const std::complex<double> v(val[2 * (size_t)q], val[2 * (size_t)q + 1]);
        int p = cur[i]++;
        A->col[p] = j;
        A->val[p] = v;
        A->kind[p] = (signed char)(j == i ? 2 : 0);
        if (fill != LQ_FILL_GENERAL && j != i) {  // mirror: row j gains (j, i) after its earlier mirrors
          p = cur[j]++;
          A->col[p] = i;
          A->val[p] = std::conj(v);
          A->kind[p] = 1;
        }
      }
    return csr_adopt(ctx, A, out);
  }
  std::vector<Entry> e;
  e.reserve((size_t)nnz * (fill == LQ_FILL_HERMITIAN_UPPER ? 2 : 1));
  for (int i = 0; i < n; ++i) {
    if (rowptr[i + 1] < rowptr[i]) {
      set_error("lq_csr_create: rowptr decreases at row %d", i);
      return LQ_ERR_ARG;
    }
    for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) {
      const int j = col[q];
      if (j < 0 || j >= n) {
        set_error("lq_csr_create: column %d out of range in row %d", j, i);
        return LQ_ERR_ARG;
      }
      const std::complex<double> v(val[2 * (size_t)q], val[2 * (size_t)q + 1]);
      if (fill == LQ_FILL_GENERAL) {
        e.push_back({i, j, v, (signed char)(j == i ? 2 : 0)});
      } else {
        if (j < i) continue;  // MKL's upper-fill descriptor ignores the lower triangle
        e.push_back({i, j, v, (signed char)(j == i ? 2 : 0)});
        if (j != i) e.push_back({j, i, std::conj(v), 1});
      }
    }
  }
  lq_csr A = new lq_csr_s();
  A->n = n;
  A->fill = fill;
  set_entries(A, e);
  return csr_adopt(ctx, A, out);
} LQ_ABI_CATCH

int lq_csr_export(lq_csr A, long long *nnz, int *rowptr, int *col, double *val) try {
  LQ_TRY(check_op(A, "lq_csr_export"));
  if (!nnz || (rowptr && (!col || !val))) {
    set_error("lq_csr_export: null argument");
    return LQ_ERR_ARG;
  }
  long long m = 0;
  for (signed char k : A->kind) m += k != 1;
  *nnz = m;
  if (!rowptr) return LQ_OK;
  long long q = 0;
  rowptr[0] = 0;
  for (int i = 0; i < A->n; ++i) {
    for (int p = A->rowptr[i]; p < A->rowptr[i + 1]; ++p) {
      if (A->kind[p] == 1) continue;
      col[q] = A->col[p];
      val[2 * q] = A->val[p].real();
      val[2 * q + 1] = A->val[p].imag();
      ++q;
    }
    rowptr[i + 1] = (int)q;
  }
  return LQ_OK;
} LQ_ABI_CATCH

int lq_csr_destroy(lq_csr A) try {
  if (!A) return LQ_OK;
  DeviceGuard g(A->ctx->device);
  (void)hipStreamSynchronize(A->ctx->stream);
  free_dev(A->dev);
  free_fused(A->fused);
  delete A;
  return LQ_OK;
} LQ_ABI_CATCH

// MKL's add C = a*U + b*I on the stored (upper) entries; the mirrored
// entries stay the conjugates of the scaled upper ones, and a row without a
// stored diagonal gains one (src/mkl_sparse_matrix.cpp:73-97).
int lq_csr_rescale(lq_csr A, double a_re, double a_im, double b_re, double b_im) try {
  LQ_TRY(check_op(A, "lq_csr_rescale"));
  const std::complex<double> a(a_re, a_im), b(b_re, b_im);
  std::vector<Entry> e = get_entries(A);
  std::vector<char> has_diag(A->n, 0);
  for (auto &x : e) {
    if (x.kind == 0) {
      x.v = a * x.v;
    } else if (x.kind == 1) {
      x.v = std::conj(a * std::conj(x.v));
    } else {
      x.v = a * x.v;
      if (!has_diag[x.row]) {
        x.v = x.v + b;
        has_diag[x.row] = 1;
      }
    }
  }
  for (int i = 0; i < A->n; ++i)
    if (!has_diag[i]) e.push_back({i, i, b, 2});
  DeviceGuard g(A->ctx->device);
  LQ_HIP_CHECK(hipStreamSynchronize(A->ctx->stream));
  set_entries(A, e);
  A->version++;
  return upload_sell(A);
} LQ_ABI_CATCH

int lq_csr_info(lq_csr A, int *n, long long *nnz, long long *nnz_stored) try {
  LQ_TRY(check_op(A, "lq_csr_info"));
  if (n) *n = A->n;
  if (nnz) *nnz = (long long)A->col.size();
  if (nnz_stored) *nnz_stored = A->dev.stored;
  return LQ_OK;
} LQ_ABI_CATCH

// ---- SpMM schedule tuning ------------------------------------------------------
// The recursion step's schedule knobs -- gather chunk (4/8/12/16 entries),
// grid of the dictionary-form launches (one set of resident blocks, or 2048
// blocks), prefetched row groups on/off -- change no arithmetic, only speed.
// lq_csr_tune times every combination on the operator at the given vector
// width and keeps the fastest in the operator; choices persist in a text file
// (LQ_TUNE_CACHE, default $HOME/.cache/linqt_amd/spmm_tune.txt) keyed by the
// GPU and the operator's shape, so a later process reads them instead of
// timing again.
namespace {

std::string tune_cache_path() {
  const char *e = getenv("LQ_TUNE_CACHE");
  if (e && *e) return e;
  const char *h = getenv("HOME");
  return std::string(h && *h ? h : "/tmp") + "/.cache/linqt_amd/spmm_tune.txt";
}

std::string tune_key(lq_csr A, int k, int mode) {
  hipDeviceProp_t prop;
  std::string arch = "unknown";
  if (hipGetDeviceProperties(&prop, A->ctx->device) == hipSuccess) arch = prop.gcnArchName;
  else (void)hipGetLastError();
  for (auto &c : arch)
    if (c == ' ') c = '_';
  std::ostringstream os;
  os << arch << ' ' << A->n << ' ' << A->dev.nslices << ' ' << A->dev.wmax << ' ' << A->dev.dwmax << ' '
     << (A->dev.ndict > 0 ? 1 : 0) << ' ' << k << ' ' << mode;
  return os.str();
}

bool tune_lookup(const std::string &key, int *u, int *gc, int *pf) {
  std::ifstream f(tune_cache_path());
  std::string line;
  bool found = false;
  while (std::getline(f, line)) {
    if (line.rfind("lqtune1 ", 0) != 0) continue;
    const std::string rest = line.substr(8);
    if (rest.size() > key.size() && rest.compare(0, key.size(), key) == 0 && rest[key.size()] == ' ') {
      std::istringstream is(rest.substr(key.size()));
      int a, b, c;
      // only schedules the tuner can choose (a damaged line is skipped)
      if ((is >> a >> b >> c) && (a == 4 || a == 8 || a == 12 || a == 16) && (b == 0 || b >= 256) &&
          (c == 0 || c == 1)) {
        *u = a;
        *gc = b;
        *pf = c;
        found = true;  // the last entry of a key wins
      }
    }
  }
  return found;
}

void tune_store(const std::string &key, int u, int gc, int pf, double us) {
  const std::string path = tune_cache_path();
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos) {
    // create the directory chain (ignore failures: the cache is optional)
    std::string dir = path.substr(0, slash);
    for (size_t p = 1; p <= dir.size(); ++p)
      if (p == dir.size() || dir[p] == '/') (void)mkdir(dir.substr(0, p).c_str(), 0755);
  }
  std::ofstream f(path, std::ios::app);
  if (f) f << "lqtune1 " << key << ' ' << u << ' ' << gc << ' ' << pf << ' ' << us << '\n';
}

}  // namespace

int lq_csr_tune(lq_csr A, int k, int mode, int *unroll, int *gridcap, int *prefetch, int *from_cache) try {
  LQ_TRY(check_op(A, "lq_csr_tune"));
  if (k < 1 || k > 64 || (mode != 0 && mode != 1)) {
    set_error("lq_csr_tune: k = %d (1..64) or mode = %d (0: step, 1: step with <phi|T>) out of range", k, mode);
    return LQ_ERR_ARG;
  }
  lq_ctx ctx = A->ctx;
  DeviceGuard dg(ctx->device);
  const std::string key = tune_key(A, k, mode);
  int u = 0, gc = 0, pf = 1, cached = 0;
  if (tune_lookup(key, &u, &gc, &pf)) {
    cached = 1;
  } else {
    const size_t vec = (size_t)A->n * k;
    int kp = 1;
    while (kp < k) kp <<= 1;
    double2 *x, *y, *phi = nullptr, *part = nullptr;
    LQ_TRY(ctx_buf(ctx, "tune_x", vec * sizeof(double2), (void **)&x));
    LQ_TRY(ctx_buf(ctx, "tune_y", vec * sizeof(double2), (void **)&y));
    LQ_HIP_CHECK(hipMemsetAsync(x, 0, vec * sizeof(double2), ctx->stream));
    LQ_HIP_CHECK(hipMemsetAsync(y, 0, vec * sizeof(double2), ctx->stream));
    if (mode == 1) {
      LQ_TRY(ctx_buf(ctx, "tune_phi", vec * sizeof(double2), (void **)&phi));
      LQ_TRY(ctx_buf(ctx, "tune_part", (size_t)spmm_dot_grid(A->n, kp) * kp * sizeof(double2), (void **)&part));
      LQ_HIP_CHECK(hipMemsetAsync(phi, 0, vec * sizeof(double2), ctx->stream));
    }
    const SellDev saved = A->dev;
    const bool dict = A->dev.ndict > 0;
    const std::vector<int> us = {4, 8, 12, 16};
    // the grid cap applies to dictionary-form launches without dot fusion
    const std::vector<int> gcs = dict && mode == 0 ? std::vector<int>{0, 2048} : std::vector<int>{saved.gridcap};
    const std::vector<int> pfs = dict ? std::vector<int>{1, 0} : std::vector<int>{saved.pf ? 1 : 0};
    hipEvent_t a, b;
    LQ_HIP_CHECK(hipEventCreate(&a));
    LQ_HIP_CHECK(hipEventCreate(&b));
    double best = 1e300;
    int rc = LQ_OK;
    for (int uu : us)
      for (int g : gcs)
        for (int p : pfs) {
          if (rc != LQ_OK) break;
          A->dev.unroll = uu;
          A->dev.gridcap = g;
          A->dev.pf = p != 0;
          VecSet sv{x, y, y, k};
          auto step = [&]() {
            return launch_spmm(ctx->stream, A->dev, d2(2, 0), d2(-1, 0), &sv, 1, mode == 1 ? kDotPhi : kDotNone, phi,
                               part, nullptr);
          };
          for (int w = 0; w < 3 && rc == LQ_OK; ++w) rc = step();
          if (rc != LQ_OK) break;
          const int reps = 20;
          (void)hipEventRecord(a, ctx->stream);
          for (int r = 0; r < reps && rc == LQ_OK; ++r) rc = step();
          (void)hipEventRecord(b, ctx->stream);
          if (hipEventSynchronize(b) != hipSuccess) {
            set_error("lq_csr_tune: timing failed");
            rc = LQ_ERR_HIP;
            break;
          }
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, a, b);
          const double t = (double)ms / reps;
          if (t < best * 0.99) {  // a change must win by more than 1 %
            best = t;
            u = uu;
            gc = g;
            pf = p;
          }
        }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    A->dev = saved;
    if (rc != LQ_OK) return rc;
    tune_store(key, u, gc, pf, best * 1e3);
  }
  A->dev.unroll = u;
  A->dev.gridcap = gc;
  A->dev.pf = pf != 0;
  if (unroll) *unroll = u;
  if (gridcap) *gridcap = gc;
  if (prefetch) *prefetch = pf;
  if (from_cache) *from_cache = cached;
  return LQ_OK;
} LQ_ABI_CATCH

int lq_spmv(lq_csr A, double a_re, double a_im, const double *x, double b_re, double b_im, double *y, int k) try {
  LQ_TRY(check_op(A, "lq_spmv"));
  if (!x || !y || k <= 0 || k > 64 || x == y) {
    set_error("lq_spmv: need distinct x, y and 1 <= k <= 64");
    return LQ_ERR_ARG;
  }
  lq_ctx ctx = A->ctx;
  DeviceGuard g(ctx->device);
  const bool use_z = !(b_re == 0.0 && b_im == 0.0);
  VecSet s{(const double2 *)x, use_z ? (const double2 *)y : nullptr, (double2 *)y, k};
  return spmm(ctx, A, d2(a_re, a_im), d2(b_re, b_im), &s, 1);
} LQ_ABI_CATCH

int lq_compute_mom_table_prec(lq_ctx ctx, const double *L, long long ldl, int ML, const double *R, long long ldr,
                              int MR, long long K, int precision, double *sub, int ldsub) try {
  if (!ctx || !L || !R || !sub || ML <= 0 || MR <= 0 || K <= 0 || ldl < K || ldr < K || ldsub < MR) {
    set_error("lq_compute_mom_table: invalid arguments");
    return LQ_ERR_ARG;
  }
  if (!precision_ok(precision)) {
    set_error("lq_compute_mom_table: unknown precision %d", precision);
    return LQ_ERR_ARG;
  }
  DeviceGuard dg(ctx->device);
  LQ_TRY(contract_table(ctx, precision, (const double2 *)L, ldl, ML, (const double2 *)R, ldr, MR, K, (double2 *)sub,
                        ldsub, false, true));
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return collect_timing(ctx);
} LQ_ABI_CATCH

int lq_compute_mom_table(lq_ctx ctx, const double *L, long long ldl, int ML, const double *R, long long ldr, int MR,
                         long long K, double *sub, int ldsub) try {
  return lq_compute_mom_table_prec(ctx, L, ldl, ML, R, ldr, MR, K, LQ_PREC_FP64, sub, ldsub);
} LQ_ABI_CATCH

// ---- 1D moments ----------------------------------------------------------------
int lq_spectral_moments(lq_csr H, lq_csr OP, const double *phi, int R, int M, int algo, double *mu) try {
  LQ_TRY(check_op(H, "lq_spectral_moments"));
  if (!phi || R <= 0 || M <= 0 || !mu || (OP && (OP->n != H->n || OP->ctx != H->ctx))) {
    set_error("lq_spectral_moments: invalid arguments");
    return LQ_ERR_ARG;
  }
  if (algo != LQ_ALGO_DIRECT && algo != LQ_ALGO_DOUBLING) {
    set_error("lq_spectral_moments: unknown algorithm %d", algo);
    return LQ_ERR_ARG;
  }
  if (algo == LQ_ALGO_DOUBLING && OP) {
    set_error("lq_spectral_moments: the doubling recursion needs OP = identity");
    return LQ_ERR_ARG;
  }
  lq_ctx ctx = H->ctx;
  DeviceGuard dg(ctx->device);
  const int n = H->n;
  const int g = std::min(R, 64);
  std::vector<std::complex<double>> acc(M, 0.0);
  for (int r0 = 0; r0 < R; r0 += g) {
    const int gg = std::min(g, R - r0);
    int kp = 1;
    while (kp < gg) kp <<= 1;
    const size_t vec = (size_t)n * gg;
    const int grid = spmm_dot_grid(n, kp);
    double2 *phi_d, *A, *B, *part0, *part1, *sums0, *sums1;
    LQ_TRY(ctx_buf(ctx, "phi", vec * sizeof(double2), (void **)&phi_d));
    LQ_TRY(ctx_buf(ctx, "T0", vec * sizeof(double2), (void **)&A));
    LQ_TRY(ctx_buf(ctx, "T1", vec * sizeof(double2), (void **)&B));
    LQ_TRY(upload_states(ctx, phi, n, r0, gg, phi_d));
    const double2 one = d2(1, 0), zero = d2(0, 0), two = d2(2, 0), mone = d2(-1, 0);
    if (OP) {
      VecSet s{phi_d, nullptr, A, gg};
      LQ_TRY(spmm(ctx, OP, one, zero, &s, 1));
    } else {
      LQ_TRY(launch_copy(ctx->stream, A, phi_d, vec));
    }
    if (algo == LQ_ALGO_DIRECT) {
      // Moment m's block partials go to slot m % RS of the partial buffer:
      // RS = M when all moments' partials fit in kPartBytes (one reduction at
      // the end), else a ring of kPartRing slots, reduced into sums0 each time
      // it fills.  Every slot is reduced by itself in the same order either way.
      const size_t slot = (size_t)grid * kp;
      const bool whole = (size_t)M * slot * sizeof(double2) <= kPartBytes;
      const int RS = whole ? M : kPartRing;
      LQ_TRY(ctx_buf(ctx, "part0", (size_t)RS * slot * sizeof(double2), (void **)&part0));
      LQ_TRY(ctx_buf(ctx, "sums0", (size_t)M * kp * sizeof(double2), (void **)&sums0));
      auto flush = [&](int m_end) -> int {  // moments [m_end - filled, m_end) are in the buffer
        const int m0 = ((m_end - 1) / RS) * RS;
        return launch_reduce_partials(ctx->stream, part0, m_end - m0, grid, kp, sums0 + (size_t)m0 * kp);
      };
      auto ring = [&](int m) { return part0 + (size_t)(m % RS) * slot; };
      LQ_TRY(launch_dot(ctx->stream, n, gg, phi_d, A, ring(0), nullptr));
      SpmvLoopGuard guard{ctx};
      LQ_TRY(spmv_loop_begin(ctx));
      if (M > 1) {
        VecSet s{A, nullptr, B, gg};
        LQ_TRY(spmm(ctx, H, one, zero, &s, 1, kDotPhi, phi_d, ring(1)));
      }
      // Captured recursion (graphs on): with all partials in one buffer the
      // whole loop m = 1 .. M-2 is one graph, captured when the same solve
      // (operator, buffers, sizes) comes again and replayed from then on;
      // with the ring, one full ring of steps (moments 64c .. 64c + 63)
      // always starts from the same A/B assignment (an even number of swaps
      // before it) and writes slots 0..63, so it is captured once and
      // replayed with the ring reduction between replays.
      const int nsteps = M - 2;
      if (whole && graph_mode() == 2 && nsteps >= 2) {
        const GraphKey key{H->uid, H->version, A, B, phi_d, part0, n, gg, nsteps};
        hipGraphExec_t g = nullptr;
        if (ctx->graph_seen == key || (ctx->graph.exec && ctx->graph.key == key)) {
          LQ_TRY(recursion_chunk_graph(ctx, key, [&]() -> int {
            double2 *a = A, *b = B;
            for (int m = 1; m + 1 < M; ++m) {
              VecSet s{b, a, a, gg};
              LQ_TRY(spmm(ctx, H, two, mone, &s, 1, kDotPhi, phi_d, ring(m + 1)));
              std::swap(a, b);
            }
            return LQ_OK;
          }, &g));
          LQ_HIP_CHECK(hipGraphLaunch(g, ctx->stream));
          ctx->loop_launches += nsteps;
          ctx->loop_bytes += (double)nsteps * spmv_bytes(H, gg, true);
          if (nsteps % 2) std::swap(A, B);
        } else {
          ctx->graph_seen = key;  // a repeat of this solve is captured
        }
        if (g) goto recursion_done;
      }
      {
        hipGraphExec_t chunk_graph = nullptr;
        const int graph_from = graphs_enabled() && !whole ? kPartRing : M;  // first moment a replay produces
        for (int m = 1; m + 1 < M; ++m) {
          if ((m + 1) % RS == 0) LQ_TRY(flush(m + 1));
          if (m + 1 >= graph_from && (m + 1) % kPartRing == 0 && m + kPartRing < M) {
            if (!chunk_graph) {
              GraphKey key{H->uid, H->version, A, B, phi_d, part0, n, gg, -kPartRing};  // < 0: a ring chunk
              LQ_TRY(recursion_chunk_graph(ctx, key, [&]() -> int {
                double2 *a = A, *b = B;
                for (int q = 0; q < kPartRing; ++q) {
                  VecSet s{b, a, a, gg};
                  LQ_TRY(spmm(ctx, H, two, mone, &s, 1, kDotPhi, phi_d, part0 + (size_t)q * slot));
                  std::swap(a, b);
                }
                return LQ_OK;
              }, &chunk_graph));
            }
            LQ_HIP_CHECK(hipGraphLaunch(chunk_graph, ctx->stream));
            ctx->loop_launches += kPartRing;
            ctx->loop_bytes += (double)kPartRing * spmv_bytes(H, gg, true);
            m += kPartRing - 1;  // an even count of steps: A and B end where they started
            continue;
          }
          // A <- 2 H~ B - A, then swap: after the step B holds T_{m+1}
          VecSet s{B, A, A, gg};
          LQ_TRY(spmm(ctx, H, two, mone, &s, 1, kDotPhi, phi_d, ring(m + 1)));
          std::swap(A, B);
        }
      }
    recursion_done:
      LQ_TRY(spmv_loop_end(ctx));
      hipEvent_t ev;
      LQ_TRY(time_begin(ctx, kCatPost, &ev));
      LQ_TRY(flush(M));
      LQ_TRY(time_end(ctx, kCatPost, ev, 0));
      std::vector<double2> h((size_t)M * kp);
      LQ_HIP_CHECK(hipMemcpyAsync(h.data(), sums0, sizeof(double2) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
      LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
      for (int m = 0; m < M; ++m) {
        double scal = 2.0 / R;
        if (m == 0) scal *= 0.5;
        for (int v = 0; v < gg; ++v) acc[m] += scal * std::complex<double>(h[(size_t)m * kp + v].x, h[(size_t)m * kp + v].y);
      }
    } else {
      // mu_0 = <T0|T0>, mu_1 = <T0|T1>; step n gives <T_n|T_n>, <T_n|T_n+1>
      const int nsteps = (M - 1) / 2;
      LQ_TRY(ctx_buf(ctx, "part0", (size_t)2 * grid * kp * sizeof(double2), (void **)&part0));
      LQ_TRY(ctx_buf(ctx, "sums0", (size_t)2 * kp * sizeof(double2), (void **)&sums0));
      // step st's partials go to ring slot (st - 1) % kPartRing, reduced as for the direct path
      LQ_TRY(ctx_buf(ctx, "part1", (size_t)kPartRing * grid * 2 * kp * sizeof(double2), (void **)&part1));
      LQ_TRY(ctx_buf(ctx, "sums1", (size_t)std::max(nsteps, 1) * 2 * kp * sizeof(double2), (void **)&sums1));
      const size_t slot0 = (size_t)grid * kp, slot1 = (size_t)grid * 2 * kp;
      LQ_TRY(launch_dot(ctx->stream, n, gg, A, A, part0, nullptr));
      if (M > 1) {
        VecSet s{A, nullptr, B, gg};
        LQ_TRY(spmm(ctx, H, one, zero, &s, 1));
        LQ_TRY(launch_dot(ctx->stream, n, gg, A, B, part0 + slot0, nullptr));
      }
      SpmvLoopGuard guard{ctx};
      LQ_TRY(spmv_loop_begin(ctx));
      auto flush1 = [&](int s_end) -> int {  // steps' slots [s0, s_end) (zero-based) are in the ring
        const int s0 = ((s_end - 1) / kPartRing) * kPartRing;
        return launch_reduce_partials(ctx->stream, part1, s_end - s0, grid, 2 * kp, sums1 + (size_t)s0 * 2 * kp);
      };
      for (int st = 1; st <= nsteps; ++st) {
        if (st > 1 && (st - 1) % kPartRing == 0) LQ_TRY(flush1(st - 1));
        // x = T_st (B), z = y = T_st-1 (A): A <- T_st+1
        VecSet s{B, A, A, gg};
        LQ_TRY(spmm(ctx, H, two, mone, &s, 1, kDotDoubling, nullptr, part1 + (size_t)((st - 1) % kPartRing) * slot1));
        std::swap(A, B);
      }
      LQ_TRY(spmv_loop_end(ctx));
      hipEvent_t ev;
      LQ_TRY(time_begin(ctx, kCatPost, &ev));
      LQ_TRY(launch_reduce_partials(ctx->stream, part0, 2, grid, kp, sums0));
      if (nsteps > 0) LQ_TRY(flush1(nsteps));
      LQ_TRY(time_end(ctx, kCatPost, ev, 0));
      std::vector<double2> h0((size_t)2 * kp), h1((size_t)std::max(nsteps, 1) * 2 * kp);
      LQ_HIP_CHECK(hipMemcpyAsync(h0.data(), sums0, sizeof(double2) * h0.size(), hipMemcpyDeviceToHost, ctx->stream));
      if (nsteps > 0)
        LQ_HIP_CHECK(hipMemcpyAsync(h1.data(), sums1, sizeof(double2) * (size_t)nsteps * 2 * kp, hipMemcpyDeviceToHost,
                                    ctx->stream));
      LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
      for (int v = 0; v < gg; ++v) {
        std::vector<std::complex<double>> m1(M);
        const std::complex<double> mu0(h0[v].x, h0[v].y), mu1(h0[kp + v].x, h0[kp + v].y);
        m1[0] = mu0;
        if (M > 1) m1[1] = mu1;
        for (int st = 1; st <= nsteps; ++st) {
          const double2 a = h1[((size_t)(st - 1) * 2 + 0) * kp + v], b = h1[((size_t)(st - 1) * 2 + 1) * kp + v];
          if (2 * st < M) m1[2 * st] = 2.0 * std::complex<double>(a.x, a.y) - mu0;
          if (2 * st + 1 < M) m1[2 * st + 1] = 2.0 * std::complex<double>(b.x, b.y) - mu1;
        }
        for (int m = 0; m < M; ++m) {
          double scal = 2.0 / R;
          if (m == 0) scal *= 0.5;
          acc[m] += scal * m1[m];
        }
      }
    }
  }
  for (int m = 0; m < M; ++m) {
    mu[2 * m] += acc[m].real();
    mu[2 * m + 1] += acc[m].imag();
  }
  return collect_timing(ctx);
} LQ_ABI_CATCH

// ---- 2D moments --------------------------------------------------------------------
static int check_2d_ops(lq_csr H, lq_csr OPL, int nopr, const lq_csr *OPR, const char *what) {
  LQ_TRY(check_op(H, what));
  if (nopr <= 0 || !OPR) {
    set_error("%s: need at least one right operator", what);
    return LQ_ERR_ARG;
  }
  if (OPL && (OPL->n != H->n || OPL->ctx != H->ctx)) {
    set_error("%s: OPL size/context mismatch", what);
    return LQ_ERR_ARG;
  }
  for (int o = 0; o < nopr; ++o)
    if (OPR[o] && (OPR[o]->n != H->n || OPR[o]->ctx != H->ctx)) {
      set_error("%s: OPR[%d] size/context mismatch", what, o);
      return LQ_ERR_ARG;
    }
  return LQ_OK;
}

int lq_correlation_moments(lq_csr H, lq_csr OPL, int nopr, const lq_csr *OPR, const double *phi, int R, int M0,
                           int M1, int batch, int precision, double *mu) try {
  LQ_TRY(check_2d_ops(H, OPL, nopr, OPR, "lq_correlation_moments"));
  if (!phi || R <= 0 || M0 <= 0 || M1 <= 0 || !mu) {
    set_error("lq_correlation_moments: invalid arguments");
    return LQ_ERR_ARG;
  }
  if (!precision_ok(precision)) {
    set_error("lq_correlation_moments: unknown precision %d", precision);
    return LQ_ERR_ARG;
  }
  lq_ctx ctx = H->ctx;
  DeviceGuard dg(ctx->device);
  const int B = batch > 0 ? batch : M0;
  const int Msq = std::max(M0, M1);
  const int Mt = ord_of(Msq - 1, B) + 1;  // true Chebyshev orders needed
  const bool padded = (M0 % B != 0) || (M1 % B != 0);
  const size_t tab = (size_t)Mt * Mt;
  double2 *raw, *init = nullptr, *outd;
  LQ_TRY(ctx_buf(ctx, "raw", tab * nopr * sizeof(double2), (void **)&raw));
  LQ_TRY(ctx_buf(ctx, "mu_out", (size_t)M0 * M1 * sizeof(double2), (void **)&outd));
  Pipe2D P{ctx, H, OPL, nopr, OPR, H->n, Mt, Mt, raw, precision};
  LQ_TRY(run_pipe2d(P, phi, R));
  for (int o = 0; o < nopr; ++o) {
    if (!padded) {
      LQ_TRY(ctx_buf(ctx, "mu_init", (size_t)M0 * M1 * sizeof(double2), (void **)&init));
      LQ_HIP_CHECK(hipMemcpyAsync(init, mu + 2 * (size_t)o * M0 * M1, (size_t)M0 * M1 * sizeof(double2),
                                  hipMemcpyHostToDevice, ctx->stream));
    }
    hipEvent_t ev;
    LQ_TRY(time_begin(ctx, kCatPost, &ev));
    LQ_TRY(launch_corr_epilogue(ctx->stream, raw + o * tab, Mt, padded ? nullptr : init, M0, M1, B, R, outd));
    LQ_TRY(time_end(ctx, kCatPost, ev, 0));
    LQ_HIP_CHECK(hipMemcpyAsync(mu + 2 * (size_t)o * M0 * M1, outd, (size_t)M0 * M1 * sizeof(double2),
                                hipMemcpyDeviceToHost, ctx->stream));
    LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
  return collect_timing(ctx);
} LQ_ABI_CATCH

// ---- Kubo FFT ------------------------------------------------------------------------
int lq_kubo_fft(lq_csr H, lq_csr OPL, int nopr, const lq_csr *OPR, const double *phi, int R, int M, int num_sections,
                int nump, int formula, int precision, double *final_data) try {
  LQ_TRY(check_2d_ops(H, OPL, nopr, OPR, "lq_kubo_fft"));
  const int n = H->n;
  if (!phi || R <= 0 || M <= 0 || nump <= 0 || !final_data || num_sections <= 0 || num_sections > n) {
    set_error("lq_kubo_fft: invalid arguments");
    return LQ_ERR_ARG;
  }
  if (M > nump) {
    set_error("lq_kubo_fft: M = %d > nump = %d overflows the reference's DFT inputs", M, nump);
    return LQ_ERR_UNSUPPORTED;
  }
  if (formula != LQ_KUBO_GREENWOOD && formula != LQ_KUBO_BASTIN) {
    set_error("lq_kubo_fft: unknown formula %d", formula);
    return LQ_ERR_ARG;
  }
  if (!precision_ok(precision)) {
    set_error("lq_kubo_fft: unknown precision %d", precision);
    return LQ_ERR_ARG;
  }
  if (num_sections > 1 && n % num_sections != 0 && R > 1) {
    set_error("lq_kubo_fft: with N %% num_sections != 0 and R > 1 the reference reads stale section tails "
              "(Greenwood_FFTs.cpp:8-11); unsupported");
    return LQ_ERR_UNSUPPORTED;
  }
  lq_ctx ctx = H->ctx;
  DeviceGuard dg(ctx->device);
  std::vector<lq_csr> opr(OPR, OPR + nopr), owned;
  if (num_sections > 1) {
    for (int o = 0; o < nopr; ++o) {
      if (!opr[o]) continue;  // identity stays block diagonal
      lq_csr m = nullptr;
      LQ_TRY(masked_copy(opr[o], num_sections, &m));
      owned.push_back(m);
      opr[o] = m;
    }
  }
  const size_t tab = (size_t)M * M;
  double2 *raw, *fd;
  double *coef;
  int rc = ctx_buf(ctx, "raw", tab * nopr * sizeof(double2), (void **)&raw);
  if (rc == LQ_OK) rc = ctx_buf(ctx, "kubo_fd", (size_t)2 * nump * nopr * sizeof(double2), (void **)&fd);
  if (rc == LQ_OK) rc = ctx_buf(ctx, "kubo_coef", (size_t)M * sizeof(double), (void **)&coef);
  if (rc == LQ_OK) {
    Pipe2D P{ctx, H, OPL, nopr, opr.data(), n, M, M, raw, precision};
    rc = run_pipe2d(P, phi, R);
  }
  if (rc == LQ_OK) {
    // c_m = (2 - delta_m0) * Jackson_kernel(m, M) (src/Greenwood_FFTs.cpp:17-18)
    std::vector<double> c(M);
    for (int m = 0; m < M; ++m)
      c[m] = (2 - (m == 0)) *
             (1 / (M + 1.0) *
              ((M - m + 1.0) * cos(kPi * m / (M + 1.0)) + sin(kPi * m / (M + 1.0)) / tan(kPi / (M + 1.0))));
    hipError_t e = hipMemcpyAsync(coef, c.data(), sizeof(double) * M, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
      set_error("lq_kubo_fft: coefficient upload: %s", hipGetErrorString(e));
      rc = LQ_ERR_HIP;
    }
    hipEvent_t ev;
    if (rc == LQ_OK) rc = time_begin(ctx, kCatPost, &ev);
    if (rc == LQ_OK) rc = launch_scale(ctx->stream, raw, tab * nopr, 1.0 / R);
    for (int o = 0; o < nopr && rc == LQ_OK; ++o)
      rc = launch_kubo_pw(ctx->stream, raw + o * tab, M, M, nump, formula, coef, fd + (size_t)o * 2 * nump);
    if (rc == LQ_OK) rc = time_end(ctx, kCatPost, ev, 0);
    if (rc == LQ_OK) {
      e = hipMemcpyAsync(final_data, fd, (size_t)2 * nump * nopr * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) {
        set_error("lq_kubo_fft: result download: %s", hipGetErrorString(e));
        rc = LQ_ERR_HIP;
      }
    }
  }
  for (auto m : owned) lq_csr_destroy(m);
  if (rc != LQ_OK) return rc;
  return collect_timing(ctx);
} LQ_ABI_CATCH

// ---- time-dependent correlations ------------------------------------------------
// chebyshev::TimeDependentCorrelations (src/chebyshev_solver.cpp:231-275) with
// MomentsTD::Evolve (src/chebyshev_momentsTD.cpp:4-45).  States run in groups;
// the left and right vectors of a group evolve as the two sets of one SpMM per
// Bessel term.
int lq_time_correlations(lq_csr H, lq_csr OPL, lq_csr OPR, const double *phi, int R, int M, int nT, double dt,
                         double W, double Ec, double *mu) try {
  LQ_TRY(check_op(H, "lq_time_correlations"));
  if (!phi || R <= 0 || M <= 0 || nT <= 0 || !mu || !(W > 0) || (OPL && (OPL->n != H->n || OPL->ctx != H->ctx)) ||
      (OPR && (OPR->n != H->n || OPR->ctx != H->ctx))) {
    set_error("lq_time_correlations: invalid arguments");
    return LQ_ERR_ARG;
  }
  lq_ctx ctx = H->ctx;
  DeviceGuard dg(ctx->device);
  const int n = H->n;
  // Evolve's coefficients: c_0 = J_0(x), c_k = 2 (-i)^k J_k(x) while |J_k| > 1e-10,
  // x = dt a / hbar (the reference's besselJ is libm's j0 / j1 / jn)
  const double kHbar = 0.6582119624, a = W / 2.0, x = dt * (a / kHbar), tol = 1e-10;
  std::vector<std::complex<double>> coef;
  {
    std::complex<double> nIp(1.0, 0.0);
    double Jn = 0.5 * j0(x), cut = 2.0 * tol;
    int k = 0;
    while (cut > tol) {
      coef.push_back(nIp * 2.0 * Jn);
      nIp *= std::complex<double>(0.0, -1.0);
      ++k;
      Jn = k == 1 ? j1(x) : jn(k, x);
      cut = fabs(Jn);
    }
  }
  const std::complex<double> e0 = std::exp(std::complex<double>(0.0, -(Ec / kHbar) * dt));
  const int g = std::min(R, 32);
  std::vector<std::complex<double>> acc((size_t)M * nT, 0.0);
  const double2 one = d2(1, 0), zero = d2(0, 0), two = d2(2, 0), mone = d2(-1, 0);
  for (int r0 = 0; r0 < R; r0 += g) {
    const int gg = std::min(g, R - r0);
    int kp = 1;
    while (kp < gg) kp <<= 1;
    const size_t vec = (size_t)n * gg;
    const int grid = spmm_dot_grid(n, kp);
    double2 *pr, *pl, *c0, *c1, *d0, *d1, *part, *sums;
    LQ_TRY(ctx_buf(ctx, "td_phiR", vec * sizeof(double2), (void **)&pr));
    LQ_TRY(ctx_buf(ctx, "td_phiL", vec * sizeof(double2), (void **)&pl));
    LQ_TRY(ctx_buf(ctx, "td_c0", 2 * vec * sizeof(double2), (void **)&c0));
    LQ_TRY(ctx_buf(ctx, "td_c1", 2 * vec * sizeof(double2), (void **)&c1));
    d0 = c0 + vec;  // right-vector halves of the evolution pair
    d1 = c1 + vec;
    LQ_TRY(ctx_buf(ctx, "td_part", (size_t)M * grid * kp * sizeof(double2), (void **)&part));
    LQ_TRY(ctx_buf(ctx, "td_sums", (size_t)M * kp * sizeof(double2), (void **)&sums));
    LQ_TRY(upload_states(ctx, phi, n, r0, gg, pr));
    if (OPL) {
      VecSet s{pr, nullptr, pl, gg};
      LQ_TRY(spmm(ctx, OPL, one, zero, &s, 1));
    } else {
      LQ_TRY(launch_copy(ctx->stream, pl, pr, vec));
    }
    const size_t slot = (size_t)grid * kp;
    std::vector<double2> h((size_t)M * kp);
    for (int t = 0; t < nT; ++t) {
      // moments at t: T_0 = OPR phi_R, mu(m, t) += scal <phi_L|T_m>
      double2 *A = c0, *B = c1;
      if (OPR) {
        VecSet s{pr, nullptr, A, gg};
        LQ_TRY(spmm(ctx, OPR, one, zero, &s, 1));
      } else {
        LQ_TRY(launch_copy(ctx->stream, A, pr, vec));
      }
      LQ_TRY(launch_dot(ctx->stream, n, gg, pl, A, part, nullptr));
      {
        SpmvLoopGuard guard{ctx};
        LQ_TRY(spmv_loop_begin(ctx));
        if (M > 1) {
          VecSet s{A, nullptr, B, gg};
          LQ_TRY(spmm(ctx, H, one, zero, &s, 1, kDotPhi, pl, part + slot));
        }
        for (int m = 1; m + 1 < M; ++m) {
          VecSet s{B, A, A, gg};
          LQ_TRY(spmm(ctx, H, two, mone, &s, 1, kDotPhi, pl, part + (size_t)(m + 1) * slot));
          std::swap(A, B);
        }
        LQ_TRY(spmv_loop_end(ctx));
      }
      LQ_TRY(launch_reduce_partials(ctx->stream, part, M, grid, kp, sums));
      LQ_HIP_CHECK(hipMemcpyAsync(h.data(), sums, sizeof(double2) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
      // evolve phi_L and phi_R by dt: Phi <- e0 sum_k coef_k T_k(H~) Phi
      LQ_TRY(launch_copy(ctx->stream, c0, pl, vec));
      LQ_TRY(launch_copy(ctx->stream, d0, pr, vec));
      {
        VecSet sets[2] = {VecSet{c0, nullptr, c1, gg}, VecSet{d0, nullptr, d1, gg}};
        LQ_TRY(spmm(ctx, H, one, zero, sets, 2));
      }
      double2 *x0 = c0, *x1 = c1, *y0 = d0, *y1 = d1;
      {
        SpmvLoopGuard guard{ctx};
        LQ_TRY(spmv_loop_begin(ctx));
        for (size_t k = 0; k < coef.size(); ++k) {
          const std::complex<double> ck = coef[k] * e0;
          LQ_TRY(launch_caxpy(ctx->stream, d2(ck.real(), ck.imag()), x0, pl, vec, k == 0));
          LQ_TRY(launch_caxpy(ctx->stream, d2(ck.real(), ck.imag()), y0, pr, vec, k == 0));
          if (k + 1 == coef.size()) break;
          VecSet sets[2] = {VecSet{x1, x0, x0, gg}, VecSet{y1, y0, y0, gg}};
          LQ_TRY(spmm(ctx, H, two, mone, sets, 2));
          std::swap(x0, x1);
          std::swap(y0, y1);
        }
        LQ_TRY(spmv_loop_end(ctx));
      }
      LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
      for (int m = 0; m < M; ++m) {
        double scal = 2.0 / R;
        if (m == 0) scal *= 0.5;
        for (int v = 0; v < gg; ++v)
          acc[(size_t)m * nT + t] += scal * std::complex<double>(h[(size_t)m * kp + v].x, h[(size_t)m * kp + v].y);
      }
    }
  }
  for (size_t i = 0; i < acc.size(); ++i) {
    mu[2 * i] += acc[i].real();
    mu[2 * i + 1] += acc[i].imag();
  }
  return collect_timing(ctx);
} LQ_ABI_CATCH

int lq_jackson_td(lq_ctx ctx, double *mu, int M, int nT, double W, double broad_meV, int *M_out) try {
  if (!ctx || !mu || M <= 0 || nT <= 0 || !(broad_meV > 0) || !(W > 0) || !M_out) {
    set_error("lq_jackson_td: invalid arguments");
    return LQ_ERR_ARG;
  }
  DeviceGuard dg(ctx->device);
  const double eta = 2.0 * broad_meV / 1000 / W;
  int maxMom = (int)ceil(kPi / eta);
  if (maxMom > M) maxMom = M;
  double2 *d;
  LQ_TRY(ctx_buf(ctx, "jk_td", (size_t)maxMom * nT * sizeof(double2), (void **)&d));
  const size_t bytes = (size_t)maxMom * nT * sizeof(double2);  // rows m < maxMom are the first maxMom*nT
  LQ_HIP_CHECK(hipMemcpyAsync(d, mu, bytes, hipMemcpyHostToDevice, ctx->stream));
  LQ_TRY(launch_jackson_td(ctx->stream, d, maxMom, nT));
  LQ_HIP_CHECK(hipMemcpyAsync(mu, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  *M_out = maxMom;
  return LQ_OK;
} LQ_ABI_CATCH

int lq_timecorr_post(lq_ctx ctx, const double *mu, int M, int nT, int n, double W, double Ec, double dt, double EF,
                     double *t, double *out) try {
  if (!ctx || !mu || M <= 0 || nT <= 0 || !(W > 0) || !t || !out) {
    set_error("lq_timecorr_post: invalid arguments");
    return LQ_ERR_ARG;
  }
  DeviceGuard dg(ctx->device);
  double2 *d;
  double *o;
  LQ_TRY(ctx_buf(ctx, "tc_mu", (size_t)M * nT * sizeof(double2), (void **)&d));
  LQ_TRY(ctx_buf(ctx, "tc_out", (size_t)2 * nT * sizeof(double), (void **)&o));
  LQ_HIP_CHECK(hipMemcpyAsync(d, mu, (size_t)M * nT * sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
  LQ_TRY(launch_timecorr_post(ctx->stream, d, M, nT, n, W, Ec, dt, EF, o, o + nT));
  LQ_HIP_CHECK(hipMemcpyAsync(t, o, nT * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LQ_HIP_CHECK(hipMemcpyAsync(out, o + nT, nT * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return LQ_OK;
} LQ_ABI_CATCH

int lq_kubo_post(lq_ctx ctx, int formula, const double *final_data, int nump, int n, double W, double Ec, double *E,
                 double *sigma, double *integrand) try {
  if (!ctx || !final_data || nump <= 1 || n <= 0 || !E || !sigma ||
      (formula != LQ_KUBO_GREENWOOD && formula != LQ_KUBO_BASTIN)) {
    set_error("lq_kubo_post: invalid arguments");
    return LQ_ERR_ARG;
  }
  DeviceGuard dg(ctx->device);
  double2 *fd;
  double *dE, *ds, *di;
  LQ_TRY(ctx_buf(ctx, "post_fd", (size_t)2 * nump * sizeof(double2), (void **)&fd));
  LQ_TRY(ctx_buf(ctx, "post_out", (size_t)3 * nump * sizeof(double), (void **)&dE));
  ds = dE + nump;
  di = dE + 2 * nump;
  LQ_HIP_CHECK(hipMemcpyAsync(fd, final_data, (size_t)2 * nump * sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
  LQ_TRY(launch_kubo_post(ctx->stream, formula, fd, nump, n, W, Ec, dE, ds, di));
  LQ_HIP_CHECK(hipMemcpyAsync(E, dE, nump * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LQ_HIP_CHECK(hipMemcpyAsync(sigma, ds, nump * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (integrand)
    LQ_HIP_CHECK(hipMemcpyAsync(integrand, di, nump * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return LQ_OK;
} LQ_ABI_CATCH

// ---- Jackson + grid post-processors -----------------------------------------------------
static int jackson_count(int M, double W, double broad) {
  const double eta = 2.0 * broad / 1000 / W;
  int maxMom = (int)ceil(kPi / eta);
  return maxMom > M ? M : maxMom;
}

int lq_jackson_1d(lq_ctx ctx, double *mu, int M, double W, double broad, int *M_out) try {
  if (!ctx || !mu || M <= 0 || !(broad > 0) || !(W > 0)) {
    set_error("lq_jackson_1d: invalid arguments");
    return LQ_ERR_ARG;
  }
  DeviceGuard dg(ctx->device);
  const int mm = jackson_count(M, W, broad);
  double2 *d;
  LQ_TRY(ctx_buf(ctx, "jk", (size_t)M * sizeof(double2), (void **)&d));
  LQ_HIP_CHECK(hipMemcpyAsync(d, mu, (size_t)mm * sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
  LQ_TRY(launch_jackson_1d(ctx->stream, d, mm));
  LQ_HIP_CHECK(hipMemcpyAsync(mu, d, (size_t)mm * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (M_out) *M_out = mm;
  return LQ_OK;
} LQ_ABI_CATCH

int lq_jackson_2d(lq_ctx ctx, double *mu, int M0, int M1, double W, double b0, double b1, int *M0_out, int *M1_out) try {
  if (!ctx || !mu || M0 <= 0 || M1 <= 0 || !(b0 > 0) || !(b1 > 0) || !(W > 0)) {
    set_error("lq_jackson_2d: invalid arguments");
    return LQ_ERR_ARG;
  }
  DeviceGuard dg(ctx->device);
  const int mm0 = jackson_count(M0, W, b0), mm1 = jackson_count(M1, W, b1);
  double2 *din, *dout;
  LQ_TRY(ctx_buf(ctx, "jk_in", (size_t)M0 * M1 * sizeof(double2), (void **)&din));
  LQ_TRY(ctx_buf(ctx, "jk_out", (size_t)mm0 * mm1 * sizeof(double2), (void **)&dout));
  LQ_HIP_CHECK(hipMemcpyAsync(din, mu, (size_t)M0 * M1 * sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
  LQ_TRY(launch_jackson_2d(ctx->stream, din, M1, dout, mm0, mm1));
  LQ_HIP_CHECK(hipMemcpyAsync(mu, dout, (size_t)mm0 * mm1 * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (M0_out) *M0_out = mm0;
  if (M1_out) *M1_out = mm1;
  return LQ_OK;
} LQ_ABI_CATCH

int lq_dos_fftgrid(lq_ctx ctx, const double *mu, int M, int n, double W, double Ec, int num_div, double *E,
                   double *rho) try {
  if (!ctx || !mu || M <= 0 || num_div <= 0 || !E || !rho) {
    set_error("lq_dos_fftgrid: invalid arguments");
    return LQ_ERR_ARG;
  }
  DeviceGuard dg(ctx->device);
  double2 *dm;
  double *dout;
  LQ_TRY(ctx_buf(ctx, "grid_mu", (size_t)M * sizeof(double2), (void **)&dm));
  LQ_TRY(ctx_buf(ctx, "grid_out", (size_t)2 * num_div * sizeof(double), (void **)&dout));
  LQ_HIP_CHECK(hipMemcpyAsync(dm, mu, (size_t)M * sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
  LQ_TRY(launch_dos_fftgrid(ctx->stream, dm, M, n, W, Ec, num_div, dout, dout + num_div));
  LQ_HIP_CHECK(hipMemcpyAsync(E, dout, num_div * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LQ_HIP_CHECK(hipMemcpyAsync(rho, dout + num_div, num_div * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return LQ_OK;
} LQ_ABI_CATCH

static int grid2d(lq_ctx ctx, int which, const double *mu, int M0, int M1, int n, double W, double Ec, int num_div,
                  double *E, double *out) {
  if (!ctx || !mu || M0 <= 0 || M1 <= 0 || num_div <= 1 || !E || !out) {
    set_error("fftgrid post-processor: invalid arguments");
    return LQ_ERR_ARG;
  }
  DeviceGuard dg(ctx->device);
  double2 *dm;
  double *scratch, *dout;
  LQ_TRY(ctx_buf(ctx, "grid_mu", (size_t)M0 * M1 * sizeof(double2), (void **)&dm));
  LQ_TRY(ctx_buf(ctx, "grid_scratch", (size_t)2 * num_div * sizeof(double), (void **)&scratch));
  LQ_TRY(ctx_buf(ctx, "grid_out", (size_t)2 * num_div * sizeof(double), (void **)&dout));
  LQ_HIP_CHECK(hipMemcpyAsync(dm, mu, (size_t)M0 * M1 * sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
  if (which == 0)
    LQ_TRY(launch_greenwood_fftgrid(ctx->stream, dm, M0, M1, n, W, Ec, num_div, scratch, dout, dout + num_div));
  else
    LQ_TRY(launch_bastin_fftgrid(ctx->stream, dm, M0, M1, n, W, Ec, num_div, scratch, dout, dout + num_div));
  LQ_HIP_CHECK(hipMemcpyAsync(E, dout, (num_div - 1) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LQ_HIP_CHECK(hipMemcpyAsync(out, dout + num_div, (num_div - 1) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return LQ_OK;
}

int lq_greenwood_fftgrid(lq_ctx ctx, const double *mu, int M0, int M1, int n, double W, double Ec, int num_div,
                         double *E, double *sigma) try {
  return grid2d(ctx, 0, mu, M0, M1, n, W, Ec, num_div, E, sigma);
} LQ_ABI_CATCH

int lq_bastin_fftgrid(lq_ctx ctx, const double *mu, int M0, int M1, int n, double W, double Ec, int num_div, double *E,
                      double *sigma) try {
  return grid2d(ctx, 1, mu, M0, M1, n, W, Ec, num_div, E, sigma);
} LQ_ABI_CATCH

int lq_linear_grid(lq_ctx ctx, int kind, const double *mu, int M0, int M1, int n, double W, double Ec, int num_div,
                   double *E, double *out) try {
  if (!ctx || !mu || kind < LQ_LIN_SPECTRAL || kind > LQ_LIN_BASTIN_SURF || M0 <= 0 ||
      (kind != LQ_LIN_SPECTRAL && M1 <= 0) || num_div <= 1 || !E || !out) {
    set_error("lq_linear_grid: invalid arguments");
    return LQ_ERR_ARG;
  }
  if (kind == LQ_LIN_SPECTRAL) M1 = 1;
  DeviceGuard dg(ctx->device);
  double2 *dm;
  double *kern, *dout;
  LQ_TRY(ctx_buf(ctx, "grid_mu", (size_t)M0 * M1 * sizeof(double2), (void **)&dm));
  LQ_TRY(ctx_buf(ctx, "grid_scratch", (size_t)2 * num_div * sizeof(double), (void **)&kern));
  LQ_TRY(ctx_buf(ctx, "grid_out", (size_t)2 * num_div * sizeof(double), (void **)&dout));
  LQ_HIP_CHECK(hipMemcpyAsync(dm, mu, (size_t)M0 * M1 * sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
  hipEvent_t ev;
  LQ_TRY(time_begin(ctx, kCatPost, &ev));
  LQ_TRY(launch_linear_grid(ctx->stream, kind, dm, M0, M1, n, W, Ec, num_div, kern, dout, dout + num_div));
  LQ_TRY(time_end(ctx, kCatPost, ev, 0));
  const int rows = kind == LQ_LIN_SPECTRAL ? num_div : num_div - 1;
  LQ_HIP_CHECK(hipMemcpyAsync(E, dout, rows * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LQ_HIP_CHECK(hipMemcpyAsync(out, dout + num_div, rows * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return collect_timing(ctx);
} LQ_ABI_CATCH

}  // extern "C"
