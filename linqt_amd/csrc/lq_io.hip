// File ingestion: operator (CSR) and moment files, text and binary.
//
// Text layouts are the reference's:
//   CSR      Sparse::OPERATOR_FromCSRFile (src/sparse_matrix.cpp:3-50), writer
//            hopping_list::save_hopping_list_as_csr
//            (utilities/wannier2sparse/src/hopping_list.cpp:227-285);
//   moments  Moments1D::saveIn (src/chebyshev_moments1D.cpp:91-103),
//            Moments2D::saveIn (src/chebyshev_moments2D.cpp:53-68),
//            MomentsTD::saveIn (src/chebyshev_momentsTD.cpp:81-93).
// The reference reads them with `ifstream >>` one token at a time.  Here the
// file is mapped and cut into chunks at whitespace; threads count the tokens
// of their chunk, a prefix sum gives every chunk its first token index, and
// the threads then parse straight into the destination arrays.  Doubles use
// std::from_chars, which rounds correctly like the strtod behind `>>`, so
// every value is bitwise what the reference reads.
//
// Binary layouts (little endian, this library's own):
//   CSR      "LQCSR001" int64 n, int64 nnz, int32 rowptr[n+1], int32 col[nnz],
//            double val[2 nnz] (re, im interleaved);
//   moments  "LQMOM001" int32 kind, int32 0, int64 n, int64 M0, int64 M1,
//            double W, Ec, dt, double mu[2 M0 M1].
// Readers detect the binary form by its magic, so every reader accepts both.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "lq_internal.h"

namespace lq {
namespace {

const char kCsrMagic[8] = {'L', 'Q', 'C', 'S', 'R', '0', '0', '1'};
const char kMomMagic[8] = {'L', 'Q', 'M', 'O', 'M', '0', '0', '1'};

struct Mapped {
  int fd = -1;
  const char *p = nullptr;
  size_t n = 0;
  ~Mapped() {
    if (p && n) munmap((void *)p, n);
    if (fd >= 0) close(fd);
  }
  bool open(const char *path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat st;
    if (fstat(fd, &st) != 0) return false;
    n = (size_t)st.st_size;
    if (n == 0) return true;
    void *m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) {
      n = 0;
      return false;
    }
    madvise(m, n, MADV_SEQUENTIAL);
    p = (const char *)m;
    return true;
  }
};

inline bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

int io_threads(size_t bytes) {
  static const int cap = [] {
    const char *e = getenv("LQ_IO_THREADS");
    int t = e ? atoi(e) : (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(t, 64));
  }();
  const size_t per = 1 << 20;  // at least 1 MiB of text per thread
  return (int)std::max<size_t>(1, std::min<size_t>((size_t)cap, bytes / per + 1));
}

// One token: optional leading '+' (accepted by `>>`, not by from_chars).
inline bool parse_double(const char *b, const char *e, double *out) {
  if (b < e && *b == '+') ++b;
  auto r = std::from_chars(b, e, *out);
  return r.ec == std::errc() && r.ptr == e;
}

inline bool parse_long(const char *b, const char *e, long long *out) {
  if (b < e && *b == '+') ++b;
  auto r = std::from_chars(b, e, *out);
  return r.ec == std::errc() && r.ptr == e;
}

// Sequential header tokens from p; advances p.
bool next_token(const char *&p, const char *end, const char *&tb, const char *&te) {
  while (p < end && is_space(*p)) ++p;
  if (p >= end) return false;
  tb = p;
  while (p < end && !is_space(*p)) ++p;
  te = p;
  return true;
}

// Token k of the body goes to sink(k, b, e), which returns false on a parse
// error.  Parses tokens [0, want) of [b, e); returns the number found (up to
// want) or -1 (with the failing token index in *bad) on a parse error.
template <class Sink>
long long parse_body(const char *b, const char *e, long long want, const Sink &sink, long long *bad) {
  const size_t bytes = (size_t)(e - b);
  const int T = io_threads(bytes);
  std::vector<const char *> cut(T + 1);
  cut[0] = b;
  cut[T] = e;
  for (int t = 1; t < T; ++t) {
    const char *q = b + bytes * t / T;
    if (q < cut[t - 1]) q = cut[t - 1];
    while (q < e && !is_space(*q)) ++q;  // a chunk starts at whitespace
    cut[t] = q;
  }
  std::vector<long long> cnt(T + 1, 0);
  auto count = [&](int t) {
    long long c = 0;
    const char *q = cut[t], *qe = cut[t + 1];
    while (q < qe) {
      while (q < qe && is_space(*q)) ++q;
      if (q >= qe) break;
      ++c;
      while (q < qe && !is_space(*q)) ++q;
    }
    cnt[t + 1] = c;
  };
  std::vector<std::thread> th;
  for (int t = 1; t < T; ++t) th.emplace_back(count, t);
  count(0);
  for (auto &x : th) x.join();
  th.clear();
  for (int t = 0; t < T; ++t) cnt[t + 1] += cnt[t];
  std::atomic<long long> first_bad{-1};
  auto parse = [&](int t) {
    long long k = cnt[t];
    const char *q = cut[t], *qe = cut[t + 1];
    while (q < qe && k < want) {
      while (q < qe && is_space(*q)) ++q;
      if (q >= qe) break;
      const char *tb = q;
      while (q < qe && !is_space(*q)) ++q;
      if (!sink(k, tb, q)) {
        long long cur = first_bad.load();
        while ((cur < 0 || k < cur) && !first_bad.compare_exchange_weak(cur, k)) {
        }
        return;
      }
      ++k;
    }
  };
  for (int t = 1; t < T; ++t) th.emplace_back(parse, t);
  parse(0);
  for (auto &x : th) x.join();
  if (first_bad.load() >= 0) {
    *bad = first_bad.load();
    return -1;
  }
  return std::min(cnt[T], want);
}

bool has_magic(const Mapped &m, const char *magic) { return m.n >= 8 && memcmp(m.p, magic, 8) == 0; }

int read_binary_csr(const Mapped &m, const char *path, int *n, long long *nnz, int *rowptr, int *col, double *val) {
  if (m.n < 24) {
    set_error("%s: truncated binary CSR header", path);
    return LQ_ERR_ARG;
  }
  long long hn, hnnz;
  memcpy(&hn, m.p + 8, 8);
  memcpy(&hnnz, m.p + 16, 8);
  if (hn <= 0 || hn > 0x7ffffffeLL || hnnz < 0 || hnnz > 0x7fffffffLL) {
    set_error("%s: bad binary CSR sizes n=%lld nnz=%lld", path, hn, hnnz);
    return LQ_ERR_ARG;
  }
  const size_t need = 24 + 4 * (size_t)(hn + 1) + 4 * (size_t)hnnz + 16 * (size_t)hnnz;
  if (m.n < need) {
    set_error("%s: truncated binary CSR (%zu of %zu bytes)", path, m.n, need);
    return LQ_ERR_ARG;
  }
  *n = (int)hn;
  *nnz = hnnz;
  if (!rowptr) return LQ_OK;
  const char *q = m.p + 24;
  memcpy(rowptr, q, 4 * (size_t)(hn + 1));
  q += 4 * (size_t)(hn + 1);
  if (hnnz) memcpy(col, q, 4 * (size_t)hnnz);
  q += 4 * (size_t)hnnz;
  if (hnnz) memcpy(val, q, 16 * (size_t)hnnz);
  return LQ_OK;
}

int write_all(FILE *f, const void *p, size_t bytes) { return fwrite(p, 1, bytes, f) == bytes ? 0 : -1; }

// Text output buffer: formats into a large buffer and writes it in blocks.
struct Out {
  FILE *f;
  std::vector<char> buf;
  size_t used = 0;
  bool bad = false;
  explicit Out(FILE *f_) : f(f_), buf(1 << 22) {}
  void room(size_t k) {
    if (used + k > buf.size()) flush();
  }
  void flush() {
    if (used && fwrite(buf.data(), 1, used, f) != used) bad = true;
    used = 0;
  }
  void str(const char *s) {
    const size_t k = strlen(s);
    room(k);
    memcpy(buf.data() + used, s, k);
    used += k;
  }
  void num(double x, int digits) {  // ostream << double at precision `digits` (%g)
    room(40);
    used += (size_t)snprintf(buf.data() + used, 40, "%.*g", digits, x);
  }
  void integer(long long x) {
    room(24);
    auto r = std::to_chars(buf.data() + used, buf.data() + used + 24, x);
    used = (size_t)(r.ptr - buf.data());
  }
  void ch(char c) {
    room(1);
    buf[used++] = c;
  }
};

const char *mom_kind_tag(int kind) {
  return kind == 1 ? ".chebmom1D" : kind == 2 ? ".chebmom2D" : kind == 3 ? ".chebmomTD" : nullptr;
}

}  // namespace
}  // namespace lq

using namespace lq;

extern "C" int lq_csr_file_read(const char *path, int *n, long long *nnz, int *rowptr, int *col, double *val) try {
  if (!path || !n || !nnz) {
    set_error("lq_csr_file_read: null argument");
    return LQ_ERR_ARG;
  }
  Mapped m;
  if (!m.open(path)) {
    set_error("ERROR: Can not open the matrix file %s", path);
    return LQ_ERR_ARG;
  }
  if (has_magic(m, kCsrMagic)) return read_binary_csr(m, path, n, nnz, rowptr, col, val);
  const char *p = m.p, *end = m.p + m.n, *tb, *te;
  long long hn = 0, hnnz = 0;
  if (!next_token(p, end, tb, te) || !parse_long(tb, te, &hn) || !next_token(p, end, tb, te) ||
      !parse_long(tb, te, &hnnz) || hn <= 0 || hn > 0x7ffffffeLL || hnnz < 0 || hnnz > 0x7fffffffLL) {
    set_error("%s: bad CSR header (expected 'dim nnz')", path);
    return LQ_ERR_ARG;
  }
  *n = (int)hn;
  *nnz = hnnz;
  if (!rowptr) return LQ_OK;
  if (hnnz > 0 && (!col || !val)) {
    set_error("lq_csr_file_read: null output array");
    return LQ_ERR_ARG;
  }
  const long long nv = 2 * hnnz, nc = hnnz, want = nv + nc + hn + 1;
  auto sink = [&](long long k, const char *b, const char *e) -> bool {
    if (k < nv) return parse_double(b, e, val + k);
    long long x;
    if (!parse_long(b, e, &x) || x < -0x80000000LL || x > 0x7fffffffLL) return false;
    if (k < nv + nc)
      col[k - nv] = (int)x;
    else
      rowptr[k - nv - nc] = (int)x;
    return true;
  };
  long long bad = -1;
  const long long got = parse_body(p, end, want, sink, &bad);
  if (got < 0) {
    set_error("%s: unreadable number at token %lld of the CSR body", path, bad + 2);
    return LQ_ERR_ARG;
  }
  if (got < want) {
    set_error("%s: truncated CSR file (%lld of %lld numbers after the header)", path, got, want);
    return LQ_ERR_ARG;
  }
  return LQ_OK;
} LQ_ABI_CATCH

extern "C" int lq_csr_file_write(const char *path, int format, int n, long long nnz, const int *rowptr,
                                 const int *col, const double *val) try {
  if (!path || n <= 0 || nnz < 0 || !rowptr || (nnz > 0 && (!col || !val))) {
    set_error("lq_csr_file_write: invalid arguments");
    return LQ_ERR_ARG;
  }
  if (format != LQ_FMT_TEXT && format != LQ_FMT_TEXT_REF && format != LQ_FMT_BINARY) {
    set_error("lq_csr_file_write: unknown format %d", format);
    return LQ_ERR_ARG;
  }
  FILE *f = fopen(path, format == LQ_FMT_BINARY ? "wb" : "w");
  if (!f) {
    set_error("lq_csr_file_write: cannot open %s", path);
    return LQ_ERR_ARG;
  }
  int bad = 0;
  if (format == LQ_FMT_BINARY) {
    const long long hdr[2] = {n, nnz};
    bad |= write_all(f, kCsrMagic, 8);
    bad |= write_all(f, hdr, 16);
    bad |= write_all(f, rowptr, 4 * (size_t)(n + 1));
    if (nnz) bad |= write_all(f, col, 4 * (size_t)nnz);
    if (nnz) bad |= write_all(f, val, 16 * (size_t)nnz);
  } else {
    // save_hopping_list_as_csr streams values at the ofstream default
    // precision (6 significant digits); LQ_FMT_TEXT keeps 17 (exact).
    const int digits = format == LQ_FMT_TEXT_REF ? 6 : 17;
    Out o(f);
    o.integer(n);
    o.ch(' ');
    o.integer(nnz);
    o.ch('\n');
    for (long long q = 0; q < nnz; ++q) {
      o.num(val[2 * q], digits);
      o.ch(' ');
      o.num(val[2 * q + 1], digits);
      o.ch(' ');
    }
    o.ch('\n');
    for (long long q = 0; q < nnz; ++q) {
      o.integer(col[q]);
      o.ch(' ');
    }
    o.ch('\n');
    for (int i = 0; i <= n; ++i) {
      o.integer(rowptr[i]);
      o.ch(' ');
    }
    o.ch('\n');
    o.flush();
    bad |= o.bad ? -1 : 0;
  }
  if (fclose(f) != 0) bad = -1;
  if (bad) {
    set_error("lq_csr_file_write: write to %s failed", path);
    return LQ_ERR_ARG;
  }
  return LQ_OK;
} LQ_ABI_CATCH

extern "C" int lq_csr_file_load(lq_ctx ctx, const char *path, int fill, lq_csr *out) try {
  int n = 0;
  long long nnz = 0;
  LQ_TRY(lq_csr_file_read(path, &n, &nnz, nullptr, nullptr, nullptr));
  std::vector<int> rowptr((size_t)n + 1), col(nnz > 0 ? nnz : 1);
  std::vector<double> val(nnz > 0 ? 2 * nnz : 2);
  LQ_TRY(lq_csr_file_read(path, &n, &nnz, rowptr.data(), col.data(), val.data()));
  return lq_csr_create(ctx, n, (int)nnz, rowptr.data(), col.data(), val.data(), fill, out);
} LQ_ABI_CATCH

extern "C" int lq_mom_file_read(const char *path, lq_mom_header *h, double *mu) try {
  if (!path || !h) {
    set_error("lq_mom_file_read: null argument");
    return LQ_ERR_ARG;
  }
  Mapped m;
  if (!m.open(path)) {
    set_error("lq_mom_file_read: cannot open %s", path);
    return LQ_ERR_ARG;
  }
  lq_mom_header r{};
  if (has_magic(m, kMomMagic)) {
    if (m.n < 64) {
      set_error("%s: truncated binary moment header", path);
      return LQ_ERR_ARG;
    }
    int kind;
    long long hn, M0, M1;
    memcpy(&kind, m.p + 8, 4);
    memcpy(&hn, m.p + 16, 8);
    memcpy(&M0, m.p + 24, 8);
    memcpy(&M1, m.p + 32, 8);
    memcpy(&r.W, m.p + 40, 8);
    memcpy(&r.Ec, m.p + 48, 8);
    memcpy(&r.dt, m.p + 56, 8);
    if (kind < 1 || kind > 3 || M0 <= 0 || M1 <= 0 || M0 > 0x7fffffffLL || M1 > 0x7fffffffLL ||
        m.n < 64 + 16 * (size_t)M0 * (size_t)M1) {
      set_error("%s: bad or truncated binary moment file", path);
      return LQ_ERR_ARG;
    }
    r.kind = kind;
    r.n = hn;
    r.M0 = (int)M0;
    r.M1 = (int)M1;
    *h = r;
    if (mu) memcpy(mu, m.p + 64, 16 * (size_t)M0 * (size_t)M1);
    return LQ_OK;
  }
  // Text: the kind comes from the file name, as the reference's readers
  // check it (src/chebyshev_moments2D.cpp:16-22 and its 1D / TD twins).
  const std::string name(path);
  int kind = 0;
  for (int k = 1; k <= 3; ++k)
    if (name.find(mom_kind_tag(k)) != std::string::npos) kind = k;
  if (!kind) {
    set_error("The first argument does not seem to be a valid .chebmom1D/.chebmom2D/.chebmomTD file: %s", path);
    return LQ_ERR_ARG;
  }
  const char *p = m.p, *end = m.p + m.n, *tb, *te;
  double hdr[7];
  const int nh = kind == 1 ? 4 : kind == 2 ? 5 : 7;
  for (int i = 0; i < nh; ++i)
    if (!next_token(p, end, tb, te) || !parse_double(tb, te, &hdr[i])) {
      set_error("%s: bad moment file header", path);
      return LQ_ERR_ARG;
    }
  r.kind = kind;
  r.n = (long long)hdr[0];
  r.W = hdr[1];
  r.Ec = hdr[2];
  if (kind == 1) {
    r.M0 = (int)hdr[3];
    r.M1 = 1;
  } else if (kind == 2) {
    r.M0 = (int)hdr[3];
    r.M1 = (int)hdr[4];
  } else {  // N W Ec nT dt / M nT
    r.dt = hdr[4];
    r.M0 = (int)hdr[5];
    r.M1 = (int)hdr[6];
  }
  if (r.M0 <= 0 || r.M1 <= 0) {
    set_error("%s: bad moment counts %d x %d", path, r.M0, r.M1);
    return LQ_ERR_ARG;
  }
  *h = r;
  if (!mu) return LQ_OK;
  const long long want = 2LL * r.M0 * r.M1;
  auto sink = [&](long long k, const char *b, const char *e) -> bool { return parse_double(b, e, mu + k); };
  long long bad = -1;
  const long long got = parse_body(p, end, want, sink, &bad);
  if (got < 0) {
    set_error("%s: unreadable number at moment value %lld", path, bad);
    return LQ_ERR_ARG;
  }
  if (got < want) {
    set_error("%s: truncated moment file (%lld of %lld numbers)", path, got, want);
    return LQ_ERR_ARG;
  }
  return LQ_OK;
} LQ_ABI_CATCH

extern "C" int lq_mom_file_write(const char *path, int format, const lq_mom_header *h, const double *mu) try {
  if (!path || !h || !mu || h->kind < 1 || h->kind > 3 || h->M0 <= 0 || h->M1 <= 0 || (h->kind == 1 && h->M1 != 1)) {
    set_error("lq_mom_file_write: invalid arguments");
    return LQ_ERR_ARG;
  }
  if (format != LQ_FMT_TEXT && format != LQ_FMT_TEXT_REF && format != LQ_FMT_BINARY) {
    set_error("lq_mom_file_write: unknown format %d", format);
    return LQ_ERR_ARG;
  }
  FILE *f = fopen(path, format == LQ_FMT_BINARY ? "wb" : "w");
  if (!f) {
    set_error("lq_mom_file_write: cannot open %s", path);
    return LQ_ERR_ARG;
  }
  const size_t count = (size_t)h->M0 * (size_t)h->M1;
  int bad = 0;
  if (format == LQ_FMT_BINARY) {
    const int kind[2] = {h->kind, 0};
    const long long sz[3] = {h->n, h->M0, h->M1};
    const double fl[3] = {h->W, h->Ec, h->dt};
    bad |= write_all(f, kMomMagic, 8);
    bad |= write_all(f, kind, 8);
    bad |= write_all(f, sz, 24);
    bad |= write_all(f, fl, 24);
    bad |= write_all(f, mu, 16 * count);
  } else {
    // the reference streams at precision(digits10) = 15
    const int digits = format == LQ_FMT_TEXT_REF ? 15 : 17;
    Out o(f);
    o.integer(h->n);
    o.ch(' ');
    o.num(h->W, digits);
    if (h->kind == 1) {
      o.ch(' ');
      o.num(h->Ec, digits);
      o.ch('\n');
      o.integer(h->M0);
      o.ch('\n');
    } else if (h->kind == 2) {
      o.str("  ");
      o.num(h->Ec, digits);
      o.ch('\n');
      o.integer(h->M0);
      o.ch(' ');
      o.integer(h->M1);
      o.str(" \n");
    } else {
      o.ch(' ');
      o.num(h->Ec, digits);
      o.ch(' ');
      o.integer(h->M1);
      o.ch(' ');
      o.num(h->dt, digits);
      o.str(" \n");
      o.integer(h->M0);
      o.ch(' ');
      o.integer(h->M1);
      o.str(" \n");
    }
    for (size_t q = 0; q < count; ++q) {
      o.num(mu[2 * q], digits);
      o.ch(' ');
      o.num(mu[2 * q + 1], digits);
      o.ch('\n');
    }
    o.flush();
    bad |= o.bad ? -1 : 0;
  }
  if (fclose(f) != 0) bad = -1;
  if (bad) {
    set_error("lq_mom_file_write: write to %s failed", path);
    return LQ_ERR_ARG;
  }
  return LQ_OK;
} LQ_ABI_CATCH
