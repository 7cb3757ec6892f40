// lq_batch.hip -- many independent spectral-moment (DOS) jobs advanced together.
//
// Each job is chebyshev::SpectralMoments with OP = identity
// (src/chebyshev_solver.cpp:176-200) for its own H~, states and M.  Instead of
// one launch per step per job, a chunk of `chunk` recursion steps is issued
// as one grouped launch per step per vector-width class (k_spmm_batch), so
// small lattices that cannot fill the GPU alone share it, and the launch count
// no longer grows with the number of jobs.
//
// * Admission: jobs queue in submission order and are admitted first-fit at
//   chunk boundaries while their device footprint (states, two recursion
//   vectors, dot partials, moments) fits the batch's HBM budget.  The budget
//   bounds all device memory the batch holds: the blocks of running jobs (a
//   reused block may be larger than its job needs) and the blocks kept for
//   reuse.
// * Determinism: a job's block count, row schedule, dot partial layout and
//   reduction order depend on the job alone, so its moments are bitwise the
//   same alone, in any mix, and through lq_spectral_moments (R <= 64).
// * Tuning: the unroll width of the gathers does not change the arithmetic;
//   the first chunks of each width class try 4/8/12/16 and keep the fastest
//   per row-step (LQ_BATCH_UNROLL fixes it).
#include <algorithm>
#include <chrono>
#include <complex>
#include <cstdlib>
#include <map>
#include <vector>

#include "lq_internal.h"

using namespace lq;

namespace {

int pad2(int k) {
  int p = 1;
  while (p < k) p <<= 1;
  return p;
}

constexpr int kUnrolls[4] = {4, 8, 12, 16};

struct Job {
  int id = 0;
  lq_csr H = nullptr;
  int n = 0, R = 0, M = 0, kp = 1, nblk = 0;
  int state = 0;  // 0 queued, 1 running, 2 done
  std::vector<double> phi_host;
  const double *phi_dev = nullptr;
  void *mem = nullptr;
  size_t bytes = 0;  // device footprint of the job
  size_t held = 0;   // size of the block it runs in (a reused block may be larger)
  BatchJobDev d{};
  std::vector<std::complex<double>> result;
};

struct Tune {
  int best = 0;           // chosen unroll (0 while trying)
  int next = 0;           // next candidate index
  double cost[4] = {0, 0, 0, 0};
};

size_t job_bytes(int n, int R, int M, int chunk) {
  const int kp = pad2(R);
  const size_t nblk = (size_t)spmm_dot_grid(n, kp);
  return (size_t)3 * n * R * sizeof(double2)           // phi, buf0, buf1
         + (size_t)chunk * nblk * kp * sizeof(double2)  // partials of one chunk
         + (size_t)M * kp * sizeof(double2);            // moments
}

}  // namespace

struct lq_batch_s {
  lq_ctx ctx = nullptr;
  size_t budget = 0, used = 0;
  int chunk = 32;
  int t = 0;  // launch index of the next chunk
  int next_id = 1;
  std::map<int, Job> jobs;
  std::vector<int> queue;  // submission order
  int *tc = nullptr;       // device copy of the chunk's first launch index
  struct Cls {
    BatchJobDev *dev = nullptr;
    size_t cap = 0;
    std::vector<BatchJobDev> host;  // source of the table upload, kept until the chunk ends
  };
  std::map<int, Cls> cls;
  std::map<int, Tune> tune;
  int fixed_unroll = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // Device blocks of finished jobs kept for reuse (best fit), so admitting and
  // harvesting jobs calls hipMalloc only when no kept block fits and hipFree
  // (which waits for the whole device, other streams included) only when more
  // than kKeepBytes would be kept or a new block would exceed the budget.
  // used (blocks of running jobs) + spare_bytes <= budget at all times.
  std::multimap<size_t, void *> spare;
  size_t spare_bytes = 0;
};

namespace {

constexpr size_t kKeepBytes = (size_t)4 << 30;

// A kept block that fits (best fit, at most twice the need + 64 MiB): it moves
// from the kept blocks to the running ones, so the batch's memory stays the same.
std::multimap<size_t, void *>::iterator spare_fit(lq_batch b, size_t bytes) {
  auto it = b->spare.lower_bound(bytes);
  if (it != b->spare.end() && it->first <= 2 * bytes + ((size_t)64 << 20)) return it;
  return b->spare.end();
}

// Whether a job of `bytes` can be admitted now without the batch holding more
// than its budget (kept blocks may be freed to make room).
bool block_admissible(lq_batch b, size_t bytes) {
  return spare_fit(b, bytes) != b->spare.end() || b->used + bytes <= b->budget;
}

int block_get(lq_batch b, size_t bytes, void **p, size_t *held) {
  auto it = spare_fit(b, bytes);
  if (it != b->spare.end()) {
    *p = it->second;
    *held = it->first;
    b->spare_bytes -= it->first;
    b->spare.erase(it);
    return LQ_OK;
  }
  while (b->used + b->spare_bytes + bytes > b->budget && !b->spare.empty()) {  // make room: largest first
    auto big = std::prev(b->spare.end());
    b->spare_bytes -= big->first;
    void *q = big->second;
    b->spare.erase(big);
    LQ_HIP_CHECK(hipFree(q));
  }
  LQ_HIP_CHECK(hipMalloc(p, bytes));
  *held = bytes;
  return LQ_OK;
}

int block_put(lq_batch b, void *p, size_t bytes) {
  b->spare.emplace(bytes, p);
  b->spare_bytes += bytes;
  while (b->spare_bytes > kKeepBytes && !b->spare.empty()) {  // drop the largest kept blocks
    auto it = std::prev(b->spare.end());
    b->spare_bytes -= it->first;
    void *q = it->second;
    b->spare.erase(it);
    LQ_HIP_CHECK(hipFree(q));
  }
  return LQ_OK;
}

int admit(lq_batch b, Job &j) {
  lq_ctx ctx = b->ctx;
  LQ_TRY(block_get(b, j.bytes, &j.mem, &j.held));
  char *p = (char *)j.mem;
  const size_t vec = (size_t)j.n * j.R * sizeof(double2);
  BatchJobDev &d = j.d;
  d.sptr = j.H->dev.sptr;
  d.col = j.H->dev.col;
  d.val = j.H->dev.val;
  d.dsptr = j.H->dev.dsptr;
  d.wcol = j.H->dev.wcol;
  d.dict = j.H->dev.dict;
  d.diag = j.H->dev.diag;
  d.ndict = j.H->dev.ndict;
  d.phi = (const double2 *)p;
  d.buf0 = (double2 *)(p + vec);
  d.buf1 = (double2 *)(p + 2 * vec);
  d.partial = (double2 *)(p + 3 * vec);
  d.mom = (double2 *)(p + 3 * vec + (size_t)b->chunk * j.nblk * j.kp * sizeof(double2));
  d.n = j.n;
  d.k = j.R;
  d.t0 = b->t;
  d.steps = j.M - 1;
  d.nblk = j.nblk;
  LQ_TRY(upload_states(ctx, j.phi_dev ? j.phi_dev : j.phi_host.data(), j.n, 0, j.R, (double2 *)d.phi));
  j.phi_host.clear();
  j.phi_host.shrink_to_fit();
  LQ_TRY(launch_copy(ctx->stream, d.buf0, d.phi, (size_t)j.n * j.R));
  // mu_0 = <phi|T_0> with the single-job dot schedule
  int grid = 0;
  LQ_TRY(launch_dot(ctx->stream, j.n, j.R, d.phi, d.buf0, d.partial, &grid));
  LQ_TRY(launch_reduce_partials(ctx->stream, d.partial, 1, grid, j.kp, d.mom));
  j.state = 1;
  b->used += j.held;
  return LQ_OK;
}

int harvest(lq_batch b, Job &j) {
  std::vector<double2> h((size_t)j.M * j.kp);
  LQ_HIP_CHECK(hipMemcpy(h.data(), j.d.mom, sizeof(double2) * h.size(), hipMemcpyDeviceToHost));
  j.result.assign(j.M, 0.0);
  for (int m = 0; m < j.M; ++m) {
    double scal = 2.0 / j.R;
    if (m == 0) scal *= 0.5;
    for (int v = 0; v < j.R; ++v) j.result[m] += scal * std::complex<double>(h[(size_t)m * j.kp + v].x, h[(size_t)m * j.kp + v].y);
  }
  void *m = j.mem;
  j.mem = nullptr;
  b->used -= j.held;
  j.state = 2;
  LQ_TRY(block_put(b, m, j.held));
  return LQ_OK;
}

int run_chunk(lq_batch b) {
  lq_ctx ctx = b->ctx;
  // admission, first fit in submission order
  for (size_t q = 0; q < b->queue.size();) {
    Job &j = b->jobs[b->queue[q]];
    if (block_admissible(b, j.bytes)) {
      LQ_TRY(admit(b, j));
      b->queue.erase(b->queue.begin() + q);
    } else {
      ++q;
    }
  }
  // running jobs by width class, in id order
  std::map<int, std::vector<Job *>> by;
  int L = 0;
  for (auto &kv : b->jobs) {
    Job &j = kv.second;
    if (j.state != 1) continue;
    const int left = j.d.t0 + j.d.steps - b->t;
    if (left > 0) {
      by[j.kp].push_back(&j);
      L = std::max(L, left);
    }
  }
  L = std::min(L, b->chunk);
  if (L > 0) {
    LQ_TRY(launch_set_int(ctx->stream, b->tc, b->t));
    for (auto &kv : by) {
      const int kp = kv.first;
      auto &c = b->cls[kp];
      std::vector<BatchJobDev> &tab = c.host;
      tab.clear();
      int grid = 0;
      double work = 0.0;
      for (Job *j : kv.second) {
        j->d.blk0 = grid;
        grid += j->nblk;
        tab.push_back(j->d);
        const int active = std::min(L, j->d.t0 + j->d.steps - b->t);
        work += (double)active * ((double)j->H->col.size() + (double)j->n * j->R);
      }
      if (c.cap < tab.size()) {
        if (c.dev) LQ_HIP_CHECK(hipFree(c.dev));
        c.cap = std::max<size_t>(tab.size(), 16);
        LQ_HIP_CHECK(hipMalloc(&c.dev, c.cap * sizeof(BatchJobDev)));
      }
      LQ_HIP_CHECK(hipMemcpyAsync(c.dev, tab.data(), tab.size() * sizeof(BatchJobDev), hipMemcpyHostToDevice,
                                  ctx->stream));
      Tune &tu = b->tune[kp];
      int unroll = b->fixed_unroll ? b->fixed_unroll : (tu.best ? tu.best : kUnrolls[tu.next]);
      const bool timing = !b->fixed_unroll && !tu.best;
      if (timing) LQ_HIP_CHECK(hipEventRecord(b->ev0, ctx->stream));
      hipEvent_t ev;
      LQ_TRY(time_begin(ctx, kCatSpmv, &ev));
      for (int s = 0; s < L; ++s)
        LQ_TRY(launch_spmm_batch(ctx->stream, kp, unroll, c.dev, (int)tab.size(), grid, b->tc, s));
      double bytes = 0.0;
      for (Job *j : kv.second) {
        const int active = std::min(L, j->d.t0 + j->d.steps - b->t);
        bytes += (double)active * (20.0 * j->H->col.size() + 4.0 * (j->n + 1) + 48.0 * j->R * (double)j->n);
      }
      LQ_TRY(time_end(ctx, kCatSpmv, ev, bytes, L));
      if (timing) {
        LQ_HIP_CHECK(hipEventRecord(b->ev1, ctx->stream));
        LQ_HIP_CHECK(hipEventSynchronize(b->ev1));
        float ms = 0.f;
        LQ_HIP_CHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        tu.cost[tu.next] = work > 0 ? ms / work : 0.0;
        if (++tu.next == 4) {
          int bi = 0;
          for (int i = 1; i < 4; ++i)
            if (tu.cost[i] < tu.cost[bi]) bi = i;
          tu.best = kUnrolls[bi];
        }
      }
      LQ_TRY(launch_batch_reduce(ctx->stream, c.dev, (int)tab.size(), b->tc, L, kp));
    }
    LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    b->t += L;
  } else {
    LQ_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
  for (auto &kv : b->jobs) {
    Job &j = kv.second;
    if (j.state == 1 && b->t >= j.d.t0 + j.d.steps) LQ_TRY(harvest(b, j));
  }
  return collect_timing(ctx);
}

}  // namespace

extern "C" {

int lq_batch_job_bytes(int n, int R, int M, int chunk, size_t *bytes) try {
  if (n <= 0 || R <= 0 || R > 64 || M <= 0 || chunk <= 0 || !bytes) {
    set_error("lq_batch_job_bytes: invalid arguments");
    return LQ_ERR_ARG;
  }
  *bytes = job_bytes(n, R, M, chunk);
  return LQ_OK;
} LQ_ABI_CATCH

int lq_batch_create(lq_ctx ctx, size_t hbm_budget, int chunk, lq_batch *out) try {
  if (!ctx || !out || chunk < 0) {
    set_error("lq_batch_create: invalid arguments");
    return LQ_ERR_ARG;
  }
  DeviceGuard dg(ctx->device);
  auto *b = new lq_batch_s;
  b->ctx = ctx;
  b->chunk = chunk > 0 ? chunk : 32;
  if (hbm_budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      delete b;
      set_error("lq_batch_create: hipMemGetInfo failed");
      return LQ_ERR_HIP;
    }
    hbm_budget = fr / 10 * 8;
  }
  b->budget = hbm_budget;
  const char *u = getenv("LQ_BATCH_UNROLL");
  if (u) {
    const int v = atoi(u);
    if (v == 4 || v == 8 || v == 12 || v == 16) b->fixed_unroll = v;
  }
  if (hipMalloc(&b->tc, sizeof(int)) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
      hipEventCreate(&b->ev1) != hipSuccess) {
    (void)hipGetLastError();
    if (b->tc) (void)hipFree(b->tc);
    delete b;
    set_error("lq_batch_create: device allocation failed");
    return LQ_ERR_NOMEM;
  }
  *out = b;
  return LQ_OK;
} LQ_ABI_CATCH

int lq_batch_destroy(lq_batch b) try {
  if (!b) return LQ_OK;
  DeviceGuard dg(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  for (auto &kv : b->jobs)
    if (kv.second.mem) (void)hipFree(kv.second.mem);
  for (auto &kv : b->spare) (void)hipFree(kv.second);
  for (auto &kv : b->cls)
    if (kv.second.dev) (void)hipFree(kv.second.dev);
  if (b->tc) (void)hipFree(b->tc);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  delete b;
  return LQ_OK;
} LQ_ABI_CATCH

int lq_batch_submit(lq_batch b, lq_csr H, const double *phi, int R, int M, int *job) try {
  if (!b || !H || !phi || R <= 0 || R > 64 || M <= 0 || !job || H->ctx != b->ctx) {
    set_error("lq_batch_submit: invalid arguments (1 <= R <= 64, H on the batch's context)");
    return LQ_ERR_ARG;
  }
  Job j;
  j.id = b->next_id++;
  j.H = H;
  j.n = H->n;
  j.R = R;
  j.M = M;
  j.kp = pad2(R);
  j.nblk = spmm_dot_grid(j.n, j.kp);
  j.bytes = job_bytes(j.n, R, M, b->chunk);
  if (j.bytes > b->budget) {
    set_error("lq_batch_submit: job needs %zu bytes, above the batch budget of %zu", j.bytes, b->budget);
    return LQ_ERR_NOMEM;
  }
  if (is_device_ptr(phi, b->ctx->device)) {
    j.phi_dev = phi;  // caller keeps it valid until the job is admitted
  } else {
    j.phi_host.assign(phi, phi + 2 * (size_t)R * j.n);
  }
  *job = j.id;
  b->queue.push_back(j.id);
  b->jobs.emplace(j.id, std::move(j));
  return LQ_OK;
} LQ_ABI_CATCH

int lq_batch_run(lq_batch b, int max_chunks, int *running, int *queued) try {
  if (!b) {
    set_error("lq_batch_run: null batch");
    return LQ_ERR_ARG;
  }
  DeviceGuard dg(b->ctx->device);
  for (int c = 0; max_chunks <= 0 || c < max_chunks; ++c) {
    int run = 0;
    for (auto &kv : b->jobs) run += kv.second.state == 1;
    if (run == 0 && b->queue.empty()) break;
    LQ_TRY(run_chunk(b));
  }
  int run = 0;
  for (auto &kv : b->jobs) run += kv.second.state == 1;
  if (running) *running = run;
  if (queued) *queued = (int)b->queue.size();
  return LQ_OK;
} LQ_ABI_CATCH

int lq_batch_poll(lq_batch b, int job, int *state) try {
  if (!b || !state) {
    set_error("lq_batch_poll: invalid arguments");
    return LQ_ERR_ARG;
  }
  auto it = b->jobs.find(job);
  if (it == b->jobs.end()) {
    set_error("lq_batch_poll: unknown job %d", job);
    return LQ_ERR_ARG;
  }
  *state = it->second.state;
  return LQ_OK;
} LQ_ABI_CATCH

int lq_batch_result(lq_batch b, int job, double *mu) try {
  if (!b || !mu) {
    set_error("lq_batch_result: invalid arguments");
    return LQ_ERR_ARG;
  }
  auto it = b->jobs.find(job);
  if (it == b->jobs.end() || it->second.state != 2) {
    set_error("lq_batch_result: job %d is not finished", job);
    return LQ_ERR_ARG;
  }
  const Job &j = it->second;
  for (int m = 0; m < j.M; ++m) {
    mu[2 * m] = j.result[m].real();
    mu[2 * m + 1] = j.result[m].imag();
  }
  b->jobs.erase(it);
  return LQ_OK;
} LQ_ABI_CATCH

int lq_batch_unroll(lq_batch b, int R, int *unroll) try {
  if (!b || !unroll || R <= 0 || R > 64) {
    set_error("lq_batch_unroll: invalid arguments");
    return LQ_ERR_ARG;
  }
  if (b->fixed_unroll) {
    *unroll = b->fixed_unroll;
    return LQ_OK;
  }
  auto it = b->tune.find(pad2(R));
  *unroll = it == b->tune.end() ? 0 : it->second.best;
  return LQ_OK;
} LQ_ABI_CATCH

}  // extern "C"
