// lq_noalloc.hip -- allocation-free variant of the 1D moment solver.
//
// lq_spectral_moments_noalloc runs the direct recursion of lq_spectral_moments
// (src/chebyshev_solver.cpp:176-200) with every buffer beyond the caller's
// inputs and outputs -- the interleaved states, the two recursion vectors, the
// ring of per-block dot partials, their per-moment sums and the running moment
// sums -- carved out of one fixed device array, `scratch`, instead of
// hipMalloc'd context buffers.  It makes no allocation of any kind and no host
// synchronisation: it only enqueues kernels (and device-to-device copies) on
// the context's stream, so callers that may not allocate (code under stream
// capture, tight serving loops) can use it.
//
// The array is a file-scope __device__ object: each GPU the library runs on
// holds its own copy for the life of the process.  Host code never uses the
// array's name as an address (on the host it is only a stand-in symbol); the
// per-device address is looked up once, with hipGetSymbolAddress, when the
// first context of that device is created (noalloc_init), and only that
// address is passed to kernels.
#include <mutex>

#include "lq_internal.h"

namespace lq {
namespace {

// 64 MiB per GPU.  A DOS solve of N = 65536 sites with R = 8 states and
// M = 1024 moments needs 26 MiB (three [N][8] blocks, a 16-slot ring of 1024
// block partials x 8, the per-moment sums and the running sums).
constexpr size_t kScratchBytes = (size_t)64 << 20;
// Ring slots of the dot partials.  Each slot is reduced by itself, so the ring
// length does not change any result; 16 keeps the ring at 256 KiB per state.
constexpr int kNaRing = 16;
static __device__ __align__(256) double2 scratch[kScratchBytes / sizeof(double2)];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double2 *g_scratch[kMaxDevices] = {};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

__global__ void k_na_clear(double2 *__restrict__ p, int count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) p[t] = make_double2(0.0, 0.0);
}

// acc[m] += scal_m * sums[m - m0][v] for v = 0..g-1 in that order, with
// scal_m = base (1/2 of it at m = 0): the host fold of lq_spectral_moments,
// with products and sums rounded separately as the host does them (no fused
// multiply-add: contraction is switched off for this function's own
// arithmetic; HIP's __dmul_rn/__dadd_rn are plain operators that the
// compiler would fuse after inlining).
__global__ void k_na_fold(const double2 *__restrict__ sums, int kp, int g, int m0, int count, double base,
                          double2 *__restrict__ acc) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int m = m0 + t;
  const double scal = m == 0 ? base * 0.5 : base;
  double2 a = acc[m];
  for (int v = 0; v < g; ++v) {
    const double2 s = sums[(size_t)t * kp + v];
    a.x = a.x + scal * s.x;
    a.y = a.y + scal * s.y;
  }
  acc[m] = a;
}

// mu[m] += acc[m] (the caller's table accumulates, as lq_spectral_moments' does).
__global__ void k_na_add(const double2 *__restrict__ acc, int M, double2 *__restrict__ mu) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M) return;
  const double2 a = acc[t];
  double2 u = mu[t];
  u.x = u.x + a.x;
  u.y = u.y + a.y;
  mu[t] = u;
}

unsigned blocks_for(size_t count) { return (unsigned)((count + 255) / 256); }

}  // namespace

int noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return LQ_OK;  // the entry point reports it
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return LQ_OK;
  void *p = nullptr;
  LQ_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch[device] = (double2 *)p;
  return LQ_OK;
}

}  // namespace lq

using namespace lq;

extern "C" {

int lq_spectral_moments_noalloc(lq_csr H, lq_csr OP, const double *phi, int R, int M, double *mu) try {
  if (!H || !H->ctx) {
    set_error("lq_spectral_moments_noalloc: null operator");
    return LQ_ERR_ARG;
  }
  lq_ctx ctx = H->ctx;
  if (!phi || !mu || R <= 0 || M <= 0 || (OP && (OP->n != H->n || OP->ctx != ctx))) {
    set_error("lq_spectral_moments_noalloc: invalid arguments");
    return LQ_ERR_ARG;
  }
  DeviceGuard dg(ctx->device);
  if (!is_device_ptr(phi, ctx->device) || !is_device_ptr(mu, ctx->device)) {
    set_error("lq_spectral_moments_noalloc: phi and mu must be device memory of the context's GPU");
    return LQ_ERR_ARG;
  }
  double2 *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_scratch[ctx->device];
  }
  if (!base) {
    set_error("lq_spectral_moments_noalloc: scratch of device %d not set up", ctx->device);
    return LQ_ERR_HIP;
  }
  const int n = H->n;
  // the groups of lq_spectral_moments: up to 64 states per recursion
  const int g = std::min(R, 64);
  int kp = 1;
  while (kp < g) kp <<= 1;
  const int grid = spmm_dot_grid(n, kp);
  const size_t vec = (size_t)n * g;
  // scratch layout (256-byte aligned pieces): acc[M] | phi [n][g] | T0 | T1 |
  // ring [kNaRing][grid][kp] | sums [kNaRing][kp]
  const size_t b_acc = align256((size_t)M * sizeof(double2));
  const size_t b_vec = align256(vec * sizeof(double2));
  const size_t b_ring = align256((size_t)kNaRing * grid * kp * sizeof(double2));
  const size_t b_sums = align256((size_t)kNaRing * kp * sizeof(double2));
  const size_t need = b_acc + 3 * b_vec + b_ring + b_sums;
  if (need > kScratchBytes) {
    set_error("lq_spectral_moments_noalloc: n = %d, R = %d, M = %d needs %zu bytes of scratch, above the "
              "maximum of %zu",
              n, R, M, need, kScratchBytes);
    return LQ_ERR_ARG;
  }
  char *p = (char *)base;
  double2 *acc = (double2 *)p;
  double2 *phi_d = (double2 *)(p + b_acc);
  double2 *A = (double2 *)(p + b_acc + b_vec);
  double2 *B = (double2 *)(p + b_acc + 2 * b_vec);
  double2 *ring = (double2 *)(p + b_acc + 3 * b_vec);
  double2 *sums = (double2 *)(p + b_acc + 3 * b_vec + b_ring);
  hipStream_t st = ctx->stream;

  hipLaunchKernelGGL(k_na_clear, dim3(blocks_for(M)), dim3(256), 0, st, acc, M);
  LQ_HIP_CHECK(hipGetLastError());
  const double2 one = make_double2(1, 0), zero = make_double2(0, 0), two = make_double2(2, 0),
                mone = make_double2(-1, 0);
  const double scal = 2.0 / R;
  for (int r0 = 0; r0 < R; r0 += g) {
    const int gg = std::min(g, R - r0);
    int kpg = 1;
    while (kpg < gg) kpg <<= 1;
    const int gridg = spmm_dot_grid(n, kpg);
    const size_t slotg = (size_t)gridg * kpg;
    const size_t vg = (size_t)n * gg;
    LQ_TRY(launch_interleave(st, (const double2 *)phi, n, r0, gg, phi_d));
    double2 *T0 = A, *T1 = B;
    if (OP) {
      VecSet s{phi_d, nullptr, T0, gg};
      LQ_TRY(launch_spmm(st, OP->dev, one, zero, &s, 1, kDotNone, nullptr, nullptr, nullptr));
    } else {
      LQ_TRY(launch_copy(st, T0, phi_d, vg));
    }
    auto at = [&](int m) { return ring + (size_t)(m % kNaRing) * slotg; };
    auto flush = [&](int m_end) -> int {
      const int m0 = ((m_end - 1) / kNaRing) * kNaRing;
      LQ_TRY(launch_reduce_partials(st, ring, m_end - m0, gridg, kpg, sums));
      hipLaunchKernelGGL(k_na_fold, dim3(blocks_for(m_end - m0)), dim3(256), 0, st, sums, kpg, gg, m0, m_end - m0,
                         scal, acc);
      LQ_HIP_CHECK(hipGetLastError());
      return LQ_OK;
    };
    LQ_TRY(launch_dot(st, n, gg, phi_d, T0, at(0), nullptr));
    if (M > 1) {
      VecSet s{T0, nullptr, T1, gg};
      LQ_TRY(launch_spmm(st, H->dev, one, zero, &s, 1, kDotPhi, phi_d, at(1), nullptr));
    }
    for (int m = 1; m + 1 < M; ++m) {
      if ((m + 1) % kNaRing == 0) LQ_TRY(flush(m + 1));
      VecSet s{T1, T0, T0, gg};
      LQ_TRY(launch_spmm(st, H->dev, two, mone, &s, 1, kDotPhi, phi_d, at(m + 1), nullptr));
      std::swap(T0, T1);
    }
    LQ_TRY(flush(M));
  }
  hipLaunchKernelGGL(k_na_add, dim3(blocks_for(M)), dim3(256), 0, st, acc, M, (double2 *)mu);
  LQ_HIP_CHECK(hipGetLastError());
  return LQ_OK;
} LQ_ABI_CATCH

}  // extern "C"
