// lq_probe.hip -- streaming-bandwidth probe of the device memory.
//
// Hand-written streaming kernels over a buffer of the probe's own: read
// (every 16-byte word loaded once, folded into one word per thread so the
// loads are not dropped), copy (read + write of two halves) and write.  They
// give the achievable HBM rate the recursion kernels are compared against; a
// runtime copy is a different kernel with its own launch shape.
#include "lq_internal.h"

namespace lq {
namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kProbeThreads = 256;
constexpr int kProbeUnroll = 4;

__global__ __launch_bounds__(kProbeThreads) void k_probe_read(const u32x4 *__restrict__ src, size_t n,
                                                             u32x4 *__restrict__ sink) {
  const size_t stride = (size_t)gridDim.x * kProbeThreads;
  size_t i = (size_t)blockIdx.x * kProbeThreads + threadIdx.x;
  u32x4 acc = {0u, 0u, 0u, 0u};
  for (; i + (kProbeUnroll - 1) * stride < n; i += kProbeUnroll * stride) {
    u32x4 v[kProbeUnroll];
#pragma unroll
    for (int u = 0; u < kProbeUnroll; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < kProbeUnroll; ++u) acc ^= v[u];
  }
  for (; i < n; i += stride) {
    acc ^= src[i];
  }
  sink[(size_t)blockIdx.x * kProbeThreads + threadIdx.x] = acc;
}

__global__ __launch_bounds__(kProbeThreads) void k_probe_copy(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst,
                                                             size_t n) {
  const size_t stride = (size_t)gridDim.x * kProbeThreads;
  size_t i = (size_t)blockIdx.x * kProbeThreads + threadIdx.x;
  for (; i + (kProbeUnroll - 1) * stride < n; i += kProbeUnroll * stride) {
    u32x4 v[kProbeUnroll];
#pragma unroll
    for (int u = 0; u < kProbeUnroll; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < kProbeUnroll; ++u) __builtin_nontemporal_store(v[u], dst + i + u * stride);
  }
  for (; i < n; i += stride) dst[i] = src[i];
}

__global__ __launch_bounds__(kProbeThreads) void k_probe_write(u32x4 *__restrict__ dst, size_t n, unsigned seed) {
  const size_t stride = (size_t)gridDim.x * kProbeThreads;
  const u32x4 v = {seed, seed + 1, seed + 2, seed + 3};
  for (size_t i = (size_t)blockIdx.x * kProbeThreads + threadIdx.x; i < n; i += stride)
    __builtin_nontemporal_store(v, dst + i);
}

}  // namespace
}  // namespace lq

using namespace lq;

extern "C" {

int lq_hbm_probe(lq_ctx ctx, size_t bytes, int kind, int reps, double *ms, double *moved) try {
  if (!ctx || !ms || !moved || bytes < ((size_t)1 << 20) || reps <= 0 || kind < 0 || kind > 2) {
    set_error("lq_hbm_probe: invalid arguments");
    return LQ_ERR_ARG;
  }
  DeviceGuard dg(ctx->device);
  int cus = 0;
  LQ_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const int grid = cus * 8;
  const size_t words = bytes / sizeof(u32x4);
  u32x4 *buf = nullptr, *sink = nullptr;
  LQ_TRY(ctx_buf(ctx, "probe", words * sizeof(u32x4), (void **)&buf));
  LQ_TRY(ctx_buf(ctx, "probe_sink", (size_t)grid * kProbeThreads * sizeof(u32x4), (void **)&sink));
  hipStream_t st = ctx->stream;
  auto launch = [&](unsigned seed) -> int {
    if (kind == 0)
      hipLaunchKernelGGL(k_probe_read, dim3(grid), dim3(kProbeThreads), 0, st, buf, words, sink);
    else if (kind == 1)
      hipLaunchKernelGGL(k_probe_copy, dim3(grid), dim3(kProbeThreads), 0, st, buf, buf + words / 2, words / 2);
    else
      hipLaunchKernelGGL(k_probe_write, dim3(grid), dim3(kProbeThreads), 0, st, buf, words, seed);
    LQ_HIP_CHECK(hipGetLastError());
    return LQ_OK;
  };
  hipLaunchKernelGGL(k_probe_write, dim3(grid), dim3(kProbeThreads), 0, st, buf, words, 7u);  // touch every page
  LQ_HIP_CHECK(hipGetLastError());
  LQ_TRY(launch(1));  // warm-up
  struct Events {  // destroyed on every return path
    hipEvent_t a = nullptr, b = nullptr;
    ~Events() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
  } ev;
  LQ_HIP_CHECK(hipEventCreate(&ev.a));
  LQ_HIP_CHECK(hipEventCreate(&ev.b));
  LQ_HIP_CHECK(hipEventRecord(ev.a, st));
  for (int r = 0; r < reps; ++r) LQ_TRY(launch(2 + r));
  LQ_HIP_CHECK(hipEventRecord(ev.b, st));
  LQ_HIP_CHECK(hipEventSynchronize(ev.b));
  float t = 0.f;
  LQ_HIP_CHECK(hipEventElapsedTime(&t, ev.a, ev.b));
  *ms = (double)t / reps;
  *moved = (double)(words * sizeof(u32x4));  // read, read + write of halves, or written
  return LQ_OK;
} LQ_ABI_CATCH

}  // extern "C"
