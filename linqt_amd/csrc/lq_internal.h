// lq_internal.h -- shared declarations of the gfx950 engine (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/linqt_amd.h"

namespace lq {

void set_error(const char *fmt, ...);

// Every C ABI entry point is a function-try-block ending in LQ_ABI_CATCH: no
// C++ exception crosses the C boundary; it becomes an error code and the text
// lq_last_error() returns.
int abi_fail(const char *what, int code);
#define LQ_ABI_CATCH                                                      \
  catch (const std::bad_alloc &) {                                        \
    return ::lq::abi_fail("host memory allocation failed", LQ_ERR_NOMEM); \
  }                                                                       \
  catch (const std::exception &e) {                                       \
    return ::lq::abi_fail(e.what(), LQ_ERR_ARG);                          \
  }                                                                       \
  catch (...) {                                                           \
    return ::lq::abi_fail("unexpected C++ exception", LQ_ERR_ARG);        \
  }

#define LQ_HIP_CHECK(call)                                                                    \
  do {                                                                                        \
    hipError_t lq_e_ = (call);                                                                \
    if (lq_e_ != hipSuccess) {                                                                \
      ::lq::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call, hipGetErrorString(lq_e_)); \
      return LQ_ERR_HIP;                                                                      \
    }                                                                                         \
  } while (0)

#define LQ_TRY(expr)              \
  do {                            \
    int lq_rc_ = (expr);          \
    if (lq_rc_ != LQ_OK) return lq_rc_; \
  } while (0)

// Rows per slice of the sliced-ELL layout.  One slice is one 64-lane
// wavefront's worth of rows when a lane owns a row.
constexpr int kSlice = 64;

// Slots of the ring that holds per-block dot partials of the 1D recursion
// (one slot per moment or step); a full ring is reduced before it is reused.
constexpr int kPartRing = 64;
// Device bytes up to which the 1D solver keeps every moment's partials (one
// reduction at the end instead of a ring).
constexpr size_t kPartBytes = (size_t)256 << 20;

// Kernel-time categories recorded with HIP events.
enum TimeCat { kCatSpmv = 0, kCatContract = 1, kCatPost = 2, kNumCats = 3 };

// Device sliced-ELL operator.  Entry j of row r of slice s lives at
// sptr[s] + j*kSlice + (r % kSlice); padding entries have col = own row and
// value 0.
//
// Dictionary form (lattice operators: few distinct off-diagonal values): the
// diagonal is a separate [n] array and only off-diagonal entries are stored,
// each as one 32-bit word col | (value index << 24) into a table of at most
// 256 values (padding: own row and a zero value).  4 bytes per stored entry
// instead of 20; built whenever the operator qualifies, next to the plain form.
struct SellDev {
  int n = 0;
  int nslices = 0;
  int *sptr = nullptr;   // [nslices+1] entry offsets (multiples of kSlice)
  int *col = nullptr;    // [stored]
  double2 *val = nullptr;// [stored]
  long long stored = 0;
  int wmax = 0;          // widest slice (entries per row)
  // dictionary form (ndict == 0: absent)
  int ndict = 0;
  int *dsptr = nullptr;      // [nslices+1]
  unsigned *wcol = nullptr;  // [dstored]
  double2 *dict = nullptr;   // [ndict]
  double2 *diag = nullptr;   // [n] (nullptr: zero diagonal)
  long long dstored = 0;
  int dwmax = 0;
  // schedule of the dictionary-form launches, fixed when the operator is
  // uploaded (dict_schedule): grid cap and prefetched row groups
  int gridcap = 0;
  bool pf = true;
  // gather chunk of every launch (4, 8, 12, 16; 0: the smallest covering the
  // widest slice), set by the schedule tuner (lq_csr_tune)
  int unroll = 0;
};

// Reads the schedule knobs LQ_SPMM_GRIDCAP (grid cap of dictionary-form
// launches, >= 256) and LQ_SPMM_PF (0: no prefetched row groups) into d.
void dict_schedule(SellDev &d);

// Arguments of one vector set of a (multi-set) SpMM launch: y = a*A*x + b*z
// for k interleaved vectors ([n][k] layout).  z may alias y; z == nullptr
// means b is ignored and y is overwritten.
struct VecSet {
  const double2 *x = nullptr;
  const double2 *z = nullptr;
  double2 *y = nullptr;
  int k = 0;
};

// Dot-product fusion of the recursion kernel.
enum DotMode {
  kDotNone = 0,
  kDotPhi = 1,      // partial = sum_rows conj(phi[i][v]) * y[i][v]
  kDotDoubling = 2  // partials = sum conj(x_i)x_i and sum conj(x_i) y_i (row's own x)
};

// Right operators applied inside the recursion step of H (at most 2; their
// patterns lie inside H's): per stored slot of H's dictionary form, byte o of
// vidx indexes operator o's value table; vdiag[o] is its diagonal (or null).
struct FusedDev {
  int nv = 0;
  unsigned short *vidx = nullptr;
  double2 *vdict[2] = {nullptr, nullptr};
  int nvdict[2] = {0, 0};
  double2 *vdiag[2] = {nullptr, nullptr};
  unsigned long long key_uid[2] = {0, 0};
  int key_ver[2] = {-1, -1};
  int hver = -1;
};

// A launch's fused right-operator products: yv[o] = V_o x for the lanes of
// vector set `set` (same [n][k] layout as that set's y).
struct FusedArgs {
  int nv = 0;
  int set = 1;
  const unsigned short *vidx = nullptr;
  const double2 *vdict[2] = {nullptr, nullptr};
  int nvdict[2] = {0, 0};
  const double2 *vdiag[2] = {nullptr, nullptr};
  double2 *yv[2] = {nullptr, nullptr};
  int nt = 0;  // nonzero: yv stored and z loaded non-temporally (set by launch_spmm)
};

// One job of a batched recursion launch (device table entry).
struct BatchJobDev {
  const int *sptr;
  const int *col;
  const double2 *val;
  // dictionary form of the operator (ndict == 0: use sptr/col/val)
  const int *dsptr;
  const unsigned *wcol;
  const double2 *dict;
  const double2 *diag;
  int ndict;
  const double2 *phi;  // [n][k] states
  double2 *buf0, *buf1;  // T_s ring
  double2 *partial;      // [chunk slot][nblk][k-width] dot partials
  double2 *mom;          // [M][k-width] per-state moments
  int n, k, t0, steps, blk0, nblk;
};

// Launchers (lq_spmv.hip).  Return LQ_OK or an error code.
int launch_spmm(hipStream_t st, const SellDev &A, double2 alpha, double2 beta, const VecSet *sets,
                int nsets, int dot_mode, const double2 *phi, double2 *partial, int *grid_used,
                const FusedArgs *fused = nullptr);
int spmm_dot_grid(int n, int kpad);  // number of blocks a dot-mode launch uses
int launch_dot(hipStream_t st, int n, int k, const double2 *a, const double2 *b, double2 *partial,
               int *grid_used);
int launch_reduce_partials(hipStream_t st, const double2 *partial, int nslots, int grid, int width,
                           double2 *out);
int launch_copy(hipStream_t st, double2 *dst, const double2 *src, size_t count);
// Batched recursion (one width class kp, grid = sum of the jobs' nblk): step
// slot of the chunk whose first launch index is *tc.
int launch_spmm_batch(hipStream_t st, int kp, int unroll, const BatchJobDev *jobs, int njobs, int grid,
                      const int *tc, int slot);
int launch_batch_reduce(hipStream_t st, const BatchJobDev *jobs, int njobs, const int *tc, int nslots, int kp);
int launch_set_int(hipStream_t st, int *p, int v);
// dst[i*g + v] = src[(r0 + v)*n + i] (state-major device states -> [n][g]).
int launch_interleave(hipStream_t st, const double2 *src, int n, int r0, int g, double2 *dst);

// Contraction (lq_contract.hip): out[m0*ldo + m1] (+)= sum_kk conj(A[m0][kk]) B[m1][kk].
int contract_workspace_bytes(int M0, int M1, long long K, size_t *bytes, int *nsplit);
// Partial tables of the nsplit K-chunks go to work; launch_contract_reduce
// sums them (deterministically) into out, adding to it when accumulate != 0.
int launch_contract_f64(hipStream_t st, const double2 *A, long long lda, int M0, const double2 *B,
                        long long ldb, int M1, long long K, double2 *work, int nsplit);
int launch_contract_reduce(hipStream_t st, const double2 *work, int nsplit, int M0, int M1, double2 *out,
                           int ldo, int accumulate);
// The same for M1 = nout*bc columns (nout <= 2 operators' blocks of bc
// columns): column o*bc + s goes to outs[o][m*ldo + s].
int launch_contract_reduce_cols(hipStream_t st, const double2 *work, int nsplit, int M0, int M1, int bc,
                                double2 *const *outs, int ldo, int accumulate);

// Split-precision contraction (lq_ozaki.hip).  launch_slice cuts rows x K
// complex values (row stride ldx) into S int8 digit planes Q[row][t][re/im][Kp]
// and per-group exponents E[row][Kp / 4096]; launch_contract_i8 writes nsplit
// partial tables to work (same layout as launch_contract_f64).
long long ozaki_kp(long long K);
int ozaki_slice_bytes(int rows, long long K, int S, size_t *qbytes, size_t *ebytes);
int ozaki_nsplit(int M0, int M1, long long K);
int launch_slice(hipStream_t st, const double2 *X, long long ldx, int rows, long long K, int S, signed char *Q,
                 int *E);
int launch_contract_i8(hipStream_t st, const signed char *QA, const int *EA, int M0, const signed char *QB,
                       const int *EB, int M1, long long K, int S, double2 *work, int nsplit);
inline bool precision_ok(int p) { return p == 0 || p == 2 || p == 3 || (p >= 4 && p <= 8); }
// Split precision with int8 digit planes (LQ_PREC_INT8X4..X8).
inline bool precision_int8(int p) { return p >= 4 && p <= 8; }
// Floating-point split forms (lq_split.hip): FP32 hi/lo or 3 x BF16 pieces,
// FP32 MFMA accumulation folded into FP64 every 256 K-elements.
int split_nsplit(int M0, int M1, long long K);
int launch_contract_split(hipStream_t st, int precision, const double2 *A, long long lda, int M0, const double2 *B,
                          long long ldb, int M1, long long K, double2 *work, int nsplit);

// Post-processing (lq_post.hip).
int launch_corr_epilogue(hipStream_t st, const double2 *raw, int ldr, const double2 *init, int M0,
                         int M1, int B, int R, double2 *out);
int launch_scale(hipStream_t st, double2 *x, size_t count, double s);
int launch_kubo_pw(hipStream_t st, const double2 *mu, int ldm, int M, int nump, int formula,
                   const double *coef, double2 *out);
int launch_kubo_post(hipStream_t st, int formula, const double2 *final_data, int nump, int n, double W,
                     double Ec, double *E, double *sigma, double *integrand);
int launch_jackson_1d(hipStream_t st, double2 *mu, int maxMom);
int launch_jackson_2d(hipStream_t st, const double2 *mu_in, int M1_in, double2 *mu_out, int mm0, int mm1);
int launch_dos_fftgrid(hipStream_t st, const double2 *mu, int M, int n, double W, double Ec, int num_div,
                       double *E, double *rho);
// scratch: 2*num_div doubles; E, out: num_div-1 doubles each.
int launch_greenwood_fftgrid(hipStream_t st, const double2 *mu, int M0, int M1, int n, double W, double Ec,
                             int num_div, double *scratch, double *E, double *out);
int launch_bastin_fftgrid(hipStream_t st, const double2 *mu, int M0, int M1, int n, double W, double Ec,
                          int num_div, double *scratch, double *E, double *out);
// Time evolution and time-correlation post-processing (lq_post.hip).
int launch_caxpy(hipStream_t st, double2 c, const double2 *x, double2 *y, size_t count, int overwrite);
int launch_jackson_td(hipStream_t st, double2 *mu, int maxMom, int nT);
int launch_timecorr_post(hipStream_t st, const double2 *mu, int M, int nT, int n, double W, double Ec, double dt,
                         double EF, double *t, double *out);
// Linear-grid tools (kind = LQ_LIN_*): kern scratch of num_div doubles; E, out
// receive num_div (spectral) or num_div - 1 rows.
int launch_linear_grid(hipStream_t st, int kind, const double2 *mu, int M0, int M1, int n, double W, double Ec,
                       int num_div, double *kern, double *E, double *out);

}  // namespace lq

// What a captured recursion graph was built for (lq_api.hip).
struct GraphKey {
  unsigned long long uid = 0;
  int version = -1;
  const void *a = nullptr, *b = nullptr, *phi = nullptr, *partial = nullptr;
  int n = 0, k = 0, steps = 0;
  bool operator==(const GraphKey &o) const {
    return uid == o.uid && version == o.version && a == o.a && b == o.b && phi == o.phi && partial == o.partial &&
           n == o.n && k == o.k && steps == o.steps;
  }
};

// Context: one device, one stream, grow-only named scratch buffers.
struct lq_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  bool timing = false;
  lq_timing tsum{};
  struct Pending {
    int cat;
    hipEvent_t a, b;
    double bytes_or_flops;
    int launches;
  };
  // SpMM loop timing: one event pair around a whole recursion loop (per-launch
  // events would add their own overhead to kernels of a few microseconds)
  int spmv_loop = 0;
  hipEvent_t loop_ev = nullptr;
  double loop_bytes = 0.0;
  int loop_launches = 0;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
  };
  std::unordered_map<std::string, Buf> bufs;
  Buf pinned;  // grow-only pinned host staging (ctx_pinned)
  // Second stream (default priority) for the moment contraction of the
  // streamed 2D pipeline, and the timing-free events that order the two.
  hipStream_t stream2 = nullptr;
  std::vector<hipEvent_t> sync_events;
  // the most recent captured recursion chunk
  struct {
    hipGraphExec_t exec = nullptr;
    GraphKey key;
  } graph;
  GraphKey graph_seen;  // the last whole-loop solve run without a graph
};

// Host-side record of the completed operator (kept for rescale and for
// section masking) plus its device sliced-ELL image.
struct lq_csr_s {
  lq_ctx ctx = nullptr;
  int n = 0;
  int fill = 0;
  // completed entries in row order, columns ascending within a row
  std::vector<int> rowptr;
  std::vector<int> col;
  std::vector<std::complex<double>> val;
  std::vector<signed char> kind;  // 0: as stored (upper or general), 1: mirrored conj, 2: diagonal
  lq::SellDev dev;
  unsigned long long uid = 0;  // unique per operator of the process
  int version = 0;             // bumped by every rescale
  lq::FusedDev fused;          // right operators fused into this operator's recursion step
};

namespace lq {
// The same CSR already in device memory (row pointers, columns, values), so
// the device forms can be laid out without uploading it again.
struct DevCsr {
  const int *rowptr = nullptr;
  const int *col = nullptr;
  const double2 *val = nullptr;
};

// Takes ownership of A (host CSR filled in) and uploads it; on error A is freed.
// src: optional device copy of A's CSR (the host arrays stay A's).
int csr_adopt(lq_ctx ctx, lq_csr A, lq_csr *out, const DevCsr *src = nullptr);

// Makes a context's device current for the scope of a call.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};
bool is_device_ptr(const void *p, int device);
// State-major phi (host, or device memory of the context's GPU) -> device
// [n][g] interleaved block of states r0..r0+g-1.
int upload_states(lq_ctx ctx, const double *phi, int n, int r0, int g, double2 *dst);
// Looks up this device's address of the allocation-free entry point's scratch
// array (lq_noalloc.hip); called once per device, with it current, when a
// context is created.
int noalloc_init(int device);
// Context helpers (lq_api.hip).
int ctx_buf(lq_ctx ctx, const char *name, size_t bytes, void **p);
// Grow-only pinned host buffer of the context (device-to-host staging).
int ctx_pinned(lq_ctx ctx, size_t bytes, void **p);
// dst[0, bytes) = src[0, bytes) on up to 16 host threads.
void host_copy(void *dst, const void *src, size_t bytes);
// Device -> pageable host copies through the context's pinned staging buffer,
// 2 x kStageHalf bytes whatever the sizes, chunk copies overlapping host
// threads (the stream must be idle or ordered before).
constexpr size_t kStageHalf = (size_t)32 << 20;
struct D2H {
  const void *dev;
  void *host;
  size_t bytes;
};
int staged_d2h(lq_ctx ctx, const std::vector<D2H> &parts);
int time_begin(lq_ctx ctx, int cat, hipEvent_t *ev);
int time_end(lq_ctx ctx, int cat, hipEvent_t ev, double amount, int launches = 1);
// The same on another stream of the context (stream2).
int time_begin_on(lq_ctx ctx, hipStream_t st, int cat, hipEvent_t *ev);
int time_end_on(lq_ctx ctx, hipStream_t st, int cat, hipEvent_t ev, double amount, int launches = 1);
// The context's second stream and n ordering events (created on first use).
int ctx_stream2(lq_ctx ctx, hipStream_t *st);
int ctx_sync_events(lq_ctx ctx, int n, hipEvent_t **ev);
// Bracket a run of SpMM launches with one event pair (bytes and launches summed).
int spmv_loop_begin(lq_ctx ctx);
int spmv_loop_end(lq_ctx ctx);
int collect_timing(lq_ctx ctx);
}  // namespace lq
