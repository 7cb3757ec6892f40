"""linqt_amd -- MI355X (gfx950) engine for LinQT's Chebyshev-moment path.

Python mirror of the reference's solver interface over the C ABI
(``include/linqt_amd.h``).  Names and argument meanings follow the
reference's C++ API:

* :class:`SparseMatrix` -- ``SparseMatrixType`` (Hermitian/upper CSR,
  ``Multiply``, ``Rescale``), ``include/sparse_matrix.hpp:47-87``;
* :func:`spectral_moments` -- ``chebyshev::SpectralMoments``
  (``src/chebyshev_solver.cpp:176-200``);
* :func:`correlation_expansion_moments` --
  ``chebyshev::CorrelationExpansionMoments`` (``src/chebyshev_solver.cpp:86-172``);
* :func:`kubo_fft` / :func:`kubo_post` -- ``Kubo_solver_FFT::compute`` and
  its post-processing (``src/Kubo_solver_FFT.cpp``,
  ``src/Kubo_solver_FFT_postProcess.cpp``);
* :func:`jackson_1d`, :func:`jackson_2d`, :func:`dos_fftgrid`,
  :func:`greenwood_fftgrid`, :func:`bastin_fftgrid` -- the Jackson kernels
  and the FFT-grid post-processors.

Every call runs on the GPU through ``liblinqt_amd.so``; nothing falls back to
the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (LQ_ALGO_DIRECT, LQ_ALGO_DOUBLING, LQ_FILL_GENERAL, LQ_FILL_HERMITIAN_UPPER,
                   LQ_KUBO_BASTIN, LQ_KUBO_GREENWOOD, LQ_PREC_FP64, LQ_PREC_FP32X2, LQ_PREC_BF16X3, LQ_PREC_INT8X4, LQ_PREC_INT8X5,
                   LQ_PREC_INT8X6, LQ_PREC_INT8X7, LQ_PREC_INT8X8, LQ_FMT_TEXT, LQ_FMT_TEXT_REF, LQ_FMT_BINARY, LinqtError)

__all__ = [
    "Context", "SparseMatrix", "DeviceStates", "Batch", "read_csr_file", "write_csr_file", "read_mom_file",
    "write_mom_file", "linear_grid", "time_correlations", "jackson_td",
    "timecorr_post", "spectral_moments", "spectral_moments_nonorth", "overlap_solve", "correlation_expansion_moments", "kubo_fft", "kubo_post",
    "jackson_1d", "jackson_2d", "dos_fftgrid", "greenwood_fftgrid", "bastin_fftgrid", "LinqtError",
    "LQ_FILL_HERMITIAN_UPPER", "LQ_FILL_GENERAL", "LQ_KUBO_GREENWOOD", "LQ_KUBO_BASTIN", "LQ_ALGO_DIRECT",
    "LQ_ALGO_DOUBLING", "LQ_PREC_FP64", "LQ_PREC_FP32X2", "LQ_PREC_BF16X3", "LQ_PREC_INT8X4", "LQ_PREC_INT8X5", "LQ_PREC_INT8X6", "LQ_PREC_INT8X7",
    "LQ_PREC_INT8X8",
]


def _dp(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    assert a.flags.c_contiguous and a.dtype == np.int32
    return a.ctypes.data_as(C.POINTER(C.c_int))


def device_count() -> int:
    n = C.c_int(0)
    rc = _lib.load().lq_device_count(C.byref(n))
    return n.value if rc == 0 else 0


class Context:
    """One GPU, one HIP stream (``lq_ctx``)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _lib.call("lq_ctx_create", device, C.byref(h))
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            _lib.load().lq_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def release(self):
        """Frees the device scratch buffers the context keeps between calls."""
        _lib.call("lq_ctx_release", self.h)

    def synchronize(self):
        _lib.call("lq_ctx_synchronize", self.h)

    def set_timing(self, on: bool = True):
        _lib.call("lq_ctx_set_timing", self.h, 1 if on else 0)

    def reset_timing(self):
        _lib.call("lq_ctx_reset_timing", self.h)

    def timing(self) -> dict:
        t = _lib.lq_timing()
        _lib.call("lq_ctx_get_timing", self.h, C.byref(t))
        return {name: getattr(t, name) for name, _ in t._fields_}

    # device buffers (for lq_spmv)
    def alloc(self, nbytes: int) -> C.c_void_p:
        p = C.c_void_p()
        _lib.call("lq_malloc", self.h, nbytes, C.byref(p))
        return p

    def free(self, p: C.c_void_p):
        _lib.call("lq_free", self.h, p)

    def upload(self, p: C.c_void_p, a: np.ndarray):
        a = np.ascontiguousarray(a)
        _lib.call("lq_memcpy_h2d", self.h, p, a.ctypes.data_as(C.c_void_p), a.nbytes)

    def download(self, p: C.c_void_p, a: np.ndarray):
        assert a.flags.c_contiguous
        _lib.call("lq_memcpy_d2h", self.h, a.ctypes.data_as(C.c_void_p), p, a.nbytes)


class SparseMatrix:
    """``SparseMatrixType``: a device operator built from host CSR arrays.

    ``fill=LQ_FILL_HERMITIAN_UPPER`` (the reference's only descriptor) treats
    the arrays as the upper triangle of a Hermitian matrix.
    """

    def __init__(self, ctx: Context, n, rowptr, col, val, fill: int = LQ_FILL_HERMITIAN_UPPER, id: str = ""):
        self.ctx = ctx
        self.n = int(n)
        self.id = id
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.complex128)
        h = C.c_void_p()
        _lib.call("lq_csr_create", ctx.h, self.n, int(col.shape[0]), _ip(rowptr), _ip(col),
                  _dp(val.view(np.float64)), fill, C.byref(h))
        self.h = h

    @classmethod
    def from_csr(cls, ctx: Context, m, fill: int = LQ_FILL_HERMITIAN_UPPER, id: str = ""):
        return cls(ctx, m.n, m.rowptr, m.col, m.val, fill=fill, id=id)

    @classmethod
    def from_file(cls, ctx: Context, path: str, id: str = ""):
        """``SparseMatrixBuilder::BuildOPFromCSRFile`` (text or binary CSR, parsed natively)."""
        return cls.from_csr(ctx, read_csr_file(path), id=id)

    @classmethod
    def from_hoppings(cls, ctx: Context, nwann: int, cell, edge, val, dims, id: str = ""):
        """wannier2sparse's ``wrap_in_supercell`` + ``save_hopping_list_as_csr``
        (utilities/wannier2sparse/src/hopping_list.cpp:156-285) generated on the
        device: the Hermitian-upper operator of a unit-cell hopping list on a
        ``dims`` supercell.  cell: (nhop, 3) ints, edge: (nhop, 2) zero-based
        orbitals (from, to), val: (nhop,) complex."""
        cell = np.ascontiguousarray(np.asarray(cell, dtype=np.int32).reshape(-1, 3))
        edge = np.ascontiguousarray(np.asarray(edge, dtype=np.int32).reshape(-1, 2))
        v = np.ascontiguousarray(np.asarray(val, dtype=np.complex128).reshape(-1))
        nh = v.shape[0]
        if nh == 0:
            cell, edge, v = np.zeros((1, 3), np.int32), np.zeros((1, 2), np.int32), np.zeros(1, np.complex128)
        d = np.ascontiguousarray(np.asarray(dims, dtype=np.int32).reshape(3))
        self = cls.__new__(cls)
        self.ctx, self.id = ctx, id
        h = C.c_void_p()
        _lib.call("lq_csr_from_hoppings", ctx.h, int(nwann), nh, _ip(cell), _ip(edge), _dp(v.view(np.float64)),
                  _ip(d), C.byref(h))
        self.h = h
        self.n = self.info()["n"]
        return self

    def export(self):
        """The stored entries (what a reference CSR file holds) as ``io.CSR``."""
        from .io import CSR

        nnz = C.c_longlong(0)
        _lib.call("lq_csr_export", self.h, C.byref(nnz), None, None, None)
        rowptr = np.zeros(self.n + 1, np.int32)
        col = np.zeros(max(nnz.value, 1), np.int32)
        val = np.zeros(max(nnz.value, 1), np.complex128)
        _lib.call("lq_csr_export", self.h, C.byref(nnz), _ip(rowptr), _ip(col), _dp(val.view(np.float64)))
        return CSR(self.n, rowptr, col[: nnz.value], val[: nnz.value])

    def isIdentity(self) -> bool:
        return self.id == "1"

    def close(self):
        # a closed context already released the device; nothing to free then
        if getattr(self, "h", None) is not None and self.h.value and self.ctx.h is not None:
            _lib.load().lq_csr_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def Rescale(self, a, b):
        """``SparseMatrixType::Rescale``: A <- a*A + b*I."""
        a, b = complex(a), complex(b)
        _lib.call("lq_csr_rescale", self.h, a.real, a.imag, b.real, b.imag)
        return self

    def tune(self, k: int, mode: int = 0) -> dict:
        """Picks the fastest recursion-step schedule for k vectors (mode 1: with
        the <phi|T> dot fused) and keeps it in this operator; see lq_csr_tune."""
        u, g, p, c = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _lib.call("lq_csr_tune", self.h, k, mode, C.byref(u), C.byref(g), C.byref(p), C.byref(c))
        return {"unroll": u.value, "gridcap": g.value, "prefetch": p.value, "from_cache": bool(c.value)}

    def info(self):
        n, nnz, st = C.c_int(), C.c_longlong(), C.c_longlong()
        _lib.call("lq_csr_info", self.h, C.byref(n), C.byref(nnz), C.byref(st))
        return {"n": n.value, "nnz": nnz.value, "stored": st.value}

    def Multiply(self, a, x: np.ndarray, b=0.0, y: np.ndarray | None = None) -> np.ndarray:
        """``SparseMatrixType::Multiply(a, x, b, y)`` on host arrays.

        x, y are (n,) or (n, k) complex; returns the new y."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        k = 1 if x.ndim == 1 else x.shape[1]
        y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, dtype=np.complex128).copy()
        px = self.ctx.alloc(x.nbytes)
        py = self.ctx.alloc(y.nbytes)
        try:
            self.ctx.upload(px, x)
            self.ctx.upload(py, y)
            a, b = complex(a), complex(b)
            _lib.call("lq_spmv", self.h, a.real, a.imag, px, b.real, b.imag, py, k)
            self.ctx.download(py, y)
        finally:
            self.ctx.free(px)
            self.ctx.free(py)
        return y


class DeviceStates:
    """R states of length n resident in device memory (state-major complex128).

    Passing one to a solver keeps the inputs in HBM: no host transfer happens
    inside the solve."""

    def __init__(self, ctx: Context, phi: np.ndarray):
        phi = np.ascontiguousarray(np.atleast_2d(phi), dtype=np.complex128)
        self.ctx = ctx
        self.R, self.n = phi.shape
        self.ptr = ctx.alloc(phi.nbytes)
        ctx.upload(self.ptr, phi)

    def close(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value and self.ctx.h is not None:
            self.ctx.free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _states(phi):
    """Returns (double pointer, R) for host arrays or DeviceStates."""
    if isinstance(phi, DeviceStates):
        return C.cast(phi.ptr, C.POINTER(C.c_double)), phi.R, None
    arr = np.ascontiguousarray(np.atleast_2d(phi), dtype=np.complex128)
    return _dp(arr.view(np.float64)), arr.shape[0], arr


def _h(op):
    return None if op is None else op.h


def spectral_moments(H: SparseMatrix, OP, phi: np.ndarray, M: int, algo: int = LQ_ALGO_DIRECT,
                     mu: np.ndarray | None = None) -> np.ndarray:
    """``chebyshev::SpectralMoments``: mu_m += sum_r (2/R)(1/2)^{delta_m0} <phi_r|T_m(H) OP|phi_r>.

    OP None (or an operator with id "1") is the identity."""
    if OP is not None and OP.isIdentity():
        OP = None
    pp, R, keep = _states(phi)
    mu = np.zeros(M, dtype=np.complex128) if mu is None else np.array(mu, dtype=np.complex128)
    _lib.call("lq_spectral_moments", H.h, _h(OP), pp, R, M, algo, _dp(mu.view(np.float64)))
    return mu


def spectral_moments_nonorth(H: SparseMatrix, OP, S: SparseMatrix, phi: np.ndarray, M: int, tol: float = 1e-4,
                             maxit: int = 2000, mu: np.ndarray | None = None, info: dict | None = None) -> np.ndarray:
    """``chebyshev::SpectralMoments_nonOrth`` (src/chebyshev_solver.cpp:203-229):
    the Chebyshev recursion of S^-1 H with BiCGSTAB overlap solves.  S should
    be built with ``fill=LQ_FILL_GENERAL`` from the overlap file's arrays, the
    matrix the reference solves with.  ``info`` (a dict) receives the total
    BiCGSTAB iterations."""
    if OP is not None and OP.isIdentity():
        OP = None
    pp, R, keep = _states(phi)
    mu = np.zeros(M, dtype=np.complex128) if mu is None else np.array(mu, dtype=np.complex128)
    it = C.c_longlong(0)
    _lib.call("lq_spectral_moments_nonorth", H.h, _h(OP), S.h, pp, R, M, tol, maxit, _dp(mu.view(np.float64)),
              C.byref(it))
    if info is not None:
        info["iterations"] = it.value
    return mu


def overlap_solve(S: SparseMatrix, b: np.ndarray, tol: float = 1e-4, maxit: int = 2000):
    """``linalg::orthogonalize`` (src/mkl_linear_algebra.cpp:100-131): x = S^-1 b
    by BiCGSTAB; returns (x, iterations, |r| / |b|)."""
    ctx = S.ctx
    b = np.ascontiguousarray(b, dtype=np.complex128)
    pb, px = ctx.alloc(b.nbytes), ctx.alloc(b.nbytes)
    try:
        ctx.upload(pb, b)
        it, err = C.c_int(0), C.c_double(0.0)
        _lib.call("lq_overlap_solve", S.h, pb, px, tol, maxit, C.byref(it), C.byref(err))
        x = np.zeros_like(b)
        ctx.download(px, x)
    finally:
        ctx.free(pb)
        ctx.free(px)
    return x, it.value, err.value


def time_correlations(H: SparseMatrix, OPL, OPR, phi, M: int, nT: int, dt: float, W: float, Ec: float,
                      mu=None) -> np.ndarray:
    """``chebyshev::TimeDependentCorrelations``: (M, nT) table mu(m, n) at t_n = n dt (fs),
    accumulated into mu when given.  H rescaled with bandwidth W and centre Ec."""
    pp, R, keep = _states(phi)
    out = np.zeros((M, nT), dtype=np.complex128) if mu is None else np.array(mu, dtype=np.complex128)
    _lib.call("lq_time_correlations", H.h, _h(OPL), _h(OPR), pp, R, M, nT, dt, W, Ec, _dp(out.view(np.float64)))
    return out


def jackson_td(ctx: Context, mu: np.ndarray, W: float, broad: float) -> np.ndarray:
    mu = np.array(mu, dtype=np.complex128, order="C")
    mo = C.c_int(0)
    _lib.call("lq_jackson_td", ctx.h, _dp(mu.view(np.float64)), mu.shape[0], mu.shape[1], W, broad, C.byref(mo))
    return mu[: mo.value].copy()


def timecorr_post(ctx: Context, mu: np.ndarray, n: int, W: float, Ec: float, dt: float, EF: float):
    mu = np.ascontiguousarray(mu, dtype=np.complex128)
    t, o = np.zeros(mu.shape[1]), np.zeros(mu.shape[1])
    _lib.call("lq_timecorr_post", ctx.h, _dp(mu.view(np.float64)), mu.shape[0], mu.shape[1], n, W, Ec, dt, EF,
              _dp(t), _dp(o))
    return t, o


def linear_grid(ctx: Context, kind: int, mu: np.ndarray, n: int, W: float, Ec: float, num_div: int | None = None):
    """The linear-grid tools (``*FromChebmom.cpp``) on Jackson-damped moments:
    kind LQ_LIN_*; num_div defaults to the tools' 30 * max moments.  Returns
    (E, value) rows as the tool prints them."""
    mu = np.ascontiguousarray(mu, dtype=np.complex128)
    M0, M1 = (mu.shape[0], 1) if mu.ndim == 1 else mu.shape
    if num_div is None:
        num_div = 30 * max(M0, M1)
    rows = num_div if kind == _lib.LQ_LIN_SPECTRAL else num_div - 1
    E, out = np.zeros(num_div), np.zeros(num_div)
    _lib.call("lq_linear_grid", ctx.h, kind, _dp(mu.view(np.float64)), M0, M1, n, W, Ec, num_div, _dp(E), _dp(out))
    return E[:rows], out[:rows]


class Batch:
    """Batched ``SpectralMoments`` jobs (identity OP) on one context.

    ``submit`` queues a job; ``run`` advances every admitted job by chunks of
    recursion steps with one grouped launch per step per state-count class,
    admitting queued jobs at chunk boundaries while their device footprint
    fits ``hbm_budget`` bytes (0: 80 % of free memory).  A job's moments are
    bitwise the same alone, in any mix, and to ``spectral_moments``."""

    def __init__(self, ctx: Context, hbm_budget: int = 0, chunk: int = 32):
        self.ctx = ctx
        self.h = C.c_void_p()
        self.chunk = chunk
        _lib.call("lq_batch_create", ctx.h, hbm_budget, chunk, C.byref(self.h))
        self._keep = {}

    @staticmethod
    def job_bytes(n: int, R: int, M: int, chunk: int = 32) -> int:
        b = C.c_size_t(0)
        _lib.call("lq_batch_job_bytes", n, R, M, chunk, C.byref(b))
        return b.value

    def submit(self, H: SparseMatrix, phi, M: int) -> int:
        pp, R, keep = _states(phi)
        job = C.c_int(0)
        _lib.call("lq_batch_submit", self.h, H.h, pp, R, M, C.byref(job))
        self._keep[job.value] = (H, M, phi if isinstance(phi, DeviceStates) else None)
        return job.value

    def run(self, max_chunks: int = 0):
        """Returns (running, queued) after at most max_chunks chunks (0: until idle)."""
        r, q = C.c_int(0), C.c_int(0)
        _lib.call("lq_batch_run", self.h, max_chunks, C.byref(r), C.byref(q))
        return r.value, q.value

    def state(self, job: int) -> int:
        s = C.c_int(0)
        _lib.call("lq_batch_poll", self.h, job, C.byref(s))
        return s.value

    def result(self, job: int) -> np.ndarray:
        M = self._keep[job][1]
        mu = np.zeros(M, dtype=np.complex128)
        _lib.call("lq_batch_result", self.h, job, _dp(mu.view(np.float64)))
        del self._keep[job]
        return mu

    def unroll(self, R: int) -> int:
        u = C.c_int(0)
        _lib.call("lq_batch_unroll", self.h, R, C.byref(u))
        return u.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value and self.ctx.h is not None:
            _lib.call("lq_batch_destroy", self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def correlation_expansion_moments(H: SparseMatrix, OPL, OPR, phi: np.ndarray, M0: int, M1: int | None = None,
                                  batch: int = 0, precision: int = LQ_PREC_FP64, mu=None) -> np.ndarray:
    """``chebyshev::CorrelationExpansionMoments``.

    OPR may be one operator or a list (sharing the left recursion); returns
    (M0, M1) or (len(OPR), M0, M1) complex tables."""
    M1 = M0 if M1 is None else M1
    single = not isinstance(OPR, (list, tuple))
    oprs = [OPR] if single else list(OPR)
    arr = (C.c_void_p * len(oprs))(*[None if o is None else o.h.value for o in oprs])
    pp, R, keep = _states(phi)
    if mu is None:
        out = np.zeros((len(oprs), M0, M1), dtype=np.complex128)
    else:
        out = np.array(mu, dtype=np.complex128).reshape(len(oprs), M0, M1).copy()
    _lib.call("lq_correlation_moments", H.h, _h(OPL), len(oprs), C.cast(arr, C.POINTER(C.c_void_p)),
              pp, R, M0, M1, batch, precision, _dp(out.view(np.float64)))
    return out[0] if single else out


def kubo_fft(H: SparseMatrix, OPL, OPR, phi: np.ndarray, M: int, nump: int, formula: int,
             num_sections: int = 1, precision: int = LQ_PREC_FP64) -> np.ndarray:
    """``Kubo_solver_FFT::compute``: the mean over states of r_data, (2*nump,) per operator."""
    single = not isinstance(OPR, (list, tuple))
    oprs = [OPR] if single else list(OPR)
    arr = (C.c_void_p * len(oprs))(*[None if o is None else o.h.value for o in oprs])
    pp, R, keep = _states(phi)
    fd = np.zeros((len(oprs), 2 * nump), dtype=np.complex128)
    _lib.call("lq_kubo_fft", H.h, _h(OPL), len(oprs), C.cast(arr, C.POINTER(C.c_void_p)),
              pp, R, M, num_sections, nump, formula, precision, _dp(fd.view(np.float64)))
    return fd[0] if single else fd


def kubo_post(ctx: Context, formula: int, final_data: np.ndarray, n: int, W: float, Ec: float):
    """``Kubo_solver_FFT_postProcess``: returns (E, sigma, integrand)."""
    fd = np.ascontiguousarray(final_data, dtype=np.complex128)
    nump = fd.shape[0] // 2
    E, s, g = np.zeros(nump), np.zeros(nump), np.zeros(nump)
    _lib.call("lq_kubo_post", ctx.h, formula, _dp(fd.view(np.float64)), nump, n, W, Ec, _dp(E), _dp(s), _dp(g))
    return E, s, g


def jackson_1d(ctx: Context, mu: np.ndarray, W: float, broad_meV: float) -> np.ndarray:
    mu = np.array(mu, dtype=np.complex128)
    m = C.c_int(0)
    _lib.call("lq_jackson_1d", ctx.h, _dp(mu.view(np.float64)), mu.shape[0], W, broad_meV, C.byref(m))
    return mu[: m.value].copy()


def jackson_2d(ctx: Context, mu: np.ndarray, W: float, b0: float, b1: float) -> np.ndarray:
    mu = np.array(mu, dtype=np.complex128)
    M0, M1 = mu.shape
    o0, o1 = C.c_int(0), C.c_int(0)
    _lib.call("lq_jackson_2d", ctx.h, _dp(mu.view(np.float64)), M0, M1, W, b0, b1, C.byref(o0), C.byref(o1))
    return mu.reshape(-1)[: o0.value * o1.value].reshape(o0.value, o1.value).copy()


def dos_fftgrid(ctx: Context, mu: np.ndarray, n: int, W: float, Ec: float, num_div: int):
    mu = np.ascontiguousarray(mu, dtype=np.complex128)
    E, rho = np.zeros(num_div), np.zeros(num_div)
    _lib.call("lq_dos_fftgrid", ctx.h, _dp(mu.view(np.float64)), mu.shape[0], n, W, Ec, num_div, _dp(E), _dp(rho))
    return E, rho


def greenwood_fftgrid(ctx: Context, mu: np.ndarray, n: int, W: float, Ec: float, num_div: int):
    mu = np.ascontiguousarray(mu, dtype=np.complex128)
    E, s = np.zeros(num_div - 1), np.zeros(num_div - 1)
    _lib.call("lq_greenwood_fftgrid", ctx.h, _dp(mu.view(np.float64)), mu.shape[0], mu.shape[1], n, W, Ec, num_div,
              _dp(E), _dp(s))
    return E, s


def bastin_fftgrid(ctx: Context, mu: np.ndarray, n: int, W: float, Ec: float, num_div: int):
    mu = np.ascontiguousarray(mu, dtype=np.complex128)
    E, s = np.zeros(num_div - 1), np.zeros(num_div - 1)
    _lib.call("lq_bastin_fftgrid", ctx.h, _dp(mu.view(np.float64)), mu.shape[0], mu.shape[1], n, W, Ec, num_div,
              _dp(E), _dp(s))
    return E, s


# ---- file ingestion (host side; no GPU needed) --------------------------------
def read_csr_file(path: str):
    """``Sparse::OPERATOR_FromCSRFile`` (src/sparse_matrix.cpp:3-50) through the
    native multi-threaded reader; text or binary CSR.  Returns ``io.CSR``."""
    from .io import CSR

    n, nnz = C.c_int(0), C.c_longlong(0)
    bp = str(path).encode()
    _lib.call("lq_csr_file_read", bp, C.byref(n), C.byref(nnz), None, None, None)
    rowptr = np.zeros(n.value + 1, dtype=np.int32)
    col = np.zeros(max(nnz.value, 1), dtype=np.int32)
    val = np.zeros(max(nnz.value, 1), dtype=np.complex128)
    _lib.call("lq_csr_file_read", bp, C.byref(n), C.byref(nnz), _ip(rowptr), _ip(col), _dp(val.view(np.float64)))
    return CSR(n.value, rowptr, col[: nnz.value], val[: nnz.value])


def write_csr_file(path: str, m, fmt: int = _lib.LQ_FMT_TEXT) -> None:
    """Writes an ``io.CSR``; ``LQ_FMT_TEXT_REF`` is byte for byte what
    ``hopping_list::save_hopping_list_as_csr`` writes (hopping_list.cpp:262-283)."""
    rowptr = np.ascontiguousarray(m.rowptr, dtype=np.int32)
    col = np.ascontiguousarray(m.col, dtype=np.int32)
    val = np.ascontiguousarray(m.val, dtype=np.complex128)
    if col.size == 0:
        col, val = np.zeros(1, np.int32), np.zeros(1, np.complex128)
    _lib.call("lq_csr_file_write", str(path).encode(), fmt, int(m.n), int(m.nnz), _ip(rowptr), _ip(col),
              _dp(val.view(np.float64)))


MOM_1D, MOM_2D, MOM_TD = 1, 2, 3


def read_mom_file(path: str):
    """Moments1D / Moments2D / MomentsTD file constructors, text or binary.
    Returns (mu, header) with mu of shape (M,), (M0, M1) or (M, nT) and header
    a dict with kind, n, W, Ec, dt."""
    h = _lib.lq_mom_header()
    bp = str(path).encode()
    _lib.call("lq_mom_file_read", bp, C.byref(h), None)
    mu = np.zeros(h.M0 * h.M1, dtype=np.complex128)
    _lib.call("lq_mom_file_read", bp, C.byref(h), _dp(mu.view(np.float64)))
    if h.kind != MOM_1D:
        mu = mu.reshape(h.M0, h.M1)
    return mu, {"kind": h.kind, "n": h.n, "W": h.W, "Ec": h.Ec, "dt": h.dt}


def write_mom_file(path: str, mu: np.ndarray, n: int, W: float, Ec: float, kind: int | None = None,
                   dt: float = 0.0, fmt: int = _lib.LQ_FMT_TEXT_REF) -> None:
    """Moments1D/2D/TD::saveIn layouts (LQ_FMT_TEXT_REF: the reference's 15
    digits), 17-digit text or binary."""
    mu = np.ascontiguousarray(mu, dtype=np.complex128)
    if kind is None:
        kind = MOM_1D if mu.ndim == 1 else MOM_2D
    M0, M1 = (mu.shape[0], 1) if mu.ndim == 1 else mu.shape
    h = _lib.lq_mom_header(kind, M0, M1, n, W, Ec, dt)
    _lib.call("lq_mom_file_write", str(path).encode(), fmt, C.byref(h), _dp(mu.reshape(-1).view(np.float64)))

