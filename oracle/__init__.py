"""ctypes binding of the CPU FP64 oracle (``liblq_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``, where it is the checker or the
timed CPU baseline -- never the thing measured or shipped.

Arrays cross as contiguous numpy complex128 (interleaved re/im doubles).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblq_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build() -> None:
    """Compiles the oracle in place (``make`` in this directory)."""
    import subprocess

    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(_LIB_PATH)
        L.lqo_op_create.restype = C.c_void_p
        L.lqo_op_create.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp]
        L.lqo_op_destroy.argtypes = [C.c_void_p]
        L.lqo_op_nnz.argtypes = [C.c_void_p]
        L.lqo_op_rescale.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double]
        L.lqo_op_mv.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp, C.c_double, C.c_double, _dp]
        L.lqo_op_dense.argtypes = [C.c_void_p, _dp]
        L.lqo_spectral_moments.argtypes = [C.c_void_p, C.c_void_p, C.c_int, _dp, C.c_int, C.c_int, _dp]
        L.lqo_correlation_moments.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, _dp, C.c_int,
                                              C.c_int, C.c_int, C.c_int, _dp]
        L.lqo_dft_backward.argtypes = [C.c_int, _dp, _dp]
        L.lqo_kubo_jackson.restype = C.c_double
        L.lqo_kubo_jackson.argtypes = [C.c_int, C.c_int]
        L.lqo_kubo_fft.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, _dp, C.c_int, C.c_int,
                                   C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.lqo_kubo_post.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]
        L.lqo_jackson_1d.argtypes = [_dp, C.c_int, C.c_double, C.c_double]
        L.lqo_jackson_2d.argtypes = [_dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _ip, _ip]
        L.lqo_time_correlations.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, _dp, C.c_int, C.c_int,
                                            C.c_int, C.c_double, C.c_double, C.c_double, _dp]
        L.lqo_jackson_td.argtypes = [_dp, C.c_int, C.c_int, C.c_double, C.c_double, _ip]
        L.lqo_timecorr_post.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double, _dp, _dp]
        L.lqo_linear_kernel.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _dp, _dp]
        L.lqo_linear_grid.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                      _dp, _dp]
        for fn in ("lqo_dos_fftgrid",):
            getattr(L, fn).argtypes = [_dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp]
        for fn in ("lqo_greenwood_fftgrid", "lqo_bastin_fftgrid"):
            getattr(L, fn).argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.lqo_supercell_csr.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _ip, _ip, C.POINTER(C.c_longlong), _ip,
                                        _ip, _dp]
        L.lqo_overlap_solve.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_double, C.c_int, _dp]
        L.lqo_spectral_moments_nonorth.argtypes = [C.c_void_p, C.c_void_p, C.c_int, _ip, _ip, _dp, _dp, C.c_int,
                                                   C.c_int, C.c_double, C.c_int, _dp]
        L.lqo_csr_read_stream.argtypes = [C.c_char_p, _ip, C.POINTER(C.c_longlong), _ip, _ip, _dp]
        L.lqo_omp_threads.restype = C.c_int
        L.lqo_omp_cheb_step.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.lqo_omp_momtable.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp]
        L.lqo_omp_spectral_moments.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, C.c_int, _dp]
        L.lqo_omp_mv.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.lqo_omp_bastin_ffts.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, _dp]
        _lib = L
    return _lib


def _d(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    assert a.flags.c_contiguous and a.dtype == np.int32
    return a.ctypes.data_as(_ip)


def _check(rc: int, what: str) -> None:
    if rc < 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


class Op:
    """An oracle operator: the reference's SparseMatrixType (Hermitian/upper)."""

    def __init__(self, n, rowptr, col, val):
        self.n = int(n)
        self._rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self._col = np.ascontiguousarray(col, dtype=np.int32)
        self._val = np.ascontiguousarray(val, dtype=np.complex128)
        self.h = lib().lqo_op_create(self.n, int(self._col.shape[0]), _i(self._rowptr), _i(self._col),
                                     _d(self._val))
        if not self.h:
            raise ValueError("invalid CSR arrays")

    @classmethod
    def from_csr(cls, m):
        return cls(m.n, m.rowptr, m.col, m.val)

    def __del__(self):
        if getattr(self, "h", None):
            lib().lqo_op_destroy(self.h)
            self.h = None

    def rescale(self, a, b):
        a, b = complex(a), complex(b)
        _check(lib().lqo_op_rescale(self.h, a.real, a.imag, b.real, b.imag), "rescale")
        return self

    def mv(self, a, x, b, y):
        a, b = complex(a), complex(b)
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.array(y, dtype=np.complex128, copy=True)
        _check(lib().lqo_op_mv(self.h, a.real, a.imag, _d(x), b.real, b.imag, _d(y)), "mv")
        return y

    def nnz(self) -> int:
        return lib().lqo_op_nnz(self.h)

    def dense(self):
        d = np.zeros((self.n, self.n), dtype=np.complex128)
        _check(lib().lqo_op_dense(self.h, _d(d)), "dense")
        return d


def _h(op):
    return None if op is None else op.h


def spectral_moments(H: Op, OP, phi: np.ndarray, M: int, mu=None) -> np.ndarray:
    phi = np.ascontiguousarray(np.atleast_2d(phi), dtype=np.complex128)
    mu = np.zeros(M, dtype=np.complex128) if mu is None else np.array(mu, dtype=np.complex128)
    _check(lib().lqo_spectral_moments(H.h, _h(OP), H.n, _d(phi), phi.shape[0], M, _d(mu)), "spectral_moments")
    return mu


def correlation_moments(H: Op, OPL, OPR, phi: np.ndarray, M0: int, M1: int, batch: int = 0,
                        mu=None) -> np.ndarray:
    phi = np.ascontiguousarray(np.atleast_2d(phi), dtype=np.complex128)
    mu = np.zeros((M0, M1), dtype=np.complex128) if mu is None else np.array(mu, dtype=np.complex128)
    _check(lib().lqo_correlation_moments(H.h, _h(OPL), _h(OPR), H.n, _d(phi), phi.shape[0], M0, M1,
                                         batch, _d(mu)), "correlation_moments")
    return mu


def dft_backward(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    _check(lib().lqo_dft_backward(x.shape[0], _d(x), _d(out)), "dft")
    return out


def kubo_jackson(m: int, M: int) -> float:
    return lib().lqo_kubo_jackson(m, M)


def kubo_fft(H: Op, OPL, OPR, phi: np.ndarray, M: int, nump: int, formula: int, num_sections: int = 1):
    """Returns (final_data[2*nump], last_r_data[2*nump])."""
    phi = np.ascontiguousarray(np.atleast_2d(phi), dtype=np.complex128)
    fd = np.zeros(2 * nump, dtype=np.complex128)
    rd = np.zeros(2 * nump, dtype=np.complex128)
    _check(lib().lqo_kubo_fft(H.h, _h(OPL), _h(OPR), H.n, _d(phi), phi.shape[0], M, num_sections, nump,
                              formula, _d(fd), _d(rd)), "kubo_fft")
    return fd, rd


def kubo_post(formula: int, final_data: np.ndarray, n: int, W: float, Ec: float):
    """Returns (E, sigma, integrand) as the reference's output files hold them."""
    fd = np.ascontiguousarray(final_data, dtype=np.complex128)
    nump = fd.shape[0] // 2
    E, s, g = np.zeros(nump), np.zeros(nump), np.zeros(nump)
    _check(lib().lqo_kubo_post(formula, _d(fd), nump, n, W, Ec, _d(E), _d(s), _d(g)), "kubo_post")
    return E, s, g


def jackson_1d(mu: np.ndarray, W: float, broad: float) -> np.ndarray:
    mu = np.array(mu, dtype=np.complex128)
    m = lib().lqo_jackson_1d(_d(mu), mu.shape[0], W, broad)
    _check(m, "jackson_1d")
    return mu[:m].copy()


def jackson_2d(mu: np.ndarray, W: float, b0: float, b1: float) -> np.ndarray:
    mu = np.array(mu, dtype=np.complex128)
    M0, M1 = mu.shape
    o0, o1 = C.c_int(0), C.c_int(0)
    _check(lib().lqo_jackson_2d(_d(mu), M0, M1, W, b0, b1, C.byref(o0), C.byref(o1)), "jackson_2d")
    return mu.reshape(-1)[: o0.value * o1.value].reshape(o0.value, o1.value).copy()


def dos_fftgrid(mu: np.ndarray, n: int, W: float, Ec: float, num_div: int):
    mu = np.ascontiguousarray(mu, dtype=np.complex128)
    E, rho = np.zeros(num_div), np.zeros(num_div)
    _check(lib().lqo_dos_fftgrid(_d(mu), mu.shape[0], n, W, Ec, num_div, _d(E), _d(rho)), "dos_fftgrid")
    return E, rho


def greenwood_fftgrid(mu: np.ndarray, n: int, W: float, Ec: float, num_div: int):
    mu = np.ascontiguousarray(mu, dtype=np.complex128)
    E, out = np.zeros(num_div - 1), np.zeros(num_div - 1)
    _check(lib().lqo_greenwood_fftgrid(_d(mu), mu.shape[0], mu.shape[1], n, W, Ec, num_div, _d(E), _d(out)),
           "greenwood_fftgrid")
    return E, out


def bastin_fftgrid(mu: np.ndarray, n: int, W: float, Ec: float, num_div: int):
    mu = np.ascontiguousarray(mu, dtype=np.complex128)
    E, out = np.zeros(num_div - 1), np.zeros(num_div - 1)
    _check(lib().lqo_bastin_fftgrid(_d(mu), mu.shape[0], mu.shape[1], n, W, Ec, num_div, _d(E), _d(out)),
           "bastin_fftgrid")
    return E, out


def time_correlations(H: Op, OPL, OPR, phi: np.ndarray, M: int, nT: int, dt: float, W: float, Ec: float,
                      mu=None) -> np.ndarray:
    """TimeDependentCorrelations: (M, nT) table, accumulated into mu when given."""
    phi = np.ascontiguousarray(np.atleast_2d(phi), dtype=np.complex128)
    out = np.zeros((M, nT), dtype=np.complex128) if mu is None else np.array(mu, dtype=np.complex128)
    _check(lib().lqo_time_correlations(H.h, None if OPL is None else OPL.h, None if OPR is None else OPR.h, H.n,
                                       _d(phi), phi.shape[0], M, nT, dt, W, Ec, _d(out)), "time_correlations")
    return out


def jackson_td(mu: np.ndarray, W: float, broad: float) -> np.ndarray:
    mu = np.array(mu, dtype=np.complex128, order="C")
    mo = C.c_int(0)
    _check(lib().lqo_jackson_td(_d(mu), mu.shape[0], mu.shape[1], W, broad, C.byref(mo)), "jackson_td")
    return mu[: mo.value].copy()


def timecorr_post(mu: np.ndarray, n: int, W: float, Ec: float, dt: float, EF: float):
    mu = np.ascontiguousarray(mu, dtype=np.complex128)
    t, o = np.zeros(mu.shape[1]), np.zeros(mu.shape[1])
    _check(lib().lqo_timecorr_post(_d(mu), mu.shape[0], mu.shape[1], n, W, Ec, dt, EF, _d(t), _d(o)),
           "timecorr_post")
    return t, o


LINEAR_KINDS = {"spectral": 0, "greenwood": 1, "bastin": 2, "bastinI": 3, "bastinII": 4, "sea": 5, "surf": 6}


def linear_kernel(kind: int, mu: np.ndarray, n: int, W: float, x: np.ndarray) -> np.ndarray:
    """The linear-grid tool's kernel (scaled, before any running sum) at normalised energies x."""
    mu = np.ascontiguousarray(mu, dtype=np.complex128)
    x = np.ascontiguousarray(x, dtype=np.float64)
    M0, M1 = (mu.shape[0], 1) if mu.ndim == 1 else mu.shape
    k = np.zeros(len(x))
    _check(lib().lqo_linear_kernel(kind, _d(mu), M0, M1, n, W, len(x), _d(x), _d(k)), "linear_kernel")
    return k


def linear_grid(kind: int, mu: np.ndarray, n: int, W: float, Ec: float, num_div: int):
    """The whole tool: (E, value) rows as printed."""
    mu = np.ascontiguousarray(mu, dtype=np.complex128)
    M0, M1 = (mu.shape[0], 1) if mu.ndim == 1 else mu.shape
    E, out = np.zeros(num_div), np.zeros(num_div)
    rows = lib().lqo_linear_grid(kind, _d(mu), M0, M1, n, W, Ec, num_div, _d(E), _d(out))
    _check(0 if rows > 0 else -1, "linear_grid")
    return E[:rows], out[:rows]


def supercell_csr(nwann: int, cell, edge, val, dims):
    """wrap_in_supercell + save_hopping_list_as_csr: the stored upper CSR
    (zero based) of the supercell operator; returns (n, rowptr, col, val)."""
    cell = np.ascontiguousarray(np.asarray(cell, dtype=np.int32).reshape(-1, 3))
    edge = np.ascontiguousarray(np.asarray(edge, dtype=np.int32).reshape(-1, 2))
    v = np.ascontiguousarray(np.asarray(val, dtype=np.complex128).reshape(-1))
    nh = v.shape[0]
    if nh == 0:
        cell, edge, v = np.zeros((1, 3), np.int32), np.zeros((1, 2), np.int32), np.zeros(1, np.complex128)
    d = np.ascontiguousarray(np.asarray(dims, dtype=np.int32))
    n, nnz = C.c_int(0), C.c_longlong(0)
    L = lib()
    _check(L.lqo_supercell_csr(nwann, nh, _i(cell), _i(edge), _d(v.view(np.float64)), _i(d), C.byref(n),
                               C.byref(nnz), None, None, None), "supercell_csr")
    rowptr = np.zeros(n.value + 1, np.int32)
    col = np.zeros(max(nnz.value, 1), np.int32)
    out = np.zeros(max(nnz.value, 1), np.complex128)
    _check(L.lqo_supercell_csr(nwann, nh, _i(cell), _i(edge), _d(v.view(np.float64)), _i(d), C.byref(n),
                               C.byref(nnz), _i(rowptr), _i(col), _d(out.view(np.float64))), "supercell_csr")
    return n.value, rowptr, col[: nnz.value], out[: nnz.value]


def overlap_solve(S, b, tol=1e-4, maxit=2000):
    """linalg::orthogonalize: x = S^-1 b by Eigen's BiCGSTAB on the stored CSR
    (io.CSR) as a general matrix.  Returns (x, iterations, error)."""
    b = np.ascontiguousarray(b, dtype=np.complex128)
    x = np.zeros_like(b)
    err = C.c_double(0)
    rp = np.ascontiguousarray(S.rowptr, np.int32)
    cl = np.ascontiguousarray(S.col, np.int32)
    v = np.ascontiguousarray(S.val, np.complex128)
    it = lib().lqo_overlap_solve(S.n, _i(rp), _i(cl), _d(v.view(np.float64)), _d(b.view(np.float64)),
                                 _d(x.view(np.float64)), tol, maxit, C.byref(err))
    _check(it, "overlap_solve")
    return x, it, err.value


def spectral_moments_nonorth(H: "Op", OP, S, phi: np.ndarray, M: int, tol=1e-4, maxit=2000, mu=None):
    """SpectralMoments_nonOrth; S is an io.CSR used as a general matrix."""
    phi = np.ascontiguousarray(np.atleast_2d(phi), dtype=np.complex128)
    R, n = phi.shape
    mu = np.zeros(M, dtype=np.complex128) if mu is None else mu
    rp = np.ascontiguousarray(S.rowptr, np.int32)
    cl = np.ascontiguousarray(S.col, np.int32)
    v = np.ascontiguousarray(S.val, np.complex128)
    _check(lib().lqo_spectral_moments_nonorth(H.h, _h(OP), n, _i(rp), _i(cl), _d(v.view(np.float64)),
                                              _d(phi.view(np.float64)), R, M, tol, maxit,
                                              _d(mu.view(np.float64))), "spectral_moments_nonorth")
    return mu


def csr_read_stream(path: str):
    """The reference's stream CSR reader; returns (n, rowptr, col, val)."""
    L = lib()
    n, nnz = C.c_int(0), C.c_longlong(0)
    _check(L.lqo_csr_read_stream(str(path).encode(), C.byref(n), C.byref(nnz), None, None, None), "csr_read_stream")
    rowptr = np.zeros(n.value + 1, np.int32)
    col = np.zeros(max(nnz.value, 1), np.int32)
    val = np.zeros(max(nnz.value, 1), np.complex128)
    _check(L.lqo_csr_read_stream(str(path).encode(), C.byref(n), C.byref(nnz), _i(rowptr), _i(col),
                                 _d(val.view(np.float64))), "csr_read_stream")
    return n.value, rowptr, col[: nnz.value], val[: nnz.value]


def omp_threads() -> int:
    return lib().lqo_omp_threads()


def omp_cheb_step(H: Op, x: np.ndarray, y: np.ndarray) -> None:
    """y <- 2 H x - y in place; x, y are (k, n) complex, C-contiguous."""
    assert x.shape == y.shape and y.flags.c_contiguous and x.flags.c_contiguous
    _check(lib().lqo_omp_cheb_step(H.h, x.shape[0], _d(x), _d(y)), "omp_cheb_step")


def omp_momtable(L: np.ndarray, Rv: np.ndarray, m0_begin: int, m0_end: int) -> np.ndarray:
    """sub[m0 - m0_begin, m1] = <L[m0]|Rv[m1]> for m0 in [m0_begin, m0_end)."""
    L = np.ascontiguousarray(L, dtype=np.complex128)
    Rv = np.ascontiguousarray(Rv, dtype=np.complex128)
    sub = np.zeros((m0_end - m0_begin, Rv.shape[0]), dtype=np.complex128)
    _check(lib().lqo_omp_momtable(L.shape[1], Rv.shape[0], _d(L), _d(Rv), m0_begin, m0_end, _d(sub)),
           "omp_momtable")
    return sub


def omp_mv(op: Op, x: np.ndarray, y: np.ndarray) -> None:
    """y <- op x in place; x, y are (k, n) complex, C-contiguous."""
    assert x.shape == y.shape and y.flags.c_contiguous and x.flags.c_contiguous
    _check(lib().lqo_omp_mv(op.h, x.shape[0], _d(x), _d(y)), "omp_mv")


def omp_bastin_ffts(L: np.ndarray, Rv: np.ndarray, nump: int) -> np.ndarray:
    """Per-site Bastin DFT stage over L, Rv of shape (M, sites); returns r_data (2*nump)."""
    L = np.ascontiguousarray(L, dtype=np.complex128)
    Rv = np.ascontiguousarray(Rv, dtype=np.complex128)
    M, size = L.shape
    rd = np.zeros(2 * nump, dtype=np.complex128)
    _check(lib().lqo_omp_bastin_ffts(_d(L), _d(Rv), M, nump, size, _d(rd)), "omp_bastin_ffts")
    return rd


def omp_spectral_moments(H: Op, phi: np.ndarray, M: int) -> np.ndarray:
    phi = np.ascontiguousarray(np.atleast_2d(phi), dtype=np.complex128)
    mu = np.zeros(M, dtype=np.complex128)
    _check(lib().lqo_omp_spectral_moments(H.h, H.n, _d(phi), phi.shape[0], M, _d(mu)), "omp_spectral_moments")
    return mu
