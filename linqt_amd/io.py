"""Readers and writers for LinQT's text file formats.

Formats (all whitespace separated):

* CSR operator (``src/sparse_matrix.cpp:3-50``, writer
  ``utilities/wannier2sparse/src/hopping_list.cpp:227-285``)::

      dim nnz
      re_0 im_0 re_1 im_1 ...        (nnz complex values)
      col_0 col_1 ...                (nnz zero-based column indices)
      rowptr_0 ... rowptr_dim        (dim+1 row pointers)

  Stored is the upper triangle (diagonal included); the operator it stands
  for is the Hermitian completion (``include/sparse_matrix.hpp:50-55``).
* ``.chebmom1D`` (``src/chebyshev_moments1D.cpp:19-51, 91-103``)::

      N W Ec
      M
      re im            (M lines)

* ``.chebmom2D`` (``src/chebyshev_moments2D.cpp:16-68``)::

      N W  Ec
      M0 M1
      re im            (M0*M1 lines, m0 outer)

* BOUNDS (``src/chebyshev_solver.cpp:8-23``): ``lowest highest``; a missing
  file means (-100, 100).
* state file (``include/quantum_states.hpp:155-228``): ``vector`` R N then
  R*N ``re im`` pairs, or ``local`` K then K site indices.

Numbers are written with 15 significant digits, as the reference's
``precision(dbl::digits10)`` streams do.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np


@dataclass
class CSR:
    """Zero-based CSR arrays of one operator file."""

    n: int
    rowptr: np.ndarray  # int32 [n+1]
    col: np.ndarray  # int32 [nnz]
    val: np.ndarray  # complex128 [nnz]

    @property
    def nnz(self) -> int:
        return int(self.col.shape[0])


def _tokens(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        text = f.read()
    # every entry of these files is a decimal number
    return np.array(text.split(), dtype=np.float64)


def read_csr(path: str) -> CSR:
    """``Sparse::OPERATOR_FromCSRFile`` (``src/sparse_matrix.cpp:3-50``)."""
    if not os.path.isfile(path):
        raise FileNotFoundError(f"ERROR: Can not open the matrix file {path}")
    t = _tokens(path)
    dim, nnz = int(t[0]), int(t[1])
    need = 2 + 2 * nnz + nnz + dim + 1
    if t.shape[0] < need:
        raise ValueError(f"{path}: truncated CSR file ({t.shape[0]} of {need} numbers)")
    v = t[2 : 2 + 2 * nnz].reshape(nnz, 2)
    val = v[:, 0] + 1j * v[:, 1]
    col = t[2 + 2 * nnz : 2 + 3 * nnz].astype(np.int32)
    rowptr = t[2 + 3 * nnz : 2 + 3 * nnz + dim + 1].astype(np.int32)
    return CSR(dim, rowptr, col, val.astype(np.complex128))


def _fmt(x: float) -> str:
    return format(float(x), ".15g")


def write_csr(path: str, m: CSR) -> None:
    """Writes the format of ``hopping_list::save_hopping_list_as_csr``."""
    with open(path, "w") as f:
        f.write(f"{m.n} {m.nnz}\n")
        vals = np.empty(2 * m.nnz)
        vals[0::2] = m.val.real
        vals[1::2] = m.val.imag
        f.write(" ".join(format(x, ".15g") for x in vals) + " \n")
        f.write(" ".join(str(int(c)) for c in m.col) + " \n")
        f.write(" ".join(str(int(r)) for r in m.rowptr) + " \n")


def read_chebmom1d(path: str):
    """``Moments1D(std::string)``; returns (mu, N, W, Ec)."""
    if ".chebmom1D" not in path:
        raise ValueError("The first argument does not seem to be a valid .chebmom1D file")
    t = _tokens(path)
    n, W, Ec, M = int(t[0]), float(t[1]), float(t[2]), int(t[3])
    v = t[4 : 4 + 2 * M].reshape(M, 2)
    return v[:, 0] + 1j * v[:, 1], n, W, Ec


def read_chebmomtd(path: str):
    """``MomentsTD(std::string)`` (src/chebyshev_momentsTD.cpp:48-79); returns
    (mu (M, nT), N, W, Ec, dt)."""
    if ".chebmomTD" not in path:
        raise ValueError("The first argument does not seem to be a valid .chebmomTD file")
    t = _tokens(path)
    n, W, Ec, _, dt = int(t[0]), float(t[1]), float(t[2]), int(t[3]), float(t[4])
    M, nT = int(t[5]), int(t[6])
    v = t[7 : 7 + 2 * M * nT].reshape(M * nT, 2)
    return (v[:, 0] + 1j * v[:, 1]).reshape(M, nT), n, W, Ec, dt


def write_chebmomtd(path: str, mu: np.ndarray, n: int, W: float, Ec: float, dt: float) -> None:
    """``MomentsTD::saveIn`` (src/chebyshev_momentsTD.cpp:81-93)."""
    M, nT = mu.shape
    with open(path, "w") as f:
        f.write(f"{n} {_fmt(W)} {_fmt(Ec)} {nT} {_fmt(dt)} \n{M} {nT} \n")
        for z in mu.reshape(-1):
            f.write(f"{_fmt(z.real)} {_fmt(z.imag)}\n")


def write_chebmom1d(path: str, mu: np.ndarray, n: int, W: float, Ec: float) -> None:
    """``Moments1D::saveIn`` (``src/chebyshev_moments1D.cpp:91-103``)."""
    with open(path, "w") as f:
        f.write(f"{n} {_fmt(W)} {_fmt(Ec)}\n{len(mu)}\n")
        for z in mu:
            f.write(f"{_fmt(z.real)} {_fmt(z.imag)}\n")


def read_chebmom2d(path: str):
    """``Moments2D(std::string)``; returns (mu[M0,M1], N, W, Ec)."""
    if ".chebmom2D" not in path:
        raise ValueError("The first argument does not seem to be a valid .chebmom2D file")
    t = _tokens(path)
    n, W, Ec, M0, M1 = int(t[0]), float(t[1]), float(t[2]), int(t[3]), int(t[4])
    v = t[5 : 5 + 2 * M0 * M1].reshape(M0, M1, 2)
    return v[..., 0] + 1j * v[..., 1], n, W, Ec


def write_chebmom2d(path: str, mu: np.ndarray, n: int, W: float, Ec: float) -> None:
    """``Moments2D::saveIn`` (``src/chebyshev_moments2D.cpp:53-68``)."""
    M0, M1 = mu.shape
    with open(path, "w") as f:
        f.write(f"{n} {_fmt(W)}  {_fmt(Ec)}\n{M0} {M1} \n")
        for z in mu.reshape(-1):
            f.write(f"{_fmt(z.real)} {_fmt(z.imag)}\n")


def read_bounds(path: str = "BOUNDS"):
    """``utility::SpectralBounds`` (``src/chebyshev_solver.cpp:8-23``)."""
    if os.path.isfile(path):
        t = _tokens(path)
        return float(t[0]), float(t[1])
    return -100.0, 100.0


def read_state_file(path: str, n: int) -> np.ndarray:
    """``qstates::LoadStateFile`` (``include/quantum_states.hpp:213-228``).

    Returns the states as an (R, n) complex array.  A missing file, or a kind
    other than ``vector``/``local``, gives the reference's default: one state
    localised at site n/2 (``quantum_states.hpp:85-93``).
    """
    kind = "random_phase"
    toks = []
    if os.path.isfile(path):
        with open(path) as f:
            toks = f.read().split()
        if toks:
            kind = toks[0]
    if kind == "vector":
        R, dim = int(toks[1]), int(toks[2])
        if dim != n:
            raise ValueError(f"state file dimension {dim} != system size {n}")
        v = np.array(toks[3 : 3 + 2 * R * dim], dtype=np.float64).reshape(R, dim, 2)
        return v[..., 0] + 1j * v[..., 1]
    if kind == "local":
        K = int(toks[1])
        sites = [int(x) for x in toks[2 : 2 + K]]
        out = np.zeros((K, n), dtype=np.complex128)
        for r, s in enumerate(sites):
            out[r, s] = 1.0
        return out
    return default_state(n)


def default_state(n: int) -> np.ndarray:
    """The reference's default "random" state: a delta at site n/2, R = 1."""
    out = np.zeros((1, n), dtype=np.complex128)
    out[0, n // 2] = 1.0
    return out


def write_vector_state_file(path: str, phi: np.ndarray) -> None:
    """Writes a ``vector`` state file (R states of length n)."""
    R, n = phi.shape
    with open(path, "w") as f:
        f.write(f"vector\n{R}\n{n}\n")
        for z in phi.reshape(-1):
            f.write(f"{_fmt(z.real)} {_fmt(z.imag)}\n")
