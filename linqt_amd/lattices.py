"""Seeded tight-binding generators for the benchmark lattices.

The reference builds operators with ``utilities/wannier2sparse`` from
Wannier90 files and writes them as Hermitian-upper CSR
(``hopping_list.cpp:227-285``).  These generators produce the same format
(entries with col >= row, rows sorted, duplicates kept) for the synthetic
systems of the benchmarks, deterministically from a seed:

* square lattice L x L, nearest-neighbour hopping -t, on-site Anderson
  disorder U[-W/2, W/2];
* honeycomb L x L cells (2 sites), hopping -t, Haldane second-neighbour
  t2 e^{+-i phi}, disorder U[-W/2, W/2], plus the velocity operators
  V_a = i[H, X_a] (minimum-image bond vectors);
* cubic lattice L^3, hopping -t, disorder U[-W/2, W/2];
* Kane-Mele-like honeycomb, 4 orbitals per cell (2 sites x 2 spins):
  hopping -t, intrinsic spin-orbit i lso nu_ij s_z on second neighbours,
  Rashba i lr (s x d)_z on first neighbours, disorder U[-W/2, W/2].

Random numbers come from a counter-based splitmix64 (stream, index) ->
uint64 -> U[0, 1) with 53 bits, so every array element is independent of
how the arrays are chunked.
"""
from __future__ import annotations

import numpy as np

from .io import CSR

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_STREAM = np.uint64(0xD1B54A32D192ED03)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, count: int, start: int = 0) -> np.ndarray:
    """U[0,1) doubles: element i of (seed, stream) is splitmix64 of a counter."""
    with np.errstate(over="ignore"):
        base = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) * _GOLD + np.uint64(stream) * _STREAM
        idx = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = _mix(base + idx * _GOLD)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def random_phase_states(n: int, R: int, seed: int) -> np.ndarray:
    """R random-phase vectors e^{2 pi i u}/sqrt(n) (as ``FillWithRandomPhase``,
    ``include/quantum_states.hpp:231-253``, with a counter-based RNG)."""
    u = uniform(seed, 7, n * R).reshape(R, n)
    return np.exp(2j * np.pi * u) / np.sqrt(n)


def _to_upper_csr(n: int, rows, cols, vals) -> CSR:
    """Keeps entries with col >= row, sorts by (row, col) stably."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.complex128)
    keep = cols >= rows
    rows, cols, vals = rows[keep], cols[keep], vals[keep]
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr)
    return CSR(n, rowptr.astype(np.int32), cols.astype(np.int32), vals)


def square_anderson(L: int, t: float = 1.0, W: float = 1.0, seed: int = 20261202) -> CSR:
    """Square lattice, periodic, site i = x + L*y."""
    if L < 3:
        raise ValueError("L >= 3 keeps periodic neighbours distinct")
    n = L * L
    idx = np.arange(n)
    x, y = idx % L, idx // L
    eps = W * (uniform(seed, 1, n) - 0.5)
    rows = [idx]
    cols = [idx]
    vals = [eps.astype(np.complex128)]
    for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
        j = (x + dx) % L + L * ((y + dy) % L)
        rows.append(idx)
        cols.append(j)
        vals.append(np.full(n, -t, dtype=np.complex128))
    return _to_upper_csr(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def cubic_anderson(L: int, t: float = 1.0, W: float = 1.0, seed: int = 20261205) -> CSR:
    """Cubic lattice L^3, periodic, site i = x + L*(y + L*z)."""
    if L < 3:
        raise ValueError("L >= 3 keeps periodic neighbours distinct")
    n = L ** 3
    idx = np.arange(n)
    x, y, z = idx % L, (idx // L) % L, idx // (L * L)
    eps = W * (uniform(seed, 1, n) - 0.5)
    rows, cols, vals = [idx], [idx], [eps.astype(np.complex128)]
    for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        j = (x + d[0]) % L + L * (((y + d[1]) % L) + L * ((z + d[2]) % L))
        rows.append(idx)
        cols.append(j)
        vals.append(np.full(n, -t, dtype=np.complex128))
    return _to_upper_csr(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def haldane_honeycomb(L: int, t: float = 1.0, t2: float = 0.1, phi: float = np.pi / 2, W: float = 2.0,
                      seed: int = 20261203):
    """Honeycomb L x L cells, periodic; returns (H, Vx, Vy) as upper CSR.

    Bravais vectors a1 = (1, 0), a2 = (1/2, sqrt(3)/2); site A of cell (m, k)
    at m a1 + k a2, site B at that + (a1 + a2)/3; index 2*(m + L*k) + s.
    Second-neighbour entries H(i, i + b) along b1 = a1, b2 = a2 - a1,
    b3 = -a2 are t2 e^{+i phi} on A and t2 e^{-i phi} on B (the reverse
    entries their conjugates).
    V_a(i, j) = i (r_j - r_i)_a H(i, j) with the bond vector r_j - r_i.
    """
    if L < 3:
        raise ValueError("L >= 3 keeps periodic neighbours distinct")
    a1 = np.array([1.0, 0.0])
    a2 = np.array([0.5, np.sqrt(3.0) / 2.0])
    dAB = (a1 + a2) / 3.0
    ncell = L * L
    n = 2 * ncell
    c = np.arange(ncell)
    m, k = c % L, c // L

    def site(mm, kk, s):
        return 2 * ((mm % L) + L * (kk % L)) + s

    rows, cols, vals, disp = [], [], [], []

    def add(i, j, v, d):
        rows.append(i)
        cols.append(j)
        vals.append(np.broadcast_to(np.asarray(v, dtype=np.complex128), i.shape).copy())
        disp.append(np.broadcast_to(d, (i.shape[0], 2)).copy())

    eps = W * (uniform(seed, 1, n) - 0.5)
    iA, iB = site(m, k, 0), site(m, k, 1)
    add(iA, iA, eps[iA], np.zeros(2))
    add(iB, iB, eps[iB], np.zeros(2))
    # nearest neighbours A(m,k) -> B(m,k), B(m-1,k), B(m,k-1), and reverse
    for (dm, dk) in ((0, 0), (-1, 0), (0, -1)):
        jB = site(m + dm, k + dk, 1)
        d = dAB + dm * a1 + dk * a2
        add(iA, jB, -t, d)
        add(jB, iA, -t, -d)
    # second neighbours
    for b, (dm, dk) in ((a1, (1, 0)), (a2 - a1, (-1, 1)), (-a2, (0, -1))):
        for s, ph in ((0, phi), (1, -phi)):
            i = site(m, k, s)
            j = site(m + dm, k + dk, s)
            v = t2 * np.exp(1j * ph)
            # j sits at r_i + b: H(i, j) with bond vector r_j - r_i = b
            add(i, j, v, b)
            add(j, i, np.conj(v), -b)
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    vals = np.concatenate(vals)
    disp = np.concatenate(disp)
    H = _to_upper_csr(n, rows, cols, vals)
    offd = rows != cols
    Vx = _to_upper_csr(n, rows[offd], cols[offd], 1j * disp[offd, 0] * vals[offd])
    Vy = _to_upper_csr(n, rows[offd], cols[offd], 1j * disp[offd, 1] * vals[offd])
    return H, _drop_zeros(Vx), _drop_zeros(Vy)


def kane_mele(L: int, t: float = 1.0, lso: float = 0.06, lr: float = 0.05, W: float = 1.0,
              seed: int = 20261206) -> CSR:
    """Honeycomb L x L cells, 2 sites x 2 spins (n = 4 L^2), periodic; upper CSR.

    Geometry as haldane_honeycomb; index 4*(m + L*k) + 2*s + sigma.  The
    second-neighbour entry H(i, i + b) is i lso nu sigma_z with nu = +1 on A
    and -1 on B (the Haldane pattern at phi = pi/2, spin-resolved); the
    Rashba block of a first-neighbour bond with unit vector d is
    i lr (sigma_x d_y - sigma_y d_x).
    """
    if L < 3:
        raise ValueError("L >= 3 keeps periodic neighbours distinct")
    a1 = np.array([1.0, 0.0])
    a2 = np.array([0.5, np.sqrt(3.0) / 2.0])
    dAB = (a1 + a2) / 3.0
    ncell = L * L
    n = 4 * ncell
    c = np.arange(ncell)
    m, k = c % L, c // L

    def site(mm, kk, s, sg):
        return 4 * ((mm % L) + L * (kk % L)) + 2 * s + sg

    rows, cols, vals = [], [], []

    def add(i, j, v):
        rows.append(i)
        cols.append(j)
        vals.append(np.broadcast_to(np.asarray(v, dtype=np.complex128), i.shape).copy())

    eps = W * (uniform(seed, 1, 2 * ncell) - 0.5)  # one value per site, both spins
    for s in (0, 1):
        for sg in (0, 1):
            i = site(m, k, s, sg)
            add(i, i, eps[2 * c + s])
    for (dm, dk) in ((0, 0), (-1, 0), (0, -1)):
        d = dAB + dm * a1 + dk * a2
        dx, dy = d / np.linalg.norm(d)
        # spin block of the bond A(m,k) -> B: i lr (sx dy - sy dx)
        blk = 1j * lr * np.array([[0.0, dy + 1j * dx], [dy - 1j * dx, 0.0]])
        for sa in (0, 1):
            for sb in (0, 1):
                iA = site(m, k, 0, sa)
                jB = site(m + dm, k + dk, 1, sb)
                v = (-t if sa == sb else 0.0) + blk[sa, sb]
                if v == 0:
                    continue
                add(iA, jB, v)
                add(jB, iA, np.conj(v))
    for (dm, dk) in ((1, 0), (-1, 1), (0, -1)):
        for s, nu in ((0, 1.0), (1, -1.0)):
            for sg, sz in ((0, 1.0), (1, -1.0)):
                i = site(m, k, s, sg)
                j = site(m + dm, k + dk, s, sg)
                v = 1j * lso * nu * sz
                add(i, j, v)
                add(j, i, np.conj(v))
    return _to_upper_csr(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def _drop_zeros(m: CSR) -> CSR:
    keep = np.abs(m.val) > 0
    if keep.all():
        return m
    rows = np.repeat(np.arange(m.n), np.diff(m.rowptr))
    return _to_upper_csr(m.n, rows[keep], m.col[keep], m.val[keep])
