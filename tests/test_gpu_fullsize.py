"""Parity and invariants at the benchmark sizes (one GPU).

* DOS config: square 256x256 (N = 65536), M = 1024, R = 8.
* Kubo-Bastin config: Haldane honeycomb 512x512 cells (N = 524288), V_x, V_y.
The oracle runs where it finishes in seconds (reduced M at full N); the full
M runs are checked through size-independent properties.
"""
import sys

import numpy as np
import pytest

from conftest import rel_inf
import linqt_amd as lq
from linqt_amd import lattices

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def square256(ctx, oracle):
    m = lattices.square_anderson(256, W=1.0, seed=20261202)
    A = lq.SparseMatrix.from_csr(ctx, m).Rescale(1 / 4.6, 0.0)
    O = oracle.Op.from_csr(m).rescale(1 / 4.6, 0.0)
    phi = lattices.random_phase_states(m.n, 8, seed=20261202)
    return m, A, O, phi


def test_dos_fullsize_vs_oracle_and_invariants(ctx, oracle, square256):
    m, A, O, phi = square256
    say("dos full N, M=64 vs oracle")
    ref = oracle.omp_spectral_moments(O, phi, 64)
    got = lq.spectral_moments(A, None, phi, 64)
    assert rel_inf(got, ref) < 1e-10
    say("dos full N, M=1024 direct vs doubling")
    d = lq.spectral_moments(A, None, phi, 1024)
    db = lq.spectral_moments(A, None, phi, 1024, algo=lq.LQ_ALGO_DOUBLING)
    assert rel_inf(db, d) < 1e-10
    assert abs(d[0] - 1.0) < 1e-12  # normalised states
    assert np.abs(d).max() <= 2.0 + 1e-9  # |<phi|T_m|phi>| <= 1, times the factor 2 for m > 0
    assert rel_inf(d[:64], ref) < 1e-10
    # the DOS integrates to N
    mu = lq.jackson_1d(ctx, d, 9.2, 1e-3)
    E, rho = lq.dos_fftgrid(ctx, mu, m.n, 9.2, 0.0, 2048)
    x = E / 4.6
    integral = (np.pi / 2048) * np.sum(np.sqrt(1 - x * x) * rho) * 4.6
    assert abs(integral / m.n - 1.0) < 1e-9


@pytest.fixture(scope="module")
def honeycomb512(ctx, oracle):
    H, Vx, Vy = lattices.haldane_honeycomb(512, seed=20261203)
    G = [lq.SparseMatrix.from_csr(ctx, H).Rescale(1 / 4.6, 0.0), lq.SparseMatrix.from_csr(ctx, Vx),
         lq.SparseMatrix.from_csr(ctx, Vy)]
    O = [oracle.Op.from_csr(H).rescale(1 / 4.6, 0.0), oracle.Op.from_csr(Vx), oracle.Op.from_csr(Vy)]
    phi = lattices.random_phase_states(H.n, 4, seed=20261203)
    return H.n, G, O, phi


def test_bastin_fullsize_vs_oracle(ctx, oracle, honeycomb512):
    n, G, O, phi = honeycomb512
    M = 16
    say("corr full N, M=16, R=1 vs oracle")
    got = lq.correlation_expansion_moments(G[0], G[1], [G[1], G[2]], phi[:1], M, M)
    for o, opr in enumerate((O[1], O[2])):
        ref = oracle.correlation_moments(O[0], O[1], opr, phi[:1], M, M)
        assert rel_inf(got[o], ref) < 1e-10, o


def test_bastin_fullsize_properties(ctx, honeycomb512):
    n, G, O, phi = honeycomb512
    M = 128
    say("kubo full N, M=128, R=4")
    fd = lq.kubo_fft(G[0], G[1], [G[1], G[2]], phi, M, 2 * M, lq.LQ_KUBO_BASTIN)
    # the state mean is linear: R=4 at once equals the mean of single-state runs
    fd1 = np.mean([lq.kubo_fft(G[0], G[1], [G[1], G[2]], phi[r:r + 1], M, 2 * M, lq.LQ_KUBO_BASTIN)
                   for r in range(4)], axis=0)
    assert rel_inf(fd, fd1) < 1e-10
    assert np.all(np.isfinite(fd))
