"""GPU parity of the non-orthogonal path (lq_overlap_solve,
lq_spectral_moments_nonorth) against the oracle's restatement of
SpectralMoments_nonOrth with Eigen's BiCGSTAB.  With a tight solver tolerance
both sides converge to the exact generalised recursion, so the moments agree
to 1e-10; at the reference's tolerance (1e-4) each side stops when its own
residual is small enough, so they agree to the tolerance the solver
guarantees.  Also at the benchmark size (N = 524288)."""
import sys

import numpy as np
import pytest

import linqt_amd as lq
from conftest import dense_herm, random_upper_csr
from linqt_amd import io, lattices

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _overlap_csr(n, eps, seed):
    """Stored triangle of a Hermitian positive overlap I + eps X (sparse X)."""
    rp, cl, v = random_upper_csr(n, 4.0 / n, seed, diag=True)
    rows = np.repeat(np.arange(n), np.diff(rp))
    v = eps * v
    v[cl == rows] += 1.0
    return io.CSR(n, rp, cl, v)


def test_overlap_solve_matches_oracle(ctx, oracle):
    say("overlap solve")
    n = 500
    S = _overlap_csr(n, 0.1, 3)
    Sg = lq.SparseMatrix(ctx, n, S.rowptr, S.col, S.val, fill=lq.LQ_FILL_GENERAL)
    b = np.random.default_rng(1).normal(size=n) + 1j * np.random.default_rng(2).normal(size=n)
    x, it, err = lq.overlap_solve(Sg, b, tol=1e-13)
    xo, ito, erro = oracle.overlap_solve(S, b, tol=1e-13)
    assert err <= 1e-13 and abs(it - ito) <= 1
    assert np.abs(x - xo).max() <= 1e-11 * np.abs(xo).max()
    U = np.zeros((n, n), complex)
    U[np.repeat(np.arange(n), np.diff(S.rowptr)), S.col] = S.val
    assert np.linalg.norm(U @ x - b) <= 1e-12 * np.linalg.norm(b)
    z, it0, _ = lq.overlap_solve(Sg, np.zeros(n, complex))
    assert it0 == 0 and not z.any()


@pytest.mark.parametrize("with_op", [False, True])
def test_nonorth_moments_match_oracle(ctx, oracle, with_op):
    say("nonorth moments", with_op)
    m = lattices.square_anderson(20, W=1.0, seed=7)
    n = m.n
    H = lq.SparseMatrix.from_csr(ctx, m).Rescale(1 / 4.6, 0.0)
    Ho = oracle.Op.from_csr(m).rescale(1 / 4.6, 0.0)
    S = _overlap_csr(n, 0.05, 5)
    Sg = lq.SparseMatrix(ctx, n, S.rowptr, S.col, S.val, fill=lq.LQ_FILL_GENERAL)
    OP = OPo = None
    if with_op:
        rp, cl, v = random_upper_csr(n, 3.0 / n, 8, diag=True)
        OP = lq.SparseMatrix(ctx, n, rp, cl, v)
        OPo = oracle.Op(n, rp, cl, v)
    phi = lattices.random_phase_states(n, 2, seed=3)
    info = {}
    mu = lq.spectral_moments_nonorth(H, OP, Sg, phi, 64, tol=1e-13, info=info)
    ref = oracle.spectral_moments_nonorth(Ho, OPo, S, phi, 64, tol=1e-13)
    assert info["iterations"] > 0
    assert np.abs(mu - ref).max() <= 1e-10 * np.abs(ref).max()
    # the reference's tolerance: both within the solver's accuracy of the exact moments
    mu4 = lq.spectral_moments_nonorth(H, OP, Sg, phi, 64)
    assert np.abs(mu4 - ref).max() <= 1e-3 * np.abs(ref).max()


def test_nonorth_identity_overlap_equals_orthogonal_solver(ctx):
    say("nonorth identity")
    m = lattices.square_anderson(32, W=1.0, seed=2)
    n = m.n
    H = lq.SparseMatrix.from_csr(ctx, m).Rescale(1 / 4.6, 0.0)
    I = lq.SparseMatrix(ctx, n, np.arange(n + 1), np.arange(n), np.ones(n), fill=lq.LQ_FILL_GENERAL)
    phi = lattices.random_phase_states(n, 2, seed=4)
    mu = lq.spectral_moments_nonorth(H, None, I, phi, 100)
    ref = lq.spectral_moments(H, None, phi, 100)
    assert np.abs(mu - ref).max() <= 1e-12 * np.abs(ref).max()


def test_nonorth_fullsize(ctx, oracle):
    """N = 524288 honeycomb with a near-identity overlap: GPU moments against
    the oracle at a tight tolerance (a few moments: each oracle solve is a
    sequential BiCGSTAB over half a million rows)."""
    say("nonorth fullsize")
    from bench import haldane_hoppings

    cell, edge, val = haldane_hoppings()
    H = lq.SparseMatrix.from_hoppings(ctx, 2, cell, edge, val, (512, 512, 1))
    Hc = H.export()
    H.Rescale(1 / 3.2, 0.0)
    Ho = oracle.Op.from_csr(Hc).rescale(1 / 3.2, 0.0)
    n = Hc.n
    # overlap: 1 on the diagonal, 0.02 on the nearest-neighbour bonds of H (a tight-binding overlap)
    rows = np.repeat(np.arange(n), np.diff(Hc.rowptr))
    v = np.where(Hc.col == rows, 1.0 + 0j, np.where(np.abs(Hc.val) > 0.5, 0.02 + 0j, 0j))
    S = io.CSR(n, Hc.rowptr, Hc.col, v)
    Sg = lq.SparseMatrix(ctx, n, S.rowptr, S.col, S.val, fill=lq.LQ_FILL_GENERAL)
    phi = lattices.random_phase_states(n, 1, seed=6)
    mu = lq.spectral_moments_nonorth(H, None, Sg, phi, 6, tol=1e-12)
    ref = oracle.spectral_moments_nonorth(Ho, None, S, phi, 6, tol=1e-12)
    assert np.abs(mu - ref).max() <= 1e-9 * np.abs(ref).max()
