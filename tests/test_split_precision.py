"""Split-precision (INT8 digit) contraction.

CPU: the numpy model of the method (tests/ozaki_model.py) against exact
complex128 products -- the error stays within the a-priori bound
(S + 1) 2^-7S / 2 sum_g 2^(eA+eB) |g|, falls by ~2^7 per extra digit, and
does not grow coherently for correlated rows (round-to-nearest digits).
GPU: lq_compute_mom_table_prec against the model (same integer sums, so the
difference is only FP64 rounding of the fold: 1e-13 of the magnitude), and
the full Kubo pipeline against the FP64 path with the error it reaches.
"""
import sys

import numpy as np
import pytest

import linqt_amd as lq
import ozaki_model as oz
from conftest import rel_inf
from linqt_amd import lattices

SLICES = [4, 5, 6, 7, 8]


def _operands(M0, M1, K, seed, spread=True):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(M0, K)) + 1j * rng.normal(size=(M0, K))
    B = rng.normal(size=(M1, K)) + 1j * rng.normal(size=(M1, K))
    if spread and K > oz.GROUP:
        # group-wise magnitudes 2^-30 .. 2^+20 and an all-zero group
        for g in range(0, K, oz.GROUP):
            A[:, g:g + oz.GROUP] *= 2.0 ** rng.integers(-30, 20)
            B[:, g:g + oz.GROUP] *= 2.0 ** rng.integers(-30, 20)
        A[:, oz.GROUP:2 * oz.GROUP] = 0
    return A, B


@pytest.mark.parametrize("S", SLICES)
def test_model_within_bound(S):
    A, B = _operands(6, 5, 3 * oz.GROUP + 100, seed=S)
    exact = A.conj() @ B.T
    got = oz.contract(A, B, S)
    # at 8 digits the bound falls below the FP64 rounding of the exact product
    # and of the model's fold: allow that rounding (a few ulps of the magnitude)
    bound = oz.error_bound(A, B, S) + 4e-15 * oz.magnitude(A, B)
    err_r, err_i = np.abs((got - exact).real), np.abs((got - exact).imag)
    assert np.all(err_r <= bound) and np.all(err_i <= bound)


@pytest.mark.parametrize("S", [4, 5, 6])
def test_model_bound_is_tight_for_coherent_errors(S):
    # constant operands: every truncation remainder has the same sign, so the
    # error grows linearly in K and reaches a sizeable fraction of the bound
    c = 0.7312345678901234 + 0.9876543210987654j
    A = np.full((2, oz.GROUP), c)
    B = np.full((2, oz.GROUP), np.conj(c) * 0.93)
    exact = A.conj() @ B.T
    d = oz.contract(A, B, S) - exact
    ratio = np.maximum(np.abs(d.real), np.abs(d.imag)) / oz.error_bound(A, B, S)
    assert 0.01 < ratio.max() <= 1.0


def test_model_correlated_rows_do_not_accumulate_bias():
    # L ~ R (as for sigma_xx): the error of the diagonal stays a random walk,
    # far below the coherent bound that truncated digits would approach
    rng = np.random.default_rng(1)
    K = 8 * oz.GROUP
    A = rng.normal(size=(4, K)) + 1j * rng.normal(size=(4, K))
    B = A + 0.3 * (rng.normal(size=(4, K)) + 1j * rng.normal(size=(4, K)))
    d = np.diag(oz.contract(A, B, 5) - A.conj() @ B.T)
    assert np.abs(d).max() < 0.02 * np.diag(oz.error_bound(A, B, 5)).min()  # coherent sums reach ~0.3


def test_model_error_falls_per_digit():
    A, B = _operands(4, 4, oz.GROUP, seed=11, spread=False)
    exact = A.conj() @ B.T
    errs = [np.abs(oz.contract(A, B, S) - exact).max() for S in (3, 4, 5)]
    assert errs[1] < errs[0] / 20 and errs[2] < errs[1] / 20


def test_model_exact_for_short_digits():
    # values with at most 7S significant bits below the group exponent are exact
    rng = np.random.default_rng(5)
    A = (rng.integers(-2**13, 2**13, size=(3, 300)) + 1j * rng.integers(-2**13, 2**13, size=(3, 300))) / 2.0**14
    B = (rng.integers(-2**13, 2**13, size=(2, 300)) + 1j * rng.integers(-2**13, 2**13, size=(2, 300))) / 2.0**14
    assert np.array_equal(oz.contract(A, B, 4), A.conj() @ B.T)


# ---- GPU -------------------------------------------------------------------------------


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _table_gpu(ctx, A, B, precision):
    from linqt_amd import _lib

    ML, K = A.shape
    MR = B.shape[0]
    A = np.ascontiguousarray(A)
    B = np.ascontiguousarray(B)
    pl, pr, ps = ctx.alloc(A.nbytes), ctx.alloc(B.nbytes), ctx.alloc(ML * MR * 16)
    try:
        ctx.upload(pl, A)
        ctx.upload(pr, B)
        _lib.call("lq_compute_mom_table_prec", ctx.h, pl, K, ML, pr, K, MR, K, precision, ps, MR)
        sub = np.zeros((ML, MR), dtype=np.complex128)
        ctx.download(ps, sub)
    finally:
        for p in (pl, pr, ps):
            ctx.free(p)
    return sub


@pytest.mark.gpu
@pytest.mark.parametrize("S", SLICES)
@pytest.mark.parametrize("ML,MR,K", [(70, 50, 3 * 4096 + 100), (130, 129, 300), (1, 3, 17), (33, 200, 2 * 4096)])
def test_mom_table_split_vs_model(ctx, S, ML, MR, K):
    say("split table", S, ML, MR, K)
    A, B = _operands(ML, MR, K, seed=ML * 7 + MR + S)
    got = _table_gpu(ctx, A, B, S)
    model = oz.contract(A, B, S)
    mag = oz.magnitude(A, B)
    d = got - model
    assert np.all(np.abs(d.real) <= 1e-13 * mag) and np.all(np.abs(d.imag) <= 1e-13 * mag)
    bound = oz.error_bound(A, B, S)
    exact = A.conj() @ B.T
    assert np.all(np.abs((got - exact).real) <= bound * (1 + 1e-9) + 1e-13 * mag)


@pytest.mark.gpu
def test_mom_table_fp64_precision_matches_plain(ctx):
    A, B = _operands(40, 30, 5000, seed=3, spread=False)
    got = _table_gpu(ctx, A, B, 0)
    assert rel_inf(got, A.conj() @ B.T) < 1e-12
    for bad in (1, 9, -1):
        with pytest.raises(lq.LinqtError):
            _table_gpu(ctx, A, B, bad)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [lq.LQ_PREC_FP32X2, lq.LQ_PREC_BF16X3])
@pytest.mark.parametrize("ML,MR,K", [(70, 50, 3000), (130, 129, 300), (1, 3, 17), (65, 200, 8192)])
def test_mom_table_float_splits(ctx, prec, ML, MR, K):
    """FP32 hi/lo and 3 x BF16: every entry within the FP32-accumulation bound
    (256-term FP32 sums, gamma_256 ~ 1.5e-5 of the summed magnitudes of the
    three real products) of the exact table, and visibly not FP64-exact."""
    say("float split table", prec, ML, MR, K)
    A, B = _operands(ML, MR, K, seed=ML * 5 + MR + prec, spread=False)
    got = _table_gpu(ctx, A, B, prec)
    exact = A.conj() @ B.T
    mag3 = (np.abs(A.real) + np.abs(A.imag)) @ (np.abs(B.real) + np.abs(B.imag)).T
    err = np.abs(got - exact)
    assert np.all(err <= 3e-5 * mag3 + 1e-300), float((err / mag3).max())
    rel = rel_inf(got, exact)
    print(f"precision {prec} {ML}x{MR}x{K}: rel err {rel:.2e}", flush=True)
    assert rel < 1e-5
    if K >= 300:
        assert rel > 1e-13  # the FP32 accumulation shows


@pytest.mark.gpu
def test_kubo_split_precision_vs_fp64(ctx):
    H, Vx, Vy = lattices.haldane_honeycomb(16, seed=4)
    G = [lq.SparseMatrix.from_csr(ctx, H).Rescale(1 / 4.6, 0.0), lq.SparseMatrix.from_csr(ctx, Vx),
         lq.SparseMatrix.from_csr(ctx, Vy)]
    phi = lattices.random_phase_states(H.n, 2, seed=9)
    M, nump = 96, 256
    ref = lq.kubo_fft(G[0], G[1], [G[1], G[2]], phi, M, nump, lq.LQ_KUBO_BASTIN)
    sig_ref = [lq.kubo_post(ctx, 1, ref[o], H.n, 9.2, 0.0)[1] for o in range(2)]
    errs = {}
    for S in (4, 5, 6, 7, 8):
        say("kubo split", S)
        got = lq.kubo_fft(G[0], G[1], [G[1], G[2]], phi, M, nump, lq.LQ_KUBO_BASTIN, precision=S)
        for o in range(2):
            sig = lq.kubo_post(ctx, 1, got[o], H.n, 9.2, 0.0)[1]
            errs[(S, o)] = rel_inf(sig, sig_ref[o])
    for prec in (lq.LQ_PREC_FP32X2, lq.LQ_PREC_BF16X3):
        say("kubo float split", prec)
        got = lq.kubo_fft(G[0], G[1], [G[1], G[2]], phi, M, nump, lq.LQ_KUBO_BASTIN, precision=prec)
        for o in range(2):
            sig = lq.kubo_post(ctx, 1, got[o], H.n, 9.2, 0.0)[1]
            errs[(prec, o)] = rel_inf(sig, sig_ref[o])
            assert np.all(np.isfinite(sig)) and errs[(prec, o)] < 1.0
    print("sigma rel err", errs, flush=True)
    for o in range(2):
        assert errs[(7, o)] < 1e-8  # the accuracy target of the split path
        assert errs[(6, o)] < 1e-7 and errs[(5, o)] < 1e-4 and errs[(4, o)] < 1e-2
        assert errs[(7, o)] < errs[(6, o)] < errs[(5, o)] < errs[(4, o)]
        assert errs[(8, o)] < min(errs[(7, o)], 1e-9)
