"""The allocation-free 1D solver, lq_spectral_moments_noalloc, against
lq_spectral_moments (direct algorithm) through the C ABI.

Both run the same groups, kernels and summation order, so the moments must be
equal bit for bit (the bar for floating point here is 1e-10 relative on the
moments; bitwise equality is the stronger statement).  Two calls in a row, the
second on a smaller input, check that the shared device scratch carries
nothing from one call into the next; a size above the scratch maximum is
rejected with LQ_ERR_ARG before any work is enqueued.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(ctx, H, OP, phi, M, mu0=None):
    """mu0 + moments from the allocation-free entry point (device phi, device mu)."""
    from linqt_amd import _lib

    phi = np.ascontiguousarray(np.atleast_2d(phi), dtype=np.complex128)
    R = phi.shape[0]
    mu = np.zeros(M, dtype=np.complex128) if mu0 is None else np.array(mu0, dtype=np.complex128)
    pphi, pmu = ctx.alloc(phi.nbytes), ctx.alloc(mu.nbytes)
    try:
        ctx.upload(pphi, phi)
        ctx.upload(pmu, mu)
        _lib.call("lq_spectral_moments_noalloc", H.h, None if OP is None else OP.h, pphi, R, M, pmu)
        ctx.synchronize()
        ctx.download(pmu, mu)
    finally:
        ctx.free(pphi)
        ctx.free(pmu)
    return mu


def _case(ctx, L, R, M, seed, with_op):
    import linqt_amd as lq
    from linqt_amd import lattices

    S = lattices.square_anderson(L, seed=seed)
    H = lq.SparseMatrix.from_csr(ctx, S).Rescale(1 / 4.6, 0.0)
    OP = None
    if with_op:
        # a diagonal operator (site occupation weights) as the OP of <phi|T_m OP|phi>
        rng = np.random.default_rng(seed + 1)
        d = rng.uniform(0.5, 1.5, S.n)
        rp = np.arange(S.n + 1, dtype=np.int32)
        OP = lq.SparseMatrix(ctx, S.n, rp, np.arange(S.n, dtype=np.int32), d.astype(np.complex128))
    phi = lattices.random_phase_states(S.n, R, seed=seed + 2)
    return H, OP, phi


@pytest.mark.parametrize("with_op", [False, True])
def test_noalloc_equals_spectral_moments_twice(with_op):
    import linqt_amd as lq

    ctx = lq.Context(0)
    try:
        # first call: N = 4096, R = 8, M = 300; second: N = 256, R = 3, M = 50
        for L, R, M, seed in ((64, 8, 300, 11), (16, 3, 50, 23)):
            H, OP, phi = _case(ctx, L, R, M, seed, with_op)
            start = np.linspace(-1, 1, M) + 0.5j  # the table is accumulated into
            ref = lq.spectral_moments(H, OP, phi, M, mu=start)
            got = _noalloc(ctx, H, OP, phi, M, mu0=start)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"noalloc L={L} R={R} M={M} op={with_op}: max rel diff {err:.2e}", flush=True)
            assert err <= 1e-10
            assert np.array_equal(got.view(np.float64), ref.view(np.float64))
    finally:
        ctx.close()


def test_noalloc_more_states_than_one_group():
    """R = 70: a group of 64 states and one of 6, summed in the entry point's order."""
    import linqt_amd as lq

    ctx = lq.Context(0)
    try:
        H, OP, phi = _case(ctx, 12, 70, 40, 5, False)
        ref = lq.spectral_moments(H, None, phi, 40)
        got = _noalloc(ctx, H, None, phi, 40)
        assert np.array_equal(got.view(np.float64), ref.view(np.float64))
    finally:
        ctx.close()


def test_noalloc_rejects_sizes_above_its_scratch():
    import linqt_amd as lq
    from linqt_amd import _lib

    ctx = lq.Context(0)
    try:
        n, R, M = 1 << 20, 8, 16  # three [n][8] blocks alone are 384 MiB
        rp = np.arange(n + 1, dtype=np.int32)
        H = lq.SparseMatrix(ctx, n, rp, np.arange(n, dtype=np.int32), np.full(n, 0.5, dtype=np.complex128))
        pphi, pmu = ctx.alloc(16 * n * R), ctx.alloc(16 * M)
        try:
            rc = _lib.load().lq_spectral_moments_noalloc(H.h, None, pphi, R, M, pmu)
            assert rc == _lib.LQ_ERR_ARG
            assert b"above the maximum" in _lib.load().lq_last_error()
            # host pointers are refused as well
            h = np.zeros(2 * n * R)
            rc = _lib.load().lq_spectral_moments_noalloc(H.h, None, h.ctypes.data_as(C.c_void_p), R, M, pmu)
            assert rc == _lib.LQ_ERR_ARG
        finally:
            ctx.free(pphi)
            ctx.free(pmu)
    finally:
        ctx.close()
