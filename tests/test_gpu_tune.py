"""The recursion-step schedule tuner (lq_csr_tune).

Every schedule it chooses between does the same arithmetic, so moments are
bitwise equal before and after tuning; the choice is kept in the operator,
persisted in the cache file and read back (without timing) for the same GPU
and operator shape, also by another operator of that shape.
"""
import os

import numpy as np
import pytest

import linqt_amd as lq
from linqt_amd import lattices

pytestmark = pytest.mark.gpu


def test_tuned_schedules_keep_results_and_persist(ctx, tmp_path, monkeypatch):
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("LQ_TUNE_CACHE", str(cache))
    S = lattices.square_anderson(96, W=1.0, seed=5)
    H = lq.SparseMatrix.from_csr(ctx, S).Rescale(1 / 4.6, 0.0)
    phi = lattices.random_phase_states(S.n, 8, seed=6)
    before = lq.spectral_moments(H, None, phi, 300)
    Hh, Vx, Vy = lattices.haldane_honeycomb(24, seed=7)
    G = lq.SparseMatrix.from_csr(ctx, Hh).Rescale(1 / 4.6, 0.0)
    V = [lq.SparseMatrix.from_csr(ctx, Vx), lq.SparseMatrix.from_csr(ctx, Vy)]
    ph2 = lattices.random_phase_states(Hh.n, 3, seed=8)
    corr_before = lq.correlation_expansion_moments(G, V[0], V, ph2, 64, 64)

    t1 = H.tune(8, mode=1)
    t0 = H.tune(8, mode=0)
    tg = G.tune(3, mode=0)
    for t in (t1, t0, tg):
        assert t["unroll"] in (4, 8, 12, 16) and not t["from_cache"]
    assert t0["gridcap"] in (0, 2048) and t0["prefetch"] in (0, 1)
    # the operator now runs the tuned schedule (the last tuning applied): same moments
    after = lq.spectral_moments(H, None, phi, 300)
    assert np.array_equal(before, after)
    assert np.array_equal(corr_before, lq.correlation_expansion_moments(G, V[0], V, ph2, 64, 64))
    # read back from the file, also for another operator of the same shape
    H2 = lq.SparseMatrix.from_csr(ctx, lattices.square_anderson(96, W=1.0, seed=9)).Rescale(1 / 4.6, 0.0)
    r = H2.tune(8, mode=0)
    assert r["from_cache"] and {k: r[k] for k in ("unroll", "gridcap", "prefetch")} == \
        {k: t0[k] for k in ("unroll", "gridcap", "prefetch")}
    lines = [ln for ln in cache.read_text().splitlines() if ln.startswith("lqtune1 ")]
    assert len(lines) == 3
    # a damaged entry for the same key (appended last) is skipped, the valid one still read
    key = " ".join(lines[1].split()[1:-4])
    with open(cache, "a") as f:
        f.write(f"lqtune1 {key} 5 -3 7 1.0\n")
    r2 = H2.tune(8, mode=0)
    assert r2["from_cache"] and {k: r2[k] for k in ("unroll", "gridcap", "prefetch")} == \
        {k: t0[k] for k in ("unroll", "gridcap", "prefetch")}
    with pytest.raises(lq.LinqtError):
        H.tune(0)


def test_ctx_release_then_solve_again(ctx):
    """lq_ctx_release frees the context's buffers and cached graphs; the next
    solves allocate again and give the same moments and tables."""
    S = lattices.square_anderson(32, W=1.0, seed=2)
    H = lq.SparseMatrix.from_csr(ctx, S).Rescale(1 / 4.6, 0.0)
    phi = lattices.random_phase_states(S.n, 4, seed=3)
    a = lq.spectral_moments(H, None, phi, 200)
    Hh, Vx, Vy = lattices.haldane_honeycomb(12, seed=4)
    G = lq.SparseMatrix.from_csr(ctx, Hh).Rescale(1 / 4.6, 0.0)
    V = lq.SparseMatrix.from_csr(ctx, Vx)
    ph = lattices.random_phase_states(Hh.n, 2, seed=5)
    t = lq.kubo_fft(G, V, V, ph, 48, 128, lq.LQ_KUBO_BASTIN)
    ctx.release()
    ctx.release()  # twice is harmless
    assert np.array_equal(lq.spectral_moments(H, None, phi, 200), a)
    assert np.array_equal(lq.kubo_fft(G, V, V, ph, 48, 128, lq.LQ_KUBO_BASTIN), t)
