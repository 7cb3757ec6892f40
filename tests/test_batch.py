"""Batched spectral-moment jobs (grouped recursion launches).

Every job of a batch must give, bit for bit, what it gives alone and what
lq_spectral_moments gives on a fresh mu; small jobs are also checked
against the oracle.  Admission under a tight budget, continuous submission
between chunks, width classes with R = 1..16 and M = 1 / 2 edge cases are
covered.
"""
import sys

import numpy as np
import pytest

import linqt_amd as lq
from conftest import rel_inf
from linqt_amd import lattices

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _jobs(ctx):
    """(H device, H host, phi, M) for a mix of lattices, state counts and orders."""
    specs = [("sq", 16, 1, 64), ("sq", 24, 3, 100), ("cub", 8, 2, 37), ("km", 6, 5, 80), ("cub", 10, 16, 50),
             ("sq", 40, 8, 120), ("km", 10, 1, 1), ("sq", 12, 4, 2), ("cub", 12, 8, 90), ("km", 8, 16, 33)]
    out = []
    for i, (kind, L, R, M) in enumerate(specs):
        if kind == "sq":
            Hh, a = lattices.square_anderson(L, W=1.0, seed=100 + i), 4.6
        elif kind == "cub":
            Hh, a = lattices.cubic_anderson(L, W=2.0, seed=100 + i), 7.2
        else:
            Hh, a = lattices.kane_mele(L, W=1.0, seed=100 + i), 4.6
        H = lq.SparseMatrix.from_csr(ctx, Hh).Rescale(1 / a, 0.0)
        phi = lattices.random_phase_states(Hh.n, R, seed=200 + i)
        out.append((H, Hh, a, phi, M))
    return out


def test_batch_matches_single_job_and_oracle(ctx, oracle):
    jobs = _jobs(ctx)
    say("batch mix", len(jobs))
    b = lq.Batch(ctx, chunk=16)
    ids = [b.submit(H, phi, M) for (H, _, _, phi, M) in jobs]
    running, queued = b.run()
    assert running == 0 and queued == 0
    for jid, (H, Hh, a, phi, M) in zip(ids, jobs):
        assert b.state(jid) == 2
        got = b.result(jid)
        single = lq.spectral_moments(H, None, phi, M)
        assert np.array_equal(got, single), (jid, np.abs(got - single).max())
        if Hh.n <= 2000:
            O = oracle.Op.from_csr(Hh).rescale(1 / a, 0.0)
            ref = oracle.spectral_moments(O, None, phi, M)
            assert rel_inf(got, ref) < 1e-10
    for R in (1, 2, 4, 8, 16):
        assert b.unroll(R) in (0, 4, 8, 12, 16)
    b.close()


def test_batch_alone_equals_mixed_and_tight_budget(ctx):
    jobs = _jobs(ctx)
    say("alone vs mixed, tight budget")
    alone = []
    for (H, _, _, phi, M) in jobs:
        b = lq.Batch(ctx, chunk=8)
        j = b.submit(H, phi, M)
        b.run()
        alone.append(b.result(j))
        b.close()
    # a budget that holds only the two largest footprints at once
    sizes = sorted(lq.Batch.job_bytes(Hh.n, phi.shape[0], M, 8) for (_, Hh, _, phi, M) in jobs)
    b = lq.Batch(ctx, hbm_budget=sizes[-1] + sizes[-2], chunk=8)
    ids = [b.submit(H, phi, M) for (H, _, _, phi, M) in jobs]
    b.run()
    for jid, ref in zip(ids, alone):
        assert np.array_equal(b.result(jid), ref)
    with pytest.raises(lq.LinqtError):
        b.submit(jobs[0][0], np.zeros((65, jobs[0][1].n), complex), 4)  # R > 64
    small = lq.Batch(ctx, hbm_budget=1024, chunk=8)
    with pytest.raises(lq.LinqtError):
        small.submit(jobs[0][0], jobs[0][3], jobs[0][4])  # above the budget
    small.close()
    b.close()


def test_batch_continuous_submission(ctx):
    jobs = _jobs(ctx)
    say("continuous submission")
    b = lq.Batch(ctx, chunk=8)
    ids = {}
    pending = list(range(len(jobs)))
    ticks = 0
    while pending or any(b.state(j) != 2 for j in ids.values()):
        if pending:  # one new job per chunk
            k = pending.pop(0)
            H, _, _, phi, M = jobs[k]
            ids[k] = b.submit(H, phi, M)
        b.run(max_chunks=1)
        ticks += 1
        assert ticks < 10000
    for k, jid in ids.items():
        H, _, _, phi, M = jobs[k]
        assert np.array_equal(b.result(jid), lq.spectral_moments(H, None, phi, M))
    b.close()


def test_batch_device_states(ctx):
    H, Hh, a, phi, M = _jobs(ctx)[5]
    say("device states")
    ds = lq.DeviceStates(ctx, phi)
    b = lq.Batch(ctx)
    j = b.submit(H, ds, M)
    b.run()
    assert np.array_equal(b.result(j), lq.spectral_moments(H, None, phi, M))
    b.close()
