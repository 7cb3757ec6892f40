"""GPU parity: every C-ABI entry point against the CPU oracle / golden files.

Bar: moments within 1e-10 relative (inf-norm over the table, as the
reference's moments are O(1) and many vanish by symmetry); the energy-grid
outputs within 1e-9 relative; the contraction within 1e-12.
"""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, dense_herm, random_upper_csr, rel_inf
import linqt_amd as lq
from linqt_amd import io, lattices

pytestmark = pytest.mark.gpu

LC = os.path.join(GOLDEN, "linear_chain")
TOL_MU = 1e-10


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("k", [1, 2, 3, 4, 8, 16])
def test_spmv_matches_dense(ctx, k):
    n = 301
    rowptr, col, val = random_upper_csr(n, 0.03, seed=20 + k)
    say("spmv", n, k)
    A = lq.SparseMatrix(ctx, n, rowptr, col, val)
    D = dense_herm(n, rowptr, col, val)
    rng = np.random.default_rng(k)
    x = rng.normal(size=(n, k)) + 1j * rng.normal(size=(n, k))
    y = rng.normal(size=(n, k)) + 1j * rng.normal(size=(n, k))
    for a, b in ((1.0, 0.0), (2.0, -1.0), (0.3 - 0.7j, 1.1 + 0.2j)):
        got = A.Multiply(a, x, b, y)
        ref = a * (D @ x) + b * y
        assert rel_inf(got, ref) < 1e-13, (a, b)


def test_spmv_general_fill_and_rescale(ctx, oracle):
    n = 150
    rowptr, col, val = random_upper_csr(n, 0.05, seed=31, diag=False)
    say("rescale", n)
    A = lq.SparseMatrix(ctx, n, rowptr, col, val)
    O = oracle.Op(n, rowptr, col, val)
    A.Rescale(1 / 4.6, -0.3)
    O.rescale(1 / 4.6, -0.3)
    x = lattices.random_phase_states(n, 1, seed=2)[0]
    assert rel_inf(A.Multiply(1.0, x), O.mv(1.0, x, 0.0, np.zeros(n))) < 1e-14
    # general fill: the arrays are used as stored
    G = lq.SparseMatrix(ctx, n, rowptr, col, val, fill=lq.LQ_FILL_GENERAL)
    Dg = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        for q in range(rowptr[i], rowptr[i + 1]):
            Dg[i, col[q]] += val[q]
    assert rel_inf(G.Multiply(1.0, x), Dg @ x) < 1e-14


@pytest.mark.parametrize("ML,MR,K", [(70, 50, 1000), (64, 64, 4096 * 9 + 5), (1, 3, 17), (130, 129, 300),
                                      (256, 200, 4096 * 5 + 3)])
def test_compute_mom_table(ctx, ML, MR, K):
    import ctypes as C
    from linqt_amd import _lib

    say("mom_table", ML, MR, K)
    rng = np.random.default_rng(ML + MR)
    L = rng.normal(size=(ML, K)) + 1j * rng.normal(size=(ML, K))
    R = rng.normal(size=(MR, K)) + 1j * rng.normal(size=(MR, K))
    pl, pr, ps = ctx.alloc(L.nbytes), ctx.alloc(R.nbytes), ctx.alloc(ML * MR * 16)
    try:
        ctx.upload(pl, L)
        ctx.upload(pr, R)
        _lib.call("lq_compute_mom_table", ctx.h, pl, K, ML, pr, K, MR, K, ps, MR)
        sub = np.zeros((ML, MR), dtype=np.complex128)
        ctx.download(ps, sub)
    finally:
        for p in (pl, pr, ps):
            ctx.free(p)
    ref = L.conj() @ R.T
    assert rel_inf(sub, ref) < 1e-13


_TABLE_SCRIPT = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import linqt_amd as lq
from linqt_amd import _lib
ML, MR, K = (int(a) for a in sys.argv[2:5])
rng = np.random.default_rng(5)
L = rng.normal(size=(ML, K)) + 1j * rng.normal(size=(ML, K))
R = rng.normal(size=(MR, K)) + 1j * rng.normal(size=(MR, K))
ctx = lq.Context(0)
pl, pr, ps = ctx.alloc(L.nbytes), ctx.alloc(R.nbytes), ctx.alloc(ML * MR * 16)
ctx.upload(pl, L)
ctx.upload(pr, R)
_lib.call("lq_compute_mom_table", ctx.h, pl, K, ML, pr, K, MR, K, ps, MR)
sub = np.zeros((ML, MR), dtype=np.complex128)
ctx.download(ps, sub)
np.save(sys.argv[5], sub)
"""


@pytest.mark.parametrize("ML,MR,K", [(320, 200, 20483), (384, 320, 65536)])
def test_contraction_kernels_bitwise_equal(tmp_path, ML, MR, K):
    """The three-product kernels -- the 128 x 128 tile kernel loaded as a code
    object (the default), the 128 x 64 kernels with direct-to-LDS loads into
    two buffers (LQ_CONTRACT=dl) or three (LQ_CONTRACT_DLBUF=3), with one wave
    per SIMD (LQ_CONTRACT=w1) or register-staged (LQ_CONTRACT_DL=0), and the
    128 x 64 kernel the default falls back to for rows too long for its 32-bit
    lane offsets (forced with a lowered limit) -- form the same products in the
    same order: equal
    tables, bit for bit, including a partial last K stage and edge tiles.  The
    four-product kernel (LQ_CONTRACT=4m) rounds differently and is held to the
    same tolerance against the product itself.  The kernel is chosen once per
    process, so each runs in a child process."""
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = {}
    for tag, extra in (("t128", {}), ("dl2", {"LQ_CONTRACT": "dl"}), ("dl3", {"LQ_CONTRACT_DLBUF": "3"}),
                       ("reg", {"LQ_CONTRACT_DL": "0"}), ("w1", {"LQ_CONTRACT": "w1"}), ("4m", {"LQ_CONTRACT": "4m"}),
                       ("longrows", {"LQ_CONTRACT_T128_MAXLD": str(K - 1)})):
        say("contraction kernels", ML, MR, K, tag)
        path = tmp_path / f"{tag}.npy"
        env = {k: v for k, v in os.environ.items() if not k.startswith("LQ_CONTRACT")}
        env.update(extra)
        r = subprocess.run([sys.executable, "-c", _TABLE_SCRIPT, root, str(ML), str(MR), str(K), str(path)],
                           env=env, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        out[tag] = np.load(path)
    for tag in ("t128", "dl2", "dl3", "w1", "longrows"):
        assert np.array_equal(out[tag], out["reg"]), tag
    rng = np.random.default_rng(5)
    L = rng.normal(size=(ML, K)) + 1j * rng.normal(size=(ML, K))
    R = rng.normal(size=(MR, K)) + 1j * rng.normal(size=(MR, K))
    ref = L.conj() @ R.T
    assert rel_inf(out["t128"], ref) < 1e-13
    assert rel_inf(out["4m"], ref) < 1e-13


_FUSED_SCRIPT = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import linqt_amd as lq
from linqt_amd import lattices
ctx = lq.Context(0)
H, Vx, Vy = lattices.haldane_honeycomb(24, seed=9)
Hg = lq.SparseMatrix.from_csr(ctx, H).Rescale(1 / 4.6, 0.0)
V = [lq.SparseMatrix.from_csr(ctx, Vx), lq.SparseMatrix.from_csr(ctx, Vy)]
phi = lattices.random_phase_states(H.n, 3, seed=4)
mu = lq.correlation_expansion_moments(Hg, V[0], V, phi, 96, 80)
r = lq.kubo_fft(Hg, V[0], V, phi, 64, 256, lq.LQ_KUBO_BASTIN)
np.savez(sys.argv[2], mu=mu, r=r)
"""


def test_nontemporal_streams_bitwise_equal(tmp_path):
    """The fused recursion step stores the operator products and loads T_{n-1}
    non-temporally unless LQ_SPMM_NT=0; only the cache policy differs, so
    correlation tables and Kubo-Bastin sums are equal bit for bit.  The switch
    is read once per process: each setting runs in a child process."""
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = {}
    for nt in ("1", "0"):
        say("non-temporal streams", "LQ_SPMM_NT=" + nt)
        path = tmp_path / f"nt{nt}.npz"
        env = {k: v for k, v in os.environ.items() if not k.startswith("LQ_SPMM")}
        env["LQ_SPMM_NT"] = nt
        r = subprocess.run([sys.executable, "-c", _FUSED_SCRIPT, root, str(path)], env=env, capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        out[nt] = np.load(path)
    assert np.array_equal(out["1"]["mu"], out["0"]["mu"])
    assert np.array_equal(out["1"]["r"], out["0"]["r"])
    assert np.all(np.isfinite(out["1"]["mu"])) and np.abs(out["1"]["mu"]).max() > 0


@pytest.mark.parametrize("M,W,name", [
    (100, 10.0, "NonLEqOpVX-VXlinear_chainKPM_M100x100.chebmom2D"),
    (300, 4.9, "NonLEqOpVX-VXlinear_chainKPM_M300x300.chebmom2D"),
])
def test_golden_chebmom2d_on_gpu(ctx, M, W, name):
    say("golden", name)
    gold, n, _, _ = io.read_chebmom2d(os.path.join(LC, name))
    H = lq.SparseMatrix.from_file(ctx, os.path.join(LC, "operators", "linear_chain.HAM.CSR"))
    V = lq.SparseMatrix.from_file(ctx, os.path.join(LC, "operators", "linear_chain.VX.CSR"))
    H.Rescale(2.0 / W, 0.0)
    mu = lq.correlation_expansion_moments(H, V, V, io.default_state(n), M, M)
    assert rel_inf(mu, gold) < 1e-12


def _square(ctx, oracle, L, seed):
    m = lattices.square_anderson(L, W=1.0, seed=seed)
    A = lq.SparseMatrix.from_csr(ctx, m).Rescale(1 / 4.6, 0.0)
    O = oracle.Op.from_csr(m).rescale(1 / 4.6, 0.0)
    return m, A, O


@pytest.mark.parametrize("R,M", [(1, 1), (1, 2), (3, 101), (8, 256), (9, 64)])
def test_spectral_moments_direct_and_doubling(ctx, oracle, R, M):
    say("spectral", R, M)
    m, A, O = _square(ctx, oracle, 24, seed=R)
    phi = lattices.random_phase_states(m.n, R, seed=40 + R)
    ref = oracle.spectral_moments(O, None, phi, M)
    got = lq.spectral_moments(A, None, phi, M)
    assert rel_inf(got, ref) < TOL_MU
    got2 = lq.spectral_moments(A, None, phi, M, algo=lq.LQ_ALGO_DOUBLING)
    assert rel_inf(got2, ref) < TOL_MU
    # accumulation into an existing table, as Moments1D does
    acc = lq.spectral_moments(A, None, phi, M, mu=np.ones(M))
    assert rel_inf(acc, ref + 1.0) < TOL_MU


def test_spectral_moments_with_operator(ctx, oracle):
    say("spectral OP")
    m, A, O = _square(ctx, oracle, 20, seed=3)
    rowptr, col, val = random_upper_csr(m.n, 0.005, seed=5)
    OPg = lq.SparseMatrix(ctx, m.n, rowptr, col, val)
    OPo = oracle.Op(m.n, rowptr, col, val)
    phi = lattices.random_phase_states(m.n, 4, seed=6)
    ref = oracle.spectral_moments(O, OPo, phi, 77)
    assert rel_inf(lq.spectral_moments(A, OPg, phi, 77), ref) < TOL_MU
    with pytest.raises(lq.LinqtError):
        lq.spectral_moments(A, OPg, phi, 77, algo=lq.LQ_ALGO_DOUBLING)


def _honeycomb(ctx, oracle, L, seed=3):
    H, Vx, Vy = lattices.haldane_honeycomb(L, seed=seed)
    G = [lq.SparseMatrix.from_csr(ctx, H).Rescale(1 / 4.6, 0.0), lq.SparseMatrix.from_csr(ctx, Vx),
         lq.SparseMatrix.from_csr(ctx, Vy)]
    O = [oracle.Op.from_csr(H).rescale(1 / 4.6, 0.0), oracle.Op.from_csr(Vx), oracle.Op.from_csr(Vy)]
    return H.n, G, O


@pytest.mark.parametrize("M,batch,R", [(24, 0, 1), (24, 0, 3), (20, 6, 2), (20, 7, 2), (33, 16, 1)])
def test_correlation_moments_vs_oracle(ctx, oracle, M, batch, R):
    say("corr", M, batch, R)
    n, G, O = _honeycomb(ctx, oracle, 8)
    phi = lattices.random_phase_states(n, R, seed=M + R)
    got = lq.correlation_expansion_moments(G[0], G[1], [G[1], G[2]], phi, M, M, batch=batch)
    for o, opr in enumerate((O[1], O[2])):
        ref = oracle.correlation_moments(O[0], O[1], opr, phi, M, M, batch=batch)
        assert rel_inf(got[o], ref) < TOL_MU, o
    # the table is Hermitian after the reference's symmetrisation
    assert np.allclose(got[0], got[0].conj().T, rtol=0, atol=1e-14 * np.abs(got[0]).max())


def test_correlation_moments_accumulate_and_identity_ops(ctx, oracle):
    say("corr accumulate / identity")
    n, G, O = _honeycomb(ctx, oracle, 6)
    phi = lattices.random_phase_states(n, 2, seed=5)
    rng = np.random.default_rng(0)
    init = rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16))
    got = lq.correlation_expansion_moments(G[0], None, None, phi, 16, 16, mu=init)
    ref = oracle.correlation_moments(O[0], None, None, phi, 16, 16, mu=init)
    assert rel_inf(got, ref) < TOL_MU


@pytest.mark.parametrize("formula,M,nump,ns,R", [(0, 16, 16, 1, 2), (1, 16, 16, 1, 2), (1, 12, 64, 1, 3),
                                                 (0, 20, 40, 2, 2), (1, 8, 32, 3, 1)])
def test_kubo_fft_vs_oracle(ctx, oracle, formula, M, nump, ns, R):
    say("kubo", formula, M, nump, ns, R)
    n, G, O = _honeycomb(ctx, oracle, 6)
    phi = lattices.random_phase_states(n, R, seed=7)
    got = lq.kubo_fft(G[0], G[1], [G[1], G[2]], phi, M, nump, formula, num_sections=ns)
    for o, opr in enumerate((O[1], O[2])):
        ref, _ = oracle.kubo_fft(O[0], O[1], opr, phi, M, nump, formula, num_sections=ns)
        assert rel_inf(got[o], ref) < 1e-10, o
        E, s, g = lq.kubo_post(ctx, formula, got[o], n, 9.2, 0.0)
        Er, sr, gr = oracle.kubo_post(formula, ref, n, 9.2, 0.0)
        assert np.allclose(E, Er, rtol=0, atol=1e-13)
        assert rel_inf(s, sr) < 1e-9
        if formula == 1:
            assert rel_inf(g, gr) < 1e-9


def test_kubo_rejects_m_above_nump(ctx, oracle):
    n, G, O = _honeycomb(ctx, oracle, 4)
    phi = lattices.random_phase_states(n, 1, seed=1)
    with pytest.raises(lq.LinqtError):
        lq.kubo_fft(G[0], G[1], G[1], phi, 32, 16, 1)


def test_jackson_and_grid_postprocessors(ctx, oracle):
    say("post-processors")
    rng = np.random.default_rng(3)
    M, n, W, Ec = 48, 1234, 7.0, 0.2
    mu1 = rng.normal(size=M) + 1j * rng.normal(size=M)
    assert rel_inf(lq.jackson_1d(ctx, mu1, W, 60.0), oracle.jackson_1d(mu1, W, 60.0)) < 1e-14
    assert len(lq.jackson_1d(ctx, mu1, W, 60.0)) == len(oracle.jackson_1d(mu1, W, 60.0))
    mu2 = rng.normal(size=(M, M - 4)) + 1j * rng.normal(size=(M, M - 4))
    j2 = lq.jackson_2d(ctx, mu2, W, 90.0, 40.0)
    j2o = oracle.jackson_2d(mu2, W, 90.0, 40.0)
    assert j2.shape == j2o.shape and rel_inf(j2, j2o) < 1e-14
    for nd in (M, 2 * M + 1):
        E, r = lq.dos_fftgrid(ctx, mu1, n, W, Ec, nd)
        Eo, ro = oracle.dos_fftgrid(mu1, n, W, Ec, nd)
        assert np.allclose(E, Eo, rtol=0, atol=1e-13) and rel_inf(r, ro) < 1e-10
        E, s = lq.greenwood_fftgrid(ctx, mu2, n, W, Ec, nd)
        Eo, so = oracle.greenwood_fftgrid(mu2, n, W, Ec, nd)
        assert np.allclose(E, Eo, rtol=0, atol=1e-13) and rel_inf(s, so) < 1e-9
        E, s = lq.bastin_fftgrid(ctx, mu2, n, W, Ec, nd)
        Eo, so = oracle.bastin_fftgrid(mu2, n, W, Ec, nd)
        assert np.allclose(E, Eo, rtol=0, atol=1e-13) and rel_inf(s, so) < 1e-9


def test_bad_arguments_raise(ctx):
    with pytest.raises(lq.LinqtError):
        lq.SparseMatrix(ctx, 3, np.array([0, 1, 2, 2]), np.array([0, 5]), np.ones(2))
    with pytest.raises(lq.LinqtError):
        lq.SparseMatrix(ctx, 3, np.array([0, 1, 2, 9]), np.array([0, 1]), np.ones(2))


def test_dictionary_form_matches_plain_form(ctx, monkeypatch):
    """Lattice operators get the dictionary form (value table + packed
    columns, separate diagonal); it must agree with the plain form."""
    ops = [lattices.square_anderson(20, W=1.0, seed=3), lattices.kane_mele(6, W=1.0, seed=4)]
    H, Vx, Vy = lattices.haldane_honeycomb(7, seed=5)
    ops += [H, Vx, Vy]
    rng = np.random.default_rng(9)
    for m in ops:
        say("dict vs plain", m.n)
        monkeypatch.setenv("LQ_SPMM_DICT", "0")
        P = lq.SparseMatrix.from_csr(ctx, m)
        monkeypatch.setenv("LQ_SPMM_DICT", "1")
        D = lq.SparseMatrix.from_csr(ctx, m)
        for k in (1, 3, 8):
            x = rng.normal(size=(m.n, k)) + 1j * rng.normal(size=(m.n, k))
            y = rng.normal(size=(m.n, k)) + 1j * rng.normal(size=(m.n, k))
            a, b = 0.7 - 0.2j, -1.0
            assert rel_inf(D.Multiply(a, x, b, y), P.Multiply(a, x, b, y)) < 1e-14
        Dd = dense_herm(m.n, m.rowptr, m.col, m.val)
        x = rng.normal(size=(m.n, 2)) + 1j * rng.normal(size=(m.n, 2))
        assert rel_inf(D.Multiply(1.0, x), Dd @ x) < 1e-13
    # rescaled operator and the 1D solver on both forms
    m = lattices.square_anderson(24, W=1.0, seed=6)
    monkeypatch.setenv("LQ_SPMM_DICT", "0")
    P = lq.SparseMatrix.from_csr(ctx, m).Rescale(1 / 4.6, 0.1)
    monkeypatch.setenv("LQ_SPMM_DICT", "1")
    D = lq.SparseMatrix.from_csr(ctx, m).Rescale(1 / 4.6, 0.1)
    phi = lattices.random_phase_states(m.n, 3, seed=1)
    assert rel_inf(lq.spectral_moments(D, None, phi, 200), lq.spectral_moments(P, None, phi, 200)) < 1e-12


def test_prefetched_row_groups_bitwise_equal(ctx, monkeypatch):
    """The prefetched row-group schedule of the dictionary form (slice offsets
    staged per block, next group's words loaded behind the gathers) leaves
    every result bit unchanged, also past the 64 staged groups of a block
    (grid cap 256: 68 and 135 groups per block here) and over slices that
    hold only diagonal entries (width 0)."""
    n = 1_100_000
    rows = np.arange(n)
    hop = (rows < n - 1) & ~((rows >= 127) & (rows < 256))  # rows 128..255 isolated
    onsite = np.where((rows >= 128) & (rows < 256), 0.5 + 0.01 * (rows % 7), 0.0)
    cnt = 1 + hop.astype(np.int64)
    rowptr = np.concatenate([[0], np.cumsum(cnt)])
    col = np.empty(rowptr[-1], np.int64)
    val = np.empty(rowptr[-1], complex)
    col[rowptr[:-1]] = rows
    val[rowptr[:-1]] = onsite
    h = rows[hop]
    col[rowptr[h] + 1] = h + 1
    val[rowptr[h] + 1] = -1.0 + 0.25j
    m = io.CSR(n, rowptr, col, val)
    monkeypatch.setenv("LQ_SPMM_DICT", "1")
    monkeypatch.setenv("LQ_SPMM_GRIDCAP", "256")
    monkeypatch.setenv("LQ_SPMM_PF", "0")
    P = lq.SparseMatrix.from_csr(ctx, m).Rescale(1 / 2.7, 0.0)
    monkeypatch.setenv("LQ_SPMM_PF", "1")
    F = lq.SparseMatrix.from_csr(ctx, m).Rescale(1 / 2.7, 0.0)
    monkeypatch.delenv("LQ_SPMM_GRIDCAP")
    D = lq.SparseMatrix.from_csr(ctx, m).Rescale(1 / 2.7, 0.0)
    rng = np.random.default_rng(17)
    for k in (3, 8):
        say("prefetch schedule", n, k)
        x = rng.normal(size=(n, k)) + 1j * rng.normal(size=(n, k))
        y = rng.normal(size=(n, k)) + 1j * rng.normal(size=(n, k))
        ref = P.Multiply(0.7 - 0.2j, x, -1.0, y)
        assert np.array_equal(F.Multiply(0.7 - 0.2j, x, -1.0, y), ref)
        assert np.array_equal(D.Multiply(0.7 - 0.2j, x, -1.0, y), ref)
        # spot rows against the chain's own formula
        for i in (0, 126, 127, 128, 200, 255, 256, 70_000, n - 1):
            e = onsite[i] * x[i]
            if hop[i]:
                e = e + (-1.0 + 0.25j) * x[i + 1]
            if i > 0 and hop[i - 1]:
                e = e + (-1.0 - 0.25j) * x[i - 1]
            assert np.allclose(ref[i], (0.7 - 0.2j) * e / 2.7 - y[i], rtol=0, atol=1e-13)
        phi = lattices.random_phase_states(n, k, seed=k)
        mu = lq.spectral_moments(P, None, phi, 24)
        assert np.array_equal(lq.spectral_moments(F, None, phi, 24), mu)


def test_dictionary_values_indexed_across_row_ranges(ctx, monkeypatch):
    """The dictionary form's value table is built over row ranges on several
    threads and merged in first-seen order; an operator with more than 2^20
    entries whose second half uses hopping values the first half never does
    (so later ranges renumber their indices) gives the plain form's products
    and the chain's own formula."""
    n = 600_000
    rows = np.arange(n)
    t = np.where(rows < n // 2, -1.0 + 0.3j, 0.7 - 0.9j)  # hop (i, i+1) value
    t[2 * n // 3:] = 0.25 + 0.5j
    onsite = 0.1 * np.cos(rows)
    hop = rows < n - 1
    cnt = 1 + hop.astype(np.int64)
    rowptr = np.concatenate([[0], np.cumsum(cnt)])
    col = np.empty(rowptr[-1], np.int64)
    val = np.empty(rowptr[-1], complex)
    col[rowptr[:-1]] = rows
    val[rowptr[:-1]] = onsite
    h = rows[hop]
    col[rowptr[h] + 1] = h + 1
    val[rowptr[h] + 1] = t[h]
    m = io.CSR(n, rowptr, col, val)
    monkeypatch.setenv("LQ_SPMM_DICT", "0")
    P = lq.SparseMatrix.from_csr(ctx, m)
    monkeypatch.setenv("LQ_SPMM_DICT", "1")
    D = lq.SparseMatrix.from_csr(ctx, m)
    rng = np.random.default_rng(23)
    x = rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2))
    say("dictionary across row ranges", n)
    yd, yp = D.Multiply(1.0, x), P.Multiply(1.0, x)
    assert rel_inf(yd, yp) < 1e-14
    for i in (0, 1, n // 2 - 1, n // 2, n // 2 + 1, 2 * n // 3 - 1, 2 * n // 3, 2 * n // 3 + 1, n - 1):
        e = onsite[i] * x[i]
        if i < n - 1:
            e = e + t[i] * x[i + 1]
        if i > 0:
            e = e + np.conj(t[i - 1]) * x[i - 1]
        assert np.allclose(yd[i], e, rtol=0, atol=1e-13), i


def test_fused_right_operators_match_separate_applies(ctx, oracle, monkeypatch):
    """V_x, V_y (patterns inside H's) are applied inside the H step; results
    equal the separate applies, the oracle, and an operator outside H's
    pattern falls back to separate applies."""
    n, G, O = _honeycomb(ctx, oracle, 6)
    phi = lattices.random_phase_states(n, 3, seed=11)
    for formula in (0, 1):
        say("fused kubo", formula)
        monkeypatch.setenv("LQ_FUSE_OPR", "1")
        fused = lq.kubo_fft(G[0], G[1], [G[1], G[2]], phi, 40, 80, formula)
        monkeypatch.setenv("LQ_FUSE_OPR", "0")
        sep = lq.kubo_fft(G[0], G[1], [G[1], G[2]], phi, 40, 80, formula)
        assert rel_inf(fused, sep) < 1e-12
        ref, _ = oracle.kubo_fft(O[0], O[1], O[2], phi, 40, 80, formula)
        assert rel_inf(fused[1], ref) < 1e-10
    monkeypatch.setenv("LQ_FUSE_OPR", "1")
    one = lq.correlation_expansion_moments(G[0], G[1], G[2], phi, 24, 24)
    ref = oracle.correlation_moments(O[0], O[1], O[2], phi, 24, 24)
    assert rel_inf(one, ref) < 1e-10
    # a right operator with entries outside H's pattern
    rowptr, col, val = random_upper_csr(n, 0.02, seed=5)
    X = lq.SparseMatrix(ctx, n, rowptr, col, val)
    OX = oracle.Op(n, rowptr, col, val)
    got = lq.correlation_expansion_moments(G[0], G[1], [G[2], X], phi, 16, 16)
    assert rel_inf(got[1], oracle.correlation_moments(O[0], O[1], OX, phi, 16, 16)) < 1e-10
    assert rel_inf(got[0], oracle.correlation_moments(O[0], O[1], O[2], phi, 16, 16)) < 1e-10


def _same_rows(got, ref, tol):
    gE, gv = got
    rE, rv = ref
    assert len(gE) == len(rE) and np.allclose(gE, rE, rtol=0, atol=1e-13 * max(1.0, np.abs(rE).max()))
    assert np.array_equal(np.isfinite(gv), np.isfinite(rv))  # +-1: NaN or inf by summation order
    f = np.isfinite(rv)
    if f.any():
        assert np.abs(gv[f] - rv[f]).max() <= tol * np.abs(rv[f]).max()


@pytest.mark.parametrize("kind", range(7))
def test_linear_grid_tools_vs_oracle(ctx, oracle, kind):
    say("linear grid", kind)
    rng = np.random.default_rng(kind)
    mu = rng.normal(size=(36, 28)) + 1j * rng.normal(size=(36, 28))
    if kind == 0:
        mu = mu[:, 0].copy()
    n, W, Ec = 123, 7.5, -0.3
    nd = 30 * (36 if kind else 36)
    _same_rows(lq.linear_grid(ctx, kind, mu, n, W, Ec, nd), oracle.linear_grid(kind, mu, n, W, Ec, nd), 1e-10)
    # golden linear-chain table, Jackson-damped as the tools do
    g, gn, gW, gEc = io.read_chebmom2d(os.path.join(LC, "NonLEqOpVX-VXlinear_chainKPM_M100x100.chebmom2D"))
    muJ = oracle.jackson_2d(g, gW, 60.0, 60.0)
    if kind == 0:
        muJ = muJ[:, 0].copy()
    nd = 30 * max(np.atleast_2d(muJ).shape)
    _same_rows(lq.linear_grid(ctx, kind, muJ, gn, gW, gEc), oracle.linear_grid(kind, muJ, gn, gW, gEc, nd), 1e-9)


def test_time_correlations_vs_oracle(ctx, oracle):
    """TimeDependentCorrelations with V_x, V_y and with identities, several states,
    a band centre != 0 (the propagator's phase), then the TD Jackson kernel and
    the time-correlation post-processor."""
    say("time correlations")
    H, Vx, Vy = lattices.haldane_honeycomb(6, seed=21)
    W, Ec = 9.4, 0.15
    a = W / 2
    G = lq.SparseMatrix.from_csr(ctx, H).Rescale(1 / a, -Ec / a)
    O = oracle.Op.from_csr(H).rescale(1 / a, -Ec / a)
    GV = [lq.SparseMatrix.from_csr(ctx, Vx), lq.SparseMatrix.from_csr(ctx, Vy)]
    OV = [oracle.Op.from_csr(Vx), oracle.Op.from_csr(Vy)]
    phi = lattices.random_phase_states(H.n, 3, seed=22)
    M, nT, dt = 30, 6, 1.3
    for (gl, gr, ol, orr) in ((GV[0], GV[1], OV[0], OV[1]), (None, None, None, None)):
        got = lq.time_correlations(G, gl, gr, phi, M, nT, dt, W, Ec)
        ref = oracle.time_correlations(O, ol, orr, phi, M, nT, dt, W, Ec)
        assert rel_inf(got, ref) < 1e-10
    tdj = lq.jackson_td(ctx, got, W, 400.0)
    tdo = oracle.jackson_td(ref, W, 400.0)
    assert tdj.shape == tdo.shape and rel_inf(tdj, tdo) < 1e-10
    t, v = lq.timecorr_post(ctx, tdj, H.n, W, Ec, dt, 0.7)
    to, vo = oracle.timecorr_post(tdo, H.n, W, Ec, dt, 0.7)
    assert np.allclose(t, to) and rel_inf(v, vo) < 1e-9


@pytest.mark.parametrize("fill", [lq.LQ_FILL_HERMITIAN_UPPER, lq.LQ_FILL_GENERAL])
def test_csr_create_counting_path_equals_sorting_path(ctx, fill):
    """Rows with sorted columns take the O(nnz) counting construction; the same
    entries with each row reversed (duplicates and lower-triangle entries
    included) take the stable sort.  Both give the same operator bitwise."""
    say("csr create paths", fill)
    rng = np.random.default_rng(12)
    n = 700
    rows, cols = [], []
    for i in range(n):
        js = sorted(rng.choice(n, size=6, replace=False).tolist() + [i, i])  # a duplicated diagonal
        if i % 7 == 0:
            js.append(js[-1])  # a duplicated off-diagonal entry
        rows += [i] * len(js)
        cols += js
    rows, cols = np.array(rows), np.array(cols, np.int32)
    vals = rng.normal(size=cols.size) + 1j * rng.normal(size=cols.size)
    rowptr = np.zeros(n + 1, np.int32)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.int32)
    A = lq.SparseMatrix(ctx, n, rowptr, cols, vals, fill=fill)
    rc, rv = cols.copy(), vals.copy()
    for i in range(n):
        s = slice(rowptr[i], rowptr[i + 1])
        rc[s], rv[s] = rc[s][::-1], rv[s][::-1]
    B = lq.SparseMatrix(ctx, n, rowptr, rc, rv, fill=fill)
    ea, eb = A.export(), B.export()
    assert A.info() == B.info() and np.array_equal(ea.rowptr, eb.rowptr)
    # the stable sort keeps equal (row, col) entries in input order, which the reversal flips:
    # compare per (row, col) sums, then the operators themselves
    x = rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2))
    ya, yb = A.Multiply(1.0, x), B.Multiply(1.0, x)
    assert np.abs(ya - yb).max() <= 1e-13 * np.abs(ya).max()
    # without duplicates the two paths agree bitwise
    keep = np.ones(cols.size, bool)
    for i in range(n):
        seen = set()
        for q in range(rowptr[i], rowptr[i + 1]):
            if cols[q] in seen:
                keep[q] = False
            seen.add(cols[q])
    r2 = rows[keep]
    rp2 = np.zeros(n + 1, np.int32)
    np.add.at(rp2, r2 + 1, 1)
    rp2 = np.cumsum(rp2).astype(np.int32)
    c2, v2 = cols[keep], vals[keep]
    A2 = lq.SparseMatrix(ctx, n, rp2, c2, v2, fill=fill)
    rc2, rv2 = c2.copy(), v2.copy()
    for i in range(n):
        s = slice(rp2[i], rp2[i + 1])
        rc2[s], rv2[s] = rc2[s][::-1], rv2[s][::-1]
    B2 = lq.SparseMatrix(ctx, n, rp2, rc2, rv2, fill=fill)
    e2a, e2b = A2.export(), B2.export()
    assert np.array_equal(e2a.col, e2b.col) and np.array_equal(e2a.val.view(np.int64), e2b.val.view(np.int64))
    assert np.array_equal(A2.Multiply(1.0, x), B2.Multiply(1.0, x))
