"""The streamed 2D pipeline: left vectors per pass, right products in blocks
contracted on a second stream while the recursion continues.

Its tables must equal the resident pipeline's (every stored vector on the
device, one contraction per operator) to 1e-13 of their maximum -- only the
split of each K sum into chunks differs -- and the reference's golden
.chebmom2D files to 1e-12.  The pipeline is chosen once per process
(LQ_PIPE, LQ_PIPE_BL, LQ_PIPE_BR), so each setting runs in a child process:
one right block per table, many blocks with a partial last one, several left
passes (the right recursion restarted per pass, as the reference's
BATCH_SIZE loop does), fused and unfused right operators, three right
operators, sections, BATCH_SIZE < M, split precision.  A Kubo-Bastin job
whose stored vectors exceed the GPU's memory (N = 2097152, M = 4096) runs
through the streamed pipeline.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, rel_inf

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_SCRIPT = r"""
import os, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import linqt_amd as lq
from linqt_amd import io, lattices
golden = sys.argv[2]
ctx = lq.Context(0)
out = {}
H, Vx, Vy = lattices.haldane_honeycomb(24, seed=9)
Hg = lq.SparseMatrix.from_csr(ctx, H).Rescale(1 / 4.6, 0.0)
V = [lq.SparseMatrix.from_csr(ctx, Vx), lq.SparseMatrix.from_csr(ctx, Vy)]
phi = lattices.random_phase_states(H.n, 3, seed=4)
out["corr"] = lq.correlation_expansion_moments(Hg, V[0], V, phi, 96, 80)
out["corr_b"] = lq.correlation_expansion_moments(Hg, V[0], V, phi, 60, 60, batch=16)
out["kubo"] = lq.kubo_fft(Hg, V[0], V, phi, 64, 256, lq.LQ_KUBO_BASTIN)
out["kubo_gw"] = lq.kubo_fft(Hg, None, V[1], phi, 50, 128, lq.LQ_KUBO_GREENWOOD, num_sections=2)
out["kubo3"] = lq.kubo_fft(Hg, V[0], [V[0], V[1], None], phi, 40, 128, lq.LQ_KUBO_BASTIN)
out["kubo_i8"] = lq.kubo_fft(Hg, V[0], V, phi, 64, 256, lq.LQ_KUBO_BASTIN, precision=lq.LQ_PREC_INT8X7)
# a right operator outside H's pattern (third-neighbour hops): not fused into the H step
rng = np.random.default_rng(3)
n = H.n
lo = np.arange(n - 7)
vals = rng.normal(size=lo.size) + 1j * rng.normal(size=lo.size)
rp = np.concatenate([np.arange(n - 6), np.full(7, n - 7)]).astype(np.int32)
W = lq.SparseMatrix(ctx, n, rp, (lo + 7).astype(np.int32), vals)
out["corr_unfused"] = lq.correlation_expansion_moments(Hg, V[0], [W], phi, 70, 70)
gold_dir = os.path.join(golden, "linear_chain")
Hc = lq.SparseMatrix.from_file(ctx, os.path.join(gold_dir, "operators", "linear_chain.HAM.CSR"))
Vc = lq.SparseMatrix.from_file(ctx, os.path.join(gold_dir, "operators", "linear_chain.VX.CSR"))
Hc.Rescale(2.0 / 4.9, 0.0)
nc = Hc.info()["n"]
out["golden300"] = lq.correlation_expansion_moments(Hc, Vc, Vc, io.default_state(nc), 300, 300)
np.savez(sys.argv[3], **out)
"""

SETTINGS = {
    "resident": {},
    "one_block": {"LQ_PIPE": "stream"},
    "blocks": {"LQ_PIPE": "stream", "LQ_PIPE_BR": "17"},
    "passes": {"LQ_PIPE": "stream", "LQ_PIPE_BR": "16", "LQ_PIPE_BL": "10"},
}


@pytest.fixture(scope="module")
def runs(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("stream")
    out = {}
    for tag, extra in SETTINGS.items():
        print("[case] pipeline setting", tag, extra, flush=True)
        path = tmp / f"{tag}.npz"
        env = {k: v for k, v in os.environ.items() if not k.startswith("LQ_PIPE")}
        env.update(extra)
        r = subprocess.run([sys.executable, "-c", _SCRIPT, ROOT, GOLDEN, str(path)], env=env, capture_output=True,
                           text=True, timeout=600)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        out[tag] = dict(np.load(path))
    return out


@pytest.mark.parametrize("tag", ["one_block", "blocks", "passes"])
def test_streamed_tables_equal_resident(runs, tag):
    res, got = runs["resident"], runs[tag]
    for key in ("corr", "corr_b", "kubo", "kubo_gw", "kubo3", "corr_unfused", "golden300"):
        err = rel_inf(got[key], res[key])
        print(f"[case] {tag} {key}: {err:.2e}", flush=True)
        assert err < 1e-13, (tag, key, err)
    assert np.all(np.isfinite(got["corr"])) and np.abs(got["corr"]).max() > 0
    # split precision: the same digit method on the same vectors, chunked differently
    assert rel_inf(got["kubo_i8"], res["kubo_i8"]) < 1e-12


@pytest.mark.parametrize("tag", list(SETTINGS))
def test_streamed_golden(runs, tag):
    from linqt_amd import io

    gold, _, _, _ = io.read_chebmom2d(os.path.join(GOLDEN, "linear_chain",
                                                   "NonLEqOpVX-VXlinear_chainKPM_M300x300.chebmom2D"))
    assert rel_inf(runs[tag]["golden300"], gold) < 1e-12


@pytest.mark.slow
def test_bastin_beyond_device_memory_completes(ctx):
    """N = 2097152 (honeycomb 1024 x 1024 cells), M = 4096: the resident
    pipeline would store (M + 2M) N 16 B = 412 GB; the streamed one keeps the
    4096 left vectors (137 GB) and two blocks of right products."""
    import linqt_amd as lq
    from linqt_amd import lattices

    H, Vx, Vy = lattices.haldane_honeycomb(1024, seed=20261207)
    Hg = lq.SparseMatrix.from_csr(ctx, H).Rescale(1 / 4.6, 0.0)
    V = [lq.SparseMatrix.from_csr(ctx, Vx), lq.SparseMatrix.from_csr(ctx, Vy)]
    phi = lattices.random_phase_states(H.n, 1, seed=5)
    print(f"[case] bastin N={H.n} M=4096 R=1 streamed", flush=True)
    fd = lq.kubo_fft(Hg, V[0], V, phi, 4096, 8192, lq.LQ_KUBO_BASTIN)
    assert np.all(np.isfinite(fd)) and np.abs(fd).max() > 0
    for o in range(2):
        E, s, _ = lq.kubo_post(ctx, lq.LQ_KUBO_BASTIN, fd[o], H.n, 9.2, 0.0)
        assert np.all(np.isfinite(s[1:-1]))
