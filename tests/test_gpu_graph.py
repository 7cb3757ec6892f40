"""Captured recursion chunks (HIP graphs) in the 1D solver.

When every moment's partials fit in one buffer the whole recursion loop of a
solve is captured when the same solve comes again and replayed from then on;
otherwise full rings of 64 recursion steps are captured once and replayed,
with the ring's reduction launched between replays.  The replayed kernels are the
eager ones with the same arguments, so the moments must be equal bit for
bit to the run with every step launched on its own (LQ_GRAPH=0; the default
captures only ring chunks, LQ_GRAPH=1 also whole loops; read once per
process: each setting runs in a child process).  Every solve of the
script replays the cached graph several times, with other work of the
program (operator applies, a different solve that replaces the cached graph,
host synchronisations) between the solves, and every result is compared.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_SCRIPT = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import linqt_amd as lq
from linqt_amd import lattices
ctx = lq.Context(0)
S = lattices.square_anderson(64, W=1.0, seed=11)
H = lq.SparseMatrix.from_csr(ctx, S).Rescale(1 / 4.6, 0.0)
rp = np.arange(S.n + 1, dtype=np.int32)
OP = lq.SparseMatrix(ctx, S.n, rp, np.arange(S.n, dtype=np.int32), np.linspace(0.5, 1.5, S.n).astype(np.complex128))
phi8 = lattices.random_phase_states(S.n, 8, seed=12)
phi70 = lattices.random_phase_states(S.n, 70, seed=13)
out = {}
x = lattices.random_phase_states(S.n, 1, seed=14)[0]
S2 = lattices.square_anderson(256, W=1.0, seed=15)
H2 = lq.SparseMatrix.from_csr(ctx, S2).Rescale(1 / 4.6, 0.0)
phi64 = lq.DeviceStates(ctx, lattices.random_phase_states(S2.n, 64, seed=16))
for rep in range(3):
    out[f"dos1024_{rep}"] = lq.spectral_moments(H, None, phi8, 1024)
    out[f"dos1024again_{rep}"] = lq.spectral_moments(H, None, phi8, 1024)  # the same solve again: whole loop captured
    out[f"ring_{rep}"] = lq.spectral_moments(H2, None, phi64, 300)   # partials above 256 MiB: 64-step ring chunks
    out[f"m200_{rep}"] = lq.spectral_moments(H, None, phi8, 200)
    y = H.Multiply(1.0, x)          # eager work between solves
    ctx.synchronize()
    out[f"op_{rep}"] = lq.spectral_moments(H, OP, phi8, 300)
    out[f"r70_{rep}"] = lq.spectral_moments(H, None, phi70, 260)
    out[f"m100_{rep}"] = lq.spectral_moments(H, None, phi8, 100)
    out[f"y_{rep}"] = y
np.savez(sys.argv[2], **out)
"""


@pytest.fixture(scope="module")
def runs(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("graph")
    out = {}
    for tag, val in (("graph", "1"), ("ring_only", ""), ("eager", "0")):
        print("[case] LQ_GRAPH =", val, flush=True)
        path = tmp / f"{tag}.npz"
        env = {k: v for k, v in os.environ.items() if k != "LQ_GRAPH"}
        if val:
            env["LQ_GRAPH"] = val
        r = subprocess.run([sys.executable, "-c", _SCRIPT, ROOT, str(path)], env=env, capture_output=True,
                           text=True, timeout=600)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        out[tag] = dict(np.load(path))
    return out


def test_graph_replays_equal_eager_launches(runs):
    g, e = runs["graph"], runs["eager"]
    assert set(g) == set(e) == set(runs["ring_only"])
    for key in sorted(g):
        assert np.array_equal(g[key], e[key]), key
        assert np.array_equal(runs["ring_only"][key], e[key]), key
    # every repetition of a solve gives the same moments (cached graph replayed)
    for name in ("dos1024", "dos1024again", "ring", "m200", "op", "r70", "m100"):
        for rep in (1, 2):
            assert np.array_equal(g[f"{name}_{rep}"], g[f"{name}_0"]), (name, rep)
    assert np.array_equal(g["dos1024again_0"], g["dos1024_0"])
    assert abs(g["dos1024_0"][0] - 1.0) < 1e-12
