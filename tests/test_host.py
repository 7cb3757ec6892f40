"""Host-side checks that need no GPU: file formats, generators, the C ABI's
exported symbols."""
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, dense_herm
from linqt_amd import io, lattices


def test_csr_roundtrip(tmp_path):
    m = lattices.square_anderson(5, seed=1)
    p = tmp_path / "x.HAM.CSR"
    io.write_csr(str(p), m)
    r = io.read_csr(str(p))
    assert r.n == m.n and np.array_equal(r.rowptr, m.rowptr) and np.array_equal(r.col, m.col)
    assert np.allclose(r.val, m.val, rtol=1e-14, atol=0)


def test_golden_csr_is_upper_triangle():
    H = io.read_csr(os.path.join(GOLDEN, "linear_chain", "operators", "linear_chain.HAM.CSR"))
    rows = np.repeat(np.arange(H.n), np.diff(H.rowptr))
    assert H.n == 20000 and H.nnz == 20000
    assert np.all(H.col >= rows)


def test_chebmom_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    mu = rng.normal(size=(10, 20)) + 1j * rng.normal(size=(10, 20))
    p = str(tmp_path / "t.chebmom2D")
    io.write_chebmom2d(p, mu, 100, 1.0, 2.0)
    r, n, W, Ec = io.read_chebmom2d(p)
    assert (n, W, Ec) == (100, 1.0, 2.0)
    assert np.allclose(r, mu, rtol=1e-14, atol=0)
    m1 = mu[0]
    p1 = str(tmp_path / "t.chebmom1D")
    io.write_chebmom1d(p1, m1, 7, 3.0, -1.0)
    r1, n1, W1, E1 = io.read_chebmom1d(p1)
    assert (n1, W1, E1) == (7, 3.0, -1.0) and np.allclose(r1, m1, rtol=1e-14, atol=0)
    with pytest.raises(ValueError):
        io.read_chebmom2d(str(tmp_path / "nope.txt"))


def test_state_files(tmp_path):
    phi = lattices.random_phase_states(11, 3, seed=2)
    p = str(tmp_path / "s.vec")
    io.write_vector_state_file(p, phi)
    assert np.allclose(io.read_state_file(p, 11), phi, rtol=1e-14, atol=0)
    q = tmp_path / "loc"
    q.write_text("local 2 3 5\n")
    loc = io.read_state_file(str(q), 11)
    assert loc.shape == (2, 11) and loc[0, 3] == 1 and loc[1, 5] == 1
    d = io.read_state_file(str(tmp_path / "missing"), 11)
    assert d.shape == (1, 11) and d[0, 5] == 1 and np.abs(d).sum() == 1


def test_random_phase_states_normalised_and_deterministic():
    a = lattices.random_phase_states(1000, 4, seed=5)
    b = lattices.random_phase_states(1000, 4, seed=5)
    assert np.array_equal(a, b)
    assert np.allclose(np.linalg.norm(a, axis=1), 1.0)
    u = lattices.uniform(1, 2, 100000)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01
    # counter based: a slice equals the tail of the full draw
    assert np.array_equal(lattices.uniform(1, 2, 10, start=50), u[50:60])


def _dense(m):
    return dense_herm(m.n, m.rowptr, m.col, m.val)


def test_square_lattice_spectrum_and_bounds():
    m = lattices.square_anderson(6, t=1.0, W=0.0)
    e = np.linalg.eigvalsh(_dense(m))
    k = 2 * np.pi * np.arange(6) / 6
    ref = np.sort((-2 * np.cos(k)[:, None] - 2 * np.cos(k)[None, :]).ravel())
    assert np.allclose(e, ref, atol=1e-12)
    d = lattices.square_anderson(8, W=1.0, seed=7)
    assert np.abs(np.linalg.eigvalsh(_dense(d))).max() < 4.6
    full = 2 * d.nnz - d.n
    assert full == 5 * d.n


def test_haldane_operators():
    L = 6
    H, Vx, Vy = lattices.haldane_honeycomb(L, t2=0.1, W=2.0, seed=3)
    n = 2 * L * L
    Dh, Dx, Dy = _dense(H), _dense(Vx), _dense(Vy)
    assert np.abs(np.linalg.eigvalsh(Dh)).max() <= 4.6
    assert 2 * H.nnz - n == 10 * n  # 3 NN + 6 NNN + diagonal
    assert 2 * Vx.nnz == 8 * n and 2 * Vy.nnz == 7 * n
    # clean, t2 = 0: spectrum of graphene is +-|f(k)|; check the max = 3t
    H0, _, _ = lattices.haldane_honeycomb(L, t2=0.0, W=0.0)
    e0 = np.linalg.eigvalsh(_dense(H0))
    assert abs(e0.max() - 3.0) < 1e-12 and abs(e0.min() + 3.0) < 1e-12
    # velocity = i[H, X] on bonds that do not wrap: compare with coordinates
    a1 = np.array([1.0, 0.0])
    a2 = np.array([0.5, np.sqrt(3) / 2])
    pos = np.zeros((n, 2))
    for c in range(L * L):
        mm, kk = c % L, c // L
        pos[2 * c] = mm * a1 + kk * a2
        pos[2 * c + 1] = pos[2 * c] + (a1 + a2) / 3
    dxm = pos[None, :, 0] - pos[:, None, 0]
    inner = (np.abs(dxm) < 1.01) & (np.abs(pos[None, :, 1] - pos[:, None, 1]) < 1.01)
    ref = 1j * dxm * Dh
    np.fill_diagonal(ref, 0)
    assert np.allclose(np.where(inner, Dx, 0), np.where(inner, ref, 0), atol=1e-12)
    assert np.allclose(Dx, Dx.conj().T) and np.allclose(Dy, Dy.conj().T)


def test_cubic_lattice():
    m = lattices.cubic_anderson(4, W=0.0)
    e = np.linalg.eigvalsh(_dense(m))
    assert abs(e.min() + 6.0) < 1e-12 and 2 * m.nnz - m.n == 7 * m.n


def test_kane_mele_hermitian_kramers_and_clean_limit():
    # time reversal (SOC and Rashba both respect it): Kramers pairs at W = 0
    e = np.sort(np.linalg.eigvalsh(_dense(lattices.kane_mele(6, W=0.0))))
    assert np.abs(e[0::2] - e[1::2]).max() < 1e-12
    # lso = lr = 0: two spin copies of graphene, spectrum within [-3, 3], E = +-3 at Gamma
    g = np.sort(np.linalg.eigvalsh(_dense(lattices.kane_mele(6, lso=0.0, lr=0.0, W=0.0))))
    assert abs(g[0] + 3.0) < 1e-12 and abs(g[-1] - 3.0) < 1e-12
    m = lattices.kane_mele(6, W=1.0)
    assert m.n == 4 * 36 and np.all(m.col >= np.repeat(np.arange(m.n), np.diff(m.rowptr)))
    assert 2 * m.nnz - m.n == 13 * m.n  # 1 on-site + 3 hops + 3 Rashba flips + 6 second neighbours


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "linqt_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(lq_\w+)\(", txt, re.M)))


def test_c_abi_library_exports_every_declared_symbol():
    from linqt_amd import _lib

    L = _lib.load()  # loading runs no GPU code
    syms = _header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(_lib.exported_symbols()) == syms
    assert L.lq_abi_version() == 1


def test_product_has_no_cpu_fallback():
    """The package never imports the oracle."""
    pkg = os.path.join(ROOT, "linqt_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in txt and "lq_oracle" not in txt, f
