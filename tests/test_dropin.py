"""Drop-in tests: the reference's CLI contract (operators/ + BOUNDS + state
files in the working directory, same output names and formats), run through
the C++ facade and the C ABI, checked against the golden files / the oracle."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, rel_inf
from linqt_amd import io, lattices

BIN = os.path.join(ROOT, "linqt_amd", "bin")
LC = os.path.join(GOLDEN, "linear_chain")


def _run(exe, args, cwd, env=None):
    e = dict(os.environ)
    e.update(env or {})
    print("[case]", exe, *args, flush=True)
    r = subprocess.run([os.path.join(BIN, exe), *map(str, args)], cwd=cwd, capture_output=True, text=True,
                       timeout=300, env=e)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


def _two_col(path):
    a = np.loadtxt(path)
    return a[:, 0], a[:, 1]


def test_drivers_print_usage_without_gpu():
    """The binaries load and link (no GPU call happens on the usage path)."""
    for exe in ("inline_compute-kpm-spectralOp", "inline_compute-kpm-nonEqOp", "inline_compute-Kubo-kpm-FFT",
                "spectralFunctionFromChebmom_FFTgrid", "kuboGreenwoodFromChebmom_FFTgrid",
                "kuboBastinFromChebmom_FFTgrid", "spectralFunctionFromChebmom", "kuboGreenwoodFromChebmom",
                "kuboBastinFromChebmom", "kuboBastinIFromChebmom", "kuboBastinIIFromChebmom",
                "kuboBastinSeaFromChebmom", "kuboBastinSurfFromChebmom", "inline_compute-kpm-CorrOp",
                "timeCorrelationsFromChebmom"):
        r = subprocess.run([os.path.join(BIN, exe)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0 and "should be called" in r.stdout, exe


@pytest.fixture()
def chain_dir(tmp_path):
    os.makedirs(tmp_path / "operators")
    for f in ("linear_chain.HAM.CSR", "linear_chain.VX.CSR"):
        shutil.copy(os.path.join(LC, "operators", f), tmp_path / "operators" / f)
    return tmp_path


@pytest.mark.gpu
@pytest.mark.parametrize("M,bounds,golden", [
    (100, "-5 5", "NonLEqOpVX-VXlinear_chainKPM_M100x100.chebmom2D"),
    (300, "-2.45 2.45", "NonLEqOpVX-VXlinear_chainKPM_M300x300.chebmom2D"),
])
def test_nonEqOp_driver_reproduces_golden(chain_dir, M, bounds, golden):
    (chain_dir / "BOUNDS").write_text(bounds + "\n")
    _run("inline_compute-kpm-nonEqOp", ["linear_chain", "VX", "VX", M], chain_dir)
    out = chain_dir / f"NonEqOpVX-VXlinear_chainKPM_M{M}x{M}_statedefault.chebmom2D"
    mu, n, W, Ec = io.read_chebmom2d(str(out))
    gold, gn, gW, gEc = io.read_chebmom2d(os.path.join(LC, golden))
    assert (n, W, Ec) == (gn, gW, gEc)
    assert rel_inf(mu, gold) < 1e-12
    # the FFT-grid post-processors on the produced file
    _run("kuboGreenwoodFromChebmom_FFTgrid", [out.name, 50], chain_dir)
    _run("kuboBastinFromChebmom_FFTgrid", [out.name, 50], chain_dir)


@pytest.mark.gpu
def test_post_processor_drivers_match_oracle(chain_dir, oracle):
    g = os.path.join(LC, "NonLEqOpVX-VXlinear_chainKPM_M100x100.chebmom2D")
    shutil.copy(g, chain_dir / "g.chebmom2D")
    _run("kuboGreenwoodFromChebmom_FFTgrid", ["g.chebmom2D", 80], chain_dir)
    _run("kuboBastinFromChebmom_FFTgrid", ["g.chebmom2D", 80], chain_dir)
    mu, n, W, Ec = io.read_chebmom2d(g)
    muJ = oracle.jackson_2d(mu, W, 80.0, 80.0)
    nd = max(muJ.shape)
    for name, fn in (("KuboGreenWood_gJACKSON_FFTgrid.dat", oracle.greenwood_fftgrid),
                     ("KuboBastin_gJACKSON_FFTgrid.dat", oracle.bastin_fftgrid)):
        E, s = _two_col(chain_dir / name)
        Eo, so = fn(muJ, n, W, Ec, nd)
        assert np.allclose(E, Eo, rtol=1e-5, atol=1e-9)
        assert rel_inf(s, so) < 1e-5  # files hold 6 significant digits


@pytest.mark.gpu
def test_spectralOp_driver_with_state_file(tmp_path, oracle):
    m = lattices.square_anderson(16, W=1.0, seed=11)
    os.makedirs(tmp_path / "operators")
    io.write_csr(str(tmp_path / "operators" / "sq.HAM.CSR"), m)
    (tmp_path / "BOUNDS").write_text("-4.6 4.6\n")
    phi = lattices.random_phase_states(m.n, 3, seed=12)
    io.write_vector_state_file(str(tmp_path / "phases"), phi)
    _run("inline_compute-kpm-spectralOp", ["sq", "1", 64, "phases"], tmp_path)
    mu, n, W, Ec = io.read_chebmom1d(str(tmp_path / "SpectralOp1sqKPM_M64_statephases.chebmom1D"))
    H = io.read_csr(str(tmp_path / "operators" / "sq.HAM.CSR"))
    Ho = oracle.Op.from_csr(H).rescale(1 / 4.6, 0.0)
    phi_file = io.read_state_file(str(tmp_path / "phases"), m.n)
    ref = oracle.spectral_moments(Ho, None, phi_file, 64)
    assert (n, W, Ec) == (m.n, 9.2, 0.0) and rel_inf(mu, ref) < 1e-12
    _run("spectralFunctionFromChebmom_FFTgrid", ["SpectralOp1sqKPM_M64_statephases.chebmom1D", 200], tmp_path)
    E, rho = _two_col(tmp_path / "meanSpectralOp1sqKPM_M64_statephasesJACKSON_FFTgrid.dat")
    muJ = oracle.jackson_1d(mu, W, 200.0)
    Eo, ro = oracle.dos_fftgrid(muJ, n, W, Ec, len(muJ))
    assert np.allclose(E, Eo, rtol=1e-5, atol=1e-9) and rel_inf(rho, ro) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("formula", ["greenwood", "bastin"])
def test_kubo_fft_driver_matches_oracle(chain_dir, oracle, formula):
    (chain_dir / "BOUNDS").write_text("-5 5\n")
    M = 48
    env = {"LQ_KUBO_FORMULA": formula} if formula == "bastin" else {}
    _run("inline_compute-Kubo-kpm-FFT", ["linear_chain", "VX", "VX", M], chain_dir, env)
    prefix = "Bastin_FFT" if formula == "bastin" else "Greenwood_FFT"
    E, s = _two_col(chain_dir / f"{prefix}VX-VXlinear_chainKPM_M{M}x{M}_statedefault.conductivity")
    H = io.read_csr(os.path.join(LC, "operators", "linear_chain.HAM.CSR"))
    V = io.read_csr(os.path.join(LC, "operators", "linear_chain.VX.CSR"))
    Ho, Vo = oracle.Op.from_csr(H).rescale(0.2, 0.0), oracle.Op.from_csr(V)
    fid = 1 if formula == "bastin" else 0
    fd, _ = oracle.kubo_fft(Ho, Vo, Vo, io.default_state(H.n), M, M, fid)
    Eo, so, _ = oracle.kubo_post(fid, fd, H.n, 10.0, 0.0)
    assert np.allclose(E, Eo, rtol=1e-5, atol=1e-9)
    assert rel_inf(s, so) < 1e-5


def _rows_with_nan(path):
    rows = [line.split() for line in open(path) if line.strip()]
    return np.array([float(r[0]) for r in rows]), np.array([float(r[1]) for r in rows])


@pytest.mark.gpu
def test_linear_grid_drivers_match_oracle(chain_dir, oracle):
    """The seven linear-grid programs on the golden linear-chain moments: same file names,
    30*M rows on the tools' grids, values to the 6 digits of the format (NaN where the
    reference's formulas divide by zero at x = +-1)."""
    g = os.path.join(LC, "NonLEqOpVX-VXlinear_chainKPM_M100x100.chebmom2D")
    shutil.copy(g, chain_dir / "g.chebmom2D")
    mu, n, W, Ec = io.read_chebmom2d(g)
    muJ = oracle.jackson_2d(mu, W, 80.0, 80.0)
    nd = 30 * max(muJ.shape)
    for exe, prefix, kind in (("kuboGreenwoodFromChebmom", "KuboGreenWood_", 1),
                              ("kuboBastinFromChebmom", "KuboBastin_", 2),
                              ("kuboBastinIFromChebmom", "KuboBastinI_", 3),
                              ("kuboBastinIIFromChebmom", "KuboBastinII_", 4),
                              ("kuboBastinSeaFromChebmom", "KuboBastinSea_", 5),
                              ("kuboBastinSurfFromChebmom", "KuboBastinSurf_", 6)):
        _run(exe, ["g.chebmom2D", 80], chain_dir)
        E, v = _rows_with_nan(chain_dir / f"{prefix}gJACKSON.dat")
        Eo, vo = oracle.linear_grid(kind, muJ, n, W, Ec, nd)
        assert len(E) == len(Eo) == nd - 1, exe
        assert np.allclose(E, Eo, rtol=1e-5, atol=1e-9), exe
        assert np.array_equal(np.isfinite(v), np.isfinite(vo)), exe
        f = np.isfinite(vo)
        if f.any():
            assert np.abs(v[f] - vo[f]).max() <= 2e-5 * np.abs(vo[f]).max(), exe
    # the 1D tool on a 1D moment file made from the table's first column
    io.write_chebmom1d(str(chain_dir / "s.chebmom1D"), mu[:, 0], n, W, Ec)
    _run("spectralFunctionFromChebmom", ["s.chebmom1D", 80], chain_dir)
    E, v = _rows_with_nan(chain_dir / "meansJACKSON.dat")
    m1 = oracle.jackson_1d(io.read_chebmom1d(str(chain_dir / "s.chebmom1D"))[0], W, 80.0)
    Eo, vo = oracle.linear_grid(0, m1, n, W, Ec, 30 * len(m1))
    assert len(E) == len(Eo) and np.array_equal(np.isfinite(v), np.isfinite(vo))
    f = np.isfinite(vo)
    assert np.abs(v[f] - vo[f]).max() <= 2e-5 * np.abs(vo[f]).max()


@pytest.mark.gpu
def test_time_correlation_drivers_match_oracle(chain_dir, oracle):
    """inline_compute-kpm-CorrOp (default state: site N/2) and timeCorrelationsFromChebmom on
    the reference's linear chain."""
    (chain_dir / "BOUNDS").write_text("-5 5\n")
    M, nT, tmax = 40, 5, 2.0
    _run("inline_compute-kpm-CorrOp", ["linear_chain", "VX", "VX", M, nT, tmax], chain_dir)
    out = chain_dir / f"TimeCorrVX-VXlinear_chainKPM_M{M}_statedefault.chebmomTD"
    mu, n, W, Ec, dt = io.read_chebmomtd(str(out))
    H = io.read_csr(os.path.join(LC, "operators", "linear_chain.HAM.CSR"))
    V = io.read_csr(os.path.join(LC, "operators", "linear_chain.VX.CSR"))
    Ho = oracle.Op.from_csr(H).rescale(1 / 5.0, 0.0)
    Vo = oracle.Op.from_csr(V)
    phi = io.default_state(H.n)
    ref = oracle.time_correlations(Ho, Vo, Vo, phi, M, nT, tmax / (nT - 1), 10.0, 0.0)
    assert (n, W, Ec) == (H.n, 10.0, 0.0) and abs(dt - tmax / (nT - 1)) < 1e-14
    assert rel_inf(mu, ref) < 1e-12
    _run("timeCorrelationsFromChebmom", [out.name, 100, 0.3], chain_dir)
    rows = sorted(chain_dir.glob("meanTimeCorr*EF0.300000JACKSON.dat"))
    assert len(rows) == 1
    t, v = _two_col(rows[0])
    to, vo = oracle.timecorr_post(oracle.jackson_td(mu, W, 100.0), n, W, Ec, dt, 0.3)
    assert np.allclose(t, to, rtol=1e-5) and rel_inf(v, vo) < 1e-5


@pytest.mark.gpu
def test_nonorth_driver_matches_oracle(tmp_path, oracle):
    """inline_compute-kpm-spectralOp-nonOrth: operators/LABEL.{HAM,S}.CSR and
    BOUNDS in, SpectralOp-nonOrth-... / Spectral-nonOrth-OP1... out; the
    overlap solves run at the reference's tolerance (1e-4), so the moments
    match the oracle's tight-tolerance ones to that accuracy."""
    os.makedirs(tmp_path / "operators")
    m = lattices.square_anderson(16, W=1.0, seed=3)
    n = m.n
    io.write_csr(str(tmp_path / "operators" / "sq.HAM.CSR"), m)
    rows = np.repeat(np.arange(n), np.diff(m.rowptr))
    sval = np.where(m.col == rows, 1.0 + 0j, 0.03 * np.sign(m.val.real) + 0j)
    S = io.CSR(n, m.rowptr, m.col, sval)
    io.write_csr(str(tmp_path / "operators" / "sq.S.CSR"), S)
    (tmp_path / "BOUNDS").write_text("-4.6 4.6\n")
    _run("inline_compute-kpm-spectralOp-nonOrth", ["sq", "1", 40], tmp_path)
    out = tmp_path / "Spectral-nonOrth-OP1sqKPM_M40_statedefault.chebmom1D-nonOrth"
    assert out.exists(), sorted(os.listdir(tmp_path))
    t = np.array(out.read_text().split(), dtype=float)
    assert int(t[0]) == n and int(t[3]) == 40
    mu = t[4:4 + 80].reshape(40, 2) @ np.array([1, 1j])
    H = oracle.Op.from_csr(io.read_csr(str(tmp_path / "operators" / "sq.HAM.CSR"))).rescale(1 / 4.6, 0.0)
    phi = io.default_state(n)
    ref = oracle.spectral_moments_nonorth(H, None, io.read_csr(str(tmp_path / "operators" / "sq.S.CSR")),
                                          phi, 40, tol=1e-14)
    assert np.abs(mu - ref).max() <= 1e-3 * np.abs(ref).max()
