"""Pins the CPU oracle: golden vectors, closed forms and identities.

Every oracle function used as a GPU checker is pinned here by something that
is not a retyping of its own formula:
  * the reference's golden .chebmom2D files (Hermitian-upper SpMV, rescale,
    recursion, zdotc convention, 4/R scaling, symmetrisation);
  * dense eigendecomposition closed forms T_m(H) = U cos(m arccos E) U^H;
  * numpy's FFT for the DFT;
  * the Jackson kernel's closed-form values and positivity;
  * the Greenwood/Bastin identities between the per-site DFT route and the
    2D-moment formulas (the Bastin one is the published Gamma-kernel formula).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, dense_herm, random_upper_csr, rel_inf
from linqt_amd import io, lattices

LC = os.path.join(GOLDEN, "linear_chain")


def _load_linear_chain(oracle, W):
    H = io.read_csr(os.path.join(LC, "operators", "linear_chain.HAM.CSR"))
    V = io.read_csr(os.path.join(LC, "operators", "linear_chain.VX.CSR"))
    Ho = oracle.Op.from_csr(H).rescale(2.0 / W, 0.0)
    Vo = oracle.Op.from_csr(V)
    return H, V, Ho, Vo


@pytest.mark.parametrize("M,W,name", [
    (100, 10.0, "NonLEqOpVX-VXlinear_chainKPM_M100x100.chebmom2D"),
    (300, 4.9, "NonLEqOpVX-VXlinear_chainKPM_M300x300.chebmom2D"),
])
def test_oracle_reproduces_golden_chebmom2d(oracle, M, W, name):
    gold, n, Wf, Ec = io.read_chebmom2d(os.path.join(LC, name))
    assert (n, Wf, Ec, gold.shape) == (20000, W, 0.0, (M, M))
    H, V, Ho, Vo = _load_linear_chain(oracle, W)
    mu = oracle.correlation_moments(Ho, Vo, Vo, io.default_state(H.n), M, M)
    # the files hold 15 significant digits
    assert rel_inf(mu, gold) < 1e-12


def test_dft_matches_numpy(oracle):
    rng = np.random.default_rng(1)
    for n in (1, 2, 8, 64, 12, 30):
        x = rng.normal(size=n) + 1j * rng.normal(size=n)
        ref = np.fft.ifft(x) * n  # e^{+2 pi i mk/n}, unnormalised
        assert np.allclose(oracle.dft_backward(x), ref, rtol=0, atol=1e-12 * max(1, np.abs(ref).max()))


def test_herm_upper_mv_and_rescale_match_dense(oracle):
    n = 37
    rowptr, col, val = random_upper_csr(n, 0.2, seed=3)
    # a stray lower entry must be ignored, as MKL's upper descriptor does
    op = oracle.Op(n, rowptr, col, val)
    D = dense_herm(n, rowptr, col, val)
    rng = np.random.default_rng(4)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    y = rng.normal(size=n) + 1j * rng.normal(size=n)
    a, b = 0.7 - 0.2j, -1.3 + 0.5j
    assert np.allclose(op.mv(a, x, b, y), a * D @ x + b * y, atol=1e-12)
    assert np.allclose(op.dense(), D)
    op.rescale(0.25, -0.5)
    assert np.allclose(op.dense(), 0.25 * D - 0.5 * np.eye(n), atol=1e-14)


def test_mv_ignores_strictly_lower_entries(oracle):
    # row 1 stores (1,0), which an upper-fill Hermitian descriptor never reads
    rowptr = np.array([0, 2, 4, 5], dtype=np.int32)
    col = np.array([0, 1, 0, 1, 2], dtype=np.int32)
    val = np.array([1.0, 2 + 1j, 99.0, 3.0, 4.0], dtype=np.complex128)
    op = oracle.Op(3, rowptr, col, val)
    D = np.array([[1, 2 + 1j, 0], [2 - 1j, 3, 0], [0, 0, 4]], dtype=np.complex128)
    assert np.allclose(op.dense(), D)


def _cheb_dense(Ht, M):
    E, U = np.linalg.eigh(Ht)
    th = np.arccos(np.clip(E, -1, 1))
    return [U @ np.diag(np.cos(m * th)) @ U.conj().T for m in range(M)]


def test_spectral_moments_match_eigendecomposition(oracle):
    n, M, R = 60, 40, 3
    rowptr, col, val = random_upper_csr(n, 0.1, seed=5)
    D = dense_herm(n, rowptr, col, val)
    Wb = 2.2 * np.abs(np.linalg.eigvalsh(D)).max()
    H = oracle.Op(n, rowptr, col, val).rescale(2.0 / Wb, 0.05)
    Ht = 2.0 / Wb * D + 0.05 * np.eye(n)
    phi = lattices.random_phase_states(n, R, seed=9)
    T = _cheb_dense(Ht, M)
    ref = np.array([sum((2.0 / R) * (0.5 if m == 0 else 1.0) * np.vdot(p, T[m] @ p) for p in phi)
                    for m in range(M)])
    assert rel_inf(oracle.spectral_moments(H, None, phi, M), ref) < 1e-12
    # with an operator: mu_m = <phi| T_m(H) OP |phi>
    r2, c2, v2 = random_upper_csr(n, 0.1, seed=6)
    OP = oracle.Op(n, r2, c2, v2)
    Dop = dense_herm(n, r2, c2, v2)
    ref2 = np.array([sum((2.0 / R) * (0.5 if m == 0 else 1.0) * np.vdot(p, T[m] @ Dop @ p) for p in phi)
                     for m in range(M)])
    assert rel_inf(oracle.spectral_moments(H, OP, phi, M), ref2) < 1e-12


def _raw_dense(Ht, OPL, OPR, phi, M):
    T = _cheb_dense(Ht, M)
    raw = np.zeros((M, M), dtype=np.complex128)
    for p in phi:
        Lv = np.array([T[m] @ (OPL @ p) for m in range(M)])
        Rv = np.array([OPR @ (T[m] @ p) for m in range(M)])
        raw += Lv.conj() @ Rv.T
    return raw


def _scale_sym(raw, R):
    P = raw.shape[0]
    s = np.full(P, 1.0)
    s[0] = 0.5
    mu = (4.0 / R) * np.outer(s, s) * (raw + raw.conj().T) / 2.0
    return mu


def test_correlation_moments_batch_semantics(oracle):
    """BATCH_SIZE < M: the reference's blocks restart one order early.

    The blocked loop of the oracle (a restatement of IterateAll) must equal the
    closed-form order map ord(p) = (p // B)(B - 1) + p % B applied to the
    eigendecomposition table -- the form the GPU path uses."""
    n, M, R = 40, 12, 2
    rowptr, col, val = random_upper_csr(n, 0.15, seed=7)
    D = dense_herm(n, rowptr, col, val)
    Wb = 2.2 * np.abs(np.linalg.eigvalsh(D)).max()
    H = oracle.Op(n, rowptr, col, val).rescale(2.0 / Wb, 0.0)
    Ht = 2.0 / Wb * D
    rl, cl, vl = random_upper_csr(n, 0.2, seed=8, diag=False)
    rr, cr, vr = random_upper_csr(n, 0.2, seed=9, diag=False)
    OPL, OPR = oracle.Op(n, rl, cl, vl), oracle.Op(n, rr, cr, vr)
    DL, DR = dense_herm(n, rl, cl, vl), dense_herm(n, rr, cr, vr)
    phi = lattices.random_phase_states(n, R, seed=10)
    raw_true = _raw_dense(Ht, DL, DR, phi, 2 * M)
    for B in (0, 12, 4, 5, 7):
        Bf = B if B > 0 else M
        P = ((M + Bf - 1) // Bf) * Bf
        ordp = np.array([(p // Bf) * (Bf - 1) + p % Bf for p in range(P)])
        ref = _scale_sym(raw_true[np.ix_(ordp, ordp)], R)[:M, :M]
        got = oracle.correlation_moments(H, OPL, OPR, phi, M, M, batch=B)
        assert rel_inf(got, ref) < 1e-12, B


def test_jackson_closed_form_and_truncation(oracle):
    M, W = 64, 2.0
    mu = np.ones(M, dtype=np.complex128)
    g = oracle.jackson_1d(mu, W, broad=1e-3)  # tiny broadening: no truncation
    assert len(g) == M
    assert abs(g[0] - 1.0) < 1e-15
    # g_m of the Jackson kernel with M moments: the next one, g_M, vanishes
    phi = np.pi / (M + 1)
    gM = ((M - M + 1) * np.cos(phi * M) + np.sin(phi * M) / np.tan(phi)) / (M + 1)
    assert abs(gM) < 1e-13
    # positivity: a Jackson-damped delta expansion is non-negative everywhere
    x0 = 0.3
    th0 = np.arccos(x0)
    coef = np.array([np.cos(m * th0) for m in range(M)], dtype=np.complex128)
    gd = oracle.jackson_1d(coef, W, broad=1e-3)
    xs = np.linspace(-0.999, 0.999, 2001)
    f = (gd[0].real + 2 * sum(gd[m].real * np.cos(m * np.arccos(xs)) for m in range(1, M)))
    assert f.min() > -1e-12
    # truncation count ceil(pi / eta), eta = 2 b / 1000 / W
    assert len(oracle.jackson_1d(np.ones(1000, dtype=np.complex128), 10.0, 50.0)) == int(np.ceil(np.pi / 0.01))


def test_jackson_2d_is_separable_product(oracle):
    rng = np.random.default_rng(11)
    mu = rng.normal(size=(30, 30)) + 1j * rng.normal(size=(30, 30))
    out = oracle.jackson_2d(mu, 4.0, 1e-3, 1e-3)
    g = oracle.jackson_1d(np.ones(30, dtype=np.complex128), 4.0, 1e-3).real
    assert np.allclose(out, mu * np.outer(g, g), rtol=1e-14, atol=0)
    t = oracle.jackson_2d(mu, 4.0, 200.0, 400.0)
    # ceil(pi / eta) per axis, capped at the stored count
    assert t.shape == (min(30, int(np.ceil(np.pi / 0.1))), int(np.ceil(np.pi / 0.2)))


def _kubo_setup(oracle, n=48, M=16, R=2, seed=12):
    rowptr, col, val = random_upper_csr(n, 0.12, seed=seed)
    D = dense_herm(n, rowptr, col, val)
    W = 2.3 * np.abs(np.linalg.eigvalsh(D)).max()
    H = oracle.Op(n, rowptr, col, val).rescale(2.0 / W, 0.0)
    Ht = 2.0 / W * D
    # V = i[H, X] with random site coordinates
    X = np.diag(np.random.default_rng(seed + 1).normal(size=n))
    V = 1j * (D @ X - X @ D)
    iu = np.triu_indices(n)
    mask = np.abs(V[iu]) > 0
    rows, cols, vals = iu[0][mask], iu[1][mask], V[iu][mask]
    rp = np.zeros(n + 1, dtype=np.int32)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    Vo = oracle.Op(n, rp, cols.astype(np.int32), vals)
    phi = lattices.random_phase_states(n, R, seed=seed + 2)
    return n, M, R, W, H, Ht, Vo, V, phi


def test_kubo_greenwood_identity(oracle):
    """Greenwood_FFTs + postProcess == pi N / a^2 sum g_m g_n delta_m delta_n Re mu_mn."""
    n, M, R, W, H, Ht, Vo, V, phi = _kubo_setup(oracle)
    nump = 64
    fd, _ = oracle.kubo_fft(H, Vo, Vo, phi, M, nump, 0)
    E, sig, _ = oracle.kubo_post(0, fd, n, W, 0.0)
    mu = oracle.correlation_moments(H, Vo, Vo, phi, M, M)
    a = W / 2
    g = np.array([oracle.kubo_jackson(m, M) for m in range(M)])
    x = E / a  # Ec = 0
    th = np.arccos(x)
    delta = np.array([np.cos(m * th) for m in range(M)]) / (np.pi * np.sqrt(1 - x * x))
    ref = np.pi * n / a ** 2 * np.einsum("mk,nk,mn->k", g[:, None] * delta, g[:, None] * delta, mu.real)
    assert rel_inf(sig, ref) < 1e-10


def test_kubo_bastin_matches_gamma_kernel(oracle):
    """Bastin_FFTs + postProcess == (N/(pi a^2)) Re sum c_m c_n Gamma_mn(E) raw_mn / R / (1-E^2)^2
    (Garcia, Covaci, Rappoport, PRL 114, 116602 (2015))."""
    n, M, R, W, H, Ht, Vo, V, phi = _kubo_setup(oracle)
    nump = 64
    fd, _ = oracle.kubo_fft(H, Vo, Vo, phi, M, nump, 1)
    E, sig, integ = oracle.kubo_post(1, fd, n, W, 0.0)
    raw = _raw_dense(Ht, V, V, phi, M) / R
    a = W / 2
    c = np.array([(2 - (m == 0)) * oracle.kubo_jackson(m, M) for m in range(M)])
    x = E / a
    s = np.sqrt(1 - x * x)
    th = np.arccos(x)
    m = np.arange(M)
    T = np.cos(np.outer(m, th))
    ex = np.exp(1j * np.outer(m, th))
    # Gamma_mn = T_m (E - i n s) e^{i n th} + (E + i m s) e^{-i m th} T_n
    G = (T[:, None, :] * ((x[None, None, :] - 1j * m[None, :, None] * s) * ex[None, :, :])
         + ((x[None, None, :] + 1j * m[:, None, None] * s) * ex.conj()[:, None, :]) * T[None, :, :])
    ref = (n / (np.pi * a * a)) * np.einsum("m,n,mnk,mn->k", c, c, G, raw).real / (1 - x * x) ** 2
    assert rel_inf(integ, ref) < 1e-9
    # sigma is the left-Riemann cumulative integral of the integrand on the ascending grid
    xs = E / a
    cum = np.concatenate([np.cumsum(integ[:-1] * np.diff(xs)), [0.0]])
    assert np.allclose(sig, cum, rtol=1e-12, atol=1e-12 * np.abs(cum).max())


def test_kubo_sections_equal_masked_operator(oracle):
    """num_sections > 1 (N % ns == 0): MultiplySliced applies OPR to the
    section-restricted vector, i.e. OPR's block-diagonal part."""
    n, M, R, W, H, Ht, Vo, V, phi = _kubo_setup(oracle, n=48, M=12, R=2)
    ns = 3
    fd3, _ = oracle.kubo_fft(H, Vo, Vo, phi, M, 32, 1, num_sections=ns)
    ss = n // ns
    sec = np.minimum(np.arange(n) // ss, ns - 1)
    Vm = np.where(sec[:, None] == sec[None, :], V, 0)
    iu = np.triu_indices(n)
    mask = np.abs(Vm[iu]) > 0
    rp = np.zeros(n + 1, dtype=np.int32)
    np.add.at(rp, iu[0][mask] + 1, 1)
    Vmo = oracle.Op(n, np.cumsum(rp).astype(np.int32), iu[1][mask].astype(np.int32), Vm[iu][mask])
    fdm, _ = oracle.kubo_fft(H, Vo, Vmo, phi, M, 32, 1, num_sections=1)
    assert rel_inf(fd3, fdm) < 1e-12


def test_dos_fftgrid_quadrature_gives_mu0(oracle):
    """Gauss-Chebyshev quadrature of the FFT-grid DOS returns N * mu_0."""
    L = 24
    H = lattices.square_anderson(L, W=1.0, seed=3)
    Ho = oracle.Op.from_csr(H).rescale(1 / 4.6, 0.0)
    phi = lattices.random_phase_states(H.n, 2, seed=4)
    M = 128
    mu = oracle.jackson_1d(oracle.spectral_moments(Ho, None, phi, M), 9.2, 1e-3)
    K = 256
    E, rho = oracle.dos_fftgrid(mu, H.n, 9.2, 0.0, K)
    x = E / 4.6
    # rho = (N/a) sum_m delta_m mu_m, and (pi/K) sum_k sqrt(1-x_k^2) delta_m(x_k) = delta_{m0}
    integral = (np.pi / K) * np.sum(np.sqrt(1 - x * x) * rho) * 4.6
    assert abs(integral / H.n - mu[0].real) < 1e-10


def test_greenwood_fftgrid_matches_kubo_route_for_real_moments(oracle):
    """For a real moment table the G^R-kernel Greenwood equals pi N/a^2 sum g g delta delta mu."""
    W = 10.0
    H, V, Ho, Vo = _load_linear_chain(oracle, W)
    M = 40
    mu = oracle.correlation_moments(Ho, Vo, Vo, io.default_state(H.n), M, M)
    assert np.abs(mu.imag).max() < 1e-14 * np.abs(mu).max() + 1e-300
    muJ = oracle.jackson_2d(mu.real.astype(np.complex128), W, 1e-3, 1e-3)
    E, s = oracle.greenwood_fftgrid(muJ, H.n, W, 0.0, M)
    a = W / 2
    x = E / a
    th = np.arccos(x)
    delta = np.array([np.cos(m * th) for m in range(M)]) / (np.pi * np.sqrt(1 - x * x))
    ref = np.pi * H.n / a ** 2 * np.einsum("mk,nk,mn->k", delta, delta, muJ.real)
    assert rel_inf(s, ref) < 1e-10


def test_bastin_fftgrid_trig_form(oracle):
    """kuboBastinFromChebmom_FFTgrid in trigonometric form:
    DgreenR_m(x) = -i e^{-i m th} (x + i m s) / s^3 (the x-derivative of
    greenR_m = -i e^{-i m th}/s), on the ascending grid, cumulative sum."""
    rng = np.random.default_rng(13)
    M, n, W, Ec = 24, 500, 6.0, 0.3
    mu = rng.normal(size=(M, M)) + 1j * rng.normal(size=(M, M))
    E, out = oracle.bastin_fftgrid(mu, n, W, Ec, M)
    a = W / 2
    k = np.arange(M)
    e0 = np.cos(np.pi * (2 * k + 0.5) / M)
    x = np.sort(e0)
    s = np.sqrt(1 - x * x)
    th = np.arccos(x)
    m = np.arange(M)
    delta = np.cos(np.outer(m, th)) / (np.pi * s)
    dg = -1j * np.exp(-1j * np.outer(m, th)) * (x + 1j * m[:, None] * s) / s ** 3
    kern = -2.0 * n / a / a * np.einsum("ak,bk,ab->k", delta, dg, mu).imag
    ref = np.cumsum(kern[:-1] * np.diff(x))
    assert np.allclose(E, x[:-1] * a + Ec)
    assert rel_inf(out, ref) < 1e-9
    # and the derivative identity itself, by central differences
    h = 1e-6
    xx = 0.37
    g = lambda y, mm: -1j * (y - 1j * np.sqrt(1 - y * y)) ** mm / np.sqrt(1 - y * y)
    for mm in (0, 3, 7):
        num = (g(xx + h, mm) - g(xx - h, mm)) / (2 * h)
        ana = -1j * np.exp(-1j * mm * np.arccos(xx)) * (xx + 1j * mm * np.sqrt(1 - xx * xx)) / (1 - xx * xx) ** 1.5
        assert abs(num - ana) < 1e-6 * max(1, abs(ana))


# ---- linear-grid tools (src/*FromChebmom.cpp) --------------------------------------


def _unit(M0, M1, m0, m1, v):
    mu = np.zeros((M0, M1), dtype=np.complex128)
    mu[m0, m1] = v
    return mu


def test_linear_coefficients_match_closed_forms(oracle):
    """delta_m = T_m(x)/(pi sqrt(1-x^2)) (scipy's Chebyshev T), gR_m = -i e^{-im theta}/sin theta and
    DgR_m = -i e^{-im theta}(cos theta + i m sin theta)/sin^3 theta (x = cos theta), read back through
    single-moment tables: kind 0 gives delta, kinds 1 and 2 give Im and Re of gR / DgR times delta_0."""
    from scipy.special import eval_chebyt

    x = np.linspace(-0.97, 0.95, 37)
    th, f0 = np.arccos(x), np.sqrt(1 - x * x)
    n, W = 7, 4.0
    a = W / 2
    for m in (0, 1, 5, 17):
        k = oracle.linear_kernel(0, _unit(m + 1, 1, m, 0, 1.0)[:, 0], n, W, x)
        ref = eval_chebyt(m, x) / (np.pi * f0) * n / a
        assert np.allclose(k, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
        d0 = 1 / (np.pi * f0)
        gR = -1j * np.exp(-1j * m * th) / f0
        DgR = -1j * np.exp(-1j * m * th) * (x + 1j * m * f0) / f0 ** 3
        for kind, g, s in ((1, gR, -n / a / a), (2, DgR, -2 * n / a / a)):
            for v in (1.0, 1j):  # Im(g v): v = 1 -> Im g, v = i -> Re g
                k = oracle.linear_kernel(kind, _unit(1, m + 1, 0, m, v), n, W, x)
                ref = s * d0 * np.imag(g * v)
                assert np.allclose(k, ref, rtol=1e-10, atol=1e-12 * np.abs(s * d0 * g).max())


def _jackson_table(oracle, M0=40, M1=32, seed=2):
    rng = np.random.default_rng(seed)
    mu = rng.normal(size=(M0, M1)) + 1j * rng.normal(size=(M0, M1))
    return mu


def test_linear_greenwood_and_bastin_match_fftgrid_forms(oracle):
    """The linear tools' Greenwood and Bastin kernels at the FFT-grid tools' energies equal
    those (independently pinned) tools; the spectral kernel's magnitude equals the DOS grid."""
    mu = _jackson_table(oracle)
    n, W, Ec, nd = 50, 6.0, 0.2, 64
    a = W / 2
    xg = np.cos(np.pi * (2 * np.arange(nd) + 0.5) / nd)[: nd - 1]
    Eg, sg = oracle.greenwood_fftgrid(mu, n, W, Ec, nd)
    assert rel_inf(oracle.linear_kernel(1, mu, n, W, xg), sg) < 1e-12
    Eb, sb = oracle.bastin_fftgrid(mu, n, W, Ec, nd)
    xb = (Eb - Ec) / a
    inc = np.diff(np.concatenate([[0.0], sb]))           # k_i (x_{i+1} - x_i)
    xs = np.sort(np.cos(np.pi * (2 * np.arange(nd) + 0.5) / nd))
    kb = oracle.linear_kernel(2, mu, n, W, xb)
    assert rel_inf(kb * np.diff(xs), inc) < 1e-10
    mu1 = mu[:, 0]
    Ed, rd = oracle.dos_fftgrid(mu1, n, W, Ec, nd)
    xd = (Ed - Ec) / a
    assert rel_inf(np.abs(oracle.linear_kernel(0, mu1, n, W, xd)), rd) < 1e-12


def test_linear_surf_sea_bastinII_identities(oracle):
    """Im gR_m = -pi delta_m gives: Surf = Greenwood for a real table; Sea = Bastin for an
    imaginary table; the BastinII integrand is antisymmetric in (m0, m1), so 0 for a symmetric table."""
    rng = np.random.default_rng(4)
    M = 24
    x = np.linspace(-0.9, 0.93, 41)
    n, W = 11, 5.0
    R = rng.normal(size=(M, M))
    assert rel_inf(oracle.linear_kernel(6, R + 0j, n, W, x), oracle.linear_kernel(1, R + 0j, n, W, x)) < 1e-11
    assert rel_inf(oracle.linear_kernel(5, 1j * R, n, W, x), oracle.linear_kernel(2, 1j * R, n, W, x)) < 1e-11
    S = R + R.T + 1j * (R - R.T) * 0
    k4 = oracle.linear_kernel(4, S + 0j, n, W, x)
    scale = np.abs(oracle.linear_kernel(2, S + 0j, n, W, x)).max()
    assert np.abs(k4).max() < 1e-10 * scale
    assert np.abs(oracle.linear_kernel(4, R + 0j, n, W, x)).max() > 1e-3 * scale


def test_linear_grid_rows_and_end_points(oracle):
    """Grids and printed columns: 30 M energies, xb = 0.99 for Bastin; NaN at x = +-1 (the
    sources divide by sqrt(1 - x^2) there) and NaN-filled running sums for BastinII/Sea."""
    mu = _jackson_table(oracle, 10, 8)
    n, W, Ec = 20, 4.0, 0.1
    nd = 300
    for kind in range(1, 7):
        E, o = oracle.linear_grid(kind, mu, n, W, Ec, nd)
        assert len(E) == nd - 1
        xb = 0.99 if kind == 2 else 1.0
        assert np.allclose(E, (-xb + np.arange(nd - 1) * 2 * xb / (nd - 1)) * W / 2 + Ec, rtol=0, atol=1e-14)
        if kind == 2:
            assert np.all(np.isfinite(o))
            k = oracle.linear_kernel(2, mu, n, W, (E - Ec) / (W / 2))
            assert rel_inf(np.diff(np.concatenate([[0], o])), k * (2 * xb / (nd - 1))) < 1e-9
        elif kind in (4, 5):
            assert np.all(np.isnan(o))
        else:
            assert np.isnan(o[0]) and np.all(np.isfinite(o[1:-1]))
    E, o = oracle.linear_grid(0, mu[:, 0], n, W, Ec, nd)
    assert len(E) == nd and np.isnan(o[0]) and np.isnan(o[-1]) and np.all(np.isfinite(o[1:-1]))


# ---- time-dependent correlations -------------------------------------------------------

HBAR = 0.6582119624


def test_time_correlations_match_eigendecomposition(oracle):
    """mu(m, n) = sum_r (2/R)(1/2)^{d_m0} (OPL phi)^H e^{iHt_n} T_m(H~) OPR e^{-iHt_n} phi with
    the exact propagator from eigh (the Bessel-Chebyshev evolution truncates at |J_n| <= 1e-10);
    a band centre Ec != 0 checks the phase factor, OPL = OPR = None the identities."""
    n = 40
    rowptr, col, val = random_upper_csr(n, 0.2, seed=8)
    Hd = dense_herm(n, rowptr, col, val)
    e = np.linalg.eigvalsh(Hd)
    Ec, a = 0.37, (e.max() - e.min()) / 2 * 1.05 + 0.1
    W = 2 * a
    H = oracle.Op(n, rowptr, col, val).rescale(1 / a, -Ec / a)
    r2, c2, v2 = random_upper_csr(n, 0.15, seed=9)
    r3, c3, v3 = random_upper_csr(n, 0.15, seed=10)
    OL, OR = oracle.Op(n, r2, c2, v2), oracle.Op(n, r3, c3, v3)
    Ld, Rd = dense_herm(n, r2, c2, v2), dense_herm(n, r3, c3, v3)
    rng = np.random.default_rng(3)
    phi = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    M, nT, dt = 12, 5, 1.7
    E, U = np.linalg.eigh(Hd)
    for OPL, OPR, Ldn, Rdn in ((OL, OR, Ld, Rd), (None, None, np.eye(n), np.eye(n))):
        got = oracle.time_correlations(H, OPL, OPR, phi, M, nT, dt, W, Ec)
        ref = np.zeros((M, nT), dtype=np.complex128)
        for t in range(nT):
            Ut = U @ np.diag(np.exp(-1j * E * t * dt / HBAR)) @ U.conj().T
            for m in range(M):
                Tm = U @ np.diag(np.cos(m * np.arccos((E - Ec) / a))) @ U.conj().T
                s = 2.0 / 2 * (0.5 if m == 0 else 1.0)
                for r in range(2):
                    pl = Ut @ (Ldn @ phi[r])
                    pr = Ut @ phi[r]
                    ref[m, t] += s * np.vdot(pl, Tm @ (Rdn @ pr))
        assert rel_inf(got, ref) < 1e-8


def test_jackson_td_and_time_post(oracle):
    rng = np.random.default_rng(6)
    mu = rng.normal(size=(50, 7)) + 1j * rng.normal(size=(50, 7))
    W, Ec, n, dt = 6.0, 0.2, 300, 0.5
    td = oracle.jackson_td(mu, W, 120.0)
    for t in range(7):
        assert rel_inf(td[:, t], oracle.jackson_1d(mu[:, t].copy(), W, 120.0)) < 1e-15
    EF = 0.9
    tt, o = oracle.timecorr_post(td, n, W, Ec, dt, EF)
    assert np.allclose(tt, dt * np.arange(7))
    x = np.array([(EF - Ec) / (W / 2)])
    for t in range(7):
        assert abs(o[t] - oracle.linear_kernel(0, td[:, t].copy(), n, W, x)[0]) < 1e-12 * abs(o[t]) + 1e-14


# ---- supercell operators (wannier2sparse wrap + save) -------------------------
def _hermitian_hoppings(rng, W, nh, reach):
    cell = rng.integers(-reach, reach + 1, size=(nh, 3))
    edge = rng.integers(0, W, size=(nh, 2))
    val = rng.normal(size=nh) + 1j * rng.normal(size=nh)
    cell = np.concatenate([cell, -cell])
    edge = np.concatenate([edge, edge[:, ::-1]])
    val = np.concatenate([val, val.conj()])
    return cell, edge, val


def test_supercell_matches_reference_wrap_fixture(oracle):
    """utilities/wannier2sparse/test/model_from_files/model_from_files.cpp:
    "WRAP IN SUPERCELL FINAL TEST" -- a two-orbital chain wrapped on a
    2 x 2 x 1 supercell must give exactly the 24 listed hoppings; the CSR
    file keeps those with to >= from."""
    hv = 1.0 + 1.0j
    cell = [(0, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, 0), (1, 0, 0), (0, 1, 0)]
    edge = [(0, 1), (0, 1), (0, 1), (1, 0), (1, 0), (1, 0)]
    val = [hv] * 3 + [np.conj(hv)] * 3
    listed = {(0, 1): hv, (0, 3): hv, (0, 5): hv, (1, 0): np.conj(hv), (1, 4): np.conj(hv), (1, 2): np.conj(hv),
              (2, 1): hv, (2, 7): hv, (2, 3): hv, (3, 2): np.conj(hv), (3, 6): np.conj(hv), (3, 0): np.conj(hv),
              (4, 5): hv, (4, 1): hv, (4, 7): hv, (5, 4): np.conj(hv), (5, 0): np.conj(hv), (5, 6): np.conj(hv),
              (6, 5): hv, (6, 7): hv, (6, 3): hv, (7, 6): np.conj(hv), (7, 2): np.conj(hv), (7, 4): np.conj(hv)}
    n, rowptr, col, v = oracle.supercell_csr(2, cell, edge, val, (2, 2, 1))
    assert n == 8
    got = {(i, int(col[q])): v[q] for i in range(n) for q in range(rowptr[i], rowptr[i + 1])}
    assert got == {k: x for k, x in listed.items() if k[1] >= k[0]}


@pytest.mark.parametrize("dims", [(3, 2, 1), (1, 1, 1), (2, 2, 2), (4, 1, 3), (5, 3, 1)])
def test_supercell_spectrum_is_bloch_union(oracle, dims):
    """The completed supercell operator of a Hermitian hopping list has the
    spectrum of the Bloch Hamiltonians H(k)_{e0,e1} = sum v e^{i k.c} over the
    supercell's k grid -- also when hoppings wrap onto each other (dims of 1
    and 2 with reach 2)."""
    rng = np.random.default_rng(sum(dims))
    W = 3
    cell, edge, val = _hermitian_hoppings(rng, W, 14, 2)
    n, rowptr, col, v = oracle.supercell_csr(W, cell, edge, val, dims)
    ev = np.linalg.eigvalsh(dense_herm(n, rowptr, col, v))
    ks = []
    for m in np.ndindex(*dims):
        k = 2 * np.pi * np.array(m) / np.array(dims)
        Hk = np.zeros((W, W), dtype=complex)
        for c, e, x in zip(cell, edge, val):
            Hk[e[0], e[1]] += x * np.exp(1j * k @ c)
        ks.append(np.linalg.eigvalsh(Hk))
    assert np.allclose(ev, np.sort(np.concatenate(ks)), atol=1e-10)


def test_supercell_duplicates_sum_in_tag_string_order(oracle):
    """Wrapped duplicates are summed in the order of the reference's
    multimap<string> (tag "c0 c1 c2 e0 e1 "): "10 ..." < "2 ..." < "9 ...",
    so ((1 + -1e16) + 1e16) = 0, where numeric cell order would give 1."""
    cell = [(9, 0, 0), (10, 0, 0), (2, 0, 0)]
    edge = [(0, 0)] * 3
    val = [1e16, 1.0, -1e16]
    n, rowptr, col, v = oracle.supercell_csr(1, cell, edge, val, (1, 1, 1))
    assert n == 1 and list(col) == [0] and v[0] == 0.0
    # input order does not matter, only the tags
    n, rowptr, col, v = oracle.supercell_csr(1, cell[::-1], edge, val[::-1], (1, 1, 1))
    assert v[0] == 0.0


def test_supercell_filters(oracle):
    """|v| / mean|v| <= 1e-10 is dropped before summing; all-zero and empty
    lists give an operator with no entries (the reference's "2 0" files)."""
    cell = [(0, 0, 0), (1, 0, 0), (1, 0, 0)]
    edge = [(0, 0), (0, 0), (0, 0)]
    n, rowptr, col, v = oracle.supercell_csr(1, cell, edge, [1.0, 1e-12, 2.0], (3, 1, 1))
    d = dense_herm(n, rowptr, col, v)
    assert d[0, 1] == 2.0 and d[0, 0] == 1.0  # the 1e-12 hopping is gone, not added
    n, rowptr, col, v = oracle.supercell_csr(2, [(1, 0, 0)], [(0, 1)], [0.0], (2, 1, 1))
    assert n == 4 and col.size == 0 and list(rowptr) == [0] * 5
    n, rowptr, col, v = oracle.supercell_csr(2, np.zeros((0, 3)), np.zeros((0, 2)), [], (2, 1, 1))
    assert n == 4 and col.size == 0


# ---- non-orthogonal basis (SpectralMoments_nonOrth, BiCGSTAB overlap solves) ----
def _overlap(rng, n, eps=0.08):
    """A Hermitian positive overlap S = I + eps X stored as its upper triangle
    (the reference's CSR files), and the matrix the reference solves with: the
    stored arrays read as a general matrix, i.e. triu(S)."""
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    X = (X + X.conj().T) / 2
    X[np.abs(X) < 1.2] = 0  # sparse
    S = np.eye(n) + eps * X
    U = np.triu(S)
    rows, cols = np.nonzero(U)
    rowptr = np.zeros(n + 1, np.int32)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.int32)
    return io.CSR(n, rowptr, cols.astype(np.int32), U[rows, cols].astype(np.complex128)), U


def test_overlap_solve_matches_dense_solve(oracle):
    rng = np.random.default_rng(11)
    n = 60
    S, U = _overlap(rng, n)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    x, it, err = oracle.overlap_solve(S, b, tol=1e-14, maxit=2000)
    assert err <= 1e-14 and 0 < it < 100
    assert np.allclose(x, np.linalg.solve(U, b), rtol=0, atol=1e-12 * np.abs(b).max())
    x4, it4, err4 = oracle.overlap_solve(S, b)  # the reference's tolerance 1e-4
    assert err4 <= 1e-4 and it4 <= it
    r = b - U @ x4
    assert np.linalg.norm(r) <= 1e-4 * np.linalg.norm(b) * (1 + 1e-12)
    z, it0, e0 = oracle.overlap_solve(S, np.zeros(n, complex))
    assert it0 == 0 and not z.any()


def test_nonorth_moments_with_identity_overlap_equal_orthogonal_ones(oracle):
    """S = I: every BiCGSTAB solve returns its right-hand side exactly
    (alpha = 1, s = 0), so the recursion is the orthogonal one."""
    m = lattices.square_anderson(6, W=1.0, seed=4)
    H = oracle.Op.from_csr(m).rescale(1 / 4.6, 0.0)
    n = m.n
    I = io.CSR(n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.ones(n, np.complex128))
    phi = lattices.random_phase_states(n, 2, seed=5)
    mu = oracle.spectral_moments_nonorth(H, None, I, phi, 40)
    ref = oracle.spectral_moments(H, None, phi, 40)
    assert np.abs(mu - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("with_op", [False, True])
def test_nonorth_moments_match_dense_chebyshev_of_generalised_problem(oracle, with_op):
    """mu_m = (2/R)(1/2 if m = 0) sum_r phi_r^H T_m(A) [OP] U^-1 phi_r with
    A = U^-1 H~ and U the stored triangle of S, computed densely (tight solver
    tolerance)."""
    rng = np.random.default_rng(3 + with_op)
    n, M, R = 48, 30, 2
    Hm = io.CSR(n, *random_upper_csr(n, 0.15, seed=21 + with_op))
    Dh = dense_herm(n, Hm.rowptr, Hm.col, Hm.val)
    a = 0.9 / np.abs(np.linalg.eigvalsh(Dh)).max()
    H = oracle.Op.from_csr(Hm).rescale(a, 0.0)
    S, U = _overlap(rng, n, eps=0.05)
    OPm = io.CSR(n, *random_upper_csr(n, 0.1, seed=31)) if with_op else None
    OP = oracle.Op.from_csr(OPm) if with_op else None
    phi = lattices.random_phase_states(n, R, seed=9)
    mu = oracle.spectral_moments_nonorth(H, OP, S, phi, M, tol=1e-15)
    A = np.linalg.solve(U, a * Dh)
    ref = np.zeros(M, complex)
    for r in range(R):
        t0 = np.linalg.solve(U, phi[r])
        if with_op:
            t0 = dense_herm(n, OPm.rowptr, OPm.col, OPm.val) @ t0
        t1 = A @ t0
        for m in range(M):
            ref[m] += (2.0 / R) * (0.5 if m == 0 else 1.0) * np.vdot(phi[r], t0)
            t0, t1 = t1, 2 * A @ t1 - t0
    assert np.abs(mu - ref).max() <= 1e-10 * np.abs(ref).max()


# ---- OpenMP helpers (CPU baseline timers and the full-size DOS checker) ----
def _lattice_ops(oracle):
    H = lattices.square_anderson(12, W=1.0, seed=7)
    Ho = oracle.Op.from_csr(H).rescale(1 / 4.6, 0.0)
    D = dense_herm(H.n, H.rowptr, H.col, H.val) / 4.6
    return H, Ho, D


def test_omp_cheb_step_and_mv_match_serial_mv_and_dense(oracle):
    """lqo_omp_cheb_step: y <- 2 H x - y and lqo_omp_mv: y <- OP x, for k
    vectors, equal the pinned serial lqo_op_mv and the dense completion."""
    H, Ho, D = _lattice_ops(oracle)
    n, k = H.n, 3
    rng = np.random.default_rng(21)
    x = rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))
    y0 = rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))
    y = y0.copy()
    oracle.omp_cheb_step(Ho, x, y)
    for v in range(k):
        serial = Ho.mv(2.0, x[v], -1.0, y0[v])
        assert rel_inf(y[v], serial) < 1e-15
        assert rel_inf(y[v], 2 * D @ x[v] - y0[v]) < 1e-13
    Hx, Vx, Vy = lattices.haldane_honeycomb(6, seed=3)
    Vo = oracle.Op.from_csr(Vx)
    Dv = dense_herm(Vx.n, Vx.rowptr, Vx.col, Vx.val)
    xv = rng.normal(size=(2, Vx.n)) + 1j * rng.normal(size=(2, Vx.n))
    out = np.zeros_like(xv)
    oracle.omp_mv(Vo, xv, out)
    for v in range(2):
        assert rel_inf(out[v], Vo.mv(1.0, xv[v], 0.0, np.zeros(Vx.n, complex))) < 1e-15
        assert rel_inf(out[v], Dv @ xv[v]) < 1e-13


def test_omp_momtable_matches_numpy_zdotc_table(oracle):
    """lqo_omp_momtable restricted to rows [m0_begin, m0_end) equals L^H R."""
    rng = np.random.default_rng(22)
    M, n = 9, 301
    L = rng.normal(size=(M, n)) + 1j * rng.normal(size=(M, n))
    Rv = rng.normal(size=(M + 2, n)) + 1j * rng.normal(size=(M + 2, n))
    full = L.conj() @ Rv.T
    assert rel_inf(oracle.omp_momtable(L, Rv, 0, M), full) < 1e-14
    assert rel_inf(oracle.omp_momtable(L, Rv, 3, 7), full[3:7]) < 1e-14
    assert oracle.omp_momtable(L, Rv, 4, 4).shape == (0, M + 2)


def test_omp_spectral_moments_match_serial_and_eigendecomposition(oracle):
    """The full-size DOS checker equals the eigendecomposition-pinned serial
    lqo_spectral_moments (reduction order differs: tolerance 1e-13) and the
    closed form directly."""
    H, Ho, D = _lattice_ops(oracle)
    n, M, R = H.n, 48, 3
    phi = lattices.random_phase_states(n, R, seed=23)
    omp = oracle.omp_spectral_moments(Ho, phi, M)
    serial = oracle.spectral_moments(Ho, None, phi, M)
    assert rel_inf(omp, serial) < 1e-13
    T = _cheb_dense(D, M)
    ref = np.array([sum((2.0 / R) * (0.5 if m == 0 else 1.0) * np.vdot(p, T[m] @ p) for p in phi)
                    for m in range(M)])
    assert rel_inf(omp, ref) < 1e-12


def test_omp_bastin_ffts_match_serial_kubo_bastin(oracle):
    """The per-site Bastin DFT stage over stored vectors L_m = T_m OPL phi,
    R_m = OPR T_m phi equals the serial Kubo solver's r_data (itself pinned by
    the Gamma-kernel formula) for one state."""
    n, M, R, W, H, Ht, Vo, V, phi = _kubo_setup(oracle, R=1)
    nump = 64
    fd, _ = oracle.kubo_fft(H, Vo, Vo, phi, M, nump, 1)
    T = _cheb_dense(Ht, M)
    p = phi[0]
    Lv = np.array([T[m] @ (V @ p) for m in range(M)])
    Rv = np.array([V @ (T[m] @ p) for m in range(M)])
    rd = oracle.omp_bastin_ffts(Lv, Rv, nump)
    assert rel_inf(rd, fd) < 1e-11
    # accumulation: a second call adds the same amount
    rd2 = oracle.omp_bastin_ffts(Lv, Rv, nump)
    assert rel_inf(rd2, rd) < 1e-15


def test_csr_read_stream_matches_numpy_reader(oracle, tmp_path):
    """The reference's stream reader restated in C (the ingestion baseline)
    reads the reference's own operator files and a written lattice operator
    exactly as the numpy reader does."""
    for path in (os.path.join(LC, "operators", "linear_chain.HAM.CSR"),
                 os.path.join(LC, "operators", "linear_chain.VX.CSR")):
        ref = io.read_csr(path)
        n, rowptr, col, val = oracle.csr_read_stream(path)
        assert n == ref.n
        assert np.array_equal(rowptr, ref.rowptr) and np.array_equal(col, ref.col)
        assert np.array_equal(val, ref.val)
    Hm, _, _ = lattices.haldane_honeycomb(5, seed=2)
    p = str(tmp_path / "h.CSR")
    io.write_csr(p, Hm)
    ref = io.read_csr(p)
    n, rowptr, col, val = oracle.csr_read_stream(p)
    assert n == ref.n and np.array_equal(rowptr, ref.rowptr) and np.array_equal(col, ref.col)
    assert np.array_equal(val, ref.val)


def test_op_nnz_and_threads(oracle):
    rowptr, col, val = random_upper_csr(20, 0.2, seed=24)
    assert oracle.Op(20, rowptr, col, val).nnz() == col.shape[0]
    assert oracle.omp_threads() >= 1


# Every exported oracle symbol and the test(s) above that pin it.
PINS = {
    "lqo_op_create": ["test_herm_upper_mv_and_rescale_match_dense"],
    "lqo_op_destroy": ["test_herm_upper_mv_and_rescale_match_dense"],
    "lqo_op_nnz": ["test_op_nnz_and_threads"],
    "lqo_op_rescale": ["test_herm_upper_mv_and_rescale_match_dense"],
    "lqo_op_mv": ["test_herm_upper_mv_and_rescale_match_dense", "test_mv_ignores_strictly_lower_entries"],
    "lqo_op_dense": ["test_herm_upper_mv_and_rescale_match_dense", "test_mv_ignores_strictly_lower_entries"],
    "lqo_spectral_moments": ["test_spectral_moments_match_eigendecomposition"],
    "lqo_correlation_moments": ["test_oracle_reproduces_golden_chebmom2d", "test_correlation_moments_batch_semantics"],
    "lqo_dft_backward": ["test_dft_matches_numpy"],
    "lqo_kubo_jackson": ["test_jackson_closed_form_and_truncation", "test_kubo_greenwood_identity"],
    "lqo_kubo_fft": ["test_kubo_greenwood_identity", "test_kubo_bastin_matches_gamma_kernel",
                     "test_kubo_sections_equal_masked_operator"],
    "lqo_kubo_post": ["test_kubo_greenwood_identity", "test_kubo_bastin_matches_gamma_kernel"],
    "lqo_jackson_1d": ["test_jackson_closed_form_and_truncation"],
    "lqo_jackson_2d": ["test_jackson_2d_is_separable_product"],
    "lqo_dos_fftgrid": ["test_dos_fftgrid_quadrature_gives_mu0"],
    "lqo_greenwood_fftgrid": ["test_greenwood_fftgrid_matches_kubo_route_for_real_moments"],
    "lqo_bastin_fftgrid": ["test_bastin_fftgrid_trig_form"],
    "lqo_time_correlations": ["test_time_correlations_match_eigendecomposition"],
    "lqo_jackson_td": ["test_jackson_td_and_time_post"],
    "lqo_timecorr_post": ["test_jackson_td_and_time_post"],
    "lqo_linear_kernel": ["test_linear_coefficients_match_closed_forms",
                          "test_linear_greenwood_and_bastin_match_fftgrid_forms"],
    "lqo_linear_grid": ["test_linear_grid_rows_and_end_points", "test_linear_surf_sea_bastinII_identities"],
    "lqo_overlap_solve": ["test_overlap_solve_matches_dense_solve"],
    "lqo_spectral_moments_nonorth": ["test_nonorth_moments_with_identity_overlap_equal_orthogonal_ones",
                                     "test_nonorth_moments_match_dense_chebyshev_of_generalised_problem"],
    "lqo_supercell_csr": ["test_supercell_matches_reference_wrap_fixture", "test_supercell_spectrum_is_bloch_union"],
    "lqo_csr_read_stream": ["test_csr_read_stream_matches_numpy_reader"],
    "lqo_omp_threads": ["test_op_nnz_and_threads"],
    "lqo_omp_cheb_step": ["test_omp_cheb_step_and_mv_match_serial_mv_and_dense"],
    "lqo_omp_mv": ["test_omp_cheb_step_and_mv_match_serial_mv_and_dense"],
    "lqo_omp_momtable": ["test_omp_momtable_matches_numpy_zdotc_table"],
    "lqo_omp_bastin_ffts": ["test_omp_bastin_ffts_match_serial_kubo_bastin"],
    "lqo_omp_spectral_moments": ["test_omp_spectral_moments_match_serial_and_eigendecomposition"],
}


def test_every_exported_oracle_function_has_a_pinning_test(oracle):
    """No oracle function serves as a checker or a baseline without a test
    that pins it: the library's exported lqo_* symbols equal the keys above,
    and every named test exists in this module."""
    import ctypes
    import subprocess

    ctypes.CDLL(oracle._LIB_PATH)  # built and loadable
    out = subprocess.run(["nm", "-D", "--defined-only", oracle._LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.split() and ln.split()[-1].startswith("lqo_")}
    assert exported == set(PINS), (sorted(exported - set(PINS)), sorted(set(PINS) - exported))
    g = globals()
    for sym, tests in PINS.items():
        for t in tests:
            assert callable(g.get(t)), f"{sym}: pinning test {t} missing"
