"""Pins the f64 oracle before anything is checked against it (CPU only).

Each oracle function is checked against something that does not restate it:
the reference's own native spec header compiled where it lies
(oracle/_ref), JavaScript-engine output for the generators
(tests/golden/rng_js.json), or closed forms (scipy Gauss-Legendre nodes and
spherical harmonics, numpy FFT, orthonormality, exact IMEX maps).  Each check
is chosen so that a plausible slip (a dropped Condon-Shortley sign, a logical
instead of arithmetic shift, a missing 2pi/nphi, a swapped re/im) fails it.
"""
import json
import os

import numpy as np
import pytest
from scipy.special import roots_legendre, sph_harm_y

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "rng_js.json")))


# ------------------------------------------------------------ layout.ts
def test_layout_against_reference_spec_header():
    ref = O.ref_spec()
    if ref is None:
        pytest.skip("reference tree absent")
    import ctypes as C

    for lmax, mmax in [(1, 0), (15, 15), (63, 63), (1023, 1023), (100, 37)]:
        assert O.nlm_calc(lmax, mmax) == ref.ref_nlm_calc(lmax, mmax)
    for lmax in (15, 31, 63, 127, 255, 511, 1023):
        for pdeg in (1, 2, 3):
            a, b = C.c_int(), C.c_int()
            ref.ref_grid_for_lmax(lmax, pdeg, C.byref(a), C.byref(b))
            assert O.grid_for_lmax(lmax, pdeg) == (a.value, b.value)


def test_lm_index_enumerates_m_major():
    """Brute force: the index walks 0..nlm-1 in (m, then l) order."""
    for lmax, mmax in [(7, 7), (9, 4)]:
        idx = [O.lm_index(lmax, l, m) for m in range(mmax + 1) for l in range(m, lmax + 1)]
        assert idx == list(range(O.nlm_calc(lmax, mmax)))


def test_validate_config_rules():
    assert O.validate_config(31, 31, 32, 64) == 0
    assert O.validate_config(0, 0, 4, 8) == 1
    assert O.validate_config(8, 9, 16, 32) == 2
    assert O.validate_config(8, 8, 8, 32) == 3
    assert O.validate_config(8, 8, 16, 16) == 4


# -------------------------------------------------------------- gauss.ts
@pytest.mark.parametrize("n", [1, 2, 7, 32, 64, 513, 1024])
def test_gauss_nodes_weights(n):
    x, w = O.gauss_nodes_weights(n)
    xr, wr = roots_legendre(n)
    assert np.all(np.diff(x) < 0)  # decreasing: north first
    assert np.max(np.abs(x - xr[::-1])) < 1e-14
    assert np.max(np.abs(w - wr[::-1])) < 1e-13 * max(1, n / 64)
    assert abs(w.sum() - 2) < 1e-13
    # exact for degree 2n-1
    k = 2 * n - 1
    assert abs(np.sum(w * x ** (k - 1)) - (2.0 / k if (k - 1) % 2 == 0 else 0.0)) < 1e-12


# ------------------------------------------------------------- coeffs.ts
def test_amm_closed_form_and_condon_shortley():
    amm, _ = O.legendre_coeffs(40, 40)
    for m in range(41):
        num = np.prod([(2 * k + 1) / (2 * k) for k in range(1, m + 1)]) if m else 1.0
        want = (-1) ** m * np.sqrt(num / (4 * np.pi))
        assert np.isclose(amm[m], want, rtol=1e-14)


@pytest.mark.parametrize("theta", [0.013, 0.7, 1.5707963, 2.9])
def test_legendre_row_is_ytilde(theta):
    """ytilde_l^m(theta) = Y_l^m(theta, 0) of scipy (orthonormal, Condon-Shortley)."""
    lmax = 60
    amm, ab = O.legendre_coeffs(lmax, lmax)
    for m in (0, 1, 2, 17, 33, 60):
        row = O.legendre_row(amm, ab, lmax, m, np.cos(theta), np.sin(theta))
        l = np.arange(m, lmax + 1)
        want = sph_harm_y(l, m, theta, 0.0).real
        assert np.allclose(row, want, rtol=1e-10, atol=1e-12 * np.max(np.abs(want)))


def test_orthonormality_on_gauss_grid():
    lmax, nlat = 24, 25
    amm, ab = O.legendre_coeffs(lmax, lmax)
    x, w = O.gauss_nodes_weights(nlat)
    for m in (0, 3, 24):
        P = np.array([O.legendre_row(amm, ab, lmax, m, xi, np.sqrt(1 - xi * xi)) for xi in x])
        G = 2 * np.pi * (P.T * w) @ P
        assert np.allclose(G, np.eye(lmax - m + 1), atol=1e-13)


# ---------------------------------------------------------- reference.ts
def _cfg(lmax=12, nlat=16, nphi=28):
    return O.Cfg(lmax, lmax, nlat, nphi)


def test_synth_single_mode_is_spherical_harmonic():
    cfg = _cfg()
    ref = O.ShtReference(cfg)
    theta = np.arccos(ref.ct)
    phi = 2 * np.pi * np.arange(cfg.nphi) / cfg.nphi
    for l, m, re, im in [(0, 0, 1.0, 0.0), (5, 2, 0.3, -0.7), (12, 12, -1.0, 0.5), (7, 0, 2.0, 0.0)]:
        q = np.zeros(2 * cfg.nlm)
        i = O.lm_index(cfg.lmax, l, m)
        q[2 * i], q[2 * i + 1] = re, im
        f = ref.synth(q).reshape(cfg.nlat, cfg.nphi)
        Y = sph_harm_y(l, m, theta[:, None], phi[None, :])
        c = re + 1j * im
        want = (c * Y).real if m == 0 else 2 * (c * Y).real
        assert np.allclose(f, want, atol=1e-13)


def test_analys_inverts_synth_and_matches_fft():
    cfg = _cfg()
    ref = O.ShtReference(cfg)
    q = O.random_spectrum(cfg, 7).astype(np.float64)
    f = ref.synth(q)
    assert O.rel_l2(ref.analys(f), q) < 1e-13
    # the Fourier half: G_m = (2 pi / nphi) rfft(f)[m]
    gm = ref.fourier_analys(f).reshape(cfg.mmax + 1, cfg.nlat, 2)
    G = np.fft.rfft(f.reshape(cfg.nlat, cfg.nphi), axis=1)[:, :cfg.mmax + 1] * 2 * np.pi / cfg.nphi
    assert np.allclose(gm[..., 0], G.real.T, atol=1e-13)
    assert np.allclose(gm[..., 1], G.imag.T, atol=1e-13)


def test_fast_transforms_match_direct_sums():
    for lmax, nlat, nphi in [(12, 16, 28), (31, 48, 64), (40, 41, 81)]:
        cfg = _cfg(lmax, nlat, nphi)
        ref = O.ShtReference(cfg)
        q = O.random_spectrum(cfg, lmax).astype(np.float64)
        f = ref.synth(q)
        assert O.rel_l2(ref.synth_fast(q), f) < 1e-13
        assert O.rel_l2(ref.analys_fast(f), ref.analys(f)) < 1e-13


def test_row_subsets_compose():
    """The *_rows entry points (bounded timing samples) cover disjoint rows."""
    cfg = _cfg(20, 24, 48)
    ref = O.ShtReference(cfg)
    q = O.random_spectrum(cfg, 3).astype(np.float64)
    full = ref.synth(q)
    part = ref.synth(q, rows=(0, 10)) + ref.synth(q, rows=(10, 24))
    assert np.array_equal(full, part)
    gm = ref.fourier_analys(full)
    qa = ref.legendre_analys(gm, rows=(0, 7)) + ref.legendre_analys(gm, rows=(7, 24))
    assert np.allclose(qa, ref.legendre_analys(gm), atol=1e-15)


# --------------------------------------------------------------- the RNGs
@pytest.mark.parametrize("seed", list(GOLDEN["xorshift_f32"]))
def test_random_spectrum_matches_javascript(seed):
    """reference.ts:122-136 evaluated by node: bit-exact f32 stream."""
    lmax = 40
    q = O.random_spectrum(O.Cfg(lmax, lmax, 41, 81), int(seed))
    js = np.array(GOLDEN["xorshift_f32"][seed], np.float32)
    n = js.size
    m0_imag = {2 * O.lm_index(lmax, l, 0) + 1 for l in range(lmax + 1)}
    keep = np.array([i not in m0_imag for i in range(n)])
    assert np.array_equal(q[:n][keep], js[keep])
    assert np.all(q[sorted(i for i in m0_imag)] == 0)


def test_logical_shift_variant_would_fail():
    """The planted mistake: a textbook xorshift32 (logical >>) differs from
    the reference's stream -- so the golden test above pins the shift."""
    s = 42
    out = []
    for _ in range(64):
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        out.append(np.float32((s / 4294967296.0) * 2 - 1))
    js = np.array(GOLDEN["xorshift_f32"]["42"][:64], np.float32)
    assert not np.array_equal(np.array(out, np.float32), js)


@pytest.mark.parametrize("seed", list(GOLDEN["mulberry32"]))
def test_mulberry32_matches_javascript_and_spec_header(seed):
    js = np.array(GOLDEN["mulberry32"][seed])
    got = O.mulberry32(int(seed), js.size)
    assert np.array_equal(got, js)
    ref = O.ref_spec()
    if ref is not None:
        import ctypes as C

        st = C.c_uint32(int(seed))
        spec = np.array([ref.ref_rand(C.byref(st)) for _ in range(js.size)])
        assert np.array_equal(spec, js)


# --------------------------------------------- model.ts / reaction / IMEX
def test_eigenvalues_closed_form():
    cfg = O.Cfg(20, 13, 21, 41)
    lam = O.eigenvalues(cfg)
    for m in range(cfg.mmax + 1):
        for l in range(m, cfg.lmax + 1):
            i = O.lm_index(cfg.lmax, l, m)
            assert lam[2 * i] == lam[2 * i + 1] == l * (l + 1)


@pytest.mark.parametrize("model,p0,p1", [(0, 0.1, 0.9), (1, 4.5, 6.96), (2, 0, 0)])
def test_reaction_and_imex_against_spec_header(model, p0, p1):
    ref = O.ref_spec()
    if ref is None:
        pytest.skip("reference tree absent")
    import ctypes as C

    rng = np.random.default_rng(model)
    for u, v in rng.uniform(-2, 3, (50, 2)):
        a, b = C.c_double(), C.c_double()
        ref.ref_react_f64(model, p0, p1, u, v, C.byref(a), C.byref(b))
        r1, r2 = O.react(model, p0, p1, np.array([u]), np.array([v]))
        assert r1[0] == a.value
        if model != 2:
            assert r2[0] == b.value
        for U, R, lam in rng.uniform(-1, 1, (3, 3)):
            want = ref.ref_imex_f64(0.01, 0.01 * 6.5e-4, U, R, abs(lam) * 100)
            assert O.imex(0.01, 0.01 * 6.5e-4, U, R, abs(lam) * 100) == want


def test_imex_trajectory_linear_mode_exact():
    """The oracle's step on linear.m's reaction reproduces the closed-form
    per-mode factor (analyticChecks.ts:73-122) to f64 round-off -- exercised
    through the same synth/analys/imex chain the trajectory oracle uses."""
    lmax = 15
    cfg = O.Cfg(lmax, lmax, 16, 32)
    ref = O.ShtReference(cfg)
    lam = O.eigenvalues(cfg).astype(np.float64)
    c, D, dt = -0.3, 0.01, 0.1
    U = np.zeros(2 * cfg.nlm)
    i = O.lm_index(lmax, 5, 2)
    U[2 * i], U[2 * i + 1] = 0.8, -0.35
    for _ in range(20):
        R = ref.analys_fast(c * ref.synth_fast(U))
        U = O.imex(dt, dt * D, U, R, lam)
    g = ((1 + dt * c) / (1 + dt * D * 30)) ** 20
    assert abs(U[2 * i] - 0.8 * g) < 1e-13 and abs(U[2 * i + 1] + 0.35 * g) < 1e-13
    U[2 * i:2 * i + 2] = 0
    assert np.max(np.abs(U)) < 1e-13


def test_imex_trajectory_logistic_uniform_exact():
    cfg = O.Cfg(15, 15, 32, 64)
    ref = O.ShtReference(cfg)
    u0 = 0.3
    U0 = ref.analys(np.full(cfg.npts, u0))
    U, _ = O.imex_trajectory(cfg, O.ALLENCAHN, {"dt": 0.05, "eps2": 0.0}, U0, None, 10, ref)
    want = u0
    for _ in range(10):
        want = want + 0.05 * (want - want ** 3)
    assert abs(U[0] / U0[0] * u0 - want) < 1e-13


# ------------------------------------------------ transposes and the adjoint
@pytest.mark.parametrize("lmax,mmax,nlat,nphi", [(15, 15, 16, 32), (12, 9, 17, 32), (7, 7, 9, 20)])
def test_transposes_dot_product(lmax, mmax, nlat, nphi):
    """<S x, y> = <x, S^T y> and <A y, x> = <y, A^T x> in f64 for random
    x (all 2 nlm reals, m=0 imaginary parts included) and y."""
    cfg = O.Cfg(lmax, mmax, nlat, nphi)
    ref = O.ShtReference(cfg)
    rng = np.random.default_rng(lmax + nlat)
    x = rng.normal(size=2 * cfg.nlm)
    y = rng.normal(size=nlat * nphi)
    a, b = np.dot(ref.synth_fast(x), y), np.dot(x, ref.synth_T(y))
    assert abs(a - b) < 1e-11 * max(1.0, abs(a))
    a, b = np.dot(ref.analys_fast(y), x), np.dot(y, ref.analys_T(x))
    assert abs(a - b) < 1e-11 * max(1.0, abs(a))
    # a wrong normalisation or a dropped m factor breaks the identity
    bad = ref.synth_T(y) * 1.0001
    assert abs(np.dot(ref.synth_fast(x), y) - np.dot(x, bad)) > 1e-9


@pytest.mark.parametrize("model", [O.SCHNAKENBERG, O.BRUSSELATOR, O.ALLENCAHN])
def test_imex_vjp_vs_finite_differences(model):
    """The f64 adjoint against central differences of the f64 trajectory, for
    the initial state (random direction) and every parameter."""
    cfg = O.Cfg(7, 7, 8, 16)
    ref = O.ShtReference(cfg)
    rng = np.random.default_rng(7 + model)
    params = {O.SCHNAKENBERG: {"a": 0.1, "b": 0.9, "D1": 4e-3, "D2": 8e-2, "dt": 0.05},
              O.BRUSSELATOR: {"A": 3.0, "B": 9.0, "D1": 3e-3, "D2": 1.7e-2, "dt": 0.01},
              O.ALLENCAHN: {"eps2": 1e-2, "dt": 0.05}}[model]
    two = model != O.ALLENCAHN
    U0 = rng.normal(size=2 * cfg.nlm) * 0.05
    U0[0] += 1.0
    V0 = (rng.normal(size=2 * cfg.nlm) * 0.05) if two else None
    steps = 6
    cu = rng.normal(size=2 * cfg.nlm)
    cv = rng.normal(size=2 * cfg.nlm) if two else None

    def J(p, U, V):
        Uf, Vf = O.imex_trajectory(cfg, model, p, U, V, steps, ref)
        return float(np.dot(cu, Uf) + (np.dot(cv, Vf) if two else 0.0))

    Ub, Vb, g = O.imex_vjp(cfg, model, params, U0, V0, steps, cu, cv, ref)
    dU = rng.normal(size=U0.size)
    h = 1e-6
    fd = (J(params, U0 + h * dU, V0) - J(params, U0 - h * dU, V0)) / (2 * h)
    assert abs(fd - np.dot(Ub, dU)) < 1e-6 * max(1.0, abs(fd))
    if two:
        dV = rng.normal(size=V0.size)
        fd = (J(params, U0, V0 + h * dV) - J(params, U0, V0 - h * dV)) / (2 * h)
        assert abs(fd - np.dot(Vb, dV)) < 1e-6 * max(1.0, abs(fd))
    for k in g:
        hp = 1e-6 * max(abs(params[k]), 1e-3)
        pp, pm = dict(params), dict(params)
        pp[k] += hp
        pm[k] -= hp
        fd = (J(pp, U0, V0) - J(pm, U0, V0)) / (2 * hp)
        assert abs(fd - g[k]) < 1e-5 * max(1.0, abs(fd)), (k, fd, g[k])
