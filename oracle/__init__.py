"""f64 CPU oracle for the spectral IMEX step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  It is the checker, never the thing
measured or shipped: the product path (``turing-sphere_amd``) never imports it
and fails loudly when its HIP library is missing.

Two layers:

* ``liboracle.so`` (``sht_oracle.c``): plain-C, f64, loop-for-loop restatement
  of the reference's CPU algorithm -- ``src/sht/layout.ts``, ``gauss.ts``,
  ``coeffs.ts``, ``reference.ts`` (direct summation), ``src/mgpu/noise.ts``,
  ``src/mgpu/model.ts:46-56``.
* numpy helpers here: the element-wise reaction / IMEX arithmetic of the
  ``models/*.m`` files (``bench/shtns/spec.h:146-169``), a fast f64 transform
  (the oracle's own C Legendre stage + numpy's FFT for the Fourier stage, used
  where the direct Fourier sum is too slow) and the f64 IMEX trajectory.

Every function is pinned by a check in ``tests/test_oracle.py`` against the
reference's own spec header compiled where it lies (``oracle/_ref``), golden
fixtures, or closed forms (numpy/scipy Gauss-Legendre, scipy spherical
harmonics, numpy FFT, orthonormality, round-trip identity).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(HERE, "liboracle.so")
_REF_PATH = os.path.join(HERE, "_ref", "libref_spec.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = C.POINTER(C.c_int)

_lib = None


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_nlm_calc.restype = C.c_longlong
        L.orc_nlm_calc.argtypes = [C.c_int, C.c_int]
        L.orc_lm_index.restype = C.c_longlong
        L.orc_lm_index.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_validate_config.restype = C.c_int
        L.orc_validate_config.argtypes = [C.c_int] * 4
        L.orc_grid_for_lmax.argtypes = [C.c_int, C.c_int, _ip, _ip]
        L.orc_gauss_nodes_weights.argtypes = [C.c_int, _dp, _dp]
        L.orc_legendre_coeffs.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.orc_legendre_row.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, _dp]
        L.orc_legendre_synth.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_synth_rows.argtypes = [C.c_int] * 4 + [_dp, _dp, C.c_int, C.c_int]
        L.orc_synth.argtypes = [C.c_int] * 4 + [_dp, _dp]
        L.orc_fourier_analys_rows.argtypes = [C.c_int] * 3 + [_dp, _dp, C.c_int, C.c_int]
        L.orc_legendre_analys_rows.argtypes = [C.c_int] * 3 + [_dp, _dp, C.c_int, C.c_int]
        L.orc_analys.argtypes = [C.c_int] * 4 + [_dp, _dp]
        L.orc_random_spectrum.argtypes = [C.c_int, C.c_int, C.c_uint32, _fp]
        L.orc_mulberry32.restype = C.c_double
        L.orc_mulberry32.argtypes = [C.POINTER(C.c_uint32)]
        L.orc_mulberry32_fill.argtypes = [C.c_uint32, C.c_longlong, _dp]
        L.orc_eigenvalues.argtypes = [C.c_int, C.c_int, _fp]
        L.orc_react.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_imex.restype = C.c_double
        L.orc_imex.argtypes = [C.c_double] * 5
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        _lib = L
    return _lib


def ref_spec():
    """The reference's bench/shtns/spec.h compiled where it lies (or None)."""
    if not os.path.exists(_REF_PATH):
        return None
    L = C.CDLL(_REF_PATH)
    L.ref_rand.restype = C.c_double
    L.ref_rand.argtypes = [C.POINTER(C.c_uint32)]
    L.ref_seeded_spectrum.argtypes = [C.c_int, C.c_int, C.c_uint32, _fp]
    L.ref_grid_for_lmax.argtypes = [C.c_int, C.c_int, _ip, _ip]
    L.ref_nlm_calc.restype = C.c_long
    L.ref_nlm_calc.argtypes = [C.c_int, C.c_int]
    L.ref_react_f64.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.ref_react_f32.argtypes = [C.c_int, C.c_float, C.c_float, _fp, C.c_void_p, _fp,
                                C.c_void_p, C.c_long]
    L.ref_imex_f32.argtypes = [C.c_float, C.c_float, _fp, _fp, _fp, _fp, C.c_long]
    L.ref_imex_f64.restype = C.c_double
    L.ref_imex_f64.argtypes = [C.c_double] * 5
    return L


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


# ---------------------------------------------------------------- layout.ts
def nlm_calc(lmax: int, mmax: int) -> int:
    return int(lib().orc_nlm_calc(lmax, mmax))


def lm_index(lmax: int, l: int, m: int) -> int:
    return int(lib().orc_lm_index(lmax, l, m))


def validate_config(lmax: int, mmax: int, nlat: int, nphi: int) -> int:
    return int(lib().orc_validate_config(lmax, mmax, nlat, nphi))


def grid_for_lmax(lmax: int, pdeg: int) -> tuple[int, int]:
    a, b = C.c_int(), C.c_int()
    lib().orc_grid_for_lmax(lmax, pdeg, C.byref(a), C.byref(b))
    return a.value, b.value


# ----------------------------------------------------------------- gauss.ts
def gauss_nodes_weights(n: int) -> tuple[np.ndarray, np.ndarray]:
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_nodes_weights(n, x, w)
    return x, w


# ---------------------------------------------------------------- coeffs.ts
def legendre_coeffs(lmax: int, mmax: int) -> tuple[np.ndarray, np.ndarray]:
    amm = np.zeros(mmax + 1)
    ab = np.zeros(2 * nlm_calc(lmax, mmax))
    lib().orc_legendre_coeffs(lmax, mmax, amm, ab)
    return amm, ab


def legendre_row(amm, ab, lmax: int, m: int, ct: float, st: float) -> np.ndarray:
    out = np.zeros(lmax - m + 1)
    lib().orc_legendre_row(np.ascontiguousarray(amm, np.float64),
                           np.ascontiguousarray(ab, np.float64), lmax, m, ct, st, out)
    return out


# ------------------------------------------------------------- reference.ts
@dataclass(frozen=True)
class Cfg:
    lmax: int
    mmax: int
    nlat: int
    nphi: int

    @property
    def nlm(self) -> int:
        return nlm_calc(self.lmax, self.mmax)

    @property
    def npts(self) -> int:
        return self.nlat * self.nphi


class ShtReference:
    """f64 direct-summation transforms (reference.ts:13-119)."""

    def __init__(self, cfg: Cfg):
        err = validate_config(cfg.lmax, cfg.mmax, cfg.nlat, cfg.nphi)
        if err:
            raise ValueError(f"invalid SHT config {cfg} (rule {err})")
        self.cfg = cfg
        self.nlm = cfg.nlm
        self.ct, self.wg = gauss_nodes_weights(cfg.nlat)
        self.st = np.sqrt(1.0 - self.ct * self.ct)

    def legendre_synth(self, qlm) -> np.ndarray:
        c = self.cfg
        q = np.ascontiguousarray(qlm, np.float64)
        fm = np.zeros(2 * (c.mmax + 1) * c.nlat)
        lib().orc_legendre_synth(c.lmax, c.mmax, c.nlat, q, fm)
        return fm

    def synth(self, qlm, rows: tuple[int, int] | None = None) -> np.ndarray:
        c = self.cfg
        q = np.ascontiguousarray(qlm, np.float64)
        assert q.size == 2 * self.nlm
        spat = np.zeros(c.nlat * c.nphi)
        r0, r1 = rows if rows else (0, c.nlat)
        lib().orc_synth_rows(c.lmax, c.mmax, c.nlat, c.nphi, q, spat, r0, r1)
        return spat

    def fourier_analys(self, spat, rows: tuple[int, int] | None = None) -> np.ndarray:
        c = self.cfg
        s = np.ascontiguousarray(spat, np.float64)
        gm = np.zeros(2 * (c.mmax + 1) * c.nlat)
        r0, r1 = rows if rows else (0, c.nlat)
        lib().orc_fourier_analys_rows(c.mmax, c.nlat, c.nphi, s, gm, r0, r1)
        return gm

    def legendre_analys(self, gm, rows: tuple[int, int] | None = None) -> np.ndarray:
        c = self.cfg
        g = np.ascontiguousarray(gm, np.float64)
        q = np.zeros(2 * self.nlm)
        r0, r1 = rows if rows else (0, c.nlat)
        lib().orc_legendre_analys_rows(c.lmax, c.mmax, c.nlat, g, q, r0, r1)
        return q

    def analys(self, spat) -> np.ndarray:
        c = self.cfg
        s = np.ascontiguousarray(spat, np.float64)
        assert s.size == c.nlat * c.nphi
        q = np.zeros(2 * self.nlm)
        lib().orc_analys(c.lmax, c.mmax, c.nlat, c.nphi, s, q)
        return q

    # -- fast f64 transforms: the oracle's own Legendre stage (C loops above)
    #    with numpy's FFT for the Fourier stage.  Pinned against the direct
    #    sums by tests/test_oracle.py::test_fast_matches_direct.
    def fourier_synth_fast(self, fm) -> np.ndarray:
        c = self.cfg
        F = np.asarray(fm, np.float64).reshape(c.mmax + 1, c.nlat, 2)
        X = np.zeros((c.nlat, c.nphi // 2 + 1), np.complex128)
        X[:, 0] = F[0, :, 0]
        X[:, 1:c.mmax + 1] = (F[1:, :, 0] + 1j * F[1:, :, 1]).T
        return (np.fft.irfft(X, n=c.nphi, axis=1) * c.nphi).reshape(-1)

    def fourier_analys_fast(self, spat) -> np.ndarray:
        c = self.cfg
        f = np.asarray(spat, np.float64).reshape(c.nlat, c.nphi)
        G = np.fft.rfft(f, axis=1)[:, :c.mmax + 1] * (2.0 * np.pi / c.nphi)
        gm = np.empty((c.mmax + 1, c.nlat, 2))
        gm[:, :, 0] = G.real.T
        gm[:, :, 1] = G.imag.T
        return gm.reshape(-1)

    def synth_fast(self, qlm) -> np.ndarray:
        return self.fourier_synth_fast(self.legendre_synth(qlm))

    def analys_fast(self, spat) -> np.ndarray:
        return self.legendre_analys(self.fourier_analys_fast(spat))

    # ---- transposes (reverse-mode gradients; no reference counterpart)
    # synth:  f_ij = Re F_0i + 2 sum_{m>=1} Re(F_mi e^{i m phi_j}),  F_mi = sum_l q_lm P_lm(x_i)
    # analys: q_lm = sum_i w_i P_lm(x_i) (2 pi / nphi) sum_j f_ij e^{-i m phi_j}
    # Both are real-linear maps on (re, im)-interleaved coefficient vectors;
    # their transposes follow term by term, and the dot-product identity
    # <S x, y> = <x, S^T y> pins them (tests/test_oracle.py).
    def synth_T(self, gbar) -> np.ndarray:
        """S^T: q_lm = sum_i P_lm(x_i) c_m sum_j g_ij e^{-i m phi_j}, c_0 = 1, c_m = 2."""
        c = self.cfg
        g = np.asarray(gbar, np.float64).reshape(c.nlat, c.nphi)
        G = np.fft.rfft(g, axis=1)[:, :c.mmax + 1]
        G[:, 1:] *= 2.0
        gm = np.empty((c.mmax + 1, c.nlat, 2))
        gm[:, :, 0] = G.real.T
        gm[:, :, 1] = G.imag.T
        # legendre_analys weights row i by the Gauss weight: divide it out
        _, w = gauss_nodes_weights(c.nlat)
        gm /= w[None, :, None]
        return self.legendre_analys(gm.reshape(-1))

    def analys_T(self, qbar) -> np.ndarray:
        """A^T: g_ij = w_i (2 pi / nphi) sum_{m>=0} Re(F_mi e^{i m phi_j})."""
        c = self.cfg
        F = self.legendre_synth(qbar).reshape(c.mmax + 1, c.nlat, 2).copy()
        F[1:] *= 0.5
        g = self.fourier_synth_fast(F.reshape(-1)).reshape(c.nlat, c.nphi)
        _, w = gauss_nodes_weights(c.nlat)
        return (g * (w * (2.0 * np.pi / c.nphi))[:, None]).reshape(-1)


def random_spectrum(cfg: Cfg, seed: int = 12345) -> np.ndarray:
    """reference.ts:122-136 (float32 output, m=0 imaginary parts zero)."""
    q = np.zeros(2 * cfg.nlm, np.float32)
    lib().orc_random_spectrum(cfg.lmax, cfg.mmax, seed & 0xFFFFFFFF, q)
    return q


# ------------------------------------------------------------------ noise.ts
def mulberry32(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n)
    lib().orc_mulberry32_fill(seed & 0xFFFFFFFF, n, out)
    return out


def randn(seed: int, n: int) -> np.ndarray:
    """mulberry32 + Box-Muller in pairs (noise.ts:27-43).  The cos/sin/log
    here are numpy's, so the last bit can differ from V8's (spec.h:216-222)."""
    out = np.empty(n)
    state = C.c_uint32(seed & 0xFFFFFFFF)
    L = lib()
    i = 0
    while i < n:
        u = 0.0
        while u == 0.0:
            u = L.orc_mulberry32(C.byref(state))
        r = np.sqrt(-2.0 * np.log(u))
        th = 2.0 * np.pi * L.orc_mulberry32(C.byref(state))
        out[i] = r * np.cos(th)
        if i + 1 < n:
            out[i + 1] = r * np.sin(th)
        i += 2
    return out


# ------------------------------------------------------------------ model.ts
def eigenvalues(cfg: Cfg) -> np.ndarray:
    lam = np.zeros(2 * cfg.nlm, np.float32)
    lib().orc_eigenvalues(cfg.lmax, cfg.mmax, lam)
    return lam


# ------------------------------------------- reaction / IMEX (models/*.m)
SCHNAKENBERG, BRUSSELATOR, ALLENCAHN = 0, 1, 2


def react(model: int, p0, p1, u, v=None, dtype=np.float64):
    """Vectorised reaction in the operand order of spec.h:146-161.  With
    dtype=float32 every operation rounds to f32 as an unfused fp32 kernel does."""
    t = dtype
    u = np.asarray(u, t)
    if model == ALLENCAHN:
        return u - u * u * u, None
    v = np.asarray(v, t)
    p0 = t(p0)
    p1 = t(p1)
    uuv = u * u * v
    if model == SCHNAKENBERG:
        return (p0 - u) + uuv, p1 - uuv
    return (p0 - (p1 + t(1)) * u) + uuv, p1 * u - uuv


def imex(dt, dtD, U, R, lam, dtype=np.float64):
    """spec.h:166-169: (U + dt*R) / (1 + dtD*lam), dtD formed by the caller."""
    t = dtype
    return (np.asarray(U, t) + t(dt) * np.asarray(R, t)) / (t(1) + t(dtD) * np.asarray(lam, t))


def imex_trajectory(cfg: Cfg, model: int, params: dict, U0, V0=None, steps: int = 1,
                    ref: ShtReference | None = None):
    """f64 IMEX Euler trajectory of one of the models/*.m steps, on the same
    grid as the GPU run, with the fast f64 transforms."""
    ref = ref or ShtReference(cfg)
    lam = eigenvalues(cfg).astype(np.float64)
    dt = params["dt"]
    U = np.array(U0, np.float64)
    V = None if V0 is None else np.array(V0, np.float64)
    for _ in range(steps):
        u = ref.synth_fast(U)
        if model == ALLENCAHN:
            r1, _ = react(model, 0, 0, u)
            U = imex(dt, dt * params["eps2"], U, ref.analys_fast(r1), lam)
            continue
        v = ref.synth_fast(V)
        p0, p1 = (params["a"], params["b"]) if model == SCHNAKENBERG else (params["A"], params["B"])
        r1, r2 = react(model, p0, p1, u, v)
        Un = imex(dt, dt * params["D1"], U, ref.analys_fast(r1), lam)
        Vn = imex(dt, dt * params["D2"], V, ref.analys_fast(r2), lam)
        U, V = Un, Vn
    return U, V


def _react_partials(model: int, params: dict, u, v):
    """d(reaction)/d(u, v) and d(reaction)/d(params) of the models/*.m
    reaction terms (the derivatives of spec.h:146-161's formulas)."""
    if model == ALLENCAHN:
        return {"uu": 1.0 - 3.0 * u * u}, {}
    uv = u * v
    uu = u * u
    if model == SCHNAKENBERG:
        d = {"uu": -1.0 + 2.0 * uv, "uv": uu, "vu": -2.0 * uv, "vv": -uu}
        dp = {"a": (np.ones_like(u), None), "b": (None, np.ones_like(u))}
        return d, dp
    B = params["B"]
    d = {"uu": -(B + 1.0) + 2.0 * uv, "uv": uu, "vu": B - 2.0 * uv, "vv": -uu}
    dp = {"A": (np.ones_like(u), None), "B": (-u, u)}
    return d, dp


def imex_vjp(cfg: Cfg, model: int, params: dict, U0, V0, steps: int, Ubar, Vbar=None,
             ref: ShtReference | None = None):
    """f64 reverse-mode gradient of `steps` IMEX Euler steps (imex_trajectory)
    for a cotangent (Ubar, Vbar) of the final state: returns
    (U0bar, V0bar, {param: grad}) with grads for the reaction parameters, the
    diffusion coefficients (D1, D2 / eps2) and dt.  All states are kept
    (f64, small sizes); each step is differentiated by hand:
      Un = (U + dt A r_u(S U, S V)) / (1 + dt D1 lam)."""
    ref = ref or ShtReference(cfg)
    lam = eigenvalues(cfg).astype(np.float64)
    dt = params["dt"]
    two = model != ALLENCAHN
    Us, Vs = [np.array(U0, np.float64)], [None if V0 is None else np.array(V0, np.float64)]
    for _ in range(steps):
        U1, V1 = imex_trajectory(cfg, model, params, Us[-1], Vs[-1], 1, ref)
        Us.append(U1)
        Vs.append(V1)
    Ub = np.array(Ubar, np.float64)
    Vb = None if not two else np.array(Vbar if Vbar is not None else np.zeros_like(Ub),
                                      np.float64)
    dkey = ("eps2",) if not two else ("D1", "D2")
    grads = {k: 0.0 for k in (list(dkey) + ["dt"] + (["a", "b"] if model == SCHNAKENBERG else
                                                      ["A", "B"] if model == BRUSSELATOR else []))}
    for k in reversed(range(steps)):
        U, V = Us[k], Vs[k]
        u = ref.synth_fast(U)
        v = ref.synth_fast(V) if two else None
        if two:
            p0, p1 = (params["a"], params["b"]) if model == SCHNAKENBERG else (params["A"],
                                                                              params["B"])
            ru, rv = react(model, p0, p1, u, v)
        else:
            ru, rv = react(model, 0, 0, u)
        d, dp = _react_partials(model, params, u, v)
        outs = [("U", U, ru, dkey[0], Ub)] + ([("V", V, rv, dkey[1], Vb)] if two else [])
        rbar = {}
        new = {}
        for name, X, r, dk, Xb in outs:
            Ar = ref.analys_fast(r)
            den = 1.0 + dt * params[dk] * lam
            Xn = (X + dt * Ar) / den
            new[name] = Xb / den
            grads[dk] += float(np.sum(Xb * (-Xn * dt * lam / den)))
            grads["dt"] += float(np.sum(Xb * (Ar / den - Xn * params[dk] * lam / den)))
            rbar[name] = ref.analys_T(dt * Xb / den)
        if two:
            ub = rbar["U"] * d["uu"] + rbar["V"] * d["vu"]
            vb = rbar["U"] * d["uv"] + rbar["V"] * d["vv"]
            for pk, (cu, cv) in dp.items():
                if cu is not None:
                    grads[pk] += float(np.sum(rbar["U"] * cu))
                if cv is not None:
                    grads[pk] += float(np.sum(rbar["V"] * cv))
            Ub = new["U"] + ref.synth_T(ub)
            Vb = new["V"] + ref.synth_T(vb)
        else:
            Ub = new["U"] + ref.synth_T(rbar["U"] * d["uu"])
    return Ub, Vb, grads


def rel_l2(a, b) -> float:
    """digest.ts:61-70"""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.sum((a - b) ** 2) / max(np.sum(b * b), 1e-300)))
