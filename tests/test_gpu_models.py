"""The .m models through front end -> HIP emitter -> planner -> GPU, checked
against closed forms (test/analyticChecks.ts A/B/C), the model checks of
test/modelChecks.ts, the f64 oracle trajectory, and eager-vs-graph replay."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import emit

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MFILES = os.path.join(ROOT, "turing-sphere_amd", "mfiles")

MODELS = {
    "schnakenberg": (["a", "b", "D1", "D2", "dt"], ["U", "V"], ["u", "v"]),
    "brusselator": (["A", "B", "D1", "D2", "dt"], ["U", "V"], ["u", "v"]),
    "allencahn": (["eps2", "dt"], ["U"], ["u"]),
    "linear": (["c", "D", "dt"], ["U"], ["u"]),
    "logistic": (["r", "D", "dt"], ["U"], ["u"]),
}
# src/mgpu/registry.ts:51-102 defaults
DEFAULTS = {
    "schnakenberg": {"a": 0.1, "b": 0.9, "D1": 4e-4, "D2": 8e-3, "dt": 0.05},
    "brusselator": {"A": 3.0, "B": 9.0, "D1": 3.33e-3, "D2": 1.67e-2, "dt": 0.02},
    "allencahn": {"eps2": 1e-3, "dt": 0.02},
}


def make(key, lmax, nlat, nphi, params, batch=True, grid=True, spectral=True,
         legendre="fp32", fold=True):
    import tsphere
    from tsphere.model import GpuModel

    cfg = tsphere.ShtConfig(lmax, lmax, nlat, nphi)
    names, state, view = MODELS[key]
    sht = tsphere.ShtPlan.create(cfg, legendre=legendre, max_fields=len(state))
    src = open(os.path.join(MFILES, key + ".m")).read()
    gm = GpuModel.create(sht, cfg, src, names, state, view, batch=batch, grid=grid,
                         spectral=spectral, fold=fold)
    gm.set_params(params)
    return sht, gm


def test_A_linear_mode_recurrence(gpu):
    """analyticChecks.ts:73-122"""
    lmax = 15
    nlat, nphi = O.grid_for_lmax(lmax, 1)
    c, D, dt, nsteps = -0.3, 0.01, 0.1, 20
    sht, gm = make("linear", lmax, nlat, nphi, {"c": c, "D": D, "dt": dt})
    nlm = O.nlm_calc(lmax, lmax)
    l, m = 5, 2
    idx = O.lm_index(lmax, l, m)
    U0 = np.zeros(2 * nlm, np.float32)
    U0[2 * idx], U0[2 * idx + 1] = 0.8, -0.35
    gm.upload("U", U0)
    gm.step(nsteps)
    U = gm.read("U")
    g = (1 + dt * c) / (1 + dt * D * l * (l + 1))
    f = g ** nsteps
    err = max(abs(U[2 * idx] - 0.8 * f), abs(U[2 * idx + 1] + 0.35 * f))
    mask = np.ones(2 * nlm, bool)
    mask[2 * idx:2 * idx + 2] = False
    leak = np.max(np.abs(U[mask]))
    emit(f"A: err {err:.2e} leak {leak:.2e}")
    assert err < 2e-6 and leak < 2e-6
    gm.destroy()
    sht.destroy()


def test_B_logistic_uniform_map(gpu):
    """analyticChecks.ts:124-181"""
    lmax = 15
    nlat, nphi = O.grid_for_lmax(lmax, 3)
    r, D, dt, nsteps, u0 = 0.7, 0.01, 0.05, 25, 0.3
    sht, gm = make("logistic", lmax, nlat, nphi, {"r": r, "D": D, "dt": dt})
    gm.init(np.full(nlat * nphi, u0, np.float32))
    Ustart = gm.read("U")
    gm.step(nsteps)
    Uend = gm.read("U")
    got = u0 * (Uend[0] / Ustart[0])
    want = u0
    for _ in range(nsteps):
        want += dt * r * want * (1 - want)
    u = gm.read("u")
    spread = float(u.max() - u.min())
    emit(f"B: err {abs(got - want):.2e} spread {spread:.2e}")
    assert abs(got - want) < 5e-6 and spread < 1e-6
    gm.destroy()
    sht.destroy()


def test_C_schnakenberg_linearised(gpu):
    """analyticChecks.ts:183-256"""
    p = DEFAULTS["schnakenberg"]
    lmax = 31
    nlat, nphi = O.grid_for_lmax(lmax, 3)
    sht, gm = make("schnakenberg", lmax, nlat, nphi, p)
    gm.init(np.zeros(nlat * nphi, np.float32))
    l, m, eps = 24, 7, 1e-6
    idx = O.lm_index(lmax, l, m)
    U0, V0 = gm.read("U"), gm.read("V")
    Up = U0.copy()
    Up[2 * idx] += eps
    gm.upload("U", Up)
    gm.upload("V", V0)
    nsteps = 40
    gm.step(nsteps)
    U, V = gm.read("U"), gm.read("V")
    us = p["a"] + p["b"]
    vs = p["b"] / (us * us)
    J = [[-1 + 2 * us * vs, us * us], [-2 * us * vs, -us * us]]
    lam = l * (l + 1)
    du = 1 / (1 + p["dt"] * p["D1"] * lam)
    dv = 1 / (1 + p["dt"] * p["D2"] * lam)
    cu, cv = eps, 0.0
    for _ in range(nsteps):
        cu, cv = ((cu + p["dt"] * (J[0][0] * cu + J[0][1] * cv)) * du,
                  (cv + p["dt"] * (J[1][0] * cu + J[1][1] * cv)) * dv)
    gu = U[2 * idx] - U0[2 * idx]
    gv = V[2 * idx] - V0[2 * idx]
    ru = abs(gu - cu) / max(abs(cu), 1e-30)
    rv = abs(gv - cv) / max(abs(cv), 1e-30)
    emit(f"C: rel ({ru:.2e}, {rv:.2e}) growth {abs(cu) / eps:.3f}x")
    assert ru < 5e-3 and rv < 5e-3
    assert abs(cu) > eps and abs(gu) > eps
    gm.destroy()
    sht.destroy()


@pytest.mark.parametrize("key", ["schnakenberg", "brusselator", "allencahn"])
def test_model_checks(gpu, key):
    """modelChecks.ts:26-77: compiles, fusion guard, 40 stable steps with
    contrast, at lmax 31 on the dealiased grid."""
    lmax = 31
    nlat, nphi = O.grid_for_lmax(lmax, 3)
    sht, gm = make(key, lmax, nlat, nphi, DEFAULTS[key])
    # the fusion guard counts generated element-wise statements (one per .m
    # line); the batched plan runs them as fewer, fused launches (a grid op
    # runs those between the transforms inside its grid kernel)
    statements = gm.step_plan.elementwise_statements()
    assert statements == {"schnakenberg": 5, "brusselator": 5, "allencahn": 2}[key]
    assert gm.step_plan.generated_kernels() <= statements
    noise = (0.01 * O.randn(1, nlat * nphi)).astype(np.float32)
    gm.init(noise)
    gm.step(40)
    for field in MODELS[key][2]:
        v = gm.read(field)
        contrast = float(v.max() - v.min())
        emit(f"{key}.{field}: finite {np.all(np.isfinite(v))} contrast {contrast:.3e}")
        assert np.all(np.isfinite(v)) and contrast > 1e-6
    gm.destroy()
    sht.destroy()


def test_graph_replay_matches_eager(gpu):
    """A captured step replayed several times (with eager work and syncs in
    between) gives exactly the eagerly recorded result every time."""
    p = DEFAULTS["brusselator"]
    lmax = 31
    nlat, nphi = O.grid_for_lmax(lmax, 3)
    sht, gm = make("brusselator", lmax, nlat, nphi, p)
    gm.init((0.01 * O.randn(5, nlat * nphi)).astype(np.float32))
    gm.snapshot_state()
    gm._step.encode_steps(gm.stream, 7)
    eager = gm.read("U")
    for rep in range(5):
        gm.restore_state()
        # 0: graphs of 4, 2 and 1 steps; 1, 3: one graph replayed, rest eager
        gm.graph_steps = (1, 3, 0, 0, 3)[rep]
        if rep == 3:
            gm.step(3)
            gm.step(4)
        else:
            gm.step(7)
        got = gm.read("U")
        emit(f"replay {rep}: max diff {np.max(np.abs(got - eager)):.1e}")
        assert np.array_equal(got, eager)
    gm.destroy()
    sht.destroy()


def test_parameter_change_without_recompile(gpu):
    """A parameter change re-records the launches only; results follow it."""
    lmax = 15
    nlat, nphi = O.grid_for_lmax(lmax, 1)
    sht, gm = make("linear", lmax, nlat, nphi, {"c": -0.3, "D": 0.01, "dt": 0.1})
    nlm = O.nlm_calc(lmax, lmax)
    U0 = np.zeros(2 * nlm, np.float32)
    U0[0] = 1.0
    module = gm._step._module
    for c in (-0.3, 0.5):
        gm.set_params({"c": c, "D": 0.01, "dt": 0.1})
        gm.upload("U", U0)
        gm.step(3)
        U = gm.read("U")
        assert abs(U[0] - (1 + 0.1 * c) ** 3) < 2e-6
    assert gm._step._module is module
    gm.destroy()
    sht.destroy()


def test_plain_method_trajectory_vs_f64(gpu):
    """Plain method: Schnakenberg, lmax 63 on 64x128, 1000 IMEX steps, against
    the f64 oracle's trajectory from the same f32 initial state."""
    lmax, nlat, nphi = 63, 64, 128
    p = {"a": 0.1, "b": 0.9, "D1": 6.5e-4, "D2": 40 * 6.5e-4, "dt": 0.01}
    npts = nlat * nphi
    r = O.mulberry32(0, 2 * npts)
    u0 = (1.0 + (0.02 * r[:npts] - 0.01)).astype(np.float32)
    v0 = (0.9 + (0.02 * r[npts:] - 0.01)).astype(np.float32)
    cfg = O.Cfg(lmax, lmax, nlat, nphi)
    ref = O.ShtReference(cfg)
    U0 = ref.analys_fast(u0.astype(np.float64)).astype(np.float32)
    V0 = ref.analys_fast(v0.astype(np.float64)).astype(np.float32)
    sht, gm = make("schnakenberg", lmax, nlat, nphi, p)
    gm.upload("U", U0)
    gm.upload("V", V0)
    gm.graph_steps = 10
    gm.step(1000)
    U, V = gm.read("U"), gm.read("V")
    emit("f64 trajectory (1000 steps)...")
    Ur, Vr = O.imex_trajectory(cfg, O.SCHNAKENBERG, p, U0, V0, 1000, ref)
    eu, ev = O.rel_l2(U, Ur), O.rel_l2(V, Vr)
    emit(f"plain method: relL2 U {eu:.2e} V {ev:.2e}")
    assert eu < 1e-4 and ev < 1e-4
    gm.destroy()
    sht.destroy()


@pytest.mark.parametrize("key", ["schnakenberg", "brusselator", "allencahn"])
def test_fused_plan_matches_statementwise(gpu, key):
    """The batched plan (transforms of all species in one launch, each run of
    element-wise lines fused into one kernel) gives the same state, bit for
    bit, as the one-kernel-per-line plan of the reference's shape: the same
    f32 operations in the same order, only fewer launches."""
    lmax = 31
    nlat, nphi = O.grid_for_lmax(lmax, 3)
    emit(f"fused vs per-line {key}")
    outs = []
    for batch in (False, True):
        sht, gm = make(key, lmax, nlat, nphi, DEFAULTS[key], batch=batch)
        rng = np.random.default_rng(4)
        for s in gm.state:
            q = (rng.normal(size=2 * sht.nlm) * 0.02).astype(np.float32)
            q[0] += 1.0
            gm.upload(s, q)
        gm.step(25)
        outs.append([gm.read(s) for s in gm.state] + [gm.read(v) for v in gm.view])
        gm.destroy()
        sht.destroy()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("groups", [True, False])
@pytest.mark.parametrize("key,lmax,nlat,nphi", [
    ("brusselator", 31, 32, 64), ("schnakenberg", 15, 17, 32), ("allencahn", 63, 64, 128),
    ("brusselator", 127, 128, 256), ("schnakenberg", 255, 256, 512),
    ("brusselator", 700, 701, 4096)])
def test_grid_fused_step_matches_separate_stages(gpu, monkeypatch, key, lmax, nlat, nphi,
                                                 groups):
    """Grid fusion (Legendre synthesis, one generated kernel running the row
    FFT synthesis, the element-wise statements and the row FFT analysis,
    Legendre analysis) gives the state and the returned grid values of the
    plan with separate Fourier stages and element-wise kernel, bit for bit:
    the row FFT is the library's own code (shared header), the statements
    the same f32 operations.  Both kernel forms (one thread group per field's
    row FFT, or one group transforming the rows in turn); several rows per
    workgroup (small nphi), an odd latitude count, and 2048-point rows with
    mmax well below the row's Nyquist index."""
    import tsphere.hipgen

    monkeypatch.setattr(tsphere.hipgen, "GRID_GROUPS", groups)
    emit(f"grid fusion {key} lmax {lmax} {nlat}x{nphi} groups {groups}")
    outs = []
    for grid in (False, True):
        sht, gm = make(key, lmax, nlat, nphi, DEFAULTS[key], grid=grid)
        kinds = [l.split()[0] for l in gm.describe()["step"]]
        assert ("grid" in kinds) == grid, kinds
        rng = np.random.default_rng(7)
        for s in gm.state:
            q = (rng.normal(size=2 * sht.nlm) * 0.02).astype(np.float32)
            q[0] += 1.0
            gm.upload(s, q)
        gm.step(20)
        outs.append([gm.read(s) for s in gm.state] + [gm.read(v) for v in gm.view])
        gm.destroy()
        sht.destroy()
    for a, b in zip(*outs):
        assert np.all(np.isfinite(b))
        d = float(np.max(np.abs(a - b)) / max(1e-30, float(np.max(np.abs(a)))))
        emit(f"  max rel diff {d:.2e}")
        assert np.array_equal(a, b), d


@pytest.mark.parametrize("key,lmax,nlat,nphi,legendre", [
    ("brusselator", 31, 32, 64, "fp32"), ("schnakenberg", 15, 17, 32, "fp32"),
    ("allencahn", 63, 64, 128, "fp32"), ("brusselator", 127, 128, 256, "f64rec"),
    ("schnakenberg", 255, 256, 512, "fp32"), ("brusselator", 255, 257, 512, "f64rec"),
    ("brusselator", 31, 32, 64, "bf16x3"), ("schnakenberg", 15, 17, 32, "bf16x3"),
    ("allencahn", 127, 128, 256, "bf16x3"), ("schnakenberg", 255, 257, 512, "bf16x3")])
def test_spectral_op_matches_grid_plan(gpu, key, lmax, nlat, nphi, legendre):
    """The spectral op (Legendre analysis, the implicit solve and the next
    step's Legendre synthesis in one kernel per step) gives the state and the
    returned grid values of the plan that runs them as separate launches,
    bit for bit: the same recurrence values, the same order of every sum and
    the implicit step rounded as the generated kernel rounds it.  Steps run
    as graph replays and as directly encoded remainders, over several calls
    (each call re-synthesises from the state, so a state upload between
    calls is honoured)."""
    emit(f"spectral op {key} lmax {lmax} {nlat}x{nphi} {legendre}")
    outs = []
    # separate launches; the implicit solve folded into the next synthesis;
    # the spectral op
    for spectral, fold in ((False, False), (False, True), (True, False)):
        sht, gm = make(key, lmax, nlat, nphi, DEFAULTS[key], spectral=spectral,
                       legendre=legendre, fold=fold)
        kinds = [l.split()[0] for l in gm.describe()["step"]]
        assert ("spectral" in kinds) == spectral and ("imex" in kinds) == fold, kinds
        gm.graph_steps = 4
        rng = np.random.default_rng(11)
        for s in gm.state:
            q = (rng.normal(size=2 * sht.nlm) * 0.02).astype(np.float32)
            q[0] += 1.0
            gm.upload(s, q)
        gm.step(9)
        gm.step(3)
        mid = [gm.read(s) for s in gm.state]
        gm.upload(gm.state[0], mid[0] * np.float32(0.5))
        gm.step(6)
        outs.append(mid + [gm.read(s) for s in gm.state] + [gm.read(v) for v in gm.view])
        gm.destroy()
        sht.destroy()
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.all(np.isfinite(b))
            d = float(np.max(np.abs(a - b)) / max(1e-30, float(np.max(np.abs(a)))))
            emit(f"  max rel diff {d:.2e}")
            assert np.array_equal(a, b), d


FORCED_SRC = open(os.path.join(ROOT, "tests", "models", "forced.m")).read()


@pytest.mark.parametrize("groups", [True, False])
def test_grid_fusion_external_grid_input(gpu, monkeypatch, groups):
    """A grid op whose statements read a grid value from a host buffer (the
    noise field) and which analyses a synthesised value directly, next to a
    produced one, matches the separate stages bit for bit."""
    import tsphere
    import tsphere.hipgen
    from tsphere.model import GpuModel

    monkeypatch.setattr(tsphere.hipgen, "GRID_GROUPS", groups)
    cfg = tsphere.ShtConfig(31, 31, 33, 64)
    outs = []
    for grid in (False, True):
        sht = tsphere.ShtPlan.create(cfg, max_fields=2)
        gm = GpuModel.create(sht, cfg, FORCED_SRC, ["c", "D", "dt"], ["U"], ["u"], grid=grid)
        assert ("grid" in [l.split()[0] for l in gm.describe()["step"]]) == grid
        gm.set_params({"c": 0.3, "D": 1e-3, "dt": 0.05})
        gm.init((0.5 + 0.1 * O.randn(9, cfg.npts)).astype(np.float32))
        gm.step(15)
        outs.append([gm.read("U"), gm.read("u")])
        gm.destroy()
        sht.destroy()
    for a, b in zip(*outs):
        assert np.all(np.isfinite(b)) and np.array_equal(a, b)
