"""ModelSession on the GPU (src/mgpu/session.ts): seeding, stepping, the
state-preserving measure(), display oversampling through synth_from on a
finer plan, and the reproducible run record of scripts/bench.ts."""
import json

import numpy as np
import pytest

import oracle as O
from conftest import emit

pytestmark = pytest.mark.gpu


def test_session_seed_step_measure(gpu):
    from tsphere.session import ModelSession, resolve_preset

    _, model, params = resolve_preset("schnak-spots")
    emit("session schnak-spots lmax 31")
    s = ModelSession.create(model, params, 31)
    s.seed(1)
    s.step(20)
    U = s.read("U")
    assert s.steps == 20 and s.t == pytest.approx(20 * params["dt"])
    ms = s.measure(10)
    assert ms > 0
    assert np.array_equal(s.read("U"), U)        # measure() leaves the state where it was
    assert s.steps == 20
    s.seed(1)
    s.step(20)
    assert np.array_equal(s.read("U"), U)        # seeding is reproducible
    s.destroy()


@pytest.mark.parametrize("os_", [2, 3])
def test_display_oversampling(gpu, os_):
    """read_species on an os x finer grid = the f64 synthesis of the state
    there (session.ts:195-207); switching back reads the solver's own field."""
    from tsphere.session import ModelSession, resolve_preset

    _, model, params = resolve_preset("allencahn")
    emit(f"oversample x{os_}")
    s = ModelSession.create(model, params, 31, oversample=os_)
    s.seed(3)
    s.step(5)
    fine = s.read_species(0)
    U = s.read("U")
    c = s.view_sht.cfg
    assert fine.size == c.nlat * c.nphi == os_ * os_ * s.cfg.nlat * s.cfg.nphi
    # test/modelChecks.ts:82-119: the GPU-to-GPU coefficient copy against the
    # finer plan's own upload path (read the state back, synthesise it), to
    # 1e-6 absolute
    expected = s.view_sht.synth(U)
    dmax = float(np.max(np.abs(fine - expected)))
    ref = O.ShtReference(O.Cfg(c.lmax, c.mmax, c.nlat, c.nphi)).synth_fast(U.astype(np.float64))
    e = O.rel_l2(fine, ref)
    emit(f"  fine grid {c.nlat}x{c.nphi}: vs upload path max |diff| {dmax:.1e}, vs f64 {e:.2e}")
    assert dmax <= 1e-6
    assert e < 2e-6
    s.set_oversample(1)
    coarse = s.read_species(0)
    assert np.array_equal(coarse, s.read("u"))
    s.destroy()


def test_run_record_and_dump(gpu, tmp_path):
    from tsphere.session import parse_args, run

    spec = parse_args(["--preset", "brussel", "--lmax", "15", "--steps", "40", "--warmup", "5",
                       "--seed", "2"])
    emit("run record brussel lmax 15")
    dump = tmp_path / "state.json"
    rec = run(spec, batch=8, latency_steps=10, dump_state=str(dump))
    for k in ("command", "spec", "model", "backend", "grid", "compiled", "digest", "throughput",
              "latency", "state"):
        assert k in rec
    # 5 element-wise lines: the reaction run is one generated grid kernel, the
    # IMEX run is folded into the next step's synthesis (a library kernel)
    assert rec["compiled"]["elementwiseStatements"] == 5 and rec["compiled"]["kernels"] == 1
    assert rec["state"]["finite"]
    d = json.loads(dump.read_text())
    assert d["digest"]["n"] == len(d["state"]) == 2 * rec["grid"]["nlm"]
    rec2 = run(spec, batch=8, latency_steps=10)
    assert rec2["digest"] == rec["digest"]       # same spec, same state


def test_seeded_default_run_vs_f64_trajectory(gpu):
    """scripts/compare-native.mjs --check (:85-91, :290-322): the default run
    (runSpec defaults: preset schnak-spots, lmax 63 on its dealiased grid,
    seed 1) after 20 steps against the f64 trajectory from the same seeded
    start, within the reference's 2e-3 relative L2.  The oracle restates the
    seeding independently (mulberry32 + Box-Muller, noise.ts:15-51; the .m
    init U = analys(a + b + noise), V = analys(vs * ones), in f64).  As
    compare-native does, the two sides must first be the same problem: grid
    and parameters."""
    from tsphere.session import ModelSession, config_for_spec, parse_args, resolve_preset

    spec = parse_args([])
    assert (spec.preset, spec.lmax, spec.seed) == ("schnak-spots", 63, 1)
    _, model, params = resolve_preset(spec.preset)
    cfg = config_for_spec(spec)
    nsteps = 20
    emit(f"default run vs f64: {spec.preset} lmax {spec.lmax} {cfg.nlat}x{cfg.nphi} "
         f"seed {spec.seed}, {nsteps} steps")
    s = ModelSession.create(model, params, spec.lmax)
    assert (s.cfg.nlat, s.cfg.nphi) == (cfg.nlat, cfg.nphi)
    s.seed(spec.seed)
    s.step(nsteps)
    U, V = s.read("U"), s.read("V")
    s.destroy()

    ocfg = O.Cfg(spec.lmax, spec.lmax, cfg.nlat, cfg.nphi)
    assert (ocfg.nlat, ocfg.nphi) == O.grid_for_lmax(spec.lmax, 3)
    ref = O.ShtReference(ocfg)
    npts = ocfg.nlat * ocfg.nphi
    noise = model.seed_amp * O.randn(int(spec.seed), npts)
    us = params["a"] + params["b"]
    vs = params["b"] / (us * us)
    U0 = ref.analys_fast(us + noise)
    V0 = ref.analys_fast(np.full(npts, vs))
    Ur, Vr = O.imex_trajectory(ocfg, O.SCHNAKENBERG, params, U0, V0, nsteps, ref)
    eu, ev = O.rel_l2(U, Ur), O.rel_l2(V, Vr)
    emit(f"  relL2 U {eu:.2e} V {ev:.2e}")
    assert eu < 2e-3 and ev < 2e-3


def test_long_run_schnakenberg_saturates(gpu):
    """scripts/longrun-node.ts:25-49: Schnakenberg (defaults, seed 1, lmax 31)
    to t = 100 saturates into O(1)-contrast spots: u's contrast > 0.5 and
    max < 10 at the end, finite throughout (read every 50 steps)."""
    from tsphere.session import ModelSession, m_model_by_key, default_params

    model = m_model_by_key("schnakenberg")
    params = default_params(model)
    nsteps = int(round(100 / params["dt"]))
    emit(f"long run: schnakenberg lmax 31 to t = 100 ({nsteps} steps)")
    s = ModelSession.create(model, params, 31)
    s.seed(1)
    lo = hi = float("nan")
    for k in range(0, nsteps, 50):
        s.step(min(50, nsteps - k))
        u = s.read("u")
        assert np.all(np.isfinite(u)), s.steps
        lo, hi = float(u.min()), float(u.max())
        if s.steps % 400 == 0:
            emit(f"  t={s.t:5.1f} u in [{lo:.4f}, {hi:.4f}] contrast {hi - lo:.4f}")
    s.destroy()
    emit(f"  final contrast {hi - lo:.4f}, max {hi:.4f} after {nsteps} steps")
    assert hi - lo > 0.5 and hi < 10
