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
    ref = O.ShtReference(O.Cfg(c.lmax, c.mmax, c.nlat, c.nphi)).synth_fast(U.astype(np.float64))
    e = O.rel_l2(fine, ref)
    emit(f"  fine grid {c.nlat}x{c.nphi}: vs f64 {e:.2e}")
    assert e < 3e-6
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
