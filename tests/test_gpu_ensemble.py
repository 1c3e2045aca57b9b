"""Ensembles (many runs of one model per launch), mixed-lmax ensembles in one
HIP graph, and Legendre tile shapes / autotuning.

Every ensemble run must reproduce the same run done alone by the
single-model path (stated tolerance 2e-6 relative: the batched launch may
group the fields into different workgroup shapes, which can reorder the
analysis reduction by an ulp), and a run of the mixed ensemble must track
the f64 oracle trajectory like a single run does."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import emit

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
AC_SRC = open(os.path.join(ROOT, "turing-sphere_amd", "mfiles", "allencahn.m")).read()
SCH_SRC = open(os.path.join(ROOT, "turing-sphere_amd", "mfiles", "schnakenberg.m")).read()
TOL_RUN = 2e-6


def _noise(npts, seed):
    return (0.2 * O.mulberry32(seed, npts) - 0.1).astype(np.float32)


def _single_allencahn(cfg, eps2, dt, noise, steps):
    import tsphere
    from tsphere.model import GpuModel

    sht = tsphere.ShtPlan.create(cfg, max_fields=1)
    gm = GpuModel.create(sht, cfg, AC_SRC, ["eps2", "dt"], ["U"], ["u"])
    gm.set_params({"eps2": eps2, "dt": dt})
    gm.init(noise)
    gm.step(steps)
    U = gm.read("U")
    gm.destroy()
    sht.destroy()
    return U


def test_ensemble_matches_single_runs(gpu):
    import tsphere
    from tsphere.model import GpuModel

    cfg = tsphere.ShtConfig(31, 31, 32, 64)
    R, steps, dt = 5, 30, 2e-3
    eps2 = np.array([1e-4, 4e-4, 9e-4, 1.6e-3, 2.5e-3])
    emit(f"ensemble allencahn lmax31 runs {R}")
    sht = tsphere.ShtPlan.create(cfg, max_fields=R)
    gm = GpuModel.create(sht, cfg, AC_SRC, ["eps2", "dt"], ["U"], ["u"], runs=R)
    gm.set_params({"eps2": eps2, "dt": dt})
    noise = [_noise(cfg.npts, r) for r in range(R)]
    gm.init(np.concatenate(noise))
    gm.graph_steps = 7
    gm.step(steps)
    U = gm.read("U").reshape(R, -1)
    for r in range(R):
        Ur = _single_allencahn(cfg, eps2[r], dt, noise[r], steps)
        e = O.rel_l2(U[r], Ur)
        emit(f"  run {r}: vs single {e:.2e}")
        assert e < TOL_RUN
    # a per-run parameter change applies without recompiling or re-capturing
    gm.set_params({"eps2": eps2[::-1].copy(), "dt": dt})
    gm.init(np.concatenate(noise))
    gm.step(steps)
    U2 = gm.read("U").reshape(R, -1)
    assert O.rel_l2(U2[0], _single_allencahn(cfg, eps2[-1], dt, noise[0], steps)) < TOL_RUN
    gm.destroy()
    sht.destroy()


def test_two_species_ensemble(gpu):
    """Schnakenberg, 3 runs x 2 species = 6 fields per batched transform."""
    import tsphere
    from tsphere.model import GpuModel

    cfg = tsphere.ShtConfig(15, 15, 16, 32)
    R, steps = 3, 25
    names = ["a", "b", "D1", "D2", "dt"]
    p = {"a": np.array([0.1, 0.12, 0.08]), "b": 0.9, "D1": 4e-4,
         "D2": np.array([8e-3, 6e-3, 1e-2]), "dt": 0.05}
    emit("ensemble schnakenberg lmax15 runs 3")
    sht = tsphere.ShtPlan.create(cfg, max_fields=2 * R)
    gm = GpuModel.create(sht, cfg, SCH_SRC, names, ["U", "V"], ["u", "v"], runs=R)
    gm.set_params(p)
    rng = np.random.default_rng(3)
    U0 = [rng.normal(size=2 * sht.nlm).astype(np.float32) * 0.01 for _ in range(R)]
    V0 = [rng.normal(size=2 * sht.nlm).astype(np.float32) * 0.01 for _ in range(R)]
    for r in range(R):
        U0[r][0] += 1.0
        V0[r][0] += 0.9
    gm.upload("U", np.concatenate(U0))
    gm.upload("V", np.concatenate(V0))
    gm.step(steps)
    U = gm.read("U").reshape(R, -1)
    V = gm.read("V").reshape(R, -1)
    ref = O.ShtReference(O.Cfg(15, 15, 16, 32))
    for r in range(R):
        pr = {k: (float(v[r]) if np.ndim(v) else float(v)) for k, v in p.items()}
        Ur, Vr = O.imex_trajectory(O.Cfg(15, 15, 16, 32), O.SCHNAKENBERG, pr, U0[r], V0[r],
                                   steps, ref)
        eu, ev = O.rel_l2(U[r], Ur), O.rel_l2(V[r], Vr)
        emit(f"  run {r}: U {eu:.2e} V {ev:.2e} vs f64")
        assert eu < 1e-5 and ev < 1e-5
    gm.destroy()
    sht.destroy()


@pytest.mark.parametrize("legendre", ["fp32", "bf16x3"])
def test_mixed_lmax_ensemble_one_graph(gpu, legendre):
    """Groups at lmax 15/31/63 stepped as parallel branches of one graph;
    each run equals its single-run counterpart (fp32) or tracks f64 (both)."""
    import tsphere
    from tsphere.ensemble import MixedEnsemble

    groups = [(tsphere.ShtConfig(L, L, L + 1, 2 * (L + 1)), n) for L, n in
              ((15, 3), (31, 2), (63, 3))]
    emit(f"mixed ensemble {legendre} groups {[(g.lmax, n) for g, n in groups]}")
    ens = MixedEnsemble.create(AC_SRC, ["eps2", "dt"], ["U"], ["u"], groups,
                               legendre=legendre)
    R = ens.runs
    eps2 = np.linspace(2e-4, 2e-3, R)
    dt, steps = 2e-3, 23
    ens.set_params({"eps2": eps2, "dt": dt})
    noise = [_noise(g.cfg.npts, g.first_run + r) for g in ens.groups for r in range(g.runs)]
    ens.init(noise)
    U0 = ens.read("U")
    ens.graph_steps = 10
    ens.step(steps)
    U = ens.read("U")
    for g in ens.groups:
        c = g.cfg
        ref = O.ShtReference(O.Cfg(c.lmax, c.mmax, c.nlat, c.nphi))
        for r in (0, g.runs - 1):
            i = g.first_run + r
            Ur, _ = O.imex_trajectory(O.Cfg(c.lmax, c.mmax, c.nlat, c.nphi), O.ALLENCAHN,
                                      {"eps2": eps2[i], "dt": dt}, U0[i], None, steps, ref)
            e = O.rel_l2(U[i], Ur)
            msg = f"  lmax {c.lmax} run {i}: vs f64 {e:.2e}"
            if legendre == "fp32":
                es = O.rel_l2(U[i], _single_allencahn(c, eps2[i], dt, noise[i], steps))
                msg += f" vs single {es:.2e}"
                assert es < TOL_RUN
            emit(msg)
            assert e < 2e-5
    ens.destroy()


def test_tile_shapes_and_tuning(gpu, tmp_path, monkeypatch):
    """Every candidate FP32 tile gives the built-in tile's result (the tile
    only regroups independent work); invalid tiles are refused; the tuner
    picks one and a second plan of the same shape reads it from the cache."""
    import tsphere

    monkeypatch.setenv("TSPHERE_TUNE_CACHE", str(tmp_path / "tiles.json"))
    cfg = tsphere.ShtConfig(127, 127, 128, 256)
    nf = 3
    plan = tsphere.ShtPlan.create(cfg, max_fields=nf)
    q = tsphere.DeviceBuffer(nf * 2 * plan.nlm)
    sp = tsphere.DeviceBuffer(nf * cfg.npts)
    back = tsphere.DeviceBuffer(nf * 2 * plan.nlm)
    rng = np.random.default_rng(1)
    qh = np.concatenate([O_spectrum(cfg, rng) for _ in range(nf)])
    q.upload(qh, plan.stream)

    def run():
        plan.encode_synth_into(plan.stream, plan.create_synth_binding(q, sp, nf, 2 * plan.nlm,
                                                                      cfg.npts))
        plan.encode_analys_into(plan.stream, plan.create_analys_binding(sp, back, nf, cfg.npts,
                                                                        2 * plan.nlm))
        return sp.download(plan.stream), back.download(plan.stream)

    s0, b0 = run()
    for nfb in (1, 2, 4):
        for nw, k, rows in ((1, 1, 2), (2, 1, 2), (1, 4, 2), (4, 1, 2), (2, 1, 1), (1, 4, 1)):
            for stage in ("synth", "analys"):
                plan.set_tile(stage, nf=nfb, nwaves=nw, k=k, impl=1, rows=rows)
            s1, b1 = run()
            es, eb = O.rel_l2(s1, s0), O.rel_l2(b1, b0)
            emit(f"  tile nf={nfb} nwaves={nw} k={k} rows={rows}: synth {es:.1e} "
                 f"analys {eb:.1e}")
            assert es < 1e-6 and eb < 1e-6
    with pytest.raises(tsphere.TsphereError):
        plan.set_tile("synth", nf=3)
    with pytest.raises(tsphere.TsphereError):
        plan.set_tile("analys", k=3)
    with pytest.raises(tsphere.TsphereError):
        plan.set_tile("synth", nwaves=9)
    with pytest.raises(tsphere.TsphereError):
        plan.set_tile("synth", rows=3)
    res = plan.tune(nf, reps=3)
    emit(f"  tuned: {res}")
    assert not res["cached"] and res["ms"][0] > 0
    s2, b2 = run()
    assert O.rel_l2(s2, s0) < 1e-6 and O.rel_l2(b2, b0) < 1e-6
    plan2 = tsphere.ShtPlan.create(cfg, max_fields=nf)
    res2 = plan2.tune(nf)
    assert res2["cached"] and res2["synth"] == res["synth"]
    assert plan2.get_tile("synth", nf) == res["synth"]
    plan2.destroy()
    plan.destroy()


def O_spectrum(cfg, rng):
    from test_gpu_sht import _random_field_spectrum

    return _random_field_spectrum(O.Cfg(cfg.lmax, cfg.mmax, cfg.nlat, cfg.nphi),
                                  int(rng.integers(1 << 30)))


def test_ensemble_grid_fusion_bit_exact(gpu):
    """The ensemble form of grid fusion (one grid kernel over every run's
    rows, per-run scalars from the device array) gives the ensemble plan with
    separate Fourier stages bit for bit."""
    import tsphere
    from tsphere.model import GpuModel

    cfg = tsphere.ShtConfig(31, 31, 32, 64)
    R, steps = 3, 20
    names = ["a", "b", "D1", "D2", "dt"]
    p = {"a": np.array([0.1, 0.12, 0.08]), "b": 0.9, "D1": 4e-4,
         "D2": np.array([8e-3, 6e-3, 1e-2]), "dt": 0.05}
    emit("ensemble grid fusion schnakenberg lmax31 runs 3")
    outs = []
    for grid in (False, True):
        sht = tsphere.ShtPlan.create(cfg, max_fields=2 * R)
        gm = GpuModel.create(sht, cfg, SCH_SRC, names, ["U", "V"], ["u", "v"], runs=R,
                             grid=grid)
        assert ("grid" in [l.split()[0] for l in gm.describe()["step"]]) == grid
        gm.set_params(p)
        rng = np.random.default_rng(5)
        U0 = rng.normal(size=2 * sht.nlm * R).astype(np.float32) * 0.01
        V0 = rng.normal(size=2 * sht.nlm * R).astype(np.float32) * 0.01
        gm.upload("U", U0 + np.tile(np.eye(1, 2 * sht.nlm, 0, np.float32)[0], R))
        gm.upload("V", V0 + 0.9 * np.tile(np.eye(1, 2 * sht.nlm, 0, np.float32)[0], R))
        gm.graph_steps = 5
        gm.step(steps)
        outs.append([gm.read(n) for n in ("U", "V", "u", "v")])
        gm.destroy()
        sht.destroy()
    for a, b in zip(*outs):
        assert np.all(np.isfinite(b))
        emit(f"  max diff {float(np.max(np.abs(a - b))):.1e}")
        assert np.array_equal(a, b)


@pytest.mark.parametrize("legendre,impl", [("fp32", 1), ("fp32", 3), ("fp32", 5), ("f64rec", 3),
                                           ("f64rec", 5), ("bf16x3", None)])
@pytest.mark.parametrize("model,R", [("schnakenberg", 3), ("allencahn", 9)])
def test_ensemble_fold_bit_exact(gpu, legendre, impl, model, R):
    """An ensemble step folds the implicit solve into the analysis
    (tsphere_leg_analys_imex_fields, each run's s1 and s2 from a device
    array; the 4x4-block and per-parity product analysis tiles, impl 3 and
    5, apply it as they form each analysed value): equal to the ensemble plan with the separate generated solve
    kernel bit for bit, over graph replays, an eager remainder and a per-run
    parameter change between calls (an upload, no re-capture).  Nine
    single-species runs put a partial group of fields in the last
    workgroup."""
    import tsphere
    from tsphere.model import GpuModel

    cfg = tsphere.ShtConfig(31, 31, 32, 64)
    if model == "schnakenberg":
        src, names, state, view = SCH_SRC, ["a", "b", "D1", "D2", "dt"], ["U", "V"], ["u", "v"]
        p1 = {"a": np.array([0.1, 0.12, 0.08]), "b": 0.9, "D1": 4e-4,
              "D2": np.array([8e-3, 6e-3, 1e-2]), "dt": 0.05}
        p2 = dict(p1, D2=np.array([1e-2, 8e-3, 6e-3]), dt=0.04)
    else:
        src, names, state, view = AC_SRC, ["eps2", "dt"], ["U"], ["u"]
        p1 = {"eps2": np.linspace(1e-4, 2.5e-3, R), "dt": 2e-3}
        p2 = {"eps2": np.linspace(2.5e-3, 1e-4, R), "dt": 3e-3}
    emit(f"ensemble fold {model} {legendre} analysis tile {impl} lmax31 runs {R}")
    outs = []
    for fold in (False, True):
        sht = tsphere.ShtPlan.create(cfg, legendre=legendre, max_fields=len(state) * R)
        if impl is not None:
            sht.set_tile("analys", impl=impl)
        gm = GpuModel.create(sht, cfg, src, names, state, view, runs=R, fold=fold)
        assert ("imex" in [l.split()[0] for l in gm.describe()["step"]]) == fold
        gm.set_params(p1)
        rng = np.random.default_rng(11)
        for k, n in enumerate(state):
            x = rng.normal(size=2 * sht.nlm * R).astype(np.float32) * 0.01
            gm.upload(n, x + (1.0 - 0.1 * k) * np.tile(np.eye(1, 2 * sht.nlm, 0, np.float32)[0], R))
        gm.graph_steps = 5
        gm.step(12)
        gm.set_params(p2)
        gm.step(7)
        outs.append([gm.read(n) for n in state + view])
        gm.destroy()
        sht.destroy()
    for a, b in zip(*outs):
        assert np.all(np.isfinite(b))
        emit(f"  max diff {float(np.max(np.abs(a - b))):.1e}")
        assert np.array_equal(a, b)
