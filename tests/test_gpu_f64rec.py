"""The FP64-recurrence form (f64 state and tables, values rounded to f32
once, FP32 sums on the VALU): it removes the FP32 form's dominant error --
the recurrence state's rounding -- at about the FP32 form's cost.  Stated
tolerance: 1e-5 relative vs FP64 per transform at every size up to lmax
1023 (the precision target), 2e-5 per round trip."""
import numpy as np
import pytest

import oracle as O
from conftest import emit
from test_gpu_sht import _random_field_spectrum

pytestmark = pytest.mark.gpu
TOL, TOL_RT = 1e-5, 2e-5


@pytest.mark.parametrize("lmax,nlat,nphi", [(15, 16, 32), (15, 17, 32), (63, 64, 128),
                                            (255, 256, 512), (511, 512, 1024),
                                            (1023, 1024, 2048)])
def test_f64rec_vs_f64(gpu, lmax, nlat, nphi):
    import tsphere

    cfg = O.Cfg(lmax, lmax, nlat, nphi)
    emit(f"f64rec lmax{lmax} {nlat}x{nphi}")
    plan = tsphere.ShtPlan.create(tsphere.ShtConfig(lmax, lmax, nlat, nphi), legendre="f64rec",
                                  max_fields=2)
    assert plan.legendre == "f64rec"
    ref = O.ShtReference(cfg)
    q = _random_field_spectrum(cfg, 23)
    spat = plan.synth(q)
    back = plan.analys(spat)
    e_s = O.rel_l2(spat, ref.synth_fast(q.astype(np.float64)))
    e_a = O.rel_l2(back, ref.analys_fast(spat.astype(np.float64)))
    e_rt = O.rel_l2(back, q)
    emit(f"  synth {e_s:.2e} analys {e_a:.2e} round trip {e_rt:.2e} vs f64")
    assert e_s < TOL and e_a < TOL and e_rt < TOL_RT
    plan.destroy()


def test_f64rec_model_step(gpu):
    """A two-species step through the planner with the FP64-recurrence plan,
    tracking the f64 trajectory."""
    import os

    import tsphere
    from tsphere.model import GpuModel

    lmax = 127
    cfg = tsphere.ShtConfig(lmax, lmax, 128, 256)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = open(os.path.join(root, "turing-sphere_amd", "mfiles", "schnakenberg.m")).read()
    sht = tsphere.ShtPlan.create(cfg, legendre="f64rec", max_fields=2)
    p = {"a": 0.1, "b": 0.9, "D1": 4e-4, "D2": 8e-3, "dt": 0.02}
    gm = GpuModel.create(sht, cfg, src, list(p), ["U", "V"], ["u", "v"])
    gm.set_params(p)
    rng = np.random.default_rng(9)
    U0 = (rng.normal(size=2 * sht.nlm) * 0.01).astype(np.float32)
    V0 = (rng.normal(size=2 * sht.nlm) * 0.01).astype(np.float32)
    U0[0] += 1.0 * np.sqrt(4 * np.pi)
    V0[0] += 0.9 * np.sqrt(4 * np.pi)
    gm.upload("U", U0)
    gm.upload("V", V0)
    gm.step(50)
    oc = O.Cfg(lmax, lmax, 128, 256)
    Ur, Vr = O.imex_trajectory(oc, O.SCHNAKENBERG, p, U0, V0, 50, O.ShtReference(oc))
    eu, ev = O.rel_l2(gm.read("U"), Ur), O.rel_l2(gm.read("V"), Vr)
    emit(f"f64rec schnakenberg lmax127 50 steps: U {eu:.2e} V {ev:.2e} vs f64")
    assert eu < 1e-5 and ev < 1e-5
    gm.destroy()
    sht.destroy()


@pytest.mark.parametrize("lmax,nlat,nphi,nf", [(63, 64, 128, 5), (255, 256, 512, 8),
                                               (1023, 1024, 2048, 2)])
def test_f64rec_on_matrix_cores(gpu, lmax, nlat, nphi, nf):
    """The FP64-recurrence form with its sums on the FP32 matrix cores
    (tile impl 2): same tolerance against FP64, and within 1e-6 of the VALU
    sums."""
    import tsphere

    cfg = O.Cfg(lmax, lmax, nlat, nphi)
    emit(f"f64rec matrix cores lmax{lmax} nf {nf}")
    plan = tsphere.ShtPlan.create(tsphere.ShtConfig(lmax, lmax, nlat, nphi), legendre="f64rec",
                                  max_fields=nf)
    qs = np.concatenate([_random_field_spectrum(cfg, 60 + f) for f in range(nf)])
    q = tsphere.DeviceBuffer(nf * 2 * cfg.nlm)
    sp = tsphere.DeviceBuffer(nf * cfg.npts)
    back = tsphere.DeviceBuffer(nf * 2 * cfg.nlm)
    q.upload(qs, plan.stream)
    res = []
    for impl in (1, 2):
        for st in ("synth", "analys"):
            plan.set_tile(st, impl=impl)
        plan.encode_synth_into(plan.stream, plan.create_synth_binding(q, sp, nf, 2 * cfg.nlm,
                                                                      cfg.npts))
        plan.encode_analys_into(plan.stream, plan.create_analys_binding(sp, back, nf, cfg.npts,
                                                                        2 * cfg.nlm))
        res.append((sp.download(plan.stream).reshape(nf, -1),
                    back.download(plan.stream).reshape(nf, -1)))
    ref = O.ShtReference(cfg)
    for f in (0, nf - 1):
        qf = qs[f * 2 * cfg.nlm:(f + 1) * 2 * cfg.nlm]
        e_s = O.rel_l2(res[1][0][f], ref.synth_fast(qf.astype(np.float64)))
        e_a = O.rel_l2(res[1][1][f], ref.analys_fast(res[1][0][f].astype(np.float64)))
        d = max(O.rel_l2(res[1][0][f], res[0][0][f]), O.rel_l2(res[1][1][f], res[0][1][f]))
        emit(f"  field {f}: synth {e_s:.2e} analys {e_a:.2e} vs f64; vs VALU {d:.1e}")
        assert e_s < TOL and e_a < TOL and d < 1e-6
    plan.destroy()
