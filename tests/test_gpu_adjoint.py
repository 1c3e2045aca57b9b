"""Reverse mode on the GPU: the transposes of synth/analys, and vector-Jacobian
products through checkpointed IMEX trajectories, against the f64 oracle's
adjoint (itself pinned by finite differences in tests/test_oracle.py).

Stated tolerances: transposes 1e-5 relative to the f64 transpose (FP32
transforms); trajectory gradients 1e-5 relative at these sizes (measured
1.0-1.8e-6: the FP32 forward trajectory the gradient is taken along differs
from the f64 one by ~1e-7 per step, and the adjoint accumulates FP32 rounding
over every step; at lmax 255 and 200 steps see tests/test_gpu_configs.py)."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import emit

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MFILES = os.path.join(ROOT, "turing-sphere_amd", "mfiles")


@pytest.mark.parametrize("legendre", ["fp32", "bf16x3"])
@pytest.mark.parametrize("lmax,nlat,nphi", [(31, 32, 64), (63, 65, 128), (127, 128, 256)])
def test_transposes_vs_f64(gpu, legendre, lmax, nlat, nphi):
    import ctypes as C

    import tsphere

    cfg = tsphere.ShtConfig(lmax, lmax, nlat, nphi)
    emit(f"transposes {legendre} lmax{lmax} {nlat}x{nphi}")
    plan = tsphere.ShtPlan.create(cfg, legendre=legendre)
    ref = O.ShtReference(O.Cfg(lmax, lmax, nlat, nphi))
    rng = np.random.default_rng(lmax)
    g = rng.normal(size=cfg.npts).astype(np.float32)
    q = rng.normal(size=2 * plan.nlm).astype(np.float32)
    gb, qb = tsphere.DeviceBuffer(cfg.npts), tsphere.DeviceBuffer(2 * plan.nlm)
    s = plan.stream
    P = lambda b: (C.c_void_p * 1)(b.ptr)  # noqa: E731
    gb.upload(g, s)
    tsphere._lib.call("tsphere_synth_adjoint", plan.handle, 1, 1, P(gb), 0, P(qb), 0, s.handle)
    st = qb.download(s)
    e1 = O.rel_l2(st, ref.synth_T(g.astype(np.float64)))
    qb.upload(q, s)
    tsphere._lib.call("tsphere_analys_adjoint", plan.handle, 1, 1, P(qb), 0, P(gb), 0, s.handle)
    at = gb.download(s)
    e2 = O.rel_l2(at, ref.analys_T(q.astype(np.float64)))
    emit(f"  S^T {e1:.2e}  A^T {e2:.2e}")
    assert e1 < 1e-5 and e2 < 1e-5
    plan.destroy()


MODELS = {
    "schnakenberg": (O.SCHNAKENBERG, ["a", "b", "D1", "D2", "dt"], ["U", "V"], ["u", "v"],
                     {"a": 0.1, "b": 0.9, "D1": 6.5e-4, "D2": 0.026, "dt": 0.01}),
    "brusselator": (O.BRUSSELATOR, ["A", "B", "D1", "D2", "dt"], ["U", "V"], ["u", "v"],
                    {"A": 3.0, "B": 9.0, "D1": 3.33e-3, "D2": 1.67e-2, "dt": 0.005}),
    "allencahn": (O.ALLENCAHN, ["eps2", "dt"], ["U"], ["u"], {"eps2": 1e-3, "dt": 0.02}),
}


def _adjoint(key, lmax, checkpoint_every):
    import tsphere
    from tsphere.adjoint import AdjointModel

    model, names, state, view, params = MODELS[key]
    cfg = tsphere.ShtConfig(lmax, lmax, lmax + 1, 2 * (lmax + 1))
    sht = tsphere.ShtPlan.create(cfg, max_fields=len(state))
    src = open(os.path.join(MFILES, key + ".m")).read()
    am = AdjointModel.create(sht, cfg, src, names, state, view, checkpoint_every)
    am.set_params(params)
    return cfg, sht, am, model, state, params


def _initial(cfg, nlm, state, seed):
    rng = np.random.default_rng(seed)
    out = {}
    for i, s in enumerate(state):
        q = (rng.normal(size=2 * nlm) * 0.02).astype(np.float32)
        q[0] += (1.0, 0.9)[i] * np.sqrt(4 * np.pi)  # mean 1.0 / 0.9 on the grid
        out[s] = q
    return out


# GPU vjp against the f64 adjoint along the f64 trajectory, relative; measured
# 1.0-1.8e-6 for the initial states and 2e-8 - 1.2e-6 for the parameters
TOL_VJP = 1e-5


@pytest.mark.parametrize("key,lmax,steps,every", [("schnakenberg", 31, 30, 7),
                                                   ("brusselator", 31, 24, 10),
                                                   ("allencahn", 63, 40, 20)])
def test_vjp_vs_f64_adjoint(gpu, key, lmax, steps, every):
    cfg, sht, am, model, state, params = _adjoint(key, lmax, every)
    emit(f"vjp {key} lmax{lmax} steps {steps} checkpoints every {every}")
    x0 = _initial(cfg, sht.nlm, state, 11)
    rng = np.random.default_rng(5)
    cot = {s: rng.normal(size=2 * sht.nlm).astype(np.float32) for s in state}
    grads, pg = am.vjp(x0, steps, cot)
    oc = O.Cfg(lmax, lmax, cfg.nlat, cfg.nphi)
    Ub, Vb, og = O.imex_vjp(oc, model, params, x0["U"], x0.get("V"), steps, cot["U"],
                            cot.get("V"), O.ShtReference(oc))
    eu = O.rel_l2(grads["U"], Ub)
    msg = f"  U0bar {eu:.2e}"
    assert eu < TOL_VJP
    if "V" in grads:
        ev = O.rel_l2(grads["V"], Vb)
        msg += f"  V0bar {ev:.2e}"
        assert ev < TOL_VJP
    for k, v in og.items():
        e = abs(pg[k] - v) / max(abs(v), 1e-30)
        msg += f"  {k} {pg[k]:.4e}/{v:.4e} ({e:.1e})"
        assert e < TOL_VJP, (k, pg[k], v)
    emit(msg)
    am.destroy()
    sht.destroy()


def test_checkpoint_interval_does_not_change_gradient(gpu):
    """Checkpointing only trades recomputation for memory: every interval
    gives the same gradient (the recomputed steps replay identical work)."""
    cfg, sht, am, model, state, params = _adjoint("schnakenberg", 15, 4)
    x0 = _initial(cfg, sht.nlm, state, 2)
    cot = {s: np.ones(2 * sht.nlm, np.float32) for s in state}
    g1, p1 = am.vjp(x0, 13, cot)
    am.checkpoint_every = 13
    g2, p2 = am.vjp(x0, 13, cot)
    am.checkpoint_every = 1
    g3, p3 = am.vjp(x0, 13, cot)
    for s in state:
        assert np.array_equal(g1[s], g2[s]) and np.array_equal(g1[s], g3[s])
    for k in p1:
        assert p1[k] == p2[k] == p3[k]
    am.destroy()
    sht.destroy()
