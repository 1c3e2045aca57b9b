"""The batching and differentiation forms at their full sizes.

* Mixed-lmax ensemble: 256 Allen-Cahn runs, 64 each at lmax 63/127/255/511
  on (L+1) x 2(L+1) Gauss grids, eps ~ U[0.01, 0.05] (mulberry32 seed 0, its
  own stream), dt = 1e-3, u0 ~ U(-0.1, 0.1) (mulberry32 seed = run index),
  all groups stepped by one HIP graph.  The first and last run of every
  group against the same run done alone (2e-6 relative: the batched launch
  may group the fields into other workgroup shapes, which reorders the
  analysis reduction by an ulp) and against the f64 trajectory.
* Gradients: Schnakenberg at lmax 255 (256 x 512), a=0.1, b=0.9,
  D1=6.5e-4, D2=40 D1, dt=0.01, initial grid u = 1 + U(+-0.01), v = 0.9 +
  U(+-0.01) (mulberry32 seed 0, u's draws then v's) analysed on the GPU;
  200 IMEX steps with checkpoints every 20; loss L = sum over the grid of
  (u_200 - u*)^2 with u* the u field of a run with (a, b, D1, D2) x 1.05
  from the same start (made by the f64 oracle).  The GPU gradient with
  respect to U0, V0, a, b, D1, D2 (FP32 and FP64-recurrence transforms)
  against the f64 adjoint of the oracle (imex_vjp) along the f64
  trajectory, and the parameter gradients against central finite
  differences of the f64 loss.  The f64 adjoint itself must meet the FD to
  1e-5 (pinning it at this size); the GPU gradients carry the FP32
  trajectory's rounding and are held to the tolerances stated below.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import emit

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MFILES = os.path.join(ROOT, "turing-sphere_amd", "mfiles")
AC_SRC = open(os.path.join(MFILES, "allencahn.m")).read()
SCH_SRC = open(os.path.join(MFILES, "schnakenberg.m")).read()

TOL_RUN = 2e-6


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


def test_mixed_ensemble_256_runs(gpu):
    import tsphere
    from tsphere.ensemble import MixedEnsemble

    lmaxes = (63, 127, 255, 511)
    groups = [(tsphere.ShtConfig(L, L, L + 1, 2 * (L + 1)), 64) for L in lmaxes]
    emit("mixed ensemble: 256 Allen-Cahn runs, 64 each at lmax 63/127/255/511")
    ens = MixedEnsemble.create(AC_SRC, ["eps2", "dt"], ["U"], ["u"], groups)
    assert ens.runs == 256
    eps = 0.01 + 0.04 * O.mulberry32(0, 256)
    eps2 = eps * eps
    dt, steps = 1e-3, 15
    ens.set_params({"eps2": eps2, "dt": dt})
    noise = [(0.2 * O.mulberry32(g.first_run + r, g.cfg.npts) - 0.1).astype(np.float32)
             for g in ens.groups for r in range(g.runs)]
    ens.init(noise)
    U0 = ens.read("U")
    ens.graph_steps = 10
    ens.step(steps)            # one replay of the 10-step graph, 5 steps encoded directly
    U = ens.read("U")
    for g in ens.groups:
        c = g.cfg
        oc = O.Cfg(c.lmax, c.mmax, c.nlat, c.nphi)
        ref = O.ShtReference(oc)
        tol64 = max(2e-5, 2e-10 * (c.lmax + 1) ** 2)
        for r in (0, g.runs - 1):
            i = g.first_run + r
            assert np.all(np.isfinite(U[i]))
            es = O.rel_l2(U[i], _single_allencahn(c, float(eps2[i]), dt, noise[i], steps))
            Ur, _ = O.imex_trajectory(oc, O.ALLENCAHN, {"eps2": float(eps2[i]), "dt": dt},
                                      U0[i], None, steps, ref)
            e64 = O.rel_l2(U[i], Ur)
            emit(f"  lmax {c.lmax} run {i}: vs single {es:.2e}, vs f64 {e64:.2e}")
            assert es < TOL_RUN and e64 < tol64
    ens.destroy()


def _mulberry_grid(npts):
    r = O.mulberry32(0, 2 * npts)
    return 0.02 * r[:npts] - 0.01, 0.02 * r[npts:] - 0.01


GL, GSTEPS, GEVERY = 255, 200, 20
GNAMES = ["a", "b", "D1", "D2", "dt"]
GP = {"a": 0.1, "b": 0.9, "D1": 6.5e-4, "D2": 40 * 6.5e-4, "dt": 0.01}
_F64 = {}


def _f64_reference(x0):
    """The f64 side, shared by both GPU forms: the target u* (f64 forward run
    with (a, b, D1, D2) x 1.05, rounded to f32 as data), the loss, the f64
    adjoint's gradients and central finite differences of the loss."""
    if _F64:
        return _F64
    oc = O.Cfg(GL, GL, GL + 1, 2 * (GL + 1))
    ref = O.ShtReference(oc)
    U0, V0 = x0["U"].astype(np.float64), x0["V"].astype(np.float64)
    p5 = {k: (v * 1.05 if k != "dt" else v) for k, v in GP.items()}
    Ut, _ = O.imex_trajectory(oc, O.SCHNAKENBERG, p5, U0, V0, GSTEPS, ref)
    tgt = ref.synth_fast(Ut).astype(np.float32).astype(np.float64)

    def loss(pp):
        Ur, _ = O.imex_trajectory(oc, O.SCHNAKENBERG, pp, U0, V0, GSTEPS, ref)
        return float(np.sum((ref.synth_fast(Ur) - tgt) ** 2)), Ur

    L64, U64 = loss(GP)
    Ubar = ref.synth_T(2.0 * (ref.synth_fast(U64) - tgt))
    U0b, V0b, g64 = O.imex_vjp(oc, O.SCHNAKENBERG, GP, U0, V0, GSTEPS, Ubar, None, ref)
    fd = {}
    for k in ("a", "b", "D1", "D2"):
        h = 1e-4 * abs(GP[k])
        fd[k] = (loss({**GP, k: GP[k] + h})[0] - loss({**GP, k: GP[k] - h})[0]) / (2 * h)
    _F64.update(target=tgt.astype(np.float32), loss=L64, U0b=U0b, V0b=V0b, g=g64, fd=fd)
    return _F64


# GPU gradient against the f64 adjoint and the central differences, relative.
# Measured with FP32 transforms: 2.8e-5 (U0, V0), 2.5e-5 (a, b), 1.3e-4 (D1),
# 1.0e-4 (D2); with the FP64 recurrence 2.4e-5 / 2.5e-5 / 1.5e-4 / 1.3e-4: the
# error is the FP32 trajectory's (state and element-wise arithmetic), not the
# transforms', and D1, D2 weight the damped high-degree modes by l(l+1)
TOL_GRAD = {"fp32": 3e-4, "f64rec": 3e-4}


@pytest.mark.parametrize("legendre", ["fp32", "f64rec"])
def test_gradient_schnakenberg_lmax255_200_steps(gpu, legendre):
    import tsphere
    from tsphere.adjoint import AdjointModel

    cfg = tsphere.ShtConfig(GL, GL, GL + 1, 2 * (GL + 1))
    emit(f"gradient {legendre}: schnakenberg lmax {GL} {cfg.nlat}x{cfg.nphi}, {GSTEPS} steps, "
         f"checkpoints every {GEVERY}")
    sht = tsphere.ShtPlan.create(cfg, legendre=legendre, max_fields=2)
    am = AdjointModel.create(sht, cfg, SCH_SRC, GNAMES, ["U", "V"], ["u", "v"], GEVERY)
    ru, rv = _mulberry_grid(cfg.npts)
    # the initial state analysed on the GPU by an FP32 plan, the same for both forms
    p32 = tsphere.ShtPlan.create(cfg, max_fields=1)
    x0 = {"U": p32.analys((1.0 + ru).astype(np.float32)),
          "V": p32.analys((0.9 + rv).astype(np.float32))}
    p32.destroy()
    emit("  f64 target, adjoint and finite differences ...")
    R = _f64_reference(x0)
    am.set_params(GP)
    uN = sht.synth(am.forward(x0, GSTEPS)["U"])
    cot = {"U": am.synth_adjoint(2.0 * (uN - R["target"])),
           "V": np.zeros(2 * sht.nlm, np.float32)}
    grads, pg = am.vjp(x0, GSTEPS, cot)
    loss_gpu = float(np.sum((uN.astype(np.float64) - R["target"]) ** 2))
    am.destroy()
    sht.destroy()

    emit(f"  loss GPU {loss_gpu:.6e} f64 {R['loss']:.6e}")
    eu, ev = O.rel_l2(grads["U"], R["U0b"]), O.rel_l2(grads["V"], R["V0b"])
    emit(f"  U0bar vs f64 adjoint {eu:.2e}, V0bar {ev:.2e}")
    worst = max(eu, ev)
    for k in ("a", "b", "D1", "D2"):
        fd, g64 = R["fd"][k], R["g"][k]
        e_adj_fd = abs(g64 - fd) / abs(fd)
        e_gpu_adj = abs(pg[k] - g64) / abs(g64)
        e_gpu_fd = abs(pg[k] - fd) / abs(fd)
        emit(f"  d/d{k}: GPU {pg[k]:.6e} f64 adjoint {g64:.6e} FD {fd:.6e} | "
             f"adjoint-FD {e_adj_fd:.1e} GPU-adjoint {e_gpu_adj:.1e} GPU-FD {e_gpu_fd:.1e}")
        # central differences with h = 1e-4 |p| carry O(h^2) truncation
        # (measured 3.4e-6 for D1)
        assert e_adj_fd < 1e-5
        worst = max(worst, e_gpu_adj, e_gpu_fd)
    emit(f"  worst GPU relative error {worst:.2e} (tolerance {TOL_GRAD[legendre]:.0e})")
    assert worst < TOL_GRAD[legendre]
