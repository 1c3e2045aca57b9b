"""The benchmarked step at its own size: Brusselator, lmax 1023 on the Gauss
grid 1024 x 2048, A=4.5, B=6.96, D1=2e-4, D2=8 D1, dt=0.002, initial state
(A, B/A) plus uniform(-0.01, 0.01) noise (mulberry32 seed 0, u's draws then
v's) analysed to spectral by the f64 oracle.

* The default plan (grid kernel + spectral op: k_leg_spectral_m4 at 512
  latitude pairs in 8-wave workgroups, k_leg_spectral_b4 in the split-BF16
  form, the generated grid kernel at nphi 2048) against the plan with the
  implicit solve folded into the next synthesis and against the plan of
  separate launches, bit for bit, over
  graph replays, directly encoded remainders and a state upload between
  calls (the step order of src/mgpu/plan.ts:436-478).
* The same step against the f64 trajectory (oracle, fast transforms): the
  state after 10 steps and its increment U_10 - U_0, to stated tolerances.
* Shapes beyond it that the planner reaches from the session's grid rule:
  768 latitude pairs (two 64-pair chains per lane in the spectral op) and
  1025 pairs (above the spectral op's limit: the planner takes the folded
  form), each against separate launches bit for bit.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import emit

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MFILES = os.path.join(ROOT, "turing-sphere_amd", "mfiles")
BRUSS = ["A", "B", "D1", "D2", "dt"]
BRUSSELATOR_PARAMS = {"A": 4.5, "B": 6.96, "D1": 2e-4, "D2": 8 * 2e-4, "dt": 0.002}
LMAX, NLAT, NPHI = 1023, 1024, 2048

# After 10 steps, against the f64 trajectory from the same f32 initial state:
# the state (relative L2), and its increment U_10 - U_0.  The increment's
# error is set by the FP32 reaction terms, not by the transforms: at the
# steady state A - (B + 1) u + u^2 v adds terms of ~31 that cancel to ~0.05,
# so each grid value carries ~2e-6 of FP32 rounding (measured 6.5e-5 relative
# with the FP64-recurrence transforms).  The per-transform errors at this
# size are checked in tests/test_gpu_m4.py and tests/test_gpu_f64rec.py.
STATE_TOL = 1e-6
INCR_TOL = 3e-4
# The split-BF16 form carries ~5e-6 per transform (its precision target is
# 1e-5 per transform; tests/test_gpu_mfma.py): the synthesised grid values
# carry that, and the reaction's cancellation (terms of ~31 summing to ~0.05)
# multiplies it by ~600 in the increment.  Measured after 10 steps: state
# 2.5e-6 (U) and 8.0e-6 (V), increment 4.5e-3 (U) and 3.9e-3 (V).
TOLS = {"fp32": (STATE_TOL, INCR_TOL), "f64rec": (STATE_TOL, INCR_TOL), "bf16x3": (2e-5, 1e-2)}

_INIT = {}


def _brusselator_initial(lmax=LMAX, nlat=NLAT, nphi=NPHI):
    key = (lmax, nlat, nphi)
    if key not in _INIT:
        npts = nlat * nphi
        r = O.mulberry32(0, 2 * npts)
        u0 = BRUSSELATOR_PARAMS["A"] + (0.02 * r[:npts] - 0.01)
        v0 = BRUSSELATOR_PARAMS["B"] / BRUSSELATOR_PARAMS["A"] + (0.02 * r[npts:] - 0.01)
        ref = O.ShtReference(O.Cfg(lmax, lmax, nlat, nphi))
        U0 = ref.analys_fast(u0.astype(np.float32).astype(np.float64)).astype(np.float32)
        V0 = ref.analys_fast(v0.astype(np.float32).astype(np.float64)).astype(np.float32)
        _INIT[key] = (U0, V0, ref)
    return _INIT[key]


def _make(lmax, nlat, nphi, legendre, spectral, fold, src="brusselator", params=BRUSSELATOR_PARAMS):
    import tsphere
    from tsphere.model import GpuModel

    cfg = tsphere.ShtConfig(lmax, lmax, nlat, nphi)
    sht = tsphere.ShtPlan.create(cfg, legendre=legendre, max_fields=2)
    text = open(os.path.join(MFILES, src + ".m")).read()
    names = BRUSS if src == "brusselator" else ["a", "b", "D1", "D2", "dt"]
    gm = GpuModel.create(sht, cfg, text, names, ["U", "V"], ["u", "v"], spectral=spectral,
                         fold=fold)
    gm.set_params(params)
    return sht, gm


def _run_plans(lmax, nlat, nphi, legendre, U0, V0, src="brusselator", params=BRUSSELATOR_PARAMS,
               forms=((True, False), (False, True), (False, False)), expect=None):
    outs = []
    for spectral, fold in forms:
        sht, gm = _make(lmax, nlat, nphi, legendre, spectral, fold, src, params)
        kinds = [l.split()[0] for l in gm.describe()["step"]]
        if expect is not None:
            want = expect[(spectral, fold)]
            assert ("spectral" in kinds) == (want == "spectral"), kinds
            assert ("imex" in kinds) == (want == "fold"), kinds
        emit(f"  plan spectral={spectral} fold={fold}: {kinds}")
        gm.graph_steps = 8
        gm.upload("U", U0)
        gm.upload("V", V0)
        gm.step(20)          # two 8-step graph replays + 4 directly encoded steps
        mid = [gm.read("U"), gm.read("V")]
        gm.upload("U", mid[0] * np.float32(0.999))
        gm.step(5)           # directly encoded after an upload
        outs.append(mid + [gm.read("U"), gm.read("V"), gm.read("u"), gm.read("v")])
        gm.destroy()
        sht.destroy()
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.all(np.isfinite(a)) and np.all(np.isfinite(b))
            d = float(np.max(np.abs(a - b)) / max(1e-30, float(np.max(np.abs(a)))))
            emit(f"  max rel diff {d:.2e}")
            assert np.array_equal(a, b), d
    return outs[0]


@pytest.mark.parametrize("legendre", ["fp32", "f64rec", "bf16x3"])
def test_headline_step_plans_bit_exact(gpu, legendre):
    emit(f"headline step plans, Brusselator lmax {LMAX} {NLAT}x{NPHI} {legendre}")
    U0, V0, _ = _brusselator_initial()
    expect = {(True, False): "spectral", (False, True): "fold", (False, False): "separate"}
    _run_plans(LMAX, NLAT, NPHI, legendre, U0, V0, expect=expect)


@pytest.mark.parametrize("legendre", ["fp32", "f64rec", "bf16x3"])
def test_headline_step_vs_f64_trajectory(gpu, legendre):
    steps = 10
    emit(f"headline step vs f64 trajectory, {steps} steps, {legendre}")
    U0, V0, ref = _brusselator_initial()
    sht, gm = _make(LMAX, NLAT, NPHI, legendre, True, False)
    assert "spectral" in [l.split()[0] for l in gm.describe()["step"]]
    gm.upload("U", U0)
    gm.upload("V", V0)
    gm.step(steps)
    U, V = gm.read("U"), gm.read("V")
    gm.destroy()
    sht.destroy()
    emit("  f64 trajectory ...")
    Ur, Vr = O.imex_trajectory(ref.cfg, O.BRUSSELATOR, BRUSSELATOR_PARAMS, U0, V0, steps, ref)
    d64 = [Ur - U0.astype(np.float64), Vr - V0.astype(np.float64)]
    eu = O.rel_l2(U.astype(np.float64) - U0, d64[0])
    ev = O.rel_l2(V.astype(np.float64) - V0, d64[1])
    su, sv = O.rel_l2(U, Ur), O.rel_l2(V, Vr)
    emit(f"  increment relL2 U {eu:.2e} V {ev:.2e}; state relL2 U {su:.2e} V {sv:.2e}")
    stol, itol = TOLS[legendre]
    assert su < stol and sv < stol
    assert eu < itol and ev < itol


@pytest.mark.parametrize("legendre", ["fp32", "f64rec"])
def test_two_chain_spectral_op_plans_bit_exact(gpu, legendre):
    """768 latitude pairs: the spectral op's lanes carry two 64-pair chains."""
    lmax, nlat, nphi = 767, 1536, 2048
    emit(f"two-chain spectral op plans lmax {lmax} {nlat}x{nphi} {legendre}")
    U0, V0, _ = _brusselator_initial(lmax, nlat, nphi)
    expect = {(True, False): "spectral", (False, True): "fold", (False, False): "separate"}
    _run_plans(lmax, nlat, nphi, legendre, U0, V0, expect=expect)


def test_beyond_spectral_limit_takes_fold(gpu):
    """1025 latitude pairs (nlat 2050, nphi 2048): a spectral request plans
    the folded form (the block-matrix synthesis at four chains per lane),
    which matches separate launches bit for bit."""
    lmax, nlat, nphi = 1023, 2050, 2048
    emit(f"beyond the spectral op's limit: lmax {lmax} {nlat}x{nphi}")
    U0, V0, _ = _brusselator_initial(lmax, nlat, nphi)
    expect = {(True, False): "fold", (False, False): "separate"}
    _run_plans(lmax, nlat, nphi, "fp32", U0, V0, forms=((True, False), (False, False)),
               expect=expect)


def test_split_bf16_beyond_block_limit_separate(gpu):
    """768 latitude pairs in the split-BF16 form: above the 4x4x4 block
    kernels' 512 pairs the step keeps separate launches on the 16x16x32
    column-tile kernels, and a spectral request plans exactly that."""
    lmax, nlat, nphi = 767, 1536, 2048
    emit(f"split-BF16 beyond the block kernels' limit: lmax {lmax} {nlat}x{nphi}")
    U0, V0, _ = _brusselator_initial(lmax, nlat, nphi)
    expect = {(True, False): "separate", (False, False): "separate"}
    _run_plans(lmax, nlat, nphi, "bf16x3", U0, V0, forms=((True, False), (False, False)),
               expect=expect)
