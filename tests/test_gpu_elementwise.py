"""The generated element-wise kernels, compiled by hipRTC and launched on the
GPU, against the reference's own float arithmetic: bench/shtns/spec.h's
shtb_react (the reaction terms, spec.h:146-161) and shtb_imex (the diagonal
implicit step, spec.h:166-169), compiled where the reference lies into
oracle/_ref/libref_spec.so.  Bit for bit, on random inputs, both for the
one-kernel-per-statement plan (the reference's shape, wgsl.ts:321-372) and
for the fused plan (each run of statements one kernel).  The element count
is not a multiple of 4, so the kernels' scalar tail runs as well as their
float4 body."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from conftest import emit

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MFILES = os.path.join(ROOT, "turing-sphere_amd", "mfiles")

CASES = {
    # key: (spec.h model id, parameter names, values, (p0, p1) of shtb_react,
    #       diffusion parameter of each species)
    "schnakenberg": (O.SCHNAKENBERG, ["a", "b", "D1", "D2", "dt"],
                     {"a": 0.1, "b": 0.9, "D1": 6.5e-4, "D2": 0.026, "dt": 0.01},
                     ("a", "b"), ("D1", "D2")),
    "brusselator": (O.BRUSSELATOR, ["A", "B", "D1", "D2", "dt"],
                    {"A": 4.5, "B": 6.96, "D1": 2e-4, "D2": 1.6e-3, "dt": 0.002},
                    ("A", "B"), ("D1", "D2")),
    "allencahn": (O.ALLENCAHN, ["eps2", "dt"], {"eps2": 2.5e-4, "dt": 1e-3},
                  None, ("eps2",)),
}
N = 8192 + 5


@pytest.mark.parametrize("batch", [False, True])
@pytest.mark.parametrize("key", sorted(CASES))
def test_generated_kernels_equal_spec_h(gpu, key, batch):
    import tsphere
    from tsphere import _lib
    from tsphere.device import DeviceBuffer
    from tsphere.model import GpuModel
    from tsphere.plan import FusedOp, KernelOp

    ref = O.ref_spec()
    if ref is None:
        pytest.skip("oracle/_ref/libref_spec.so (the reference's spec.h) was not built")
    model, names, params, react_p, diff = CASES[key]
    emit(f"generated kernels vs spec.h: {key} batch={batch}")
    cfg = tsphere.ShtConfig(15, 15, 16, 32)
    state = ["U", "V"] if model != O.ALLENCAHN else ["U"]
    view = [s.lower() for s in state]
    sht = tsphere.ShtPlan.create(cfg, max_fields=2)
    src = open(os.path.join(MFILES, key + ".m")).read()
    gm = GpuModel.create(sht, cfg, src, names, state, view, batch=batch, grid=False,
                         spectral=False, fold=False)
    gm.set_params(params)
    plan = gm.step_plan
    s = gm.stream
    rng = np.random.default_rng(17)
    host = {}

    def value(c):
        if c not in host:  # an input the test provides
            if c.startswith("lam"):
                ell = rng.integers(0, 64, N).astype(np.float32)
                host[c] = (ell * (ell + np.float32(1))).astype(np.float32)
            else:
                host[c] = rng.uniform(0.5, 3.0, N).astype(np.float32)
        return host[c]

    bufs = {}

    def dev(c):
        if c not in bufs:
            bufs[c] = DeviceBuffer(N, c)
            bufs[c].upload(value(c), s)
        return bufs[c]

    pv = (C.c_float * len(names))(*[float(np.float32(params[n])) for n in names])
    L = _lib.load()
    kops = [op for op in plan.ops if isinstance(op, (KernelOp, FusedOp))]
    assert kops, plan.describe()
    for op in kops:
        outs = op.outs if isinstance(op, FusedOp) else [op.out]
        ins = [dev(c) for c in op.ins]
        for c in outs:
            bufs[c] = DeviceBuffer(N, c)
        ptrs = (C.c_void_p * (len(outs) + len(ins)))(*([bufs[c].ptr for c in outs] +
                                                       [b.ptr for b in ins]))
        _lib.check(L.tsphere_kernel_launch(op.handle, N, len(ptrs), ptrs, len(names), pv,
                                           s.handle), op.label)
        for c in outs:
            host[c] = bufs[c].download(s)

    # the reaction terms are what the step analyses, in species order
    analysed = [c for op in plan.ops if getattr(op, "kind", "") == "analys" for c in op.srcs]
    assert len(analysed) == len(state), plan.describe()
    u = value(view[0])
    v = value(view[1]) if len(view) > 1 else None
    p0 = np.float32(params[react_p[0]]) if react_p else np.float32(0)
    p1 = np.float32(params[react_p[1]]) if react_p else np.float32(0)
    r1 = np.zeros(N, np.float32)
    r2 = np.zeros(N, np.float32)
    ref.ref_react_f32(model, p0, p1, u, None if v is None else v.ctypes.data, r1,
                      None if v is None else r2.ctypes.data, N)
    want = [r1, r2]
    for i, c in enumerate(analysed):
        got = host[c]
        emit(f"  reaction {i} ({c}): max |diff| {np.max(np.abs(got - want[i])):.1e}")
        assert np.array_equal(got, want[i]), c

    # the implicit step of each species: X_new = (X + dt R) / (1 + (dt D) lam),
    # R the analysed value the kernel reads, (dt D) one f32 product
    dt = np.float32(params["dt"])
    for i, x in enumerate(state):
        xn = [op for op in kops for c, stmt_ins in _members(op) if c == f"{x}n"]
        assert xn, (x, plan.describe())
        op = xn[0]
        ins = dict(_members(op))[f"{x}n"]
        X, R, lam = (host[c] for c in ins)
        dtD = np.float32(dt * np.float32(params[diff[i]]))
        exp = np.zeros(N, np.float32)
        ref.ref_imex_f32(dt, dtD, X, R, lam, exp, N)
        got = host[f"{x}n"]
        emit(f"  implicit step {x}: max |diff| {np.max(np.abs(got - exp)):.1e}")
        assert np.array_equal(got, exp), x
    for b in bufs.values():
        b.free()
    gm.destroy()
    sht.destroy()


def _members(op):
    """(output c_name, its statement's input c_names in X, R, lam order)."""
    from tsphere.plan import FusedOp

    mems = op.members if isinstance(op, FusedOp) else [op]
    return [(m.out, list(m.ins)) for m in mems]
