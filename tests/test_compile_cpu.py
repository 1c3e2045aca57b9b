"""Front end, HIP emitter and planner on the CPU (no device): what the
.m files compile to (the fusion guard of test/modelChecks.ts:17-21,41-45),
that the generated HIP compiles for gfx950, the emitter's refusals
(wgsl.ts:142-304), and that the generated arithmetic equals the reference's
own fp32 reaction/IMEX code (bench/shtns/spec.h:146-169) bit for bit."""
import os

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MFILES = os.path.join(ROOT, "turing-sphere_amd", "mfiles")

MODELS = {  # key -> (params, state, view)
    "schnakenberg": (["a", "b", "D1", "D2", "dt"], ["U", "V"], ["u", "v"]),
    "brusselator": (["A", "B", "D1", "D2", "dt"], ["U", "V"], ["u", "v"]),
    "allencahn": (["eps2", "dt"], ["U"], ["u"]),
    "linear": (["c", "D", "dt"], ["U"], ["u"]),
    "logistic": (["r", "D", "dt"], ["U"], ["u"]),
}
EXPECTED_KERNELS = {"schnakenberg": 5, "brusselator": 5, "allencahn": 2}


def source(key):
    return open(os.path.join(MFILES, key + ".m")).read()


def compile_fns(key, nlm=2080, npts=8192):
    from tsphere.mfront import CompiledModel, model_bindings

    params, state, view = MODELS[key]
    cm = CompiledModel(source(key), model_bindings(state, params, npts, nlm),
                       {"npts": npts, "nlm": nlm})
    n = len(state) + len(view)
    return cm.specialize("init", n), cm.specialize("step", n), state, view


def plan_cpu(key, batch=True):
    from tsphere.plan import HostBuffers, ModelPlan, PlanSpec

    init_fn, step_fn, state, view = compile_fns(key)
    host = HostBuffers(allocate=False)
    host.ensure_group(state, 2 * 2080)
    fb = state + [None] * len(view)
    return (ModelPlan.create(None, PlanSpec(init_fn, fb), host, batch, "init"),
            ModelPlan.create(None, PlanSpec(step_fn, fb), host, batch, "step"))


@pytest.mark.parametrize("key", sorted(EXPECTED_KERNELS))
def test_fusion_guard_kernel_counts(key):
    """modelChecks.ts:17-21: one generated kernel per element-wise line."""
    _, step = plan_cpu(key, batch=False)
    d = step.describe()
    kernels = [l for l in d if l.startswith("kernel")]
    assert len(kernels) == EXPECTED_KERNELS[key], d


def test_schnakenberg_op_list():
    """README.md:85: 11 ops per step as the reference records them.  Here,
    unbatched: one kernel per element-wise line (the reference's shape);
    batched: the two syntheses and the two analyses each become one launch,
    feedback copies fold into the kernels that produce the new state, and
    each run of element-wise lines becomes one fused kernel -- 4 launches."""
    _, step = plan_cpu("schnakenberg", batch=False)
    kinds = [l.split()[0] for l in step.describe()]
    # unbatched: 2 synth, 5 kernels, 2 analys; feedback copies folded into the
    # IMEX kernels (they run after the last read of U and V)
    assert sorted(kinds) == sorted(["synth"] * 2 + ["kernel"] * 5 + ["analys"] * 2)
    _, step = plan_cpu("schnakenberg", batch=True)
    kinds = [l.split()[0] for l in step.describe()]
    assert kinds == ["synth", "kernel", "analys", "kernel"], step.describe()
    d = step.describe()
    assert d[0].count(";") == 1 and d[2].count(";") == 1   # both species per transform
    assert d[1].count(";") == 2 and d[3].count(";") == 1   # 3 reaction lines, 2 IMEX lines
    fused = [op for op in step.ops if op.kind == "kernel"]
    # uuv is used only inside the reaction group: kept in registers, not stored
    assert len(fused[0].outs) == 2 and len(fused[1].outs) == 2


def _expand(expr: str, args: list, env: dict) -> str:
    """Substitute argument registers a<i> of a generated expression by the
    expanded expressions of the values they hold."""
    import re

    return re.sub(r"\ba(\d+)\b", lambda m: env.get(args[int(m.group(1))],
                                                  args[int(m.group(1))]), expr)


def _statement_trees(plan) -> dict:
    """c_name -> the statement's f32 expression with every value produced by
    an earlier generated statement expanded in place, in terms of the step's
    inputs (state, grid values, lam) and the parameter arguments p<k>."""
    import re

    from tsphere.plan import FusedOp, KernelOp

    env = {}
    for op in plan.ops:
        if isinstance(op, KernelOp):
            env[op.out] = _expand(op.kernel.body, list(op.ins), env)
        elif isinstance(op, FusedOp):
            # the element function: a0.. are the kernel inputs, then one
            # register per statement in program order
            lines = re.findall(r"const float a(\d+) = (.*);", op.kernel.code)
            regs = list(op.ins)
            for (idx, expr), mem in zip(lines, op.members):
                assert int(idx) == len(regs), (idx, regs)
                env[mem.out] = _expand(expr, regs, env)
                regs.append(mem.out)
            assert len(lines) == len(op.members)
    return env


@pytest.mark.parametrize("key", ["schnakenberg", "brusselator", "allencahn"])
def test_fused_kernel_matches_statementwise_f32(key):
    """The fused kernels evaluate every statement as the same f32 operation
    tree as the one-kernel-per-line form (the reference's shape), with the
    values kept in registers substituted by their own trees: the same
    operations on the same operands in the same order, under the same
    contraction-off pragma."""
    _, one = plan_cpu(key, batch=False)
    _, fused = plan_cpu(key, batch=True)
    t1, t2 = _statement_trees(one), _statement_trees(fused)
    assert set(t1) == set(t2) and len(t1) == EXPECTED_KERNELS[key]
    for c in t1:
        assert t1[c] == t2[c], (c, t1[c], t2[c])
    for p in (one, fused):
        src = p.kernel_source()
        assert src.count("#pragma clang fp contract(off)") == src.count("__global__")


def test_init_plan_and_feedback():
    init, _ = plan_cpu("schnakenberg", batch=False)
    d = init.describe()
    assert sum(l.startswith("kernel") for l in d) == 2   # us + noise, vs * ones(...)
    assert not any(l.startswith("copy") for l in d)      # U, V written in place
    init, _ = plan_cpu("schnakenberg")                   # batched: one fused kernel
    d = init.describe()
    assert sum(l.startswith("kernel") for l in d) == 1 and not any(l.startswith("copy") for l in d)


def test_generated_source_compiles_for_gfx950():
    import tsphere._lib as L

    for key in MODELS:
        init, step = plan_cpu(key)
        for p in (init, step):
            src = p.kernel_source()
            assert L.load().tsphere_rtc_check(src.encode(), None) == L.OK, L.last_error()
            assert "expf" not in src  # models use no transcendental


def test_scalar_params_are_kernel_arguments():
    _, step = plan_cpu("schnakenberg")
    assert step.param_names == ["a", "b", "D1", "D2", "dt"]
    src = step.kernel_source()
    # dt*D1 is formed from arguments in-kernel, never folded into a literal
    assert "(p4 * p2)" in src and "(p4 * p3)" in src


def _stmt(key, target, fn="step"):
    init_fn, step_fn, _, _ = compile_fns(key)
    f = step_fn if fn == "step" else init_fn
    return [s for s in f.body if getattr(s, "name", None) == target][0]


@pytest.mark.parametrize("model,p0,p1", [(0, 0.1, 0.9), (1, 4.5, 6.96)])
def test_reaction_arithmetic_equals_reference_fp32(model, p0, p1):
    """The emitted reaction expressions, evaluated with one f32 rounding per
    operation in tree order (what the -ffp-contract=off kernels compute),
    equal spec.h's float shtb_react bit for bit."""
    from tsphere.hipgen import eval_f32

    ref = O.ref_spec()
    if ref is None:
        pytest.skip("reference tree absent: oracle/_ref not built")
    key = "schnakenberg" if model == 0 else "brusselator"
    init_fn, step_fn, _, _ = compile_fns(key)
    body = step_fn.body
    pnames = {"a": p0, "b": p1} if model == 0 else {"A": p0, "B": p1}
    rng = np.random.default_rng(3)
    u = (rng.uniform(0.5, 3.0, 4096)).astype(np.float32)
    v = (rng.uniform(0.5, 3.0, 4096)).astype(np.float32)
    env = {"u": u, "v": v}
    uuv = eval_f32([s for s in body if s.name == "uuv"][0].expr, env, pnames)
    env["uuv"] = uuv
    # the two reaction terms are the hoisted analys arguments, in order
    temps = [s for s in body if s.name.startswith("t") and s.name.endswith("_")
             and s.expr.kind != "Call"]
    r1 = eval_f32(temps[0].expr, env, pnames)
    r2 = eval_f32(temps[1].expr, env, pnames)
    e1 = np.zeros_like(u)
    e2 = np.zeros_like(u)
    ref.ref_react_f32(model, p0, p1, u, v.ctypes.data, e1, e2.ctypes.data, u.size)
    assert np.array_equal(r1, e1) and np.array_equal(r2, e2)


def test_imex_arithmetic_equals_reference_fp32():
    from tsphere.hipgen import eval_f32

    ref = O.ref_spec()
    if ref is None:
        pytest.skip("reference tree absent: oracle/_ref not built")
    st = _stmt("schnakenberg", "Un")
    rng = np.random.default_rng(4)
    n = 4160
    U = rng.normal(size=n).astype(np.float32)
    R = rng.normal(size=n).astype(np.float32)
    lam = np.tile(np.arange(n // 2, dtype=np.float32) % 64, 2)
    lam = (lam * (lam + 1)).astype(np.float32)
    dt, D1 = 0.01, 6.5e-4
    from tsphere.ir import walk

    rname = [x.c_name for x in walk(st.expr) if x.kind == "Var" and x.c_name.endswith("_")][0]
    env = {"U": U, rname: R, "lam": lam}
    got = eval_f32(st.expr, env, {"dt": dt, "D1": D1})
    want = np.zeros(n, np.float32)
    dtD = np.float32(np.float32(dt) * np.float32(D1))
    ref.ref_imex_f32(np.float32(dt), dtD, U, R, lam, want, n)
    assert np.array_equal(got, want)


def test_emitter_refusals():
    """wgsl.ts: matrix product, runtime exponent, broadcasting, user
    functions, control flow -- all refused with a span."""
    from tsphere.hipgen import UnsupportedOnGpu
    from tsphere.mfront import CompiledModel, ModelCompileError, model_bindings
    from tsphere.plan import HostBuffers, ModelPlan, PlanSpec

    def plan(step_src):
        src = "function [U, u] = init(noise)\n U = analys(noise);\n u = synth(U);\nend\n" + step_src
        cm = CompiledModel(src, model_bindings(["U"], ["c", "dt"], 64, 8), {"npts": 64, "nlm": 8})
        fn = cm.specialize("step", 2)
        return ModelPlan.create(None, PlanSpec(fn, ["U", None]), HostBuffers(allocate=False))

    with pytest.raises(UnsupportedOnGpu, match="matrix '\\*' is not supported"):
        plan("function [Un, u] = step(U, lam, c, dt)\n u = synth(U);\n Un = U * lam;\nend\n")
    with pytest.raises(UnsupportedOnGpu, match="literal integer exponent"):
        plan("function [Un, u] = step(U, lam, c, dt)\n u = synth(U);\n"
             " Un = U + dt * analys(u .^ c);\nend\n")
    # a 2 x nlm value times an npts x 1 grid value (wgsl.ts:268-304): refused
    # with the statement's position in the source
    body = " u = synth(U);\n Un = U .* u;\n"
    with pytest.raises(UnsupportedOnGpu, match="implicit expansion: an operand is 2x8") as ei:
        plan("function [Un, u] = step(U, lam, c, dt)\n" + body + "end\n")
    sp = ei.value.span
    full = ("function [U, u] = init(noise)\n U = analys(noise);\n u = synth(U);\nend\n"
            "function [Un, u] = step(U, lam, c, dt)\n" + body + "end\n")
    assert full[sp.start:sp.end] == "U" and full[:sp.start].endswith("Un = ")
    with pytest.raises(UnsupportedOnGpu, match="an operand is 64x1 but the result is 64x8"):
        plan("function [Un, u] = step(U, lam, c, dt)\n u = synth(U);\n"
             " Un = U + dt * analys(u);\n w = u .* U;\nend\n")
    with pytest.raises(ModelCompileError, match="only contain assignments"):
        plan("function [Un, u] = step(U, lam, c, dt)\n u = synth(U);\n if c\n Un = U;\n end\nend\n")
    with pytest.raises(ModelCompileError, match="does not provide"):
        plan("function [Un, u] = step(U, lam, c, dt, zz)\n u = synth(U);\n Un = U;\nend\n")
    with pytest.raises(ModelCompileError, match="requires a 64x1 array"):
        plan("function [Un, u] = step(U, lam, c, dt)\n u = synth(U);\n Un = analys(U);\nend\n")


def test_power_and_literals():
    from tsphere.hipgen import f32_lit

    assert f32_lit(1.0) == "1.0f"
    assert f32_lit(0.1) == "0.1f"
    assert np.float32(float(f32_lit(6.5e-4)[:-1])) == np.float32(6.5e-4)
    _, step = plan_cpu("allencahn")
    assert "pow_i3(" in step.kernel_source()


def test_ensemble_plan_kernels():
    """Ensemble form (many runs per launch): per-run scalars from a device
    array, the shared eigenvalues indexed modulo one run, runs-times larger
    buffers, and the generated HIP compiles for gfx950."""
    from tsphere.plan import HostBuffers, ModelPlan, PlanSpec

    R = 4
    init_fn, step_fn, state, view = compile_fns("allencahn")
    host = HostBuffers(allocate=False)
    host.ensure_group(state, 2 * 2080 * R)
    host.ensure("lam", 2 * 2080)
    fb = state + [None] * len(view)
    step = ModelPlan.create(None, PlanSpec(step_fn, fb), host, True, "step", runs=R,
                            shared=("lam",))
    src = step.kernel_source()
    assert "const float* __restrict__ prm" in src
    assert "% 4160LL" in src            # lam, shared by the runs
    assert "prm[run * 2 + " in src      # eps2, dt per run
    kernels = [op for op in step.ops if op.kind == "kernel"]
    assert len(kernels) == 2 and all(op.count == 4160 * R or op.count == 8192 * R
                                     for op in kernels)
    assert host.get("U").count == 2 * 2080 * R and host.get("lam").count == 2 * 2080
    with pytest.raises(ValueError):
        ModelPlan.create(None, PlanSpec(step_fn, fb), host, True, "step", runs=0)


@pytest.mark.parametrize("key", ["schnakenberg", "brusselator", "allencahn"])
def test_adjoint_kernels_compile(key):
    """Reverse mode: every element-wise statement of the step gets
    generated adjoint code (symbolic derivatives of its IR tree) that
    compiles for gfx950; derivatives of the reaction terms appear.  One
    kernel per statement, and one per fused kernel of the step tape (its
    statements in reverse order, each in its own scope)."""
    import re

    from tsphere import _lib
    from tsphere.adjoint import build_adjoint_kernel
    from tsphere.plan import FusedOp, HostBuffers, KernelOp, ModelPlan, PlanSpec

    _, step_fn, state, view = compile_fns(key)
    for fuse in (False, True):
        host = HostBuffers(allocate=False)
        host.ensure_group(state, 2 * 2080)
        tape = ModelPlan.create(None, PlanSpec(step_fn, [None] * len(step_fn.outputs)), host,
                                True, "tape", fuse=fuse, grid=False, keep_all=True)
        srcs, nstmt = [], 0
        for i, op in enumerate(tape.ops):
            if isinstance(op, (KernelOp, FusedOp)):
                members = list(reversed(op.members)) if isinstance(op, FusedOp) else [op]
                ops_gis = [(m, [c for c in m.ins if c in tape._bufs and not c.startswith("lam")])
                           for m in members]
                code, ins, bars = build_adjoint_kernel(ops_gis, f"adj{i}", len(tape.param_names))
                assert code.count("// adjoint of:") == len(members)
                nstmt += len(members)
                srcs.append(code)
        assert nstmt == EXPECTED_KERNELS[key]
        src = ("__device__ __forceinline__ float ts_sign(float x) { return x; }\n" +
               "\n".join(srcs))
        assert re.search(r"bar\d+\[i\] \+= ob \*", src) and "partials[" in src
        assert _lib.load().tsphere_rtc_check(src.encode(), None) == _lib.OK, _lib.last_error()


@pytest.mark.parametrize("key", ["brusselator", "schnakenberg", "allencahn"])
def test_grid_fused_plan(key):
    """Grid fusion: with an FFT-sized grid, the step's synth -> element-wise
    -> analys run becomes one grid op (Legendre synthesis, one generated
    kernel running the row FFTs around the element-wise statements, Legendre
    analysis) and the generated source -- the shared row-FFT header included
    -- compiles for gfx950.  Grid values that the step returns are stored;
    values only analysed are not."""
    import types

    import tsphere
    from tsphere.plan import GridOp, HostBuffers, ModelPlan, PlanSpec

    init_fn, step_fn, state, view = compile_fns(key)
    cfg = tsphere.ShtConfig(63, 63, 64, 128)
    stub = types.SimpleNamespace(cfg=cfg, fourier_mode="fft", max_fields=8)
    host = HostBuffers(allocate=False)
    host.ensure_group(state, 2 * 2080)
    fb = state + [None] * len(view)
    step = ModelPlan.create(stub, PlanSpec(step_fn, fb), host, True, "step")
    kinds = [op.kind for op in step.ops]
    assert kinds[0] == "grid" and "synth" not in kinds and "analys" not in kinds, kinds
    g = step.ops[0]
    assert isinstance(g, GridOp) and len(g.srcs) == len(state) and len(g.dsts) >= 1
    names = {c.split("_")[0] for c in g.stores}
    assert g.nstmts >= 1 and step.elementwise_statements() >= g.nstmts
    src = step.kernel_source()
    assert "tsfft::first_pass<+1, 6>" in src and "tsfft::first_pass<-1, 6>" in src
    # the returned grid views (u, v, ...) are stored by the grid kernel
    assert len(g.stores) == len(view), (g.stores, view)
    del names
    # a DFT-sized grid (nphi not a power of two) keeps the separate stages
    stub2 = types.SimpleNamespace(cfg=tsphere.ShtConfig(63, 63, 64, 130), fourier_mode="dft",
                                  max_fields=8)
    step2 = ModelPlan.create(stub2, PlanSpec(step_fn, fb), host, True, "step")
    assert "grid" not in [op.kind for op in step2.ops]
    # and so does grid=False
    step3 = ModelPlan.create(stub, PlanSpec(step_fn, fb), host, True, "step", grid=False)
    assert [op.kind for op in step3.ops][0] == "synth"


def test_grid_fused_ensemble_plan():
    """Ensemble grid fusion: the grid kernel runs every run's rows (grid y =
    run), reads its scalars per run from the device array and compiles."""
    import types

    import tsphere
    from tsphere.plan import GridOp, HostBuffers, ModelPlan, PlanSpec

    R = 4
    init_fn, step_fn, state, view = compile_fns("schnakenberg")
    stub = types.SimpleNamespace(cfg=tsphere.ShtConfig(63, 63, 64, 128), fourier_mode="fft",
                                 max_fields=8)
    host = HostBuffers(allocate=False)
    host.ensure_group(state, 2 * 2080 * R)
    host.ensure("lam", 2 * 2080)
    fb = state + [None] * len(view)
    step = ModelPlan.create(stub, PlanSpec(step_fn, fb), host, True, "step", runs=R,
                            shared=("lam",))
    g = step.ops[0]
    assert isinstance(g, GridOp)
    src = step.kernel_source()
    assert "const int run = blockIdx.y;" in src and "prm[run * 5 + " in src
    assert "(g * R + run)" in src and "R = 4" in src


def test_grid_fused_plan_external_input():
    """A grid op that reads a host grid buffer (noise) inside its statements
    and analyses a synthesised value directly compiles, with the buffer as
    a kernel input."""
    import types

    import tsphere
    from tsphere.mfront import CompiledModel, model_bindings
    from tsphere.plan import GridOp, HostBuffers, ModelPlan, PlanSpec

    forced = open(os.path.join(ROOT, "tests", "models", "forced.m")).read()
    nlm, npts = 528, 33 * 64
    cm = CompiledModel(forced, model_bindings(["U"], ["c", "D", "dt"], npts, nlm),
                       {"npts": npts, "nlm": nlm})
    step_fn = cm.specialize("step", 2)
    cm.finish()
    stub = types.SimpleNamespace(cfg=tsphere.ShtConfig(31, 31, 33, 64), fourier_mode="fft",
                                 max_fields=2)
    host = HostBuffers(allocate=False)
    host.ensure_group(["U"], 2 * nlm)
    host.ensure("noise", npts)
    step = ModelPlan.create(stub, PlanSpec(step_fn, ["U", None]), host, True, "step")
    g = step.ops[0]
    assert isinstance(g, GridOp) and len(g.ins) == 1 and len(g.dsts) == 2
    assert "const float* gin0" in step.kernel_source()


@pytest.mark.parametrize("key", sorted(EXPECTED_KERNELS))
def test_keep_all_fused_kernels_store_every_value(key):
    """The reverse pass's step tape (ModelPlan.create(keep_all=True), no
    feedback): consecutive element-wise statements fuse into one kernel
    that writes every value it computes -- each statement's result keeps a
    buffer the reverse pass reads -- while the default fused plan writes
    only what is read after it."""
    from tsphere.plan import FusedOp, HostBuffers, KernelOp, ModelPlan, PlanSpec

    _, step_fn, state, view = compile_fns(key)
    plans = {}
    for keep in (False, True):
        host = HostBuffers(allocate=False)
        host.ensure_group(state, 2 * 2080)
        plans[keep] = ModelPlan.create(None, PlanSpec(step_fn, [None] * len(step_fn.outputs)),
                                       host, True, "tape", grid=False, keep_all=keep)
    tape = plans[True]
    members = [m for op in tape.ops for m in (op.members if isinstance(op, FusedOp) else
                                              [op] if isinstance(op, KernelOp) else [])]
    assert len(members) == EXPECTED_KERNELS[key]
    stored = {c for op in tape.ops if isinstance(op, FusedOp) for c in op.outs}
    stored |= {op.out for op in tape.ops if isinstance(op, KernelOp)}
    assert stored == {m.out for m in members}
    # the same fusion as the default plan: only the stores differ
    assert tape.generated_kernels() == plans[False].generated_kernels()
    lean = {c for op in plans[False].ops if isinstance(op, FusedOp) for c in op.outs}
    assert lean <= stored
