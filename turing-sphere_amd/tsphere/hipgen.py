"""IR expression tree -> one HIP kernel for gfx950 (the counterpart of
``buildKernel`` in src/mgpu/wgsl.ts:321-372).

For an ``Assign`` whose right-hand side is element-wise over operands of the
target's shape, emit one kernel computing ``out[i] = expr`` -- the same
contract, operator table, refusals and messages as the reference's WGSL
emitter, with these MI355X-specific choices:

* runtime scalars are kernel arguments (``p<slot>``) rather than a params
  buffer, so a parameter change only re-records the launch (never
  recompiles); exact scalars fold to f32 literals; scalars the model derives
  from parameters are ``const float`` prologue lines (wgsl.ts:107-127);
* each thread handles four consecutive elements with 16-byte loads and one
  16-byte store (HBM-bound kernels), falling back to scalar accesses for the
  tail or an unaligned buffer;
* the kernel signature is ``(long long n, float* out, const float* in0..,
  float p0..)``, which ``tsphere_kernel_launch`` packs; the pointers are not
  ``__restrict__``: a kernel that produces a state's next value may write the
  buffer it reads (copy-free feedback), and every element is read before it
  is written in the same thread;
* element-wise math calls the HIP device library (``expf``, ``sinf``...);
* hipRTC compiles with ``-ffp-contract=off``, so every f32 operation rounds
  separately, in the expression tree's order.
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

from .ir import Assign, IRExpr, NumericType, is_multi_element, is_numeric, is_tensor


class UnsupportedOnGpu(Exception):
    """wgsl.ts:18-28: a construct the GPU backend cannot express, with the
    source span of the offending node."""

    def __init__(self, message: str, span=None):
        super().__init__(message)
        self.span = span


# wgsl.ts:33-43
BINARY_OPS = {"plus": "+", "minus": "-", "times": "*", "rdivide": "/", "mtimes": "*",
              "mrdivide": "/"}
# wgsl.ts:46
UNARY_OPS = {"uminus": "-", "uplus": "+"}
# wgsl.ts:49-71 -> HIP device math (float overloads)
CALL_FNS = {
    "abs": "fabsf", "acos": "acosf", "asin": "asinf", "atan": "atanf", "atan2": "atan2f",
    "ceil": "ceilf", "cos": "cosf", "cosh": "coshf", "exp": "expf", "floor": "floorf",
    "log": "logf", "log2": "log2f", "max": "fmaxf", "min": "fminf",
    # WGSL round() rounds half to even, and the reference maps MATLAB round onto it
    "round": "rintf",
    "sign": "ts_sign", "sin": "sinf", "sinh": "sinhf", "sqrt": "sqrtf", "tan": "tanf",
    "tanh": "tanhf",
}
# Builtins numbl marks element-wise (getBuiltin(..).elementwise, wgsl.ts:196-197).
ELEMENTWISE_BUILTINS = set(CALL_FNS)


def f32_lit(v: float) -> str:
    """An f32 literal holding exactly the f32 rounding of v (wgsl.ts:75-84).
    The shortest repr of the double, parsed as float: the same decimal the
    reference hands its WGSL compiler."""
    if not math.isfinite(v):
        raise UnsupportedOnGpu(f"cannot emit non-finite literal {v}")
    if float(v).is_integer() and abs(v) < 1e21:
        return f"{int(v)}.0f"
    r = repr(float(v))
    return f"{r}f" if ("e" in r or "." in r) else f"{r}.0f"


@dataclass
class KernelInputs:
    """wgsl.ts:86-96: how operands are read inside the kernel."""
    tensors: dict            # c_name -> input slot (in0, in1, ...)
    params: dict             # c_name -> param slot (p0, p1, ...), runtime scalars
    scalars: dict            # c_name -> (name, expr): scalars derived in the .m


@dataclass
class _Ctx:
    io: KernelInputs
    prologue: list = field(default_factory=list)
    bound: dict = field(default_factory=dict)


def _scalar_ident(c_name: str) -> str:
    return "s_" + re.sub(r"[^A-Za-z0-9_]", "_", c_name)


def _bind_scalar(c_name: str, ctx: _Ctx) -> str:
    """wgsl.ts:116-127: bind a derived scalar once, after its dependencies."""
    if c_name in ctx.bound:
        return ctx.bound[c_name]
    _, expr = ctx.io.scalars[c_name]
    ident = _scalar_ident(c_name)
    ctx.bound[c_name] = ident
    rhs = _emit(expr, ctx)
    ctx.prologue.append(f"  const float {ident} = {rhs};")
    return ident


def _emit(e: IRExpr, ctx: _Ctx) -> str:
    """wgsl.ts:133-217"""
    io = ctx.io
    k = e.kind
    if k == "NumLit":
        return f32_lit(e.value)
    if k == "Var":
        if is_tensor(e.ty):
            slot = io.tensors.get(e.c_name)
            if slot is None:
                raise UnsupportedOnGpu(f"no buffer bound for '{e.name}'", e.span)
            return f"a{slot}"
        if is_numeric(e.ty) and isinstance(e.ty.exact, (int, float)):
            return f32_lit(float(e.ty.exact))
        if e.c_name in io.params:
            return f"p{io.params[e.c_name]}"
        if e.c_name in io.scalars:
            return _bind_scalar(e.c_name, ctx)
        raise UnsupportedOnGpu(
            f"scalar '{e.name}' is not a constant, a parameter, or computed in this model",
            e.span)
    if k == "Binary":
        if e.builtin in ("mtimes", "mrdivide") and is_tensor(e.left.ty) and is_tensor(e.right.ty):
            op = "*" if e.builtin == "mtimes" else "/"
            raise UnsupportedOnGpu(
                f"matrix '{op}' is not supported; use the element-wise form ('.{op}')", e.span)
        if e.builtin in ("power", "mpower"):
            return _emit_power(e.left, e.right, ctx, e.span)
        op = BINARY_OPS.get(e.builtin)
        if op is None:
            raise UnsupportedOnGpu(f"operator '{e.builtin}' is not supported", e.span)
        return f"({_emit(e.left, ctx)} {op} {_emit(e.right, ctx)})"
    if k == "Unary":
        op = UNARY_OPS.get(e.builtin)
        if op is None:
            raise UnsupportedOnGpu(f"unary '{e.builtin}' is not supported", e.span)
        return f"({op}{_emit(e.operand, ctx)})"
    if k == "Call":
        if e.name == "ones":
            return "1.0f"
        if e.name == "zeros":
            return "0.0f"
        fn = CALL_FNS.get(e.name)
        if fn is None or e.name not in ELEMENTWISE_BUILTINS:
            is_user = e.c_name != e.name
            raise UnsupportedOnGpu(
                (f"'{e.name}' is a function defined in this model. Only init and step are "
                 f"compiled — inline its body into the caller.") if is_user else
                f"'{e.name}' cannot be evaluated element-wise on the GPU", e.span)
        return f"{fn}({', '.join(_emit(a, ctx) for a in e.args)})"
    raise UnsupportedOnGpu(f"'{k}' is not supported on the GPU", getattr(e, "span", None))


def _emit_power(base: IRExpr, exponent: IRExpr, ctx: _Ctx, span) -> str:
    """wgsl.ts:224-246: small integer exponents only, as repeated multiply."""
    kval = None
    if exponent.kind == "NumLit":
        kval = exponent.value
    elif is_numeric(exponent.ty) and isinstance(exponent.ty.exact, (int, float)):
        kval = float(exponent.ty.exact)
    b = _emit(base, ctx)
    if kval is not None and float(kval).is_integer() and 0 <= kval <= 8:
        kk = int(kval)
        if kk == 0:
            return "1.0f"
        return f"pow_i{kk}({b})"
    if kval is not None and float(kval).is_integer() and -8 <= kval < 0:
        return f"(1.0f / pow_i{-int(kval)}({b}))"
    what = "a runtime value" if kval is None else repr(kval)
    raise UnsupportedOnGpu(
        f"'.^' needs a literal integer exponent in [-8, 8] (got {what}); a negative base "
        f"makes a real pow() undefined", span)


def _pow_helpers(used) -> str:
    """wgsl.ts:249-257: x * x * ... (left-associated)."""
    out = []
    for k in sorted(used):
        body = "x" if k == 1 else "x" + " * x" * (k - 1)
        out.append(f"__device__ __forceinline__ float pow_i{k}(float x) {{ return {body}; }}")
    return "\n".join(out)


def check_shapes(e: IRExpr, target: NumericType, name: str) -> None:
    """wgsl.ts:268-304: reject implicit expansion."""
    want = target.shape

    def same(t):
        return want is not None and t.shape is not None and tuple(t.shape) == tuple(want)

    def walk(x):
        if is_numeric(x.ty) and is_multi_element(x.ty) and not same(x.ty):
            got = "x".join(map(str, x.ty.shape)) if x.ty.shape else "dynamic"
            res = "x".join(map(str, want)) if want else "dynamic"
            raise UnsupportedOnGpu(
                f"'{name}' would need implicit expansion: an operand is {got} but the result "
                f"is {res}. Expand it explicitly (the GPU kernel walks one index across every "
                f"operand).", x.span)
        if x.kind == "Binary":
            walk(x.left)
            walk(x.right)
        elif x.kind == "Unary":
            walk(x.operand)
        elif x.kind == "Call" and x.name not in ("ones", "zeros"):
            for a in x.args:
                walk(a)

    walk(e)


WORKGROUP_SIZE = 256
ELEMS_PER_THREAD = 4

_SIGN = ("__device__ __forceinline__ float ts_sign(float x) "
         "{ return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }")


@dataclass
class Kernel:
    """wgsl.ts:308-313, plus what the launch needs."""
    code: str        # the kernel's definition (part of a module's source)
    name: str        # extern "C" symbol
    count: int       # output elements
    label: str
    ninputs: int     # tensor inputs, in slot order
    nparams: int     # scalar arguments p0.. (the params map's size)
    body: str = ""   # the per-element expression (for tests / describe)


def build_kernel(stmt: Assign, io: KernelInputs, count: int, label: str, name: str,
                 batch: "Batch | None" = None) -> Kernel:
    """buildKernel (wgsl.ts:321-372): one element-wise Assign -> one kernel.

    With ``batch`` the kernel runs an ensemble: ``count`` elements per run,
    ``batch.runs`` runs back to back in every per-run buffer, runtime scalars
    read per run from a device array (last pointer argument, ``npar`` floats
    per run), and inputs listed in ``batch.shared`` (one copy for all runs,
    e.g. the eigenvalues) indexed modulo ``count``."""
    if not is_numeric(stmt.ty):
        raise UnsupportedOnGpu(f"'{stmt.name}' is not a numeric array", stmt.span)
    if stmt.ty.is_complex:
        raise UnsupportedOnGpu(
            f"'{stmt.name}' is complex; the GPU backend is real-only (a spectral field is "
            f"carried as a real 2 x nlm array)", stmt.span)
    check_shapes(stmt.expr, stmt.ty, stmt.name)
    ctx = _Ctx(io)
    body = _emit(stmt.expr, ctx)
    emitted = "\n".join(ctx.prologue + [body])
    used = {int(m) for m in re.findall(r"\bpow_i(\d+)\(", emitted)}
    nin = len(io.tensors)
    npar = len(io.params)
    args = ["a" + str(i) for i in range(nin)]
    params = ["float p" + str(i) for i in range(npar)]
    elem_params = ", ".join([f"float {a}" for a in args] + params) or "void"
    kin = "".join(f", const float* in{i}" for i in range(nin))
    pro = "\n".join("  " + line.strip() for line in ctx.prologue)
    helpers = _pow_helpers(used)
    elem = f"""{helpers}
// {label}
__device__ __forceinline__ float {name}_elem({elem_params}) {{
#pragma clang fp contract(off)
{pro}
  return {body};
}}
"""
    aligned = " && ".join(["ts_al(out)"] + [f"ts_al(in{i})" for i in range(nin)])
    if batch is None:
        kpar = "".join(f", float p{i}" for i in range(npar))
        vec_loads = "\n".join(
            f"    const float4 v{i} = *reinterpret_cast<const float4*>(in{i} + i0);"
            for i in range(nin))

        def lane(c):
            return ", ".join([f"v{i}.{c}" for i in range(nin)] + [f"p{i}" for i in range(npar)])

        scalar_args = ", ".join([f"in{i}[i]" for i in range(nin)] + [f"p{i}" for i in range(npar)])
        code = elem + f"""extern "C" __global__ __launch_bounds__({WORKGROUP_SIZE}) void {name}(long long n, float* out{kin}{kpar}) {{
  const long long i0 = ((long long)blockIdx.x * {WORKGROUP_SIZE} + threadIdx.x) * {ELEMS_PER_THREAD};
  if (i0 >= n) return;
  if (i0 + {ELEMS_PER_THREAD} <= n && {aligned}) {{
{vec_loads}
    float4 r;
    r.x = {name}_elem({lane('x')});
    r.y = {name}_elem({lane('y')});
    r.z = {name}_elem({lane('z')});
    r.w = {name}_elem({lane('w')});
    *reinterpret_cast<float4*>(out + i0) = r;
  }} else {{
    for (long long i = i0; i < n && i < i0 + {ELEMS_PER_THREAD}; ++i) out[i] = {name}_elem({scalar_args});
  }}
}}
"""
        return Kernel(code=code, name=name, count=count, label=label, ninputs=nin, nparams=npar,
                      body=body)

    per = int(count)
    shared = {io.tensors[c] for c in batch.shared if c in io.tensors}

    def idx(i, var):
        return f"{var} % {per}LL" if i in shared else var

    vec_loads = "\n".join(
        f"    const float4 v{i} = *reinterpret_cast<const float4*>(in{i} + {idx(i, 'i0')});"
        for i in range(nin))
    pload = "".join(f"    const float p{k} = prm[run * {npar} + {k}];\n" for k in range(npar))

    def lane(c):
        return ", ".join([f"v{i}.{c}" for i in range(nin)] + [f"p{i}" for i in range(npar)])

    scalar_args = ", ".join([f"in{i}[{idx(i, 'i')}]" for i in range(nin)] +
                            [f"p{i}" for i in range(npar)])
    if per % ELEMS_PER_THREAD == 0:
        # the four elements of a thread always belong to one run
        vec = f"""  if (i0 + {ELEMS_PER_THREAD} <= n && {aligned}) {{
    const long long run = i0 / {per}LL;
{pload}{vec_loads}
    float4 r;
    r.x = {name}_elem({lane('x')});
    r.y = {name}_elem({lane('y')});
    r.z = {name}_elem({lane('z')});
    r.w = {name}_elem({lane('w')});
    *reinterpret_cast<float4*>(out + i0) = r;
    return;
  }}
"""
    else:
        vec = "  // runs are not a multiple of 4 elements long: element by element\n"
    code = elem + f"""// ensemble form: {batch.runs} runs x {per} elements, {npar} scalar(s) per run in prm
extern "C" __global__ __launch_bounds__({WORKGROUP_SIZE}) void {name}(long long n, float* out{kin}, const float* __restrict__ prm) {{
  const long long i0 = ((long long)blockIdx.x * {WORKGROUP_SIZE} + threadIdx.x) * {ELEMS_PER_THREAD};
  if (i0 >= n) return;
{vec}  for (long long i = i0; i < n && i < i0 + {ELEMS_PER_THREAD}; ++i) {{
    const long long run = i / {per}LL;
{pload}    out[i] = {name}_elem({scalar_args});
  }}
}}
"""
    return Kernel(code=code, name=name, count=count * batch.runs, label=label, ninputs=nin,
                  nparams=npar, body=body)


def _emit_group(stmts, ios, outs, name: str, label: str):
    """The per-element device function of a group of element-wise statements:
    ``NAME_elem(float a<ext>.., float p<k>.., float& o<j>..)`` evaluates the
    statements in program order (values produced inside the group are
    registers) and sets o<j> to the value of outs[j].  Returns (code, ext,
    npar, emitted): ext lists the external tensor inputs in slot order."""
    produced = [st.c_name for st in stmts]
    if len(set(produced)) != len(produced):
        raise UnsupportedOnGpu("a fused group may assign each value once")
    ext = []                                   # external tensor inputs, slot order
    done = set()
    for st, io in zip(stmts, ios):
        for c in sorted(io.tensors, key=lambda c: io.tensors[c]):
            if c not in done and c not in ext:
                ext.append(c)
        done.add(st.c_name)
    slots = {c: i for i, c in enumerate(ext)}  # reads resolve through this map
    params = ios[0].params
    scalars = {}
    for io in ios:
        scalars.update(io.scalars)
    ctx = _Ctx(KernelInputs(slots, params, scalars))
    body = []
    nslot = len(ext)
    for st in stmts:
        if not is_numeric(st.ty) or st.ty.is_complex:
            raise UnsupportedOnGpu(f"'{st.name}' is not a real numeric array", st.span)
        check_shapes(st.expr, st.ty, st.name)
        rhs = _emit(st.expr, ctx)
        # the produced value is a new register; later reads of it resolve there
        slots[st.c_name] = nslot
        body.append(f"  const float a{nslot} = {rhs};")
        nslot += 1
    emitted = "\n".join(ctx.prologue + body)
    used = {int(m) for m in re.findall(r"\bpow_i(\d+)\(", emitted)}
    nin, npar, nout = len(ext), len(params), len(outs)
    elem_params = ", ".join([f"float a{i}" for i in range(nin)] +
                            [f"float p{i}" for i in range(npar)] +
                            [f"float& o{j}" for j in range(nout)])
    pro = "\n".join("  " + line.strip() for line in ctx.prologue)
    sets = "\n".join(f"  o{j} = a{slots[c]};" for j, c in enumerate(outs))
    code = f"""{_pow_helpers(used)}
// {label}
__device__ __forceinline__ void {name}_elem({elem_params}) {{
#pragma clang fp contract(off)
{pro}
{chr(10).join(body)}
{sets}
}}
"""
    return code, ext, npar, emitted


def build_fused_kernel(stmts, ios, count: int, stores, label: str, name: str,
                       batch: "Batch | None" = None) -> Kernel:
    """Several consecutive element-wise statements over the same element
    count as ONE kernel (the planner's fusion of a step's element-wise run):
    each thread evaluates the statements in program order for its elements,
    values produced inside the group are registers, and only the c_names in
    `stores` (values read after the group or leaving the step) are written.
    External inputs are read before anything is stored, so a statement that
    reads a buffer a later statement overwrites sees the old value, as the
    sequential kernels would.  Signature: (long long n, float* out0..,
    const float* in0.., float p0.. | const float* prm).  Returns the kernel
    and, via Kernel.body, the input c_names in slot order."""
    code, ext, npar, emitted = _emit_group(stmts, ios, stores, name, label)
    slots = {c: i for i, c in enumerate(ext)}
    nin, nout = len(ext), len(stores)
    kout = ", ".join(f"float* out{j}" for j in range(nout))
    kin = "".join(f", const float* in{i}" for i in range(nin))
    aligned = " && ".join([f"ts_al(out{j})" for j in range(nout)] +
                          [f"ts_al(in{i})" for i in range(nin)])
    per = int(count)
    shared = set()
    if batch is not None:
        shared = {slots[c] for c in batch.shared if c in slots and slots[c] < nin}

    def idx(i, var):
        return f"{var} % {per}LL" if i in shared else var

    vec_loads = "\n".join(
        f"    const float4 v{i} = *reinterpret_cast<const float4*>(in{i} + {idx(i, 'i0')});"
        for i in range(nin))
    rdecl = "\n".join(f"    float4 r{j};" for j in range(nout))

    def lane(c):
        return ", ".join([f"v{i}.{c}" for i in range(nin)] + [f"p{i}" for i in range(npar)] +
                         [f"r{j}.{c}" for j in range(nout)])

    vstores = "\n".join(f"    *reinterpret_cast<float4*>(out{j} + i0) = r{j};"
                         for j in range(nout))
    sdecl = " ".join(f"float s{j};" for j in range(nout))
    sargs = ", ".join([f"in{i}[{idx(i, 'i')}]" for i in range(nin)] +
                      [f"p{i}" for i in range(npar)] + [f"s{j}" for j in range(nout)])
    sstores = " ".join(f"out{j}[i] = s{j};" for j in range(nout))
    if batch is None:
        kpar = "".join(f", float p{i}" for i in range(npar))
        code += f"""extern "C" __global__ __launch_bounds__({WORKGROUP_SIZE}) void {name}(long long n, {kout}{kin}{kpar}) {{
  const long long i0 = ((long long)blockIdx.x * {WORKGROUP_SIZE} + threadIdx.x) * {ELEMS_PER_THREAD};
  if (i0 >= n) return;
  if (i0 + {ELEMS_PER_THREAD} <= n && {aligned}) {{
{vec_loads}
{rdecl}
    {name}_elem({lane('x')});
    {name}_elem({lane('y')});
    {name}_elem({lane('z')});
    {name}_elem({lane('w')});
{vstores}
  }} else {{
    for (long long i = i0; i < n && i < i0 + {ELEMS_PER_THREAD}; ++i) {{
      {sdecl}
      {name}_elem({sargs});
      {sstores}
    }}
  }}
}}
"""
        ktotal = count
    else:
        pload = "".join(f"    const float p{k} = prm[run * {npar} + {k}];\n" for k in range(npar))
        vec = (f"""  if (i0 + {ELEMS_PER_THREAD} <= n && {aligned}) {{
    const long long run = i0 / {per}LL;
{pload}{vec_loads}
{rdecl}
    {name}_elem({lane('x')});
    {name}_elem({lane('y')});
    {name}_elem({lane('z')});
    {name}_elem({lane('w')});
{vstores}
    return;
  }}
""" if per % ELEMS_PER_THREAD == 0 else
               "  // runs are not a multiple of 4 elements long: element by element\n")
        code += f"""// ensemble form: {batch.runs} runs x {per} elements, {npar} scalar(s) per run in prm
extern "C" __global__ __launch_bounds__({WORKGROUP_SIZE}) void {name}(long long n, {kout}{kin}, const float* __restrict__ prm) {{
  const long long i0 = ((long long)blockIdx.x * {WORKGROUP_SIZE} + threadIdx.x) * {ELEMS_PER_THREAD};
  if (i0 >= n) return;
{vec}  for (long long i = i0; i < n && i < i0 + {ELEMS_PER_THREAD}; ++i) {{
    const long long run = i / {per}LL;
{pload}    {sdecl}
    {name}_elem({sargs});
    {sstores}
  }}
}}
"""
        ktotal = count * batch.runs
    k = Kernel(code=code, name=name, count=ktotal, label=label, ninputs=nin, nparams=npar,
               body=emitted)
    k.inputs = list(ext)
    k.stores = list(stores)
    return k


FFT_ROW_HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                              "csrc", "fft_row.h")


def fft_row_source() -> str:
    """The row-FFT device code the library's Fourier kernels are built from,
    embedded in generated grid kernels (hipRTC gets it as text)."""
    with open(FFT_ROW_HEADER) as f:
        return f.read()


def grid_lds_bytes(log2h: int, n_rows: int) -> int:
    """LDS of a grid kernel: two ping-pong rows plus `n_rows` analysis rows,
    each RB rows of H complex values per workgroup (the row FFT's twiddles
    are held in registers)."""
    h = 1 << log2h
    tpr = min(h // 4, 256)
    rb = 256 // tpr
    return (2 + n_rows) * rb * h * 8


def build_grid_kernel(name: str, label: str, log2h: int, nlat: int, mmax: int, synth, stmts,
                      ios, stores, analys, batch: "Batch | None" = None) -> Kernel:
    """One latitude row's FFT synthesis -> element-wise statements -> FFT
    analysis as ONE kernel (the planner's grid fusion).

    ``synth``: c_names of the synthesised grid values, in Fourier-scratch
    field order (field s of ``fm`` holds value synth[s] after the Legendre
    synthesis); ``stmts``/``ios``: the element-wise statements; ``stores``:
    grid values written to their buffers (synthesised or produced); ``analys``:
    c_names whose rows are analysed into Fourier-scratch field a, in order.
    The row transforms are the library's (fft_row.h), so the grid values and
    Fourier coefficients are those of the unfused sequence.  Each workgroup
    owns whole rows and reads every input row of ``fm`` before it writes any
    output row, so the analysis overwrites the scratch in place.

    Signature: (float2* fm, const float2* trig, float* st0.., const float*
    gin0.., float p0.. | const float* prm): gin are external grid inputs
    (neither synthesised nor produced), read from global memory.

    With ``batch`` (an ensemble of R runs) the grid is (row tiles, R):
    workgroup (x, run) takes the rows of one run, Fourier-scratch field
    s * R + run for value s (the batched Legendre stages' field order), grid
    values at run * nlat * nphi (one copy for inputs in ``batch.shared``),
    scalars from prm[run * npar + k]."""
    produced = [st.c_name for st in stmts]
    # element-wise outputs the kernel needs: stored values and analysis rows
    eouts = []
    for c in list(stores) + list(analys):
        if c in produced and c not in eouts:
            eouts.append(c)
    code, ext, npar, emitted = _emit_group(stmts, ios, eouts, name, label)
    gins = [c for c in ext if c not in synth]
    gins += [c for c in analys if c not in produced and c not in synth and c not in gins]
    for c in stores:
        if c not in produced and c not in synth:
            raise UnsupportedOnGpu(f"grid kernel cannot store '{c}', which it does not compute")
    arow = {c: q for q, c in enumerate(analys)}  # LDS row of each analysed value
    h = 1 << log2h
    ns = len(synth)
    R = batch.runs if batch is not None else 1
    shared = set(batch.shared) if batch is not None else set()
    groups = grid_groups(log2h, ns, len(analys))
    if groups:
        body = _grid_groups_kernel(name, label, log2h, nlat, mmax, synth, ext, gins, eouts,
                                   npar, stores, analys, R, shared, batch is not None, groups)
        k = Kernel(code=code + body, name=name, count=nlat * 2 * h * R, label=label,
                   ninputs=len(gins), nparams=npar, body=emitted)
        k.inputs, k.stores, k.grid, k.block = gins, list(stores), (log2h, nlat), 256 * groups
        return k
    lines = []
    ap = lines.append
    ap(f"// {label}")
    sig = (["float2* __restrict__ fm", "const float2* __restrict__ trig"] +
           [f"float* st{j}" for j in range(len(stores))] +
           [f"const float* gin{k}" for k in range(len(gins))] +
           ([f"float p{i}" for i in range(npar)] if batch is None else
            ["const float* __restrict__ prm"]))
    ap(f'extern "C" __global__ __launch_bounds__(256) void {name}({", ".join(sig)}) {{')
    ap(f"  using S = tsfft::FftShape<{log2h}>;")
    ap(f"  constexpr int NLAT = {nlat}, MMAX = {mmax}, NP = 4 * S::BPT, R = {R};")
    ap("  const int run = blockIdx.y;")
    ap("  const long long grun = (long long)run * NLAT * S::H;  // a run's grid, float2 units")
    if batch is not None:
        for i in range(npar):
            ap(f"  const float p{i} = prm[run * {npar} + {i}];")
    ap("  __shared__ float2 lds[2][S::RB * S::H];")
    ap(f"  __shared__ float2 arow[{max(1, len(analys))}][S::RB * S::H];")
    ap("  const int r = threadIdx.x / S::TPR, t = threadIdx.x % S::TPR;")
    ap(f"  tsfft::PassTwiddles<{log2h}> tw;")
    ap("  tw.load(trig, t);")
    ap("  const int ilat = blockIdx.x * S::RB + r;")
    ap("  const bool row = ilat < NLAT;")
    ap("  const int il = row ? ilat : NLAT - 1;")
    ap("  float2* a = lds[0] + r * S::H;")
    ap("  float2* b = lds[1] + r * S::H;")
    ap("  const long long g0 = (long long)il * S::H;  // this row, in complex (float2) units")
    for si in range(ns):
        ap(f"  float2 sv{si}[NP];")
        ap("  {")
        ap(f"    const float2* res = tsfft::synth_row<{log2h}>(fm + ((long long)({si} * R + run) "
           f"* NLAT + il) * (MMAX + 1), MMAX, 1.0f, trig, tw, a, b, t);")
        ap("#pragma unroll")
        ap(f"    for (int i = 0; i < NP; ++i) sv{si}[i] = res[t + S::TPR * i];")
        ap("    __syncthreads();  // the next row's first pass overwrites the ping-pong rows")
        ap("  }")
    def goff(cn):
        return "g0 + c" if cn in shared else "grun + g0 + c"

    # element-wise over this thread's points (complex index c: reals 2c, 2c+1)
    ap("#pragma unroll")
    ap("  for (int i = 0; i < NP; ++i) {")
    ap("    const int c = t + S::TPR * i;")
    for k, cn in enumerate(gins):
        ap(f"    const float2 gv{k} = reinterpret_cast<const float2*>(gin{k})[{goff(cn)}];")

    def operand(cn, comp):
        if cn in synth:
            return f"sv{synth.index(cn)}[i].{comp}"
        return f"gv{gins.index(cn)}.{comp}"

    for j in range(len(eouts)):
        ap(f"    float2 e{j};")
    for comp in ("x", "y"):
        args = ([operand(cn, comp) for cn in ext] + [f"p{i}" for i in range(npar)] +
                [f"e{j}.{comp}" for j in range(len(eouts))])
        ap(f"    {name}_elem({', '.join(args)});")

    def value(cn):
        if cn in eouts:
            return f"e{eouts.index(cn)}"
        if cn in synth:
            return f"sv{synth.index(cn)}[i]"
        return f"gv{gins.index(cn)}"

    ap("    if (row) {")
    for j, cn in enumerate(stores):
        ap(f"      reinterpret_cast<float2*>(st{j})[grun + g0 + c] = {value(cn)};")
    ap("    }")
    for cn, q in arow.items():
        ap(f"    arow[{q}][r * S::H + c] = {value(cn)};")
    ap("  }")
    ap("  __syncthreads();")
    for q, cn in enumerate(analys):
        ap("  {")
        ap(f"    const float2* src = arow[{q}] + r * S::H;")
        ap(f"    const float2* z = tsfft::analys_row<{log2h}>([&](int j) {{ return src[j]; }}, "
           f"tw, a, b, t);")
        ap("    if (row) {")
        ap(f"      float2* G = fm + ((long long)({q} * R + run) * NLAT + ilat) * (MMAX + 1);")
        ap("      for (int k = t; k <= MMAX; k += S::TPR)")
        ap(f"        G[k] = tsfft::analys_coef<{log2h}>(z, k, trig);")
        ap("    }")
        ap("    __syncthreads();")
        ap("  }")
    ap("}")
    k = Kernel(code=code + "\n".join(lines) + "\n", name=name, count=nlat * 2 * h * R,
               label=label, ninputs=len(gins), nparams=npar, body=emitted)
    k.inputs = gins
    k.stores = list(stores)
    k.grid = (log2h, nlat)
    k.block = 256
    return k


GRID_GROUPS = True  # False: always the one-group form (tests compare the two)


def grid_groups(log2h: int, ns: int, na: int) -> int:
    """Thread groups of a group-parallel grid kernel (one 256-thread group per
    transformed row, each with its own two LDS rows: the fields' row FFTs run
    side by side), or 0 when they do not fit a workgroup (then the rows are
    transformed one after another by one group)."""
    if not GRID_GROUPS:
        return 0
    g = max(ns, na, 1)
    g = 1 if g == 1 else 2 if g == 2 else 4 if g <= 4 else 0
    if not g:
        return 0
    h = 1 << log2h
    rb = 256 // min(h // 4, 256)
    return g if g * 2 * rb * h * 8 <= 65536 else 0


def _grid_groups_kernel(name, label, log2h, nlat, mmax, synth, ext, gins, eouts, npar, stores,
                        analys, R, shared, ensemble, G) -> str:
    """Group-parallel grid kernel: group g (threads 256 g .. 256 g + 255)
    synthesises value synth[g] and later analyses analys[g], each in its own
    ping-pong rows; between the two, every thread of the workgroup evaluates
    the element-wise statements over a share of the rows' points, reading the
    synthesised rows and writing the analysed ones in LDS (in place where a
    synthesised row is no longer needed: each point is read and written by
    one thread)."""
    h = 1 << log2h
    q0 = 1 if log2h & 1 else 2
    npass = (log2h - q0) // 2
    res = npass % 2          # ping-pong row holding a transform's result
    ns, na = len(synth), len(analys)
    lines = []
    ap = lines.append
    ap(f"// {label}")
    sig = (["float2* __restrict__ fm", "const float2* __restrict__ trig"] +
           [f"float* st{j}" for j in range(len(stores))] +
           [f"const float* gin{k}" for k in range(len(gins))] +
           ([f"float p{i}" for i in range(npar)] if not ensemble else
            ["const float* __restrict__ prm"]))
    # rows of up to 1024 complex points take G x 16 KiB of LDS, which leaves
    # room for 8 waves per SIMD: registers are capped to match (64 VGPRs;
    # without the cap the pass twiddles held through the element-wise part
    # take the two-field 1024 x 2048 kernel to 70 and a CU holds 3 workgroups instead of 4)
    occ = " __attribute__((amdgpu_waves_per_eu(8)))" if log2h <= 10 else ""
    ap(f'extern "C" __global__ __launch_bounds__({256 * G}){occ} void {name}({", ".join(sig)}) {{')
    ap(f"  using S = tsfft::FftShape<{log2h}>;")
    ap(f"  constexpr int NLAT = {nlat}, MMAX = {mmax}, R = {R}, G = {G};")
    ap("  constexpr int NPT = S::RB * S::H / (256 * G);  // complex points per thread")
    ap("  __shared__ float2 lds[G][2][S::RB * S::H];")
    ap("  const int run = blockIdx.y;")
    ap("  const long long grun = (long long)run * NLAT * S::H;  // a run's grid, float2 units")
    if ensemble:
        for i in range(npar):
            ap(f"  const float p{i} = prm[run * {npar} + {i}];")
    ap("  const int g = threadIdx.x >> 8, lt = threadIdx.x & 255;")
    ap("  const int r = lt / S::TPR, t = lt % S::TPR;")
    ap(f"  tsfft::PassTwiddles<{log2h}> tw;")
    ap("  tw.load(trig, t);")
    ap("  const int ilat = blockIdx.x * S::RB + r;")
    ap("  const int il = ilat < NLAT ? ilat : NLAT - 1;")
    ap("  float2* a = lds[g][0] + r * S::H;")
    ap("  float2* b = lds[g][1] + r * S::H;")
    # synthesis: group g < ns
    ap(f"  const bool sact = g < {ns};")
    ap("  if (sact) {")
    ap(f"    const tsfft::SynthInput<{log2h}> in{{fm + ((long long)(g * R + run) * NLAT + il) * "
       f"(MMAX + 1), trig, MMAX, 1.0f}};")
    ap(f"    tsfft::first_pass<+1, {log2h}>(in, a, t);")
    ap("  }")
    ap(f"  tsfft::rest_passes_masked<+1, {log2h}>(sact, a, b, tw, t);")
    # element-wise over the block's points

    def src_of(cn):  # LDS row of a synthesised value
        return f"lds[{synth.index(cn)}][{res}]"

    def dst_of(q):   # LDS row an analysed value is written to / read from
        return f"lds[{q}][{res}]" if q < ns else f"lds[{q}][0]"

    def goff(cn):
        return "g0 + c" if cn in shared else "grun + g0 + c"

    ap("#pragma unroll")
    ap("  for (int i = 0; i < NPT; ++i) {")
    ap("    const int ca = threadIdx.x + 256 * G * i;   // point index in the block's rows")
    ap("    const int rr = ca / S::H, c = ca % S::H;")
    ap("    const int lat = blockIdx.x * S::RB + rr;")
    ap("    const long long g0 = (long long)(lat < NLAT ? lat : NLAT - 1) * S::H;")
    for k, cn in enumerate(gins):
        ap(f"    const float2 gv{k} = reinterpret_cast<const float2*>(gin{k})[{goff(cn)}];")
    for cn in synth:
        ap(f"    const float2 sv{synth.index(cn)} = {src_of(cn)}[ca];")

    def operand(cn, comp):
        if cn in synth:
            return f"sv{synth.index(cn)}.{comp}"
        return f"gv{gins.index(cn)}.{comp}"

    for j in range(len(eouts)):
        ap(f"    float2 e{j};")
    for comp in ("x", "y"):
        args = ([operand(cn, comp) for cn in ext] + [f"p{i}" for i in range(npar)] +
                [f"e{j}.{comp}" for j in range(len(eouts))])
        ap(f"    {name}_elem({', '.join(args)});")

    def value(cn):
        if cn in eouts:
            return f"e{eouts.index(cn)}"
        if cn in synth:
            return f"sv{synth.index(cn)}"
        return f"gv{gins.index(cn)}"

    ap("    if (lat < NLAT) {")
    for j, cn in enumerate(stores):
        ap(f"      reinterpret_cast<float2*>(st{j})[grun + g0 + c] = {value(cn)};")
    ap("    }")
    for q, cn in enumerate(analys):
        ap(f"    {dst_of(q)}[ca] = {value(cn)};")
    ap("  }")
    ap("  __syncthreads();")
    # analysis: group g < na reads its row, transforms into the other buffer
    ap(f"  const bool aact = g < {na};")
    ap(f"  float2* ain = {'lds[g][' + str(res) + ']' if ns >= na else '(g < ' + str(ns) + ' ? lds[g][' + str(res) + '] : lds[g][0])'} + r * S::H;")
    ap(f"  float2* aout = (ain == lds[g][0] + r * S::H ? lds[g][1] : lds[g][0]) + r * S::H;")
    ap("  if (aact) {")
    ap(f"    tsfft::first_pass<-1, {log2h}>([&](int j) {{ return ain[j]; }}, aout, t);")
    ap("  }")
    ap(f"  const float2* z = tsfft::rest_passes_masked<-1, {log2h}>(aact, aout, ain, tw, t);")
    ap("  if (aact && ilat < NLAT) {")
    ap("    float2* G2 = fm + ((long long)(g * R + run) * NLAT + ilat) * (MMAX + 1);")
    ap("    for (int k = t; k <= MMAX; k += S::TPR)")
    ap(f"      G2[k] = tsfft::analys_coef<{log2h}>(z, k, trig);")
    ap("  }")
    ap("}")
    return "\n".join(lines) + "\n"


@dataclass
class Batch:
    """Ensemble shape of a generated kernel: `runs` runs; inputs whose
    c_names are in `shared` hold one copy for all runs."""
    runs: int
    shared: frozenset = frozenset()


MODULE_PRELUDE = f"""// generated element-wise kernels (tsphere.hipgen)
// (hipRTC provides the HIP device headers implicitly)
__device__ __forceinline__ bool ts_al(const void* p) {{ return (reinterpret_cast<unsigned long long>(p) & 15ull) == 0ull; }}
{_SIGN}
"""


def module_source(kernels) -> str:
    """One hipRTC translation unit for all kernels of a model function: a
    model compiles with one hipRTC invocation."""
    seen = set()
    parts = [MODULE_PRELUDE]
    if any(getattr(k, "grid", None) for k in kernels):
        parts.append(fft_row_source())
    for k in kernels:
        code = k.code
        # pow helpers are shared: emit each once per module
        lines = []
        for line in code.splitlines():
            m = re.match(r"__device__ __forceinline__ float (pow_i\d+)\(", line)
            if m:
                if m.group(1) in seen:
                    continue
                seen.add(m.group(1))
            lines.append(line)
        parts.append("\n".join(lines))
    return "\n".join(parts)


# --------------------------------------------------------- f32 evaluation
def eval_f32(e: IRExpr, env: dict, params: dict, scalars: dict | None = None):
    """Evaluate an element-wise IR tree with numpy float32, one rounding per
    operation in tree order -- what the generated kernel computes (it is
    compiled without contraction).  Used by tests as the expected output."""
    f = np.float32
    scalars = scalars or {}

    def ev(x):
        k = x.kind
        if k == "NumLit":
            return f(x.value)
        if k == "Var":
            if is_tensor(x.ty):
                return np.asarray(env[x.c_name], np.float32)
            if is_numeric(x.ty) and isinstance(x.ty.exact, (int, float)):
                return f(x.ty.exact)
            if x.c_name in params:
                return f(params[x.c_name])
            return ev(scalars[x.c_name][1])
        if k == "Binary":
            if x.builtin in ("power", "mpower"):
                kk = int(x.right.value if x.right.kind == "NumLit" else x.right.ty.exact)
                b = ev(x.left)
                if kk == 0:
                    return f(1)
                r = b
                for _ in range(abs(kk) - 1):
                    r = (r * b).astype(np.float32) if isinstance(r, np.ndarray) else f(r * b)
                return (f(1) / r) if kk < 0 else r
            a, b = ev(x.left), ev(x.right)
            op = BINARY_OPS[x.builtin]
            with np.errstate(all="ignore"):
                r = {"+": np.add, "-": np.subtract, "*": np.multiply, "/": np.divide}[op](a, b)
            return np.float32(r) if np.isscalar(r) else r.astype(np.float32)
        if k == "Unary":
            v = ev(x.operand)
            return -v if x.builtin == "uminus" else v
        if k == "Call":
            if x.name == "ones":
                return f(1)
            if x.name == "zeros":
                return f(0)
            fn = {"abs": np.abs, "sqrt": np.sqrt, "floor": np.floor, "ceil": np.ceil,
                  "round": np.rint, "sign": np.sign, "max": np.maximum, "min": np.minimum}
            args = [ev(a) for a in x.args]
            return fn[x.name](*args).astype(np.float32)
        raise ValueError(k)

    return ev(e)
