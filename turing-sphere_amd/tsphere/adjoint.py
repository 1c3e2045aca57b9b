"""Reverse-mode gradients through IMEX steps of a compiled .m model.

No reference counterpart: turing-sphere only runs models forward.  The step
function is compiled as for the forward run (src/mgpu/plan.ts), but as a
*tape*: every value keeps its own buffer (no in-place feedback), so after
one step all of the step's intermediates are on the device.  Reverse mode
then walks the op list backwards:

* element-wise statement ``out = f(in..., p...)``: one generated HIP kernel
  per statement adds ``out_bar * df/din`` to each input's cotangent and
  reduces ``out_bar * df/dp`` per workgroup into a per-parameter partial
  slot (deterministic: no atomics; the slots are summed in f64 at the end).
  The derivatives are taken symbolically from the statement's IR tree;
* ``x = synth(y)``: ``y_bar += S^T x_bar`` (``tsphere_synth_adjoint``);
* ``x = analys(y)``: ``y_bar += A^T x_bar`` (``tsphere_analys_adjoint``).

Trajectories are checkpointed: the forward pass keeps the state every
``checkpoint_every`` steps; the backward pass re-runs each segment from its
checkpoint, keeping the segment's step inputs, and differentiates its steps
in reverse, re-running each step once more to regenerate its intermediates.
Steps that only advance the state run through the forward plan (fused, with
the spectral op where the step has its form; bit-identical to the tape's
step), the initial pass as HIP-graph replays; each segment of the backward
pass is one captured graph (keep every step's input while advancing, then
regenerate and differentiate the steps in reverse), so a gradient is a short
sequence of graph launches.
"""
from __future__ import annotations

import ctypes as C
import math
import re

import numpy as np

from . import _lib
from .device import DeviceBuffer, Stream
from .hipgen import UnsupportedOnGpu, _Ctx, _emit, f32_lit
from .ir import is_numeric, is_tensor
from .layout import ShtConfig
from .mfront import CompiledModel, model_bindings
from .model import eigenvalues
from .plan import FusedOp, GridOp, HostBuffers, KernelOp, ModelPlan, PlanSpec, XformOp, _compile

WG = 256
EPT = 4  # elements per thread (tsphere_kernel_launch launches count / 4 threads)


# ------------------------------------------------------ symbolic derivative
def _add(a, b):
    if a is None:
        return b
    if b is None:
        return a
    return f"({a} + {b})"


def _sub(a, b):
    if b is None:
        return a
    if a is None:
        return f"(-{b})"
    return f"({a} - {b})"


def _mul(a, b):
    if a is None or b is None:
        return None
    if a == "1.0f":
        return b
    if b == "1.0f":
        return a
    return f"({a} * {b})"


def _div(a, b):
    if a is None:
        return None
    return f"({a} / {b})"


def derivative(e, wrt: str, ctx: _Ctx):
    """HIP code for d(e)/d(wrt) (None = identically zero).  `wrt` is a tensor
    c_name or a runtime-parameter c_name; derived scalars are differentiated
    through their defining expressions."""
    io = ctx.io
    k = e.kind
    if k == "NumLit":
        return None
    if k == "Var":
        if is_tensor(e.ty):
            return "1.0f" if e.c_name == wrt else None
        if is_numeric(e.ty) and isinstance(e.ty.exact, (int, float)):
            return None
        if e.c_name in io.params:
            return "1.0f" if e.c_name == wrt else None
        if e.c_name in io.scalars:
            return derivative(io.scalars[e.c_name][1], wrt, ctx)
        raise UnsupportedOnGpu(f"cannot differentiate '{e.name}'", e.span)
    if k == "Unary":
        d = derivative(e.operand, wrt, ctx)
        if e.builtin == "uminus":
            return None if d is None else f"(-{d})"
        return d
    if k == "Binary":
        b = e.builtin
        if b in ("power", "mpower"):
            kv = e.right.value if e.right.kind == "NumLit" else e.right.ty.exact
            kk = int(kv)
            d = derivative(e.left, wrt, ctx)
            if d is None or kk == 0:
                return None
            x = _emit(e.left, ctx)
            if kk == 1:
                return d
            pw = x if kk - 1 == 1 else (f"pow_i{kk - 1}({x})" if kk - 1 > 0 else
                                        f"(1.0f / pow_i{1 - kk}({x}))" if kk - 1 < -1 else
                                        f"(1.0f / {x})")
            return _mul(f"{f32_lit(float(kk))} * {pw}", d)
        dl = derivative(e.left, wrt, ctx)
        dr = derivative(e.right, wrt, ctx)
        if b == "plus":
            return _add(dl, dr)
        if b == "minus":
            return _sub(dl, dr)
        if b in ("times", "mtimes"):
            return _add(_mul(dl, _emit(e.right, ctx)), _mul(_emit(e.left, ctx), dr))
        if b in ("rdivide", "mrdivide"):
            r = _emit(e.right, ctx)
            t1 = _div(dl, r)
            t2 = None if dr is None else f"({_emit(e.left, ctx)} * {dr} / ({r} * {r}))"
            return _sub(t1, t2)
        raise UnsupportedOnGpu(f"cannot differentiate operator '{b}'", e.span)
    if k == "Call":
        if e.name in ("ones", "zeros", "sign", "floor", "ceil", "round"):
            return None
        args = [_emit(a, ctx) for a in e.args]
        ds = [derivative(a, wrt, ctx) for a in e.args]
        x = args[0] if args else None
        d = ds[0] if ds else None
        rules = {
            "exp": lambda: _mul(f"expf({x})", d),
            "log": lambda: _div(d, x),
            "log2": lambda: _div(d, f"({x} * 0.6931471805599453f)"),
            "sqrt": lambda: _div(d, f"(2.0f * sqrtf({x}))"),
            "sin": lambda: _mul(f"cosf({x})", d),
            "cos": lambda: _mul(f"(-sinf({x}))", d),
            "tan": lambda: _div(d, f"(cosf({x}) * cosf({x}))"),
            "sinh": lambda: _mul(f"coshf({x})", d),
            "cosh": lambda: _mul(f"sinhf({x})", d),
            "tanh": lambda: _mul(f"(1.0f - tanhf({x}) * tanhf({x}))", d),
            "abs": lambda: _mul(f"ts_sign({x})", d),
            "atan": lambda: _div(d, f"(1.0f + {x} * {x})"),
            "asin": lambda: _div(d, f"sqrtf(1.0f - {x} * {x})"),
            "acos": lambda: None if d is None else f"(-{d} / sqrtf(1.0f - {x} * {x}))",
        }
        if e.name in rules:
            return rules[e.name]()
        if e.name in ("max", "min"):
            op = ">=" if e.name == "max" else "<="
            a0 = ds[0] or "0.0f"
            a1 = ds[1] or "0.0f"
            if ds[0] is None and ds[1] is None:
                return None
            return f"(({args[0]} {op} {args[1]}) ? {a0} : {a1})"
        if e.name == "atan2":
            y, xx = args
            dy, dx = ds
            num = _sub(_mul(xx, dy), _mul(y, dx))
            return _div(num, f"({xx} * {xx} + {y} * {y})")
        raise UnsupportedOnGpu(f"cannot differentiate '{e.name}'", e.span)
    raise UnsupportedOnGpu(f"cannot differentiate '{k}'", getattr(e, "span", None))


def build_adjoint_kernel(ops_gis, name: str, nparams: int):
    """HIP source of the adjoint of consecutive element-wise statements (the
    statements of one fused kernel, or one statement), taken per element in
    the order given -- the reverse of the step's.

    ``ops_gis``: [(KernelOp, grad_inputs)], grad_inputs the c_names whose
    cotangents the statement adds to.  Signature: (long long n, const
    float* in0.. (``ins``), float* bar0.. (``bars``), float* partials, float
    p0..).  For each element i and statement k: ob = bar[out_k][i];
    bar[c][i] += ob * d out_k / d c for c in its grad inputs;
    partials[block * nparams + j] += the block's sum of ob * d out_k / d p_j.
    A statement's output cotangent may have been added to by a later
    statement of the step, processed before it here for the same element,
    so the cotangent pointers carry no __restrict__.
    Returns (code, ins, bars)."""
    ins, bars, blocks = [], [], []
    dps_all = {}  # parameter slot -> used by some statement
    helpers_src = ""
    for k, (op, gi) in enumerate(ops_gis):
        io = op.io
        ctx = _Ctx(io)
        _emit(op.stmt.expr, ctx)  # binds derived scalars used by the value (prologue)
        dins = [derivative(op.stmt.expr, c, ctx) for c in gi]
        dps = {io.params[c]: derivative(op.stmt.expr, c, ctx) for c in io.params}
        ren = lambda t, k=k: re.sub(r"\ba(\d+)\b", rf"s{k}a\1", t)
        for c in sorted(io.tensors, key=lambda c: io.tensors[c]):
            if c not in ins:
                ins.append(c)
        for c in [op.out] + list(gi):
            if c not in bars:
                bars.append(c)
        loads = "\n".join(f"        const float s{k}a{io.tensors[c]} = in{ins.index(c)}[i];"
                          for c in sorted(io.tensors, key=lambda c: io.tensors[c]))
        pro = "\n".join("        " + ren(line.strip()) for line in ctx.prologue)
        upd = "\n".join(f"        bar{bars.index(c)}[i] += ob * {ren(d)};"
                        for c, d in zip(gi, dins) if d is not None)
        acc = "\n".join(f"        c{j} += ob * {ren(d)};"
                        for j, d in sorted(dps.items()) if d is not None)
        for j, d in dps.items():
            if d is not None:
                dps_all[j] = True
        helpers_src += upd + acc + pro
        blocks.append(f"""      {{  // adjoint of: {op.label}
        const float ob = bar{bars.index(op.out)}[i];
{loads}
{pro}
{upd}
{acc}
      }}""")
    np_ = max(1, nparams)
    kin = "".join(f", const float* __restrict__ in{i}" for i in range(len(ins)))
    kbar = "".join(f", float* bar{j}" for j in range(len(bars)))
    kpar = "".join(f", float p{i}" for i in range(nparams))
    cdecl = " ".join(f"float c{j} = 0.0f;" for j in range(nparams))
    live = sorted(dps_all)  # parameters some statement uses
    red = ""
    for j in live:
        red += f"""  for (int o = 32; o > 0; o >>= 1) c{j} += __shfl_xor(c{j}, o);
  if ((threadIdx.x & 63) == 0) red[{j}][threadIdx.x >> 6] = c{j};
"""
    fin = "\n".join(
        f"    partials[(long long)blockIdx.x * {np_} + {j}] += "
        f"(red[{j}][0] + red[{j}][1]) + (red[{j}][2] + red[{j}][3]);" for j in live)
    used = {int(m) for m in re.findall(r"\bpow_i(\d+)\(", helpers_src)}
    helpers = "\n".join(
        f"__device__ __forceinline__ float pow_i{k}(float x) {{ return {'x' + ' * x' * (k - 1)}; }}"
        for k in sorted(used))
    body = "\n".join(blocks)
    code = f"""{helpers}
extern "C" __global__ __launch_bounds__({WG}) void {name}(long long n{kin}{kbar}, float* __restrict__ partials{kpar}) {{
  __shared__ float red[{np_}][4];
  // {EPT} consecutive elements per thread: the launch contract of tsphere_kernel_launch
  const long long i0 = ((long long)blockIdx.x * {WG} + threadIdx.x) * {EPT};
  {cdecl}
  for (long long i = i0; i < n && i < i0 + {EPT}; ++i) {{
{body}
  }}
{red}{"  __syncthreads();" if live else ""}
  if ({"threadIdx.x == 0" if live else "false"}) {{
{fin}
  }}
}}
"""
    return code, ins, bars


# ------------------------------------------------------------- the engine
class AdjointModel:
    """Forward trajectories and vector-Jacobian products of a .m model's
    step, on one spherical-harmonic plan."""

    def __init__(self):
        self.stream = Stream()
        self._module = None
        self._kernels = []
        self._owned = []

    @staticmethod
    def create(sht, cfg: ShtConfig, source: str, param_names, state, view,
               checkpoint_every: int = 20, file_name: str = "model.m") -> "AdjointModel":
        self = AdjointModel()
        self.sht, self.cfg = sht, cfg
        self.state, self.param_names = list(state), list(param_names)
        self.checkpoint_every = max(1, int(checkpoint_every))
        npts, nlm = cfg.nlat * cfg.nphi, sht.nlm
        self.npts, self.nlm = npts, nlm
        cm = CompiledModel(source, model_bindings(state, param_names, npts, nlm),
                           {"npts": npts, "nlm": nlm}, file_name)
        fn = cm.specialize("step", len(state) + len(view))
        cm.finish()
        assigned = [s.c_name for s in fn.body if s.kind == "Assign"]
        if len(assigned) != len(set(assigned)):
            raise UnsupportedOnGpu("reverse mode needs each value of the step assigned once")
        host = HostBuffers()
        host.ensure_group(self.state, 2 * nlm)
        host.ensure("lam", 2 * nlm)
        self._host = host
        # the step tape: every value stored for the reverse pass, consecutive
        # element-wise statements in one kernel each and a synth ->
        # element-wise -> analys run in one grid kernel (fewer graph nodes per
        # regenerated step)
        tape = ModelPlan.create(sht, PlanSpec(fn, [None] * len(fn.outputs)), host, True, "tape",
                                keep_all=True)
        self.tape = tape
        # state <- step(state) only: the fused plan, writing the state in place
        # (the spectral op where the step has its form, else the fold: the
        # same states as the separate launches, bit for bit)
        self.fwd = ModelPlan.create(sht, PlanSpec(fn, list(self.state) +
                                                  [None] * (len(fn.outputs) - len(self.state))),
                                    host, True, "forward", spectral=True, fold=True)
        self._graphs = {}
        self._seg = self._carry = None
        host.upload("lam", eigenvalues(cfg, nlm), self.stream)
        # state outputs: the value each state variable takes after the step
        self.out_of = {}
        for i, s in enumerate(self.state):
            self.out_of[s] = fn.outputs[i].c_name
        in_c = {p.name: p.c_name for p in fn.params}
        self.in_of = {s: in_c[s] for s in self.state}
        # cotangent buffers for every tensor value except host constants, in
        # one arena: zeroed by one fill, with the state outputs' and the state
        # inputs' cotangents each contiguous in state order, so that the
        # carried cotangent moves in and out with one copy each
        const_c = {in_c[n] for n in ("lam", "noise") if n in in_c}
        names = [c for c in tape._bufs if c not in const_c]
        first = [self.out_of[s] for s in self.state] + [self.in_of[s] for s in self.state]
        order = [c for c in first if c in names] + [c for c in names if c not in first]
        self._bar_arena = DeviceBuffer(max(1, sum(tape._bufs[c].count for c in order)), "bars")
        self._bar = {}
        off = 0
        for c in order:
            self._bar[c] = DeviceBuffer.view(self._bar_arena, off, tape._bufs[c].count, f"bar-{c}")
            off += tape._bufs[c].count
        self._tmp = {}
        # adjoint kernels, in reverse op order
        self._adj = []
        srcs = []
        self.param_slots = list(tape.param_names)
        npar = len(self.param_slots)
        rev = []
        for op in reversed(tape.ops):  # a grid op: its Legendre/element-wise/Legendre ops
            rev.extend(reversed(op.members) if isinstance(op, GridOp) else [op])
        for idx, op in enumerate(rev):
            if isinstance(op, (KernelOp, FusedOp)):
                # a fused kernel's statements: one adjoint kernel, reversed
                members = list(reversed(op.members)) if isinstance(op, FusedOp) else [op]
                ops_gis = [(m, [c for c in m.ins if c in self._bar]) for m in members]
                name = f"adj{idx}"
                code, ins, bars = build_adjoint_kernel(ops_gis, name, npar)
                count = members[0].count
                nblk = (count + WG * EPT - 1) // (WG * EPT)
                part = DeviceBuffer(nblk * max(1, npar), f"partials-{name}")
                self._owned.append(part)
                gi = sorted({c for _, g in ops_gis for c in g})
                self._adj.append(("kernel", (ins, bars, count), name, gi, part, nblk))
                srcs.append(code)
            elif isinstance(op, XformOp):
                if not any(s_ in self._tmp for s_ in op.srcs):
                    # the transpose's outputs side by side: one accumulate for all
                    tot = sum(tape._bufs[s_].count for s_ in op.srcs)
                    arena = DeviceBuffer(tot, f"tmp-{op.kind}{idx}")
                    self._owned.append(arena)
                    off = 0
                    for s_ in op.srcs:
                        self._tmp[s_] = DeviceBuffer.view(arena, off, tape._bufs[s_].count)
                        off += tape._bufs[s_].count
                for s_ in op.srcs:
                    if s_ not in self._tmp:
                        self._tmp[s_] = DeviceBuffer(tape._bufs[s_].count, f"tmp-{s_}")
                self._adj.append((op.kind, op, set()))
            else:
                raise UnsupportedOnGpu("unexpected op in the step tape")
        # a transpose whose result is the first contribution to a cotangent
        # (zeroed, nothing added before it in the reverse pass) writes it
        # directly: no temporary, no accumulate launch
        touched = {self.out_of[s] for s in self.state}  # the carried cotangent
        for a in self._adj:
            if a[0] == "kernel":
                touched |= set(a[3])
            else:
                a[2].update(s_ for s_ in a[1].srcs if s_ in self._bar and s_ not in touched)
                touched |= {s_ for s_ in a[1].srcs if s_ in self._bar}
        pre = ("__device__ __forceinline__ float ts_sign(float x) "
               "{ return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }\n")
        self.adjoint_source = pre + "\n".join(_dedupe_helpers(srcs))
        if srcs:
            self._module = _compile(self.adjoint_source, "adjoint")
            for i, a in enumerate(self._adj):
                if a[0] == "kernel":
                    h = C.c_void_p()
                    _lib.call("tsphere_module_kernel", self._module, a[2].encode(), C.byref(h))
                    self._kernels.append(h.value)
                    self._adj[i] = a + (h.value,)
        self.state_count = 2 * nlm * len(self.state)
        self._params = np.zeros(max(1, npar), np.float32)
        return self

    # -------------------------------------------------------------- utils
    def set_params(self, values: dict) -> None:
        self.tape.set_params(values)
        self.fwd.set_params(values)
        self._params[:len(self.param_slots)] = self.tape._params[:len(self.param_slots)]
        self._drop_graphs()  # their launches hold the old scalars

    def _drop_graphs(self) -> None:
        for g in self._graphs.values():
            _lib.call("tsphere_graph_destroy", g)
        self._graphs = {}

    def _replay(self, key, record) -> None:
        """Launch the graph captured under `key`, capturing `record()` on
        first use (every pointer it touches is fixed until the graphs are
        dropped)."""
        g = self._graphs.get(key)
        if g is None:
            _lib.call("tsphere_graph_begin", self.stream.handle)
            try:
                record()
            finally:
                h = C.c_void_p()
                rc = _lib.load().tsphere_graph_end(self.stream.handle, C.byref(h))
            _lib.check(rc, "graph capture")
            g = self._graphs[key] = h.value
        _lib.call("tsphere_graph_launch", g, self.stream.handle)

    def _state_buf(self, s):
        return self._host.get(s)

    def _contiguous(self, bufs) -> bool:
        """Whether the per-species buffers follow one another in memory."""
        n = 2 * self.nlm
        return all(b.ptr == bufs[0].ptr + 4 * i * n and b.count == n for i, b in enumerate(bufs))

    def _copy_species(self, dsts, srcs) -> None:
        """Copy one state vector's species: one copy when both sides are
        contiguous (the state arena, the cotangent arena, the carry)."""
        L = _lib.load()
        n = 2 * self.nlm
        if self._contiguous(dsts) and self._contiguous(srcs):
            _lib.check(L.tsphere_memcpy_dtod(dsts[0].ptr, srcs[0].ptr, 4 * n * len(dsts),
                                             self.stream.handle), "copy state")
            return
        for d, s_ in zip(dsts, srcs):
            _lib.check(L.tsphere_memcpy_dtod(d.ptr, s_.ptr, 4 * n, self.stream.handle),
                       "copy state")

    def _at(self, buf: DeviceBuffer, off: int):
        """The species slices of one state vector in `buf` at float offset off."""
        n = 2 * self.nlm
        return [DeviceBuffer.view(buf, off + i * n, n) for i in range(len(self.state))]

    def _load_state(self, src: DeviceBuffer, off: int = 0) -> None:
        """Copy one state vector (all species) from `src` at float offset off."""
        self._copy_species([self._state_buf(s) for s in self.state], self._at(src, off))

    def _store_state(self, dst: DeviceBuffer, off: int, from_outputs: bool) -> None:
        srcs = [self.tape._bufs[self.out_of[s]] if from_outputs else self._state_buf(s)
                for s in self.state]
        self._copy_species(self._at(dst, off), srcs)

    def _step_once(self) -> None:
        self.tape.encode_steps(self.stream, 1)

    # ------------------------------------------------------------ forward
    def forward(self, state0: dict, steps: int) -> dict:
        """Run `steps` steps from state0 (host arrays); returns the final state."""
        for s in self.state:
            self._state_buf(s).upload(np.asarray(state0[s], np.float32), self.stream)
        self.fwd.run(self.stream, steps, max(1, min(steps, self.checkpoint_every)))
        return {s: self._state_buf(s).download(self.stream) for s in self.state}

    # ------------------------------------------------------------ backward
    def _zero_bars(self):
        self._bar_arena.fill(0.0, self.stream)

    def _backward_step(self) -> None:
        """With the step's intermediates on the device and the cotangents of
        its outputs in their bar buffers, add the cotangents of its inputs."""
        L = _lib.load()
        sh = self.stream.handle
        pv = self._params.ctypes.data_as(C.POINTER(C.c_float))
        npar = len(self.param_slots)
        nq, ns = 2 * self.nlm, self.npts
        for a in self._adj:
            if a[0] == "kernel":
                _, (ins, bars, count), name, gi, part, nblk, h = a
                ptrs = ([self.tape._bufs[c].ptr for c in ins] + [self._bar[c].ptr for c in bars] +
                        [part.ptr])
                arr = (C.c_void_p * len(ptrs))(*ptrs)
                _lib.check(L.tsphere_kernel_launch(h, count, len(ptrs), arr, npar, pv, sh),
                           f"adjoint kernel {name}")
                continue
            kind, op, direct = a
            n = len(op.srcs)
            obar = (C.c_void_p * n)(*[self._bar[d].ptr for d in op.dsts])
            tmp = (C.c_void_p * n)(*[(self._bar if s in direct else self._tmp)[s].ptr
                                     for s in op.srcs])
            if kind == "synth":   # x = S y: y_bar += S^T x_bar
                rc = L.tsphere_synth_adjoint(self.sht.handle, n, 1, obar, 0, tmp, 0, sh)
            else:                 # x = A y: y_bar += A^T x_bar
                rc = L.tsphere_analys_adjoint(self.sht.handle, n, 1, obar, 0, tmp, 0, sh)
            _lib.check(rc, f"transpose of {op.label}")
            acc = [s_ for s_ in op.srcs if s_ in self._bar and s_ not in direct]
            if len(acc) > 1 and all(
                    self._bar[c].ptr == self._bar[acc[0]].ptr + 4 * k * self._bar[c].count and
                    self._tmp[c].ptr == self._tmp[acc[0]].ptr + 4 * k * self._tmp[c].count and
                    self._tmp[c].count == self._tmp[acc[0]].count for k, c in enumerate(acc)):
                _lib.check(L.tsphere_accumulate(self._bar[acc[0]].ptr, self._tmp[acc[0]].ptr,
                                                len(acc) * self._tmp[acc[0]].count, sh),
                           "accumulate")
                continue
            for s_ in acc:
                _lib.check(L.tsphere_accumulate(self._bar[s_].ptr, self._tmp[s_].ptr,
                                                self._tmp[s_].count, sh), "accumulate")

    def vjp(self, state0: dict, steps: int, cot: dict):
        """Cotangents of the initial state and of the parameters for a
        cotangent `cot` of the state after `steps` steps.  Returns
        ({state: array}, {param: float})."""
        n1 = self.state_count
        C_ = self.checkpoint_every
        nseg = max(1, math.ceil(steps / C_))
        chk = DeviceBuffer(nseg * n1, "checkpoints")
        if self._seg is None or self._seg.count < C_ * n1:
            # the segment graphs hold the segment and carry buffers: keep
            # them for the model's life, re-capturing only when they grow
            self._drop_graphs()
            if self._seg is not None:
                self._seg.free()
            self._seg = DeviceBuffer(C_ * n1, "segment-inputs")
        if self._carry is None:
            self._carry = DeviceBuffer(n1, "state-cotangent")
        seg, carry = self._seg, self._carry
        try:
            for s in self.state:
                self._state_buf(s).upload(np.asarray(state0[s], np.float32), self.stream)
            for g in range(nseg):
                self._store_state(chk, g * n1, from_outputs=False)
                self.fwd.run(self.stream, min(C_, steps - g * C_), C_)
            carry.upload(np.concatenate([np.asarray(cot[s], np.float32).reshape(-1)
                                         for s in self.state]), self.stream)
            for a in self._adj:
                if a[0] == "kernel":
                    a[4].fill(0.0, self.stream)
            nq = 2 * self.nlm

            def keep_and_advance(j):
                self._store_state(seg, j * n1, from_outputs=False)
                self.fwd.encode_steps(self.stream, 1)

            def differentiate(j):
                self._load_state(seg, j * n1)
                self._step_once()          # regenerate this step's intermediates
                self._zero_bars()
                self._copy_species([self._bar[self.out_of[s]] for s in self.state],
                                   self._at(carry, 0))      # cotangent in
                self._backward_step()
                self._copy_species(self._at(carry, 0),
                                   [self._bar[self.in_of[s]] for s in self.state])  # out

            def segment(n):
                # from the segment's checkpoint (already in the state): keep
                # each step's input while advancing, then differentiate the
                # steps in reverse
                for j in range(n - 1):
                    keep_and_advance(j)
                self._store_state(seg, (n - 1) * n1, from_outputs=False)
                for j in reversed(range(n)):
                    differentiate(j)

            for g in reversed(range(nseg)):
                n = min(steps, g * C_ + C_) - g * C_
                self._load_state(chk, g * n1)
                self._replay(("segment", C_, n), lambda n=n: segment(n))
            out = carry.download(self.stream)
            grads = {s: out[i * nq:(i + 1) * nq].copy() for i, s in enumerate(self.state)}
            pg = np.zeros(len(self.param_slots))
            for a in self._adj:
                if a[0] == "kernel" and self.param_slots:
                    pp = a[4].download(self.stream).astype(np.float64)
                    pg += pp.reshape(-1, max(1, len(self.param_slots)))[:, :len(pg)].sum(0)
            return grads, {n: float(pg[i]) for i, n in enumerate(self.param_slots)}
        finally:
            self.stream.sync()
            chk.free()

    # ------------------------------------------------------------ helpers
    def synth_adjoint(self, gbar) -> np.ndarray:
        """S^T gbar for one grid field (host in, host out)."""
        g = DeviceBuffer(self.npts, "gbar")
        q = DeviceBuffer(2 * self.nlm, "qbar")
        try:
            g.upload(np.asarray(gbar, np.float32), self.stream)
            _lib.call("tsphere_synth_adjoint", self.sht.handle, 1, 1,
                      (C.c_void_p * 1)(g.ptr), 0, (C.c_void_p * 1)(q.ptr), 0, self.stream.handle)
            return q.download(self.stream)
        finally:
            g.free()
            q.free()

    def destroy(self) -> None:
        self.stream.sync()
        self._drop_graphs()
        self.fwd.destroy()
        for b in (self._seg, self._carry):
            if b is not None:
                b.free()
        self._seg = self._carry = None
        for h in self._kernels:
            _lib.call("tsphere_kernel_destroy", h)
        self._kernels = []
        if self._module:
            _lib.call("tsphere_module_destroy", self._module)
            self._module = None
        for b in list(self._tmp.values()) + self._owned:
            b.free()
        self._bar_arena.free()
        self._bar, self._tmp, self._owned = {}, {}, []
        self.tape.destroy()
        self._host.destroy()



def _dedupe_helpers(srcs):
    seen = set()
    out = []
    for code in srcs:
        lines = []
        for line in code.splitlines():
            if line.startswith("__device__ __forceinline__ float pow_i"):
                key = line.split("(")[0]
                if key in seen:
                    continue
                seen.add(key)
            lines.append(line)
        out.append("\n".join(lines))
    return out
