"""Statement list -> a replayable sequence of GPU operations
(src/mgpu/plan.ts:1-534, re-designed for one HIP stream and HIP graphs).

As in the reference, everything expensive happens once at plan time:
kernel compilation (one hipRTC module per function), buffer allocation, and
the op list.  A step is then pure launch recording -- and here it is recorded
once into a HIP graph and replayed, instead of re-encoded every step
(plan.ts:436-478).

Differences from the WebGPU planner, all consequences of the hardware:

* transform batching: ready ``synth`` (or ``analys``) ops with no dependency
  between them are scheduled together and run as ONE batched transform
  (``tsphere_synth_multi``), so the Legendre recurrence -- the dominant cost
  -- is generated once for all species.  Element-wise kernels are scheduled
  as early as their inputs allow so that transforms meet in batches
  (Schnakenberg: 2 synths -> 1 launch pair, 2 analyses -> 1);
* in-place kernels need no scratch + copy-back (plan.ts:371-405): HIP has no
  binding-aliasing rule and an element-wise kernel reads each element before
  writing it, in the same thread;
* feedback without copies: when the value that feeds a state buffer
  (``Un -> U``) is produced by a kernel after the last read of ``U``, that
  kernel writes ``U`` directly and the copy op (plan.ts:257-281) disappears.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .device import DeviceBuffer, Stream
from .hipgen import (Batch, KernelInputs, UnsupportedOnGpu, build_fused_kernel,
                     build_grid_kernel, build_kernel, grid_lds_bytes, module_source)
from .ir import Assign, CompiledFunction, is_numeric, is_tensor, numel, walk

EXTERNAL_OPS = ("synth", "analys")  # externals.ts:95


class Placeholder:
    """A buffer's size without device memory: planning on a machine without a
    GPU (the fusion and scheduling tests) allocates these instead."""

    def __init__(self, count: int, label: str = ""):
        self.count, self.label, self.ptr, self.nbytes = int(count), label, None, 4 * int(count)

    def free(self) -> None:
        pass


class HostBuffers:
    """Buffers for host-bound values, shared by a model's plans (plan.ts:48-92).
    ``ensure_group`` allocates several same-sized buffers in one allocation
    (the spectral state), which keeps them adjacent in HBM."""

    def __init__(self, allocate: bool = True):
        self.allocate = allocate
        self._slots: dict = {}
        self._arenas: list = []

    def ensure(self, name: str, count: int) -> DeviceBuffer:
        b = self._slots.get(name)
        if b is not None:
            if b.count != count:
                raise UnsupportedOnGpu(
                    f"'{name}' is {b.count} elements in one program and {count} in another")
            return b
        b = DeviceBuffer(count, f"mgpu-{name}") if self.allocate else Placeholder(count, name)
        self._slots[name] = b
        return b

    def ensure_group(self, names, count: int) -> None:
        new = [n for n in names if n not in self._slots]
        if not new:
            return
        if not self.allocate:
            for n in new:
                self._slots[n] = Placeholder(count, n)
            return
        arena = DeviceBuffer(count * len(new), "mgpu-" + "+".join(new))
        self._arenas.append(arena)
        for i, n in enumerate(new):
            self._slots[n] = DeviceBuffer.view(arena, i * count, count, f"mgpu-{n}")

    def get(self, name: str):
        return self._slots.get(name)

    def names(self):
        return list(self._slots)

    def upload(self, name: str, data, stream: Stream | None = None) -> None:
        b = self._slots.get(name)
        if b is None:
            raise KeyError(f"upload: no buffer named '{name}'")
        a = np.asarray(data, np.float32).reshape(-1)
        if a.size != b.count:
            raise ValueError(f"upload '{name}': expected {b.count} elements, got {a.size}")
        b.upload(a, stream)

    def destroy(self) -> None:
        for b in self._slots.values():
            b.free()
        for a in self._arenas:
            a.free()
        self._slots.clear()
        self._arenas.clear()


@dataclass
class PlanSpec:
    """plan.ts:108-114"""
    fn: CompiledFunction
    feedback: list  # output index -> host binding name (or None)


# ------------------------------------------------------------------- ops
@dataclass
class KernelOp:
    kind: str
    label: str
    kernel: object          # hipgen.Kernel
    out: str                # c_name of the destination
    ins: list               # c_names, slot order
    handle: object = None   # tsphere_kernel
    count: int = 0
    stmt: object = None     # the IR Assign (reverse mode differentiates it)
    io: object = None       # its KernelInputs


@dataclass
class FusedOp:
    """Consecutive element-wise statements run as one kernel (batched plans):
    writes `outs`, reads `ins` (external values only)."""
    kind: str
    label: str
    kernel: object
    outs: list
    ins: list
    members: list           # the KernelOps it replaces
    handle: object = None
    count: int = 0


@dataclass
class GridOp:
    """Legendre synthesis of `srcs` into the plan's Fourier scratch, ONE
    generated grid kernel (per latitude row: FFT synthesis, the element-wise
    statements, FFT analysis), Legendre analysis into `dsts`: the grid fusion
    of a synth -> element-wise -> analys run.  The kernel writes `stores`
    and reads `ins` (external grid values)."""
    kind: str
    label: str
    srcs: list
    dsts: list
    kernel: object
    stores: list
    ins: list
    members: list           # the ops it replaces
    nstmts: int = 0
    handle: object = None
    blocks: int = 0
    block: int = 256


@dataclass
class SpectralOp:
    """A step whose ops are one GridOp and the diagonal implicit solve of every
    species, X = (X + s1 * analys(..)) ./ (1 + s2 * lam): each step runs the
    grid op's kernel, then ONE library kernel (``tsphere_leg_spectral_imex``)
    doing the grid op's Legendre analysis, the implicit solve and the
    Legendre synthesis of the new state for the next step -- the recurrence
    is generated once per step for both contractions.  A call of
    ``encode_steps`` starts with the grid op's own Legendre synthesis, so it
    needs nothing from earlier calls."""
    kind: str
    label: str
    grid: GridOp
    imex: object            # the KernelOp / FusedOp it replaces
    species: list           # state c_names, in the grid op's synthesis order
    spc: list               # analysed field f -> species index
    s1: list                # scalar IR expressions per analysed field
    s2: list
    lam: str                # c_name of the eigenvalue array
    mode: str = "spectral"  # or "fold": the implicit solve folded into the
                            # next synthesis (single runs) or into the
                            # analysis (ensembles)


@dataclass
class XformOp:
    kind: str               # 'synth' | 'analys'
    label: str
    srcs: list              # c_names
    dsts: list


@dataclass
class CopyOp:
    kind: str
    label: str
    src: str
    dst_host: str
    dst_keys: tuple = ()    # c_names the host buffer is read under in this function


def _reads(op) -> set:
    if isinstance(op, SpectralOp):
        return _reads(op.grid) | _reads(op.imex)
    if isinstance(op, KernelOp):
        return set(op.ins)
    if isinstance(op, FusedOp):
        return set(op.ins)
    if isinstance(op, GridOp):
        return set(op.srcs) | set(op.ins)
    if isinstance(op, XformOp):
        return set(op.srcs)
    return {op.src}


def _writes(op) -> set:
    if isinstance(op, SpectralOp):
        return _writes(op.grid) | _writes(op.imex)
    if isinstance(op, KernelOp):
        return {op.out}
    if isinstance(op, FusedOp):
        return set(op.outs)
    if isinstance(op, GridOp):
        return set(op.dsts) | set(op.stores)
    if isinstance(op, XformOp):
        return set(op.dsts)
    return {"@host:" + op.dst_host, *op.dst_keys}


class ModelPlan:
    """A compiled .m function, ready to run on the GPU (plan.ts:173)."""

    def __init__(self):
        self.param_names: list[str] = []
        self.ops: list = []
        self._bufs: dict[str, DeviceBuffer] = {}     # c_name -> buffer
        self._by_name: dict[str, DeviceBuffer] = {}  # source name -> buffer
        self._owned: list[DeviceBuffer] = []
        self._module = None
        self._kernels: list = []
        self._params = np.zeros(1, np.float32)
        self._sht = None
        self._graphs: dict[int, int] = {}      # steps per graph -> hipGraphExec
        self._ptr_cache = {}
        self.runs = 1
        self._prm = None       # ensembles: device copy of the per-run scalars
        self._coef = None      # ensembles with a folded implicit solve: (s1, s2) per field
        self._fm = None        # grid ops: Fourier scratch between the stages
        self._trig = None      # grid ops: the SHT plan's device twiddle table

    # --------------------------------------------------------- creation
    @staticmethod
    def create(sht, spec: PlanSpec, host: HostBuffers, batch: bool = True,
               label: str = "step", runs: int = 1, shared=(), fuse: bool = True,
               grid: bool = True, spectral: bool = False,
               fold: bool = False, keep_all: bool = False) -> "ModelPlan":
        """ModelPlan.create (plan.ts:206-407).

        ``runs`` > 1 compiles an ensemble: every value gets `runs` copies back
        to back in one buffer (host bindings named in ``shared`` keep one
        copy, e.g. the eigenvalues), each transform moves all runs of all its
        species in one batched launch per stage, and element-wise kernels
        read their scalars per run from a device array -- so a parameter
        change is a small upload, not a re-capture.

        ``grid`` (with ``fuse``, FFT Fourier stages): a synth ->
        element-wise -> analys run of ops becomes a GridOp, whose generated
        kernel keeps the grid values between the row FFTs in LDS and
        registers.

        ``spectral`` (single runs, with ``grid``): a step made of one grid op
        and the implicit solve of its species becomes a SpectralOp (the
        analysis, the solve and the next step's synthesis in one kernel).
        ``keep_all`` (with ``fuse``): a fused element-wise kernel, or a grid
        kernel, writes every value it computes to its buffer (the synthesised
        grid values included), not only those read after it (a step tape
        whose values a reverse pass reads).
        ``fold`` (with ``grid``, when not ``spectral``; ensembles too): a
        single run keeps the analysis as its own launch and folds the
        implicit solve into the next step's synthesis launch; an ensemble
        folds it into the analysis launch (tsphere_leg_analys_imex_fields)
        and synthesises the next step's state on its own, with each run's s1,
        s2 in a device array that a parameter change re-uploads."""
        self = ModelPlan()
        self._sht = sht
        self.runs = runs = int(runs)
        if runs < 1:
            raise ValueError("runs must be >= 1")
        shared = set(shared)
        fn = spec.fn
        host_of: dict[str, str] = {}       # c_name -> host name, for host-bound args
        param_slots: dict[str, int] = {}
        for p in fn.params:
            if p.binding.kind == "tensor":
                count = int(np.prod(p.binding.shape))
                buf = host.ensure(p.name, count * (1 if p.name in shared else runs))
                self._bufs[p.c_name] = buf
                self._by_name[p.name] = buf
                host_of[p.c_name] = p.name
            elif p.binding.kind == "param":
                param_slots[p.c_name] = len(self.param_names)
                self.param_names.append(p.name)
            # 'const' arguments are exact in the IR and fold into the kernels
        self._params = np.zeros(max(1, len(self.param_names)), np.float32)
        if runs > 1:
            self._params = np.zeros((runs, max(1, len(self.param_names))), np.float32)
        shared_c = frozenset(c for c, n in host_of.items() if n in shared)

        derived: dict[str, tuple] = {}
        counts: dict[str, int] = {}
        names: dict[str, str] = {}
        pending_kernels = []
        ops = []
        for stmt in fn.body:
            if stmt.kind == "ReturnFromFunction":
                continue
            if stmt.kind != "Assign":
                raise UnsupportedOnGpu(
                    f"a model function body may only contain assignments (found '{stmt.kind}')",
                    stmt.span)
            if not is_numeric(stmt.ty):
                raise UnsupportedOnGpu(f"'{stmt.name}' is not a numeric value", stmt.span)
            if not is_tensor(stmt.ty):
                derived[stmt.c_name] = (stmt.name, stmt.expr)
                continue
            count = numel(stmt.ty)
            if stmt.c_name in counts and counts[stmt.c_name] != count:
                raise UnsupportedOnGpu(f"'{stmt.name}' changes size between assignments",
                                       stmt.span)
            if stmt.c_name in self._bufs and self._bufs[stmt.c_name].count != count * (
                    1 if stmt.c_name in shared_c else runs):
                raise UnsupportedOnGpu(f"'{stmt.name}' changes size between assignments",
                                       stmt.span)
            counts[stmt.c_name] = count
            names[stmt.c_name] = stmt.name
            ext = _external_call(stmt)
            if ext is not None:
                fname, arg_c, arg_name = ext
                if arg_c not in self._bufs and arg_c not in counts:
                    raise UnsupportedOnGpu(f"'{fname}' reads '{arg_name}', which has no buffer",
                                           stmt.span)
                ops.append(XformOp(fname, f"{stmt.name} = {fname}({arg_name})", [arg_c],
                                   [stmt.c_name]))
                continue
            tensors: dict[str, int] = {}
            for node in walk(stmt.expr):
                if node.kind == "Var" and is_tensor(node.ty) and node.c_name not in tensors:
                    if node.c_name not in self._bufs and node.c_name not in counts:
                        raise UnsupportedOnGpu(f"'{stmt.name}' reads a value with no buffer",
                                               stmt.span)
                    tensors[node.c_name] = len(tensors)
            kname = f"k{len(pending_kernels)}_{_ident(stmt.name)}"
            lbl = f"{stmt.name} = <{count} elements, element-wise>"
            if stmt.c_name in shared_c:
                raise UnsupportedOnGpu(f"'{stmt.name}' is shared by all runs and cannot be "
                                       f"assigned", stmt.span)
            kern = build_kernel(stmt, KernelInputs(tensors, param_slots, derived), count, lbl,
                                kname, Batch(runs, shared_c) if runs > 1 else None)
            kio = KernelInputs(tensors, param_slots, derived)
            op = KernelOp("kernel", lbl, kern, stmt.c_name, list(tensors), count=count * runs,
                          stmt=stmt, io=kio)
            pending_kernels.append(op)
            ops.append(op)

        # Feedback (plan.ts:257-281): declared outputs -> the host buffers they replace.
        for i, out in enumerate(fn.outputs):
            to = spec.feedback[i] if i < len(spec.feedback) else None
            if not to:
                continue
            if out.c_name not in counts and out.c_name not in self._bufs:
                raise UnsupportedOnGpu(
                    f"'{fn.name}' declares the output '{out.name}' but never assigns it")
            dst = host.get(to)
            if dst is None:
                raise UnsupportedOnGpu(f"'{to}' is not a host binding")
            n_src = (counts[out.c_name] * runs if out.c_name in counts else
                     self._bufs[out.c_name].count if out.c_name in self._bufs else 0)
            if n_src != dst.count:
                raise UnsupportedOnGpu(
                    f"'{out.name}' ({n_src} elements) cannot feed '{to}' ({dst.count})")
            if self._bufs.get(out.c_name) is dst:
                continue  # the function updated the host buffer in place
            keys = tuple(c for c, n in host_of.items() if n == to)
            ops.append(CopyOp("copy", f"{out.name} -> {to}", out.c_name, to, keys))

        if batch:
            cap = sht.max_fields // runs if sht is not None else 8
            ops = _schedule(ops, max_batch=max(1, min(8, cap)))
        ops = self._fold_feedback(ops, host, host_of)
        if batch and fuse:
            ops = self._fuse(ops, {o.c_name for o in fn.outputs}, host, runs, shared_c, keep_all)
            if grid:
                ops = self._fuse_grid(ops, {o.c_name for o in fn.outputs}, host, runs, shared_c,
                                      keep_all)
                mode = "spectral" if spectral and runs == 1 else "fold" if fold else None
                if mode:
                    ops = self._fuse_spectral(ops, {o.c_name for o in fn.outputs}, param_slots,
                                              derived, host_of, mode)
        self._param_slots = param_slots
        self._derived = derived

        # Buffers for every value the .m computes (reused across steps).
        for c, n in counts.items():
            if c not in self._bufs:
                b = (DeviceBuffer(n * runs, f"mgpu-{names[c]}") if host.allocate
                     else Placeholder(n * runs, names[c]))
                self._owned.append(b)
                self._bufs[c] = b
            self._by_name[names[c]] = self._bufs[c]
        self.ops = ops

        flat = [x for op in ops for x in ((op.grid,) if isinstance(op, SpectralOp) else (op,))]
        pending_kernels = [op for op in flat if isinstance(op, (KernelOp, FusedOp, GridOp))]
        if any(isinstance(op, GridOp) for op in flat) and host.allocate:
            nfm = runs * max(max(len(op.srcs), len(op.dsts)) for op in flat
                             if isinstance(op, GridOp))
            cfg = sht.cfg
            self._fm = DeviceBuffer(nfm * cfg.nlat * (cfg.mmax + 1) * 2, f"mgpu-{label}-fm")
            trig = C.c_void_p()
            _lib.call("tsphere_plan_fourier_table", sht.handle, C.byref(trig))
            self._trig = trig.value
        if pending_kernels:
            src = module_source([op.kernel for op in pending_kernels])
        if pending_kernels and not host.allocate:
            if _lib.load().tsphere_rtc_check(src.encode(), None) != _lib.OK:
                raise UnsupportedOnGpu(f"generated HIP failed to compile for '{label}':\n"
                                       f"{_lib.last_error()}\n--- source ---\n{src}")
        elif pending_kernels:
            self._module = _compile(src, label)
            for op in pending_kernels:
                h = C.c_void_p()
                _lib.call("tsphere_module_kernel", self._module, op.kernel.name.encode(),
                          C.byref(h))
                op.handle = h.value
                self._kernels.append(h.value)
        self._host = host
        if runs > 1 and host.allocate:
            self._prm = DeviceBuffer(self._params.size, f"mgpu-{label}-params")
            spec = [op for op in ops if isinstance(op, SpectralOp)]
            if spec:
                self._coef = DeviceBuffer(2 * len(spec[0].species) * runs, f"mgpu-{label}-imex")
                self._coef.upload(self._imex_coefs(spec[0]), None)
        return self

    def _imex_coefs(self, op: "SpectralOp") -> np.ndarray:
        """An ensemble's folded implicit solve: (s1, s2) of field
        f * runs + run for analysed field f, each evaluated from that run's
        scalars as the generated kernel rounds them."""
        out = np.zeros((len(op.spc), self.runs, 2), np.float32)
        for f in range(len(op.spc)):
            for r in range(self.runs):
                out[f, r, 0] = self._scalar(op.s1[f], r)
                out[f, r, 1] = self._scalar(op.s2[f], r)
        return out.reshape(-1)

    def _fuse(self, ops, outputs: set, host: HostBuffers, runs: int, shared_c,
              keep_all: bool = False) -> list:
        """Merge each run of consecutive element-wise kernels over the same
        element count into one kernel; values used only inside a run stay in
        registers."""
        host_bufs = {id(b) for b in host._slots.values()}
        out = []
        i = 0
        while i < len(ops):
            j = i
            while (j < len(ops) and isinstance(ops[j], KernelOp) and
                   ops[j].count == ops[i].count):
                j += 1
            group = ops[i:j] if j > i else []
            if len(group) < 2 or len({g.out for g in group}) != len(group):
                out.append(ops[i])
                i += 1
                continue
            produced = {g.out for g in group}
            later = set()
            for o in ops[:i] + ops[j:]:
                later |= _reads(o)
            stores = [g.out for g in group
                      if keep_all or g.out in outputs or g.out in later or
                      id(self._bufs.get(g.out)) in host_bufs]
            if not stores:
                stores = [group[-1].out]
            name = f"f{len(out)}_" + "_".join(_ident(g.stmt.name) for g in group)[:40]
            lbl = " ; ".join(g.label for g in group)
            per = group[0].count // runs
            kern = build_fused_kernel([g.stmt for g in group], [g.io for g in group], per,
                                      stores, lbl, name,
                                      Batch(runs, shared_c) if runs > 1 else None)
            out.append(FusedOp("kernel", lbl, kern, stores, kern.inputs, list(group),
                               count=group[0].count))
            del produced
            i = j
        return out

    def _fuse_grid(self, ops, outputs: set, host: HostBuffers, runs: int, shared_c,
                   keep_all: bool = False) -> list:
        """Turn each synth -> element-wise -> analys run into a GridOp: the
        element-wise statements run between the row FFTs of one generated
        kernel, so the synthesised grid values are written only if something
        else reads them (a function output, a later op, a host buffer) and the
        analysed values never leave the workgroup."""
        sht = self._sht
        if sht is None or getattr(sht, "fourier_mode", None) != "fft":
            return ops
        cfg = sht.cfg
        h = cfg.nphi // 2
        log2h = h.bit_length() - 1
        if (1 << log2h) != h or not 2 <= log2h <= 11:
            return ops
        host_bufs = {id(b) for b in host._slots.values()}
        out = []
        i = 0
        while i < len(ops):
            S = ops[i]
            E = ops[i + 1] if i + 1 < len(ops) else None
            A = ops[i + 2] if i + 2 < len(ops) else None
            if not (isinstance(S, XformOp) and S.kind == "synth" and
                    isinstance(E, (KernelOp, FusedOp)) and E.count == cfg.npts * runs and
                    isinstance(A, XformOp) and A.kind == "analys" and
                    grid_lds_bytes(log2h, len(A.srcs)) <= 65536):
                out.append(S)
                i += 1
                continue
            members = E.members if isinstance(E, FusedOp) else [E]
            stmts = [m.stmt for m in members]
            ios = [m.io for m in members]
            produced = [st.c_name for st in stmts]
            later = set()
            for o in ops[:i] + ops[i + 3:]:
                later |= _reads(o)
            stores = [c for c in list(S.dsts) + produced
                      if keep_all or c in outputs or c in later or
                      id(self._bufs.get(c)) in host_bufs]
            name = f"g{len(out)}_" + "_".join(_ident(st.name) for st in stmts)[:40]
            lbl = f"{S.label} => {E.label} => {A.label}"
            kern = build_grid_kernel(name, lbl, log2h, cfg.nlat, cfg.mmax, list(S.dsts), stmts,
                                     ios, stores, list(A.srcs),
                                     Batch(runs, shared_c) if runs > 1 else None)
            rb = 256 // min(h // 4, 256)
            out.append(GridOp("grid", lbl, list(S.srcs), list(A.dsts), kern, stores,
                              list(kern.inputs), [S, E, A], nstmts=len(stmts),
                              blocks=(cfg.nlat + rb - 1) // rb, block=kern.block))
            i += 3
        return out

    def _fuse_spectral(self, ops, outputs: set, param_slots: dict, derived: dict,
                       host_of: dict, mode: str = "spectral") -> list:
        """[GridOp, implicit solve] -> [SpectralOp] when every statement of the
        solve is X = (X + s1 * A) ./ (1 + s2 * lam) with X a species the grid
        op synthesises (the value written back to X's own buffer), A one of
        its analysed values (read by nothing else), lam a host array and s1,
        s2 scalar expressions."""
        sht = self._sht
        if mode == "spectral" and sht is not None and (sht.cfg.nlat + 1) // 2 > 1024:
            mode = "fold"  # the spectral kernel covers up to 1024 latitude pairs
        if (sht is not None and getattr(sht, "legendre", None) == "bf16x3" and
                (sht.cfg.nlat + 1) // 2 > 512):
            return ops  # the split-BF16 block kernels cover up to 512 latitude pairs
        if (sht is None or getattr(sht, "legendre", None) not in ("fp32", "f64rec", "bf16x3") or
                len(ops) != 2 or not isinstance(ops[0], GridOp) or
                not isinstance(ops[1], (KernelOp, FusedOp))):
            return ops
        G, E = ops
        members = E.members if isinstance(E, FusedOp) else [E]
        if not 1 <= len(G.dsts) <= 4 or len(members) != len(G.dsts) or \
                len(G.srcs) != len(G.dsts):
            return ops
        scal = set(param_slots) | set(derived)
        spc, s1, s2, lam = [None] * len(G.dsts), [None] * len(G.dsts), [None] * len(G.dsts), None
        for mem in members:
            m = _imex_pattern(mem.stmt.expr, scal)
            if m is None:
                return ops
            x, a, lm, e1, e2 = m
            if (x not in G.srcs or a not in G.dsts or a in outputs or lm not in host_of or
                    self._bufs.get(mem.out) is not self._bufs.get(x) or
                    (lam is not None and lm != lam)):
                return ops
            f = G.dsts.index(a)
            if spc[f] is not None:
                return ops
            spc[f], s1[f], s2[f], lam = G.srcs.index(x), e1, e2, lm
        if sorted(spc) != list(range(len(G.srcs))):
            return ops
        lbl = f"{G.label} => {E.label}"
        return [SpectralOp("spectral", lbl, G, E, list(G.srcs), spc, s1, s2, lam, mode)]

    def _scalar(self, e, run: int | None = None) -> np.float32:
        """A scalar IR expression of params and literals in f32, rounded per
        operation as the generated kernels compute it (an ensemble: with run
        `run`'s scalars)."""
        if e.kind == "NumLit":
            return np.float32(e.value)
        if e.kind == "Var":
            if e.c_name in self._param_slots:
                row = self._params if run is None else self._params[run]
                return np.float32(row[self._param_slots[e.c_name]])
            return self._scalar(self._derived[e.c_name][1], run)
        if e.kind == "Unary" and e.builtin in ("uminus",):
            return np.float32(-self._scalar(e.operand, run))
        a, b = self._scalar(e.left, run), self._scalar(e.right, run)
        with np.errstate(all="ignore"):
            if e.builtin == "plus":
                return np.float32(a + b)
            if e.builtin == "minus":
                return np.float32(a - b)
            if e.builtin in ("mtimes", "times"):
                return np.float32(a * b)
            return np.float32(a / b)  # rdivide / mrdivide (checked by _imex_pattern)

    def _fold_feedback(self, ops, host: HostBuffers, host_of: dict) -> list:
        """Let the kernel that produces a fed-back value write the host buffer
        directly when nothing after it reads that buffer's old contents."""
        out = list(ops)
        for cp in [o for o in ops if isinstance(o, CopyOp)]:
            idx = [i for i, o in enumerate(out) if cp.src in _writes(o) and o is not cp]
            if len(idx) != 1 or cp.src in host_of:
                continue
            k = idx[0]
            prod = out[k]
            keys = set(cp.dst_keys)
            # a transform cannot read and write one buffer; an element-wise
            # kernel may (each element is read before it is written)
            if isinstance(prod, XformOp) and keys & _reads(prod):
                continue
            # nothing after the producer may still want the old contents
            if any(keys & _reads(o) for o in out[k + 1:] if o is not cp):
                continue
            if any(isinstance(o, CopyOp) and o.dst_host == cp.dst_host and o is not cp
                   for o in out):
                continue
            # the producer writes the host buffer; its c_name now names that buffer
            self._bufs[cp.src] = host.get(cp.dst_host)
            out.remove(cp)
        return out

    # --------------------------------------------------------- params
    def set_params(self, values: dict, stream: Stream | None = None) -> None:
        """plan.ts:410-420.  Scalars are launch arguments: a change only
        invalidates the recorded graph; nothing recompiles.  Ensembles take a
        scalar (all runs) or one value per run, uploaded on `stream` (ordered
        with the replays already queued there; the graph stays valid)."""
        if self.runs == 1:
            for i, name in enumerate(self.param_names):
                v = values.get(name)
                self._params[i] = v if (v is not None and np.isfinite(v)) else 0.0
            self.invalidate()
            return
        for i, name in enumerate(self.param_names):
            v = values.get(name)
            a = np.zeros(self.runs) if v is None else np.broadcast_to(
                np.asarray(v, np.float64), (self.runs,))
            self._params[:, i] = np.where(np.isfinite(a), a, 0.0)
        if self._prm is not None:
            self._prm.upload(self._params, stream)
        if self._coef is not None:
            spec = [op for op in self.ops if isinstance(op, SpectralOp)]
            self._coef.upload(self._imex_coefs(spec[0]), stream)

    def invalidate(self) -> None:
        for g in self._graphs.values():
            _lib.call("tsphere_graph_destroy", g)
        self._graphs = {}

    def buffer(self, name: str):
        return self._by_name.get(name)

    def element_count(self, name: str):
        b = self._by_name.get(name)
        return None if b is None else b.count

    # --------------------------------------------------------- recording
    def _ptrs(self, cnames):
        key = tuple(cnames)
        arr = self._ptr_cache.get(key)
        if arr is None:
            special = {"@prm": lambda: self._prm.ptr, "@fm": lambda: self._fm.ptr,
                       "@trig": lambda: self._trig}
            arr = (C.c_void_p * len(cnames))(*[
                special[c]() if c in special else self._bufs[c].ptr for c in cnames])
            self._ptr_cache[key] = arr
        return arr

    def encode_steps(self, stream: Stream, steps: int = 1) -> None:
        """plan.ts:436-478: enqueue `steps` timesteps on `stream` (no sync)."""
        L = _lib.load()
        sh = stream.handle
        pvals = self._params.ctypes.data_as(C.POINTER(C.c_float))
        np_ = len(self.param_names)
        ens = self.runs > 1
        if ens:
            np_ = 0
            nq, ns = 2 * self._sht.nlm, self._sht.cfg.npts
        spec = [op for op in self.ops if isinstance(op, SpectralOp)]
        if spec:
            # the grid op's own synthesis once; then each step's spectral
            # kernel leaves the next step's synthesis in the scratch
            G = spec[0].grid
            if ens:
                rc = L.tsphere_leg_synth_fields(self._sht.handle, len(G.srcs), self.runs,
                                                self._ptrs(G.srcs), nq, self._fm.ptr, sh)
            else:
                rc = L.tsphere_leg_synth_multi(self._sht.handle, len(G.srcs), self._ptrs(G.srcs),
                                               self._fm.ptr, sh)
            _lib.check(rc, G.label)
        for k in range(steps):
            for op in self.ops:
                if isinstance(op, SpectralOp) and ens:
                    # fold, every run at once: the analysis of all runs with
                    # the implicit solve behind it (scalars per run from the
                    # device array), then the next step's synthesis (the last
                    # step's is left to the next call, which starts with its
                    # own)
                    self.launch_grid(op.grid, stream)
                    G, h = op.grid, self._sht.handle
                    st = [op.species[sp] for sp in op.spc]
                    _lib.check(L.tsphere_leg_analys_imex_fields(
                        h, len(G.dsts), self.runs, self._fm.ptr, self._ptrs(G.dsts),
                        self._ptrs(st), nq, self._bufs[op.lam].ptr, self._coef.ptr, sh), op.label)
                    if k < steps - 1:
                        _lib.check(L.tsphere_leg_synth_fields(h, len(G.srcs), self.runs,
                                                              self._ptrs(G.srcs), nq,
                                                              self._fm.ptr, sh), G.label)
                    continue
                if isinstance(op, SpectralOp):
                    self.launch_grid(op.grid, stream)
                    n = len(op.species)
                    v1 = [float(self._scalar(e)) for e in op.s1]
                    v2 = [float(self._scalar(e)) for e in op.s2]
                    if op.mode == "spectral":
                        _lib.check(L.tsphere_leg_spectral_imex(
                            self._sht.handle, n, self._fm.ptr, self._ptrs(op.species),
                            self._bufs[op.lam].ptr, (C.c_float * n)(*v1), (C.c_float * n)(*v2),
                            (C.c_int * n)(*op.spc), sh), op.label)
                        continue
                    # fold: the analysis, then the implicit solve of each
                    # species (its analysed value, its scalars) in front of
                    # the next step's synthesis
                    G = op.grid
                    _lib.check(L.tsphere_leg_analys_multi(self._sht.handle, len(G.dsts),
                                                          self._fm.ptr, self._ptrs(G.dsts), sh),
                               G.label)
                    by = {sp: f for f, sp in enumerate(op.spc)}
                    ana = [G.dsts[by[sp]] for sp in range(n)]
                    _lib.check(L.tsphere_leg_imex_synth(
                        self._sht.handle, n, self._ptrs(ana), self._ptrs(op.species),
                        self._bufs[op.lam].ptr, (C.c_float * n)(*[v1[by[sp]] for sp in range(n)]),
                        (C.c_float * n)(*[v2[by[sp]] for sp in range(n)]), self._fm.ptr, sh),
                        op.label)
                    continue
                if isinstance(op, (KernelOp, FusedOp)):
                    outs = op.outs if isinstance(op, FusedOp) else [op.out]
                    ptrs = self._ptrs(outs + op.ins + (["@prm"] if ens else []))
                    _lib.check(L.tsphere_kernel_launch(op.handle, op.count, len(ptrs), ptrs, np_,
                                                       pvals, sh), op.label)
                elif isinstance(op, GridOp):
                    h = self._sht.handle
                    if ens:
                        rc = L.tsphere_leg_synth_fields(h, len(op.srcs), self.runs,
                                                        self._ptrs(op.srcs), nq, self._fm.ptr, sh)
                    else:
                        rc = L.tsphere_leg_synth_multi(h, len(op.srcs), self._ptrs(op.srcs),
                                                       self._fm.ptr, sh)
                    _lib.check(rc, op.label)
                    self.launch_grid(op, stream)
                    if ens:
                        rc = L.tsphere_leg_analys_fields(h, len(op.dsts), self.runs, self._fm.ptr,
                                                         self._ptrs(op.dsts), nq, sh)
                    else:
                        rc = L.tsphere_leg_analys_multi(h, len(op.dsts), self._fm.ptr,
                                                        self._ptrs(op.dsts), sh)
                    _lib.check(rc, op.label)
                elif isinstance(op, XformOp):
                    srcs = self._ptrs(op.srcs)
                    dsts = self._ptrs(op.dsts)
                    if ens:
                        if op.kind == "synth":
                            rc = L.tsphere_synth_fields(self._sht.handle, len(op.srcs), self.runs,
                                                        srcs, nq, dsts, ns, sh)
                        else:
                            rc = L.tsphere_analys_fields(self._sht.handle, len(op.srcs),
                                                         self.runs, srcs, ns, dsts, nq, sh)
                        _lib.check(rc, op.label)
                        continue
                    fn = L.tsphere_synth_multi if op.kind == "synth" else L.tsphere_analys_multi
                    _lib.check(fn(self._sht.handle, len(op.srcs), srcs, dsts, sh), op.label)
                else:
                    src = self._bufs[op.src]
                    dst = self._host.get(op.dst_host)
                    _lib.check(L.tsphere_memcpy_dtod(dst.ptr, src.ptr, src.nbytes, sh), op.label)

    def launch_grid(self, op: "GridOp", stream: Stream) -> None:
        """Enqueue a grid op's generated kernel alone (between its Legendre
        stages; also what a timing of that kernel runs)."""
        ens = self.runs > 1
        ptrs = self._ptrs(["@fm", "@trig"] + op.stores + op.ins + (["@prm"] if ens else []))
        pvals = self._params.ctypes.data_as(C.POINTER(C.c_float))
        _lib.check(_lib.load().tsphere_kernel_launch_grid(
            op.handle, op.blocks, self.runs, op.block, len(ptrs), ptrs,
            0 if ens else len(self.param_names), pvals, stream.handle), op.label)

    def _graph(self, stream: Stream, k: int) -> int:
        """The HIP graph of `k` timesteps (captured on first use or after a
        parameter change)."""
        g = self._graphs.get(k)
        if g is None:
            _lib.call("tsphere_graph_begin", stream.handle)
            try:
                self.encode_steps(stream, k)
            finally:
                h = C.c_void_p()
                rc = _lib.load().tsphere_graph_end(stream.handle, C.byref(h))
            _lib.check(rc, "graph capture")
            g = self._graphs[k] = h.value
        return g

    def prepare(self, stream: Stream, steps: int = 1, graph_steps: int = 0) -> None:
        """Capture the graphs that ``run(stream, steps, graph_steps)`` replays,
        without running them (a capture records work, it executes none), so
        that a timed ``run`` only launches."""
        if steps <= 0:
            return
        if graph_steps > 0:
            if any(k != graph_steps for k in self._graphs):
                self.invalidate()
            if steps >= graph_steps:
                self._graph(stream, graph_steps)
            return
        while steps > 0:
            k = 1 << (min(steps, 64).bit_length() - 1)
            self._graph(stream, k)
            steps -= k

    def run(self, stream: Stream, steps: int = 1, graph_steps: int = 0) -> None:
        """Replay `steps` timesteps through HIP graphs.  `graph_steps` > 0:
        one graph of that many steps, the remainder recorded directly;
        0: graphs of power-of-two step counts up to 64, largest first (at most
        seven kept), so that `steps` steps cost O(log) graph launches and a
        step count that repeats replays the same graphs."""
        if steps <= 0:
            return
        if graph_steps > 0:
            if any(k != graph_steps for k in self._graphs):
                self.invalidate()
            full, rest = divmod(steps, graph_steps)
            if full:
                g = self._graph(stream, graph_steps)
                for _ in range(full):
                    _lib.call("tsphere_graph_launch", g, stream.handle)
            if rest:
                self.encode_steps(stream, rest)
            return
        while steps > 0:
            k = 1 << (min(steps, 64).bit_length() - 1)
            _lib.call("tsphere_graph_launch", self._graph(stream, k), stream.handle)
            steps -= k

    def describe(self) -> list[str]:
        """plan.ts:486-488: '<kind> <label>' per op (the fusion guard counts
        'kernel' lines)."""
        out = []
        for op in self.ops:
            if isinstance(op, SpectralOp) and op.mode == "fold":
                out.append(f"{op.grid.kind:<7} {op.grid.label}")
                out.append(f"{'imex':<7} {op.label.split(' => ', 3)[-1]} (folded into the "
                           + ("analysis)" if self.runs > 1 else "next step's synthesis)"))
            elif isinstance(op, SpectralOp):
                out.append(f"{op.grid.kind:<7} {op.grid.label}")
                out.append(f"{op.kind:<7} {op.label.split(' => ', 3)[-1]} (with the "
                           f"analysis before it and the next step's synthesis)")
            else:
                out.append(f"{op.kind:<7} {op.label}")
        return out

    def _flat_ops(self):
        return [x for op in self.ops
                for x in ((op.grid,) if isinstance(op, SpectralOp) else (op,))]

    def kernel_source(self) -> str:
        return module_source([op.kernel for op in self._flat_ops()
                              if isinstance(op, (KernelOp, FusedOp, GridOp))])

    def generated_kernels(self) -> int:
        """Generated kernels launched per call."""
        return sum(isinstance(op, (KernelOp, FusedOp, GridOp)) for op in self._flat_ops())

    def elementwise_statements(self) -> int:
        """Element-wise .m statements a step evaluates (a spectral op's
        implicit solve runs inside its library kernel)."""
        def count(op):
            if isinstance(op, SpectralOp):
                return count(op.grid) + count(op.imex)
            return (1 if isinstance(op, KernelOp) else len(op.members) if isinstance(op, FusedOp)
                    else op.nstmts if isinstance(op, GridOp) else 0)
        return sum(count(op) for op in self.ops)

    def destroy(self) -> None:
        self.invalidate()
        for h in self._kernels:
            _lib.call("tsphere_kernel_destroy", h)
        self._kernels.clear()
        if self._module:
            _lib.call("tsphere_module_destroy", self._module)
            self._module = None
        for b in self._owned:
            b.free()
        self._owned.clear()
        if self._prm is not None:
            self._prm.free()
            self._prm = None
        if self._coef is not None:
            self._coef.free()
            self._coef = None
        if self._fm is not None:
            self._fm.free()
            self._fm = None


def _is_scalar_expr(e, scal: set) -> bool:
    if e.kind == "NumLit":
        return True
    if e.kind == "Var":
        return e.c_name in scal and not is_tensor(e.ty)
    if e.kind == "Unary":
        return e.builtin == "uminus" and _is_scalar_expr(e.operand, scal)
    if e.kind == "Binary":
        return (e.builtin in ("plus", "minus", "mtimes", "times", "rdivide", "mrdivide") and
                _is_scalar_expr(e.left, scal) and _is_scalar_expr(e.right, scal))
    return False


def _imex_pattern(e, scal: set):
    """(X + s1 * A) ./ (1 + s2 * lam) -> (X, A, lam, s1, s2) c_names and
    scalar expressions (products in either order), else None: the diagonal
    implicit step of bench/shtns/spec.h:166-169 as the models write it."""
    def tvar(v):
        return v.c_name if v.kind == "Var" and is_tensor(v.ty) else None

    def prod(b):  # scalar * tensor in either order -> (scalar expr, tensor c_name)
        if b.kind != "Binary" or b.builtin not in ("mtimes", "times"):
            return None
        if _is_scalar_expr(b.left, scal) and tvar(b.right):
            return b.left, tvar(b.right)
        if _is_scalar_expr(b.right, scal) and tvar(b.left):
            return b.right, tvar(b.left)
        return None

    if e.kind != "Binary" or e.builtin != "rdivide":
        return None
    n, d = e.left, e.right
    if n.kind != "Binary" or n.builtin != "plus" or d.kind != "Binary" or d.builtin != "plus":
        return None
    x, pa = tvar(n.left), prod(n.right)
    if x is None or pa is None:
        return None
    if not (d.left.kind == "NumLit" and float(d.left.value) == 1.0):
        return None
    pl = prod(d.right)
    if pl is None:
        return None
    return x, pa[1], pl[1], pa[0], pl[0]


def _ident(name: str) -> str:
    return "".join(ch if ch.isalnum() else "_" for ch in name)


def _compile(src: str, label: str):
    h = C.c_void_p()
    rc = _lib.load().tsphere_module_compile(src.encode(), None, C.byref(h))
    if rc != _lib.OK:
        raise UnsupportedOnGpu(f"generated HIP failed to compile for '{label}':\n"
                               f"{_lib.last_error()}\n--- source ---\n{src}")
    return h.value


def _external_call(stmt: Assign):
    """plan.ts:498-511: `x = synth(y)` / `x = analys(y)`."""
    e = stmt.expr
    if e.kind != "Call" or e.name not in EXTERNAL_OPS:
        return None
    if len(e.args) != 1 or e.args[0].kind != "Var":
        raise UnsupportedOnGpu(f"'{e.name}' must be applied to a single variable", stmt.span)
    a = e.args[0]
    return e.name, a.c_name, a.name


def _schedule(ops: list, max_batch: int) -> list:
    """List scheduling over the dependency DAG (read-after-write,
    write-after-read, write-after-write between ops): element-wise kernels run
    as soon as they are ready (program order among them), and when only
    transforms are ready, every ready transform of the earliest one's kind
    (up to max_batch) merges into one batched op."""
    n = len(ops)
    preds = [set() for _ in range(n)]
    for j in range(n):
        rj, wj = _reads(ops[j]), _writes(ops[j])
        for i in range(j):
            ri, wi = _reads(ops[i]), _writes(ops[i])
            if (wi & rj) or (ri & wj) or (wi & wj):
                preds[j].add(i)
    done: set = set()
    out = []
    while len(done) < n:
        ready = [i for i in range(n) if i not in done and preds[i] <= done]
        kern = [i for i in ready if isinstance(ops[i], KernelOp)]
        if kern:
            i = kern[0]
            out.append(ops[i])
            done.add(i)
            continue
        xf = [i for i in ready if isinstance(ops[i], XformOp)]
        if xf:
            kind = ops[xf[0]].kind
            group = [i for i in xf if ops[i].kind == kind][:max_batch]
            if len(group) == 1:
                out.append(ops[group[0]])
            else:
                out.append(XformOp(kind, " ; ".join(ops[i].label for i in group),
                                   [s for i in group for s in ops[i].srcs],
                                   [d for i in group for d in ops[i].dsts]))
            done.update(group)
            continue
        i = ready[0]
        out.append(ops[i])
        done.add(i)
    return out
