"""A .m model, compiled and running on the GPU (src/mgpu/model.ts:1-290).

Two plans -- ``init`` and ``step`` -- share the host-bound buffers (spectral
state, eigenvalues, noise).  ``step(n)`` replays the step's HIP graph; the
only blocking calls are ``read`` and ``sync`` (model.ts:261).
"""
from __future__ import annotations

import numpy as np

from .device import DeviceBuffer, Stream
from .layout import ShtConfig, lm_index
from .mfront import CompiledModel, model_bindings
from .plan import HostBuffers, ModelPlan, PlanSpec


def eigenvalues(cfg: ShtConfig, nlm: int | None = None) -> np.ndarray:
    """model.ts:46-56: l(l+1) per coefficient, duplicated across re/im."""
    nlm = nlm if nlm is not None else cfg.nlm
    lam = np.zeros(2 * nlm, np.float32)
    for m in range(cfg.mmax + 1):
        l = np.arange(m, cfg.lmax + 1)
        i = np.array([lm_index(cfg.lmax, int(x), m) for x in l])
        lam[2 * i] = l * (l + 1)
        lam[2 * i + 1] = l * (l + 1)
    return lam


class GpuModel:
    """model.ts:58-290"""

    def __init__(self):
        self.param_names: list[str] = []
        self.state: list[str] = []
        self.view: list[str] = []
        self.npts = 0
        self.nlm = 0
        self._last_ran = "init"
        self._stashed_ran = "init"
        self._destroyed = False
        self.graph_steps = 0           # 0: power-of-two graphs up to 64 steps
        self.runs = 1

    @staticmethod
    def create(sht, cfg: ShtConfig, source: str, param_names, state, view,
               stream: Stream | None = None, batch: bool = True,
               file_name: str = "model.m", runs: int = 1, grid: bool = True,
               spectral: bool = True, fold: bool = True) -> "GpuModel":
        """GpuModel.create (model.ts:103-163).  ``runs`` > 1: an ensemble of
        that many independent runs of the model (same grid), each value
        holding the runs back to back; the plan needs
        max_fields >= runs * len(state).  ``grid`` lets the planner fuse the
        grid stage (FFT synthesis, element-wise statements, FFT analysis)
        into one kernel where it can; ``spectral`` (single runs, the
        default) lets it fuse the Legendre analysis, the implicit solve and
        the next step's Legendre synthesis into one kernel where the step has
        that form and the Legendre form is FP32 or FP64-recurrence
        (bit-identical to the separate launches; at lmax 1023 129 against
        131-133 us per FP32 step, 140 against 163 with the FP64 recurrence);
        ``fold`` (used where the spectral op is off or does not apply, e.g.
        above 1024 latitude pairs, and for ensembles) folds the implicit
        solve into another launch: a single run's into the next step's
        synthesis, an ensemble's into the analysis (each run's scalars in a
        device array; with the 4x4-block and per-parity product analysis
        tiles the solve runs in the analysis kernel: 1711 against 1648
        batched steps/s on 256 runs of lmax 63-511).  Folding an ensemble's solve into the
        synthesis measured slower (at 64 runs of lmax 511 the synthesis took
        234 instead of 169 us to save a 36 us solve kernel)."""
        self = GpuModel()
        self.runs = int(runs)
        npts = cfg.nlat * cfg.nphi
        nlm = sht.nlm
        self.sht, self.cfg = sht, cfg
        self.param_names, self.state, self.view = list(param_names), list(state), list(view)
        self.npts, self.nlm = npts, nlm
        self.stream = stream or Stream()
        bindings = model_bindings(state, param_names, npts, nlm)
        compiled = CompiledModel(source, bindings, {"npts": npts, "nlm": nlm}, file_name)
        nargout = len(state) + len(view)
        init_fn = compiled.specialize("init", nargout)
        step_fn = compiled.specialize("step", nargout)
        compiled.finish()
        feedback = list(state) + [None] * len(view)

        R = self.runs
        if R > 1 and R * max(1, len(state)) > sht.max_fields:
            raise ValueError(f"an ensemble of {R} runs x {len(state)} species needs a plan "
                             f"with max_fields >= {R * len(state)} (has {sht.max_fields})")
        host = HostBuffers()
        host.ensure_group(list(state), 2 * nlm * R)
        host.ensure("lam", 2 * nlm)
        host.ensure("noise", npts * R)
        self._host = host
        shared = ("lam",)
        self._init = ModelPlan.create(sht, PlanSpec(init_fn, feedback), host, batch, "init",
                                      runs=R, shared=shared, grid=grid)
        self._step = ModelPlan.create(sht, PlanSpec(step_fn, feedback), host, batch, "step",
                                      runs=R, shared=shared, grid=grid,
                                      spectral=spectral and R == 1,
                                      fold=fold)
        host.upload("lam", eigenvalues(cfg, nlm), self.stream)
        self._stash = DeviceBuffer(max(1, len(state)) * 2 * nlm * R, "mgpu-state-stash")
        return self

    @property
    def step_plan(self) -> ModelPlan:
        return self._step

    def set_params(self, params: dict) -> None:
        """Scalars; an ensemble also takes one value per run."""
        self._init.set_params(params, self.stream)
        self._step.set_params(params, self.stream)

    def upload(self, name: str, data) -> None:
        """model.ts:175-177: write a host-owned value (state or input field)."""
        self.stream.sync()
        self._host.upload(name, data, self.stream)

    def init(self, noise) -> None:
        """model.ts:180-186"""
        self.upload("noise", noise)
        self._init.encode_steps(self.stream, 1)
        self._last_ran = "init"

    def snapshot_state(self) -> None:
        """model.ts:195-198: copy the spectral state aside (device to device)."""
        self._stashed_ran = self._last_ran
        self._copy_state("save")

    def restore_state(self) -> None:
        self._copy_state("restore")
        self._last_ran = self._stashed_ran

    def _copy_state(self, direction: str) -> None:
        if self._destroyed:
            return
        off = 0
        for name in self.state:
            b = self._host.get(name)
            part = DeviceBuffer.view(self._stash, off, b.count)
            if direction == "save":
                part.copy_from(b, self.stream)
            else:
                b.copy_from(part, self.stream)
            off += b.count

    def step(self, steps: int = 1) -> None:
        """model.ts:229-234: enqueue `steps` steps (graph replay), no sync."""
        self._step.run(self.stream, steps, self.graph_steps)
        self._last_ran = "step"

    def prepare(self, steps: int = 1) -> None:
        """Capture the graphs a ``step(steps)`` call replays, running nothing."""
        self._step.prepare(self.stream, steps, self.graph_steps)

    def sync(self) -> None:
        self.stream.sync()

    def _locate(self, name: str):
        first, second = ((self._init, self._step) if self._last_ran == "init"
                         else (self._step, self._init))
        b = first.buffer(name) or second.buffer(name) or self._host.get(name)
        return b

    def value_buffer(self, name: str):
        return self._locate(name)

    def read(self, name: str) -> np.ndarray:
        """model.ts:262-275: the only blocking read."""
        b = self._locate(name)
        if b is None:
            raise KeyError(f"read: the model has no value named '{name}'")
        return b.download(self.stream)

    def describe(self) -> dict:
        return {"init": self._init.describe(), "step": self._step.describe()}

    def destroy(self) -> None:
        self._destroyed = True
        self.stream.sync()
        self._init.destroy()
        self._step.destroy()
        self._host.destroy()
        self._stash.free()
