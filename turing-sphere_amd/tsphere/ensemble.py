"""Many runs of one .m model at mixed resolutions, stepped as one HIP graph.

The reference runs one model per ``GpuModel`` (src/mgpu/model.ts:103-290);
a parameter sweep is many separate sessions.  Here runs that share a grid
form a *group*: one ``GpuModel`` ensemble whose values hold all its runs back
to back, so each transform stage of a step is one launch for every run and
species of the group, and each element-wise statement is one kernel reading
its scalars per run.  Groups of different lmax run as parallel branches of
ONE captured HIP graph (each group on its own stream, forked from and joined
back to the ensemble's stream), so the small groups fill the machine while
the large one runs; a replay of the graph advances every run by
``graph_steps`` steps.

Each group's plan can be autotuned (Legendre tile shapes for its field
count) at creation; see :mod:`tsphere.tune`.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .device import Event, Stream
from .layout import ShtConfig
from .model import GpuModel
from .sht import ShtPlan


@dataclass
class Group:
    cfg: ShtConfig
    runs: int
    plan: ShtPlan
    model: GpuModel
    stream: Stream
    done: Event
    first_run: int     # index of this group's first run in the ensemble


class MixedEnsemble:
    def __init__(self):
        self.groups: list[Group] = []
        self.stream = Stream()
        self._fork = Event()
        self._graph = None
        self._graph_steps = 0
        # 20-step graphs: 1739 batched steps/s on 256 Allen-Cahn runs at lmax
        # 63-511, against 1714 with 10 and 1725 with 40
        self.graph_steps = 20
        self.tuning: list = []

    @staticmethod
    def create(source: str, param_names, state, view, groups, legendre: str = "fp32",
               tune: bool = False, file_name: str = "model.m") -> "MixedEnsemble":
        """``groups``: [(ShtConfig, runs), ...] in run order."""
        self = MixedEnsemble()
        nspecies = max(1, len(state))
        first = 0
        for cfg, runs in groups:
            plan = ShtPlan.create(cfg, legendre=legendre, max_fields=runs * nspecies)
            if tune:
                self.tuning.append(plan.tune(runs * nspecies))
            s = Stream()
            model = GpuModel.create(plan, cfg, source, param_names, state, view, stream=s,
                                    file_name=file_name, runs=runs)
            self.groups.append(Group(cfg, runs, plan, model, s, Event(), first))
            first += runs
        self.runs = first
        return self

    # ----------------------------------------------------------- values
    def _split(self, values):
        """Per-run values (length = total runs) or one scalar -> per group."""
        a = np.asarray(values, np.float64)
        if a.ndim == 0:
            return [a] * len(self.groups)
        if a.shape[0] != self.runs:
            raise ValueError(f"expected {self.runs} per-run values, got {a.shape[0]}")
        return [a[g.first_run:g.first_run + g.runs] for g in self.groups]

    def set_params(self, params: dict) -> None:
        """Scalars or one value per run (ensemble order); no re-capture."""
        parts = {k: self._split(v) for k, v in params.items()}
        for i, g in enumerate(self.groups):
            g.model.set_params({k: v[i] for k, v in parts.items()})

    def init(self, noise_per_run) -> None:
        """Run each run's .m ``init`` on its own noise field (list of grids)."""
        if len(noise_per_run) != self.runs:
            raise ValueError(f"expected {self.runs} noise fields")
        self.stream.sync()
        for g in self.groups:
            block = np.concatenate([np.asarray(noise_per_run[g.first_run + r], np.float32)
                                    .reshape(-1) for r in range(g.runs)])
            g.model.init(block)
        self.invalidate()

    def upload(self, name: str, per_run) -> None:
        self.stream.sync()
        for g in self.groups:
            g.model.upload(name, np.concatenate(
                [np.asarray(per_run[g.first_run + r], np.float32).reshape(-1)
                 for r in range(g.runs)]))

    def read(self, name: str) -> list:
        """One array per run (blocking)."""
        self.stream.sync()  # graph replays run on the ensemble stream
        out = []
        for g in self.groups:
            a = g.model.read(name)
            out.extend(np.split(a, g.runs))
        return out

    # ----------------------------------------------------------- stepping
    def invalidate(self) -> None:
        if self._graph is not None:
            _lib.call("tsphere_graph_destroy", self._graph)
        self._graph = None
        self._graph_steps = 0

    def _capture(self, k: int) -> None:
        self.invalidate()
        L = _lib.load()
        main = self.stream.handle
        for g in self.groups:  # nothing from before may still run on a group stream
            g.stream.sync()
        _lib.call("tsphere_graph_begin", main)
        rc = _lib.OK
        try:
            self._fork.record(self.stream)
            for g in self.groups:
                _lib.call("tsphere_stream_wait", g.stream.handle, self._fork.handle)
            for g in self.groups:
                g.model._step.encode_steps(g.stream, k)
            for g in self.groups:
                g.done.record(g.stream)
                _lib.call("tsphere_stream_wait", main, g.done.handle)
        finally:
            h = C.c_void_p()
            rc = L.tsphere_graph_end(main, C.byref(h))
        _lib.check(rc, "ensemble graph capture")
        self._graph = h.value
        self._graph_steps = k

    def step(self, steps: int = 1) -> None:
        """Enqueue `steps` steps of every run on the ensemble stream (no sync)."""
        if steps <= 0:
            return
        for g in self.groups:  # init / uploads queued on a group's own stream
            g.stream.sync()
        k = self.graph_steps
        if self._graph is None or self._graph_steps != k:
            self._capture(k)
        full, rest = divmod(steps, k)
        for _ in range(full):
            _lib.call("tsphere_graph_launch", self._graph, self.stream.handle)
        if rest:
            self._capture(rest)
            _lib.call("tsphere_graph_launch", self._graph, self.stream.handle)
            self.invalidate()
        for g in self.groups:
            g.model._last_ran = "step"

    def sync(self) -> None:
        self.stream.sync()
        for g in self.groups:
            g.stream.sync()

    def destroy(self) -> None:
        self.sync()
        self.invalidate()
        for g in self.groups:
            g.model.destroy()
            g.plan.destroy()
        self.groups = []
