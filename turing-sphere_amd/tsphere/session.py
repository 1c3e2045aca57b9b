"""Runs described the way the reference describes them: model registry and
presets (src/mgpu/registry.ts), the RunSpec command grammar
(src/bench/runSpec.ts), the seeded perturbation (src/mgpu/noise.ts),
``ModelSession`` (src/mgpu/session.ts) and state digests
(src/mgpu/digest.ts), so that one run can be reproduced -- and compared --
across implementations from the same command line.
"""
from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass, field

import numpy as np

from .layout import ShtConfig, grid_for_lmax

MFILES = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mfiles")


# ----------------------------------------------------------------- registry
@dataclass(frozen=True)
class ParamSpec:
    """registry.ts:20-28"""
    key: str
    label: str
    value: float
    min: float
    max: float
    step: float


@dataclass(frozen=True)
class MModel:
    """registry.ts:30-45 (the source is read from the package's .m files)."""
    key: str
    label: str
    species: tuple
    state: tuple
    params: tuple
    pdeg: int
    seed_amp: float

    @property
    def source(self) -> str:
        with open(os.path.join(MFILES, self.key + ".m")) as f:
            return f.read()

    @property
    def param_keys(self) -> list:
        return [p.key for p in self.params]


def _state_for(species):
    return tuple(s.upper() for s in species)


# registry.ts:50-112: same keys, defaults, ranges, dealiasing degree and seed amplitude
M_MODELS = (
    MModel("schnakenberg", "Schnakenberg", ("u", "v"), _state_for(("u", "v")), (
        ParamSpec("a", "a", 0.1, 0.01, 0.5, 0.01),
        ParamSpec("b", "b", 0.9, 0.1, 2, 0.05),
        ParamSpec("D1", "D1", 4e-4, 1e-5, 5e-3, 1e-5),
        ParamSpec("D2", "D2", 8e-3, 1e-4, 5e-2, 1e-4),
        ParamSpec("dt", "dt", 0.05, 0.005, 0.5, 0.005)), 3, 1e-2),
    MModel("brusselator", "Brusselator", ("u", "v"), _state_for(("u", "v")), (
        ParamSpec("A", "A", 3, 0.5, 6, 0.1),
        ParamSpec("B", "B", 9, 1, 15, 0.25),
        ParamSpec("D1", "D1", 3.33e-3, 1e-4, 2e-2, 1e-4),
        ParamSpec("D2", "D2", 1.67e-2, 1e-3, 1e-1, 1e-3),
        ParamSpec("dt", "dt", 0.02, 0.002, 0.1, 0.002)), 3, 1e-2),
    MModel("allencahn", "Allen-Cahn", ("u",), _state_for(("u",)), (
        ParamSpec("eps2", "eps^2", 1e-3, 1e-4, 1e-2, 1e-4),
        ParamSpec("dt", "dt", 0.02, 0.002, 0.2, 0.002)), 3, 1e-2),
)


def m_model_by_key(key: str):
    return next((m for m in M_MODELS if m.key == key), None)


def default_params(m: MModel) -> dict:
    return {p.key: p.value for p in m.params}


@dataclass(frozen=True)
class Preset:
    """registry.ts:118-125"""
    key: str
    label: str
    model_key: str
    params: tuple = ()  # (key, value) overrides


PRESETS = (  # registry.ts:127-141
    Preset("schnak-spots", "Schnakenberg - spots", "schnakenberg"),
    Preset("schnak-coarse", "Schnakenberg - coarse spots", "schnakenberg",
           (("D1", 1e-3), ("D2", 2e-2))),
    Preset("schnak-fine", "Schnakenberg - fine spots", "schnakenberg",
           (("D1", 1.6e-4), ("D2", 3.2e-3))),
    Preset("brussel", "Brusselator - stripes & spots", "brusselator"),
    Preset("allencahn", "Allen-Cahn - coarsening", "allencahn"),
)


def resolve_preset(key: str):
    """runSpec.ts:47-61: (preset, model, params), same error texts."""
    preset = next((p for p in PRESETS if p.key == key), None)
    if preset is None:
        raise ValueError(f"unknown preset '{key}' (have: {', '.join(p.key for p in PRESETS)})")
    model = m_model_by_key(preset.model_key)
    if model is None:
        raise ValueError(f"preset '{key}' names unknown model '{preset.model_key}'")
    params = default_params(model)
    params.update(dict(preset.params))
    return preset, model, params


# ------------------------------------------------------------------ RunSpec
BENCH_COMMAND = "python -m tsphere.run"
DEFAULT_LMAX = 63
DEFAULT_SEED = 1
DEFAULT_STEPS = 2000
DEFAULT_WARMUP = 100


@dataclass
class RunSpec:
    """runSpec.ts:19-31"""
    preset: str
    lmax: int
    seed: float
    steps: int
    warmup: int
    params: dict = field(default_factory=dict)


def model_for_spec(spec: RunSpec) -> MModel:
    return resolve_preset(spec.preset)[1]


def config_for_spec(spec: RunSpec) -> ShtConfig:
    """runSpec.ts:67-71: the dealiased grid of the model's reaction degree."""
    nlat, nphi = grid_for_lmax(spec.lmax, model_for_spec(spec).pdeg)
    return ShtConfig(spec.lmax, spec.lmax, nlat, nphi)


def _num(v) -> str:
    """Number::toString of the JS engine the reference's command lines come
    from: shortest round-trip digits, fixed notation for 1e-7 <= |v| < 1e21,
    exponential ("1e-7", "1.5e+21") outside."""
    from decimal import Decimal

    v = float(v)
    if v == 0:
        return "0"
    sign = "-" if v < 0 else ""
    t = Decimal(repr(abs(v))).as_tuple()
    digits = "".join(map(str, t.digits)).rstrip("0") or "0"
    k = len("".join(map(str, t.digits))) + t.exponent  # value = 0.digits x 10^k
    n = len(digits)
    if n <= k <= 21:
        out = digits + "0" * (k - n)
    elif 0 < k <= 21:
        out = digits[:k] + "." + digits[k:]
    elif -6 < k <= 0:
        out = "0." + "0" * (-k) + digits
    else:
        e = k - 1
        out = digits[0] + ("." + digits[1:] if n > 1 else "") + ("e+" if e >= 0 else "e-") + str(abs(e))
    return sign + out


def format_command(spec: RunSpec) -> str:
    """runSpec.ts:75-89"""
    model = model_for_spec(spec)
    parts = [BENCH_COMMAND, f"--preset {spec.preset}", f"--lmax {spec.lmax}",
             f"--steps {spec.steps}", f"--seed {_num(spec.seed)}"]
    parts += [f"--{p.key} {_num(spec.params[p.key])}" for p in model.params]
    if spec.warmup != DEFAULT_WARMUP:
        parts.append(f"--warmup {spec.warmup}")
    return " ".join(parts)


def parse_args(argv) -> RunSpec:
    """runSpec.ts:93-146: `--key value` or `--key=value`, any order; the
    reference's error messages."""
    flags = {}
    i = 0
    argv = list(argv)
    while i < len(argv):
        arg = argv[i]
        if not arg.startswith("--"):
            raise ValueError(f"unexpected argument '{arg}'")
        eq = arg.find("=")
        if eq >= 0:
            key, value = arg[2:eq], arg[eq + 1:]
        else:
            key = arg[2:]
            i += 1
            value = argv[i] if i < len(argv) else None
        if value is None:
            raise ValueError(f"--{key} needs a value")
        if not key:
            raise ValueError(f"bad option '{arg}'")
        flags[key] = value
        i += 1

    def take(key):
        return flags.pop(key, None)

    def number(key, dflt):
        raw = take(key)
        if raw is None:
            return dflt
        try:
            v = float(raw)
        except ValueError:
            v = math.nan
        if not math.isfinite(v):
            raise ValueError(f"--{key} must be a number (got '{raw}')")
        return v

    def count(key, dflt, lo):
        v = number(key, dflt)
        if not float(v).is_integer() or v < lo:
            raise ValueError(f"--{key} must be an integer >= {lo} (got '{_num(v)}')")
        return int(v)

    preset_key = take("preset") or PRESETS[0].key
    _, model, params = resolve_preset(preset_key)
    spec = RunSpec(preset=preset_key, lmax=count("lmax", DEFAULT_LMAX, 1),
                   seed=number("seed", DEFAULT_SEED), steps=count("steps", DEFAULT_STEPS, 1),
                   warmup=count("warmup", DEFAULT_WARMUP, 0), params=params)
    for p in model.params:
        raw = take(p.key)
        if raw is None:
            continue
        try:
            v = float(raw)
        except ValueError:
            v = math.nan
        if not math.isfinite(v):
            raise ValueError(f"--{p.key} must be a number (got '{raw}')")
        params[p.key] = v
    if flags:
        raise ValueError(
            f"unknown option(s): {', '.join('--' + k for k in flags)}\n"
            f"parameters of {model.label}: {' '.join('--' + p.key for p in model.params)}")
    return spec


# -------------------------------------------------------------------- noise
def make_rand(seed):
    """noise.ts:15-24: mulberry32, uniform in [0, 1)."""
    s = int(seed) & 0xFFFFFFFF

    def rand():
        nonlocal s
        s = (s + 0x6D2B79F5) & 0xFFFFFFFF
        t = s
        t = ((t ^ (t >> 15)) * (t | 1)) & 0xFFFFFFFF
        t ^= (t + (((t ^ (t >> 7)) * (t | 61)) & 0xFFFFFFFF)) & 0xFFFFFFFF
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296.0
    return rand


def seeded_noise(npts: int, amp: float, seed) -> np.ndarray:
    """noise.ts:27-51: amp * N(0,1) per grid point (mulberry32 + Box-Muller,
    pairs: cos first, the sin spare second).  Box-Muller goes through libm's
    log/cos/sin, so the last bit may differ from the JS engine's."""
    rand = make_rand(seed)
    out = np.empty(npts, np.float32)
    i = 0
    while i < npts:
        u = 0.0
        while u == 0.0:
            u = rand()
        r = math.sqrt(-2.0 * math.log(u))
        th = 2.0 * math.pi * rand()
        out[i] = amp * (r * math.cos(th))
        if i + 1 < npts:
            out[i + 1] = amp * (r * math.sin(th))
        i += 2
    return out


# ------------------------------------------------------------------- digest
@dataclass
class StateDigest:
    """digest.ts:16-31"""
    n: int
    min: float
    max: float
    mean: float
    rms: float
    fourier: str
    adapter: str

    def as_dict(self) -> dict:
        return dict(self.__dict__)


def digest_of(values, fourier: str, adapter: str) -> StateDigest:
    """digest.ts:33-56 (sums in f64, as the JS numbers are)."""
    a = np.asarray(values, np.float64).reshape(-1)
    n = a.size
    return StateDigest(n, float(a.min()), float(a.max()), float(a.sum() / n),
                       float(math.sqrt(float(np.dot(a, a)) / n)), fourier, adapter)


def rel_l2(a, b) -> float:
    """digest.ts:59-70"""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(math.sqrt(float(np.sum((a - b) ** 2)) / max(float(np.sum(b * b)), 1e-300)))


def _prec9(v: float) -> str:
    """JS toPrecision(9)."""
    if v == 0:
        return "0.00000000"
    e = math.floor(math.log10(abs(v)))
    if e < -6 or e >= 9:
        m = f"{v:.8e}"
        mant, ex = m.split("e")
        return f"{mant}e{'+' if int(ex) >= 0 else '-'}{abs(int(ex))}"
    return f"{v:.{max(0, 8 - e)}f}"


def format_digest(d: StateDigest) -> str:
    """digest.ts:72-79"""
    return (f"n={d.n} min={_prec9(d.min)} max={_prec9(d.max)} mean={_prec9(d.mean)} "
            f"rms={_prec9(d.rms)} fourier={d.fourier}")


def digest_drift(a: StateDigest, b: StateDigest) -> float:
    """digest.ts:82-90"""
    def rel(x, y):
        return abs(x - y) / max(abs(x), abs(y), 1e-30)
    return max(rel(a.min, b.min), rel(a.max, b.max), rel(a.mean, b.mean), rel(a.rms, b.rms))


# ------------------------------------------------------------------ session
class ModelSession:
    """session.ts:31-218: one running model -- grid, transforms, compiled .m,
    seeded state -- with an optional finer display grid."""

    def __init__(self):
        self.t = 0.0
        self.steps = 0
        self._display = None
        self._oversample = 1

    @staticmethod
    def create(model: MModel, params: dict, lmax: int, source: str | None = None,
               oversample: int = 1, legendre: str = "fp32") -> "ModelSession":
        """session.ts:83-120"""
        from .model import GpuModel
        from .sht import ShtPlan

        self = ModelSession()
        os_ = max(1, int(round(oversample)))
        nlat, nphi = grid_for_lmax(lmax, model.pdeg)
        cfg = ShtConfig(lmax, lmax, nlat, nphi)
        self.model, self.cfg = model, cfg
        self.sht = ShtPlan.create(cfg, legendre=legendre, max_fields=len(model.state))
        try:
            if os_ > 1:
                self._display = ShtPlan.create(ShtConfig(lmax, lmax, os_ * nlat, os_ * nphi))
            self.gpu = GpuModel.create(self.sht, cfg, source or model.source,
                                       model.param_keys, list(model.state),
                                       list(model.species))
            self.gpu.set_params(params)
        except Exception:
            if self._display is not None:
                self._display.destroy()
            self.sht.destroy()
            raise
        self._params = dict(params)
        self._oversample = os_
        self.npts = nlat * nphi
        return self

    @property
    def oversample(self) -> int:
        return self._oversample

    @property
    def view_sht(self):
        return self._display or self.sht

    def set_oversample(self, oversample: int) -> None:
        """session.ts:126-140: display only; the state is untouched."""
        from .sht import ShtPlan

        os_ = max(1, int(round(oversample)))
        if os_ == self._oversample:
            return
        nxt = (ShtPlan.create(ShtConfig(self.cfg.lmax, self.cfg.mmax, os_ * self.cfg.nlat,
                                        os_ * self.cfg.nphi)) if os_ > 1 else None)
        old, self._display, self._oversample = self._display, nxt, os_
        if old is not None:
            old.destroy()

    def seed(self, seed) -> None:
        """session.ts:143-147"""
        self.gpu.init(seeded_noise(self.npts, self.model.seed_amp, seed))
        self.t = 0.0
        self.steps = 0

    def set_params(self, params: dict) -> None:
        self._params = dict(params)
        self.gpu.set_params(params)

    def step(self, n: int = 1) -> None:
        """session.ts:154-159: enqueue, nothing read back."""
        self.gpu.step(n)
        self.t += n * self._params.get("dt", 0.0)
        self.steps += n

    def sync(self) -> None:
        self.gpu.sync()

    def measure(self, n: int) -> float:
        """session.ts:172-180: ms/step of a batch of n steps, leaving the
        simulation where it was (state snapshotted and restored)."""
        self.gpu.snapshot_state()
        t0 = time.perf_counter()
        self.gpu.step(n)
        self.sync()
        ms = (time.perf_counter() - t0) * 1e3 / n
        self.gpu.restore_state()
        return ms

    def read(self, name: str) -> np.ndarray:
        return self.gpu.read(name)

    def read_species(self, k: int) -> np.ndarray:
        """session.ts:200-207: the species on the display grid -- synthesised
        from the spectral state on the finer plan when oversampling."""
        if self._display is None:
            return self.read(self.model.species[k])
        buf = self.gpu.value_buffer(self.model.state[k])
        if buf is None:
            raise KeyError(f"readSpecies: no buffer for state '{self.model.state[k]}'")
        self.gpu.sync()
        return self._display.synth_from(buf)

    def describe(self) -> dict:
        return self.gpu.describe()

    def destroy(self) -> None:
        self.gpu.destroy()
        if self._display is not None:
            self._display.destroy()
        self.sht.destroy()


# ---------------------------------------------------------------- one run
def run(spec: RunSpec, batch: int = 16, latency_steps: int = 200, digest: bool = True,
        dump_state: str | None = None, legendre: str = "fp32") -> dict:
    """scripts/bench.ts:150-316: warmup, batched throughput, one-step
    latencies, and (from a fresh seed) the digest of the state after exactly
    `steps` steps.  Returns the reference's JSON record."""
    import json

    from . import _lib

    model = model_for_spec(spec)
    s = ModelSession.create(model, spec.params, spec.lmax, legendre=legendre)
    try:
        s.seed(spec.seed)
        plan = s.describe()
        kernels = s.gpu.step_plan.generated_kernels()           # generated kernels launched
        statements = s.gpu.step_plan.elementwise_statements()   # element-wise .m lines
        s.step(spec.warmup)
        s.sync()
        batches = max(1, math.ceil(spec.steps / batch))
        run_n = 0
        t0 = time.perf_counter()
        enc = 0.0
        for _ in range(batches):
            n = min(batch, spec.steps - run_n)
            e0 = time.perf_counter()
            s.step(n)
            enc += time.perf_counter() - e0
            s.sync()
            run_n += n
        thr = (time.perf_counter() - t0) * 1e3 / run_n
        lat_n = min(spec.steps, latency_steps)
        samples = np.empty(lat_n)
        for i in range(lat_n):
            a = time.perf_counter()
            s.step(1)
            s.sync()
            samples[i] = (time.perf_counter() - a) * 1e3
        srt = np.sort(samples)
        latency = {"meanMs": float(samples.mean()), "medianMs": float(np.median(samples)),
                   "p05Ms": float(srt[int(0.05 * (lat_n - 1))]),
                   "p95Ms": float(srt[int(0.95 * (lat_n - 1))]), "minMs": float(srt[0])}
        field_ = s.read(model.species[0])
        finite = bool(np.all(np.isfinite(field_)))
        dg, state = None, None
        if digest:
            s.seed(spec.seed)
            s.step(spec.steps)
            s.sync()
            state = s.read(model.state[0])
            dg = digest_of(state, s.sht.fourier_mode, "MI355X (gfx950)")
        rec = {
            "command": format_command(spec),
            "spec": dict(spec.__dict__),
            "model": model.key,
            "backend": {"adapter": "MI355X (gfx950)",
                        "runtime": f"libtsphere {_lib.load().tsphere_version()} (HIP)",
                        "precision": "fp32" if legendre == "fp32" else "bf16x3"},
            "grid": {"lmax": s.cfg.lmax, "nlat": s.cfg.nlat, "nphi": s.cfg.nphi,
                     "nlm": s.sht.nlm},
            "compiled": {"opsPerStep": len(plan["step"]), "kernels": kernels,
                         "elementwiseStatements": statements},
            "digest": dg.as_dict() if dg else None,
            "throughput": {"batch": batch, "msPerStep": thr, "stepsPerSec": 1e3 / thr,
                           "encodeMsPerStep": enc * 1e3 / run_n},
            "latency": latency,
            "state": {"t": s.t, "steps": s.steps, "species": model.species[0],
                      "min": float(field_.min()), "max": float(field_.max()),
                      "contrast": float(field_.max() - field_.min()), "finite": finite},
        }
        if dump_state and state is not None:
            with open(dump_state, "w") as f:
                json.dump({"command": rec["command"], "spec": rec["spec"],
                           "digest": rec["digest"], "state": [float(x) for x in state]}, f)
        return rec
    finally:
        s.destroy()
