#!/usr/bin/env python3
"""Benchmark: IMEX steps/s + SHT GFLOP/s at lmax = 1023 on one MI355X.

Workload (the configuration the metric is quoted on): Brusselator, 2 species,
lmax 1023 on the Gauss grid 1024 x 2048, FP32 state, A=4.5, B=6.96,
D1=2e-4, D2=8*D1, dt=0.002; initial state (A, B/A) + uniform(-0.01, 0.01)
noise (mulberry32 seed 0, u's draws then v's), analysed to spectral on the
GPU.  A step is the full IMEX Euler step compiled from the Brusselator .m
file (2 syntheses, 2 analyses, 5 element-wise statements), planned as two
launches: one generated grid kernel (row FFT synthesis of both species, the
reaction statements, row FFT analysis) and the spectral op (Legendre
analysis, the diagonal implicit solve and the next step's Legendre synthesis
in one kernel).  The timed steps replay HIP graphs whose step count divides
`--steps` (at most `--graph-steps`), captured before the timed region; inputs
are resident in HBM; the timed region is `--steps` steps bracketed by device
synchronisation.  Before timing, one synthesis -> analysis round trip of the
initial state must come back within 1e-3 relative L2 (else exit status 3).
The same step is also timed with the FP64-recurrence and split-BF16 Legendre
forms and as the plain FP32 step (Legendre sums on the vector ALU, separate
launches, no matrix cores); `split_bf16_speedup` compares the split-BF16 step
with both FP32 steps.

Printed: ONE JSON line with the metric, a `roofline` object for the dominant
kernel (algorithmic flops per launch / its mean duration measured with HIP
events on the stream it runs on) and a `cpu_baseline` object (the f64 oracle
-- the reference's direct-summation algorithm -- timed on a bounded sample of
the same step on this host's cores, extrapolated to steps/s).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--legendre fp32|bf16x3]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "turing-sphere_amd"))

LMAX, NLAT, NPHI = 1023, 1024, 2048
PARAMS = {"A": 4.5, "B": 6.96, "D1": 2e-4, "D2": 8 * 2e-4, "dt": 0.002}
PEAK_FP32 = 157.3e12   # MI355X FP32 vector (packed) TFLOP/s, datasheet
PEAK_BF16 = 2.5e15     # dense BF16 MFMA
PEAK_HBM = 8.0e12


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def nlm_calc(lmax, mmax):
    return (mmax + 1) * (lmax + 1) - (mmax * (mmax + 1)) // 2


def flops_per_transform(lmax, nlat, nphi):
    """Algorithmic flops of one transform of one field: 2*nlm*nlat for the
    Legendre contraction (4 flops per complex x real MAC per latitude pair)
    + 2.5 nlat nphi log2 nphi for the FFT."""
    nlm = nlm_calc(lmax, lmax)
    return 2.0 * nlm * nlat, 2.5 * nlat * nphi * np.log2(nphi)


def initial_grid(npts, seed=0):
    """mulberry32 uniform deviates (src/mgpu/noise.ts:15-24) -> 0.02 r - 0.01."""
    out = np.empty(2 * npts)
    s = seed & 0xFFFFFFFF
    # vectorised mulberry32: the state sequence is s_k = seed + k * 0x6d2b79f5
    k = np.arange(1, 2 * npts + 1, dtype=np.uint64)
    t = ((np.uint64(s) + k * np.uint64(0x6D2B79F5)) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    with np.errstate(over="ignore"):
        t = (t ^ (t >> np.uint32(15))) * (t | np.uint32(1))
        t ^= t + (t ^ (t >> np.uint32(7))) * (t | np.uint32(61))
        out[:] = (t ^ (t >> np.uint32(14))).astype(np.float64) / 4294967296.0
    r = 0.02 * out - 0.01
    return r[:npts], r[npts:]


def time_stage(fn, stream, reps):
    """Mean duration (ms) of `fn()` launches, HIP events on `stream`."""
    from tsphere import Event

    fn()
    stream.sync()
    e0, e1 = Event(), Event()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    return e0.elapsed_ms(e1) / reps


def cpu_sample(rows: int, threads: int, nfield: int = 2):
    """The oracle (f64 direct summation, reference.ts:37-118) on `rows`
    latitude rows of one synthesis and one analysis per field, on `threads`
    host threads; every row is an independent slice of the transform, so the
    full step costs nlat/rows times the sample."""
    import oracle as O

    cfg = O.Cfg(LMAX, LMAX, NLAT, NPHI)
    ref = O.ShtReference(cfg)
    rng = np.random.default_rng(0)
    q = rng.normal(size=2 * cfg.nlm) / 100
    spat = rng.normal(size=NLAT * NPHI)
    prev = O.get_threads()
    O.set_threads(threads)
    try:
        t0 = time.perf_counter()
        for _ in range(nfield):
            ref.synth(q, rows=(0, rows))
            gm = ref.fourier_analys(spat, rows=(0, rows))
            ref.legendre_analys(gm, rows=(0, rows))
        dt = time.perf_counter() - t0
    finally:
        O.set_threads(prev)
    step_s = dt * (NLAT / rows)
    return {"value": 1.0 / step_s, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": (f"f64 direct-summation synth+analys of {nfield} fields on {rows} of "
                       f"{NLAT} latitude rows ({dt:.1f} s on {threads} threads), scaled by "
                       f"{NLAT / rows:.0f} to one IMEX step; reaction/IMEX element-wise "
                       f"work (<0.1% of it) not included")}


def cpu_baseline(rows: int, rows_1t: int):
    """All host threads (OpenMP) on `rows` rows, and one thread -- the
    reproducible per-core figure -- on `rows_1t` rows."""
    import oracle as O

    out = cpu_sample(rows, O.get_threads())
    out["single_thread"] = cpu_sample(rows_1t, 1)
    return out


def graph_size(steps: int, cap: int) -> int:
    """The largest step count <= cap that divides `steps` (the timed steps are
    whole graph replays)."""
    return max(g for g in range(1, max(1, min(cap, steps)) + 1) if steps % g == 0)


class SanityError(RuntimeError):
    pass


def run_variant(legendre: str, steps: int, warmup: int, graph_steps: int, err: bool,
                latency_steps: int = 0, plain: bool = False):
    """Time the IMEX step with one Legendre implementation; per-kernel means;
    error of one synthesis of the initial state vs the f64 oracle.  With
    `plain` (FP32 only), the plain FP32 step: Legendre sums on the vector ALU
    (tile impl 1, its built-in shape) as separate synthesis and analysis
    launches, the implicit solve as its generated kernel, no matrix cores."""
    import tsphere
    from tsphere import ShtConfig, ShtPlan, _lib
    from tsphere.model import GpuModel

    cfg = ShtConfig(LMAX, LMAX, NLAT, NPHI)
    npts, nlm = cfg.npts, cfg.nlm
    tag = legendre + (" (plain: vector-ALU Legendre sums)" if plain else "")
    log(f"[bench] plan lmax {LMAX} {NLAT}x{NPHI} legendre={tag}")
    sht = ShtPlan.create(cfg, legendre=legendre, max_fields=2)
    if plain:
        for st in ("synth", "analys"):
            sht.set_tile(st, impl=1)
        tuned = {st: sht.get_tile(st, 2) for st in ("synth", "analys")}
    else:
        tuned = sht.tune(2, reps=5)
    log(f"[bench] {tag} tuned tiles: {tuned}")
    src = open(os.path.join(ROOT, "turing-sphere_amd", "mfiles", "brusselator.m")).read()
    gm = GpuModel.create(sht, cfg, src, ["A", "B", "D1", "D2", "dt"], ["U", "V"], ["u", "v"],
                         spectral=not plain, fold=not plain)
    gm.set_params(PARAMS)
    ru, rv = initial_grid(npts, 0)
    u0 = (PARAMS["A"] + ru).astype(np.float32)
    v0 = (PARAMS["B"] / PARAMS["A"] + rv).astype(np.float32)
    U0 = sht.analys(u0)
    V0 = sht.analys(v0)
    # sanity gate (scripts/bench-sht.ts:252-285): a round trip of the
    # initial state through this plan must come back before anything is timed
    rt = float(np.sqrt(np.sum((sht.analys(sht.synth(U0)) - U0.astype(np.float64)) ** 2) /
                       np.sum(U0.astype(np.float64) ** 2)))
    log(f"[bench] {tag} round-trip check: relL2 {rt:.2e}")
    if not rt < 1e-3:
        raise SanityError(f"{legendre}: round trip relL2 {rt:.3e} >= 1e-3; not timing")
    gm.upload("U", U0)
    gm.upload("V", V0)
    log(f"[bench] step ops: {gm.describe()['step']}")
    gm.graph_steps = graph_size(steps, graph_steps)
    gm.step(warmup)
    gm.prepare(steps)   # capture the timed graphs outside the timed region
    gm.sync()
    t0 = time.perf_counter()
    gm.step(steps)
    gm.sync()
    elapsed = time.perf_counter() - t0
    U = gm.read("U")
    finite = bool(np.all(np.isfinite(U)))
    log(f"[bench] {tag}: {steps} steps ({steps // gm.graph_steps} replays of a "
        f"{gm.graph_steps}-step graph) in {elapsed * 1e3:.1f} ms -> {steps / elapsed:.1f} steps/s")
    latency = None
    if latency_steps > 0:
        # one step per submit, each waited for (scripts/bench.ts:188-228)
        gm.graph_steps = 0
        gm.step(1)
        gm.sync()
        samples = np.empty(latency_steps)
        for i in range(latency_steps):
            t1 = time.perf_counter()
            gm.step(1)
            gm.sync()
            samples[i] = (time.perf_counter() - t1) * 1e3
        latency = {"samples": latency_steps, "meanMs": round(float(samples.mean()), 4),
                   "medianMs": round(float(np.median(samples)), 4),
                   "p90Ms": round(float(np.percentile(samples, 90)), 4),
                   "minMs": round(float(samples.min()), 4)}
        log(f"[bench] {tag} one-step latency: {latency}")

    L = _lib.load()
    s = gm.stream
    Ub = gm._host.get("U")
    fm = tsphere.DeviceBuffer(2 * 2 * (LMAX + 1) * NLAT, "bench-fm")
    qo = tsphere.DeviceBuffer(2 * 2 * nlm, "bench-q")
    sp = tsphere.DeviceBuffer(2 * npts, "bench-spat")
    h = sht.handle
    kms = {
        "leg_synth": time_stage(lambda: _lib.check(L.tsphere_leg_synth(
            h, 2, Ub.ptr, 2 * nlm, fm.ptr, s.handle)), s, 20),
        "fft_synth": time_stage(lambda: _lib.check(L.tsphere_fourier_synth(
            h, 2, fm.ptr, sp.ptr, npts, s.handle)), s, 20),
        "fft_analys": time_stage(lambda: _lib.check(L.tsphere_fourier_analys(
            h, 2, sp.ptr, npts, fm.ptr, s.handle)), s, 20),
        "leg_analys": time_stage(lambda: _lib.check(L.tsphere_leg_analys(
            h, 2, fm.ptr, qo.ptr, 2 * nlm, s.handle)), s, 20),
    }
    grid_ops = [op.grid if op.kind == "spectral" else op for op in gm.step_plan.ops
                if op.kind in ("grid", "spectral")]
    if grid_ops:
        # the step runs the Fourier stages inside its generated grid kernel
        kms["grid_kernel"] = time_stage(lambda: gm.step_plan.launch_grid(grid_ops[0], s), s, 20)
    if any(getattr(op, "mode", "") == "spectral" for op in gm.step_plan.ops):
        # the step's spectral kernel (analysis, implicit solve, next synthesis)
        # on scratch copies of the state
        lam = gm._host.get("lam")
        st = tsphere.DeviceBuffer(2 * 2 * nlm, "bench-state")
        tsphere.DeviceBuffer.view(st, 0, 2 * nlm).copy_from(gm._host.get("U"), s)
        tsphere.DeviceBuffer.view(st, 2 * nlm, 2 * nlm).copy_from(gm._host.get("V"), s)
        ptrs = (C.c_void_p * 2)(st.ptr, st.ptr + 4 * 2 * nlm)
        dt = np.float32(PARAMS["dt"])
        s1 = (C.c_float * 2)(dt, dt)
        s2 = (C.c_float * 2)(dt * np.float32(PARAMS["D1"]), dt * np.float32(PARAMS["D2"]))
        spc = (C.c_int * 2)(0, 1)
        kms["leg_spectral"] = time_stage(lambda: _lib.check(L.tsphere_leg_spectral_imex(
            h, 2, fm.ptr, ptrs, lam.ptr, s1, s2, spc, s.handle)), s, 20)
    log(f"[bench] {tag} kernel ms: {kms}")
    out = {"steps_per_s": steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
           "graph_steps": graph_size(steps, graph_steps), "round_trip_rel_l2": rt,
           "kernel_ms": kms, "finite": finite,
           "tiles": {k: tuned[k] for k in ("synth", "analys")}}
    if latency is not None:
        out["latency"] = latency
    if err:
        import oracle as O

        ref = O.ShtReference(O.Cfg(LMAX, LMAX, NLAT, NPHI))
        spat = sht.synth(U0)
        out["synth_rel_err_vs_f64"] = O.rel_l2(spat, ref.synth_fast(U0.astype(np.float64)))
        q = sht.analys(u0)
        out["analys_rel_err_vs_f64"] = O.rel_l2(q, ref.analys_fast(u0.astype(np.float64)))
        log(f"[bench] {tag} errors vs f64: synth {out['synth_rel_err_vs_f64']:.2e} "
            f"analys {out['analys_rel_err_vs_f64']:.2e}")
    gm.destroy()
    sht.destroy()
    return out


def batched_sht(lmax: int, nf: int, legendre: str, reps: int = 10):
    """Round trips of `nf` fields at `lmax` (Gauss lmax+1 x 2(lmax+1)) per second."""
    import tsphere

    cfg = tsphere.ShtConfig(lmax, lmax, lmax + 1, 2 * (lmax + 1))
    p = tsphere.ShtPlan.create(cfg, legendre=legendre, max_fields=nf)
    tiles = p.tune(nf, reps=3)
    q = tsphere.DeviceBuffer(nf * 2 * cfg.nlm)
    sb = tsphere.DeviceBuffer(nf * cfg.npts)
    q.fill(0.01, p.stream)
    bs = p.create_synth_binding(q, sb, nf, 2 * cfg.nlm, cfg.npts)
    ba = p.create_analys_binding(sb, q, nf, cfg.npts, 2 * cfg.nlm)
    ms = time_stage(lambda: (p.encode_synth_into(p.stream, bs), p.encode_analys_into(p.stream, ba)),
                    p.stream, reps)
    p.destroy()
    f_leg, f_fft = flops_per_transform(lmax, lmax + 1, 2 * (lmax + 1))
    return {"round_trips_per_s": nf / (ms * 1e-3), "ms_per_batch": ms,
            "sht_gflops": 2 * nf * (f_leg + f_fft) / (ms * 1e-3) / 1e9,
            "tiles": {k: tiles[k] for k in ("synth", "analys")}}


def reference_sht_bench(lmax: int = 63, steps: int = 2000, warmup: int = 100, batch: int = 16,
                        latency_steps: int = 200):
    """The reference's transform benchmark (scripts/bench-sht.ts:187-360) on
    this library: one field at lmax 63 on the grid its default run uses
    (gridForLmax(63, 3) = 128 x 256, src/sht/layout.ts:53-59), a round trip
    = one synthesis and one analysis, alternating two spectral buffers.
    Throughput as batches of `batch` round trips per submission awaited once
    each (one captured HIP graph per submission here), latency as single
    round trips awaited each; a first round trip must come back within 1e-3
    relative L2 before anything is timed.  Published for this command
    (README.md:350-356, RTX PRO 6000 Blackwell): 0.084 ms per round trip
    through WebGPU (11,848/s), 0.021 ms through upstream SHTNS's CUDA path
    (48,009/s)."""
    import tsphere
    from tsphere import Event, _lib
    from tsphere.layout import grid_for_lmax

    nlat, nphi = grid_for_lmax(lmax, 3)  # the default preset's cubic reaction
    cfg = tsphere.ShtConfig(lmax, lmax, nlat, nphi)
    p = tsphere.ShtPlan.create(cfg, max_fields=1)
    s = p.stream
    rng = np.random.default_rng(1)
    nlm = cfg.nlm
    lsig = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)])
    q0 = (rng.normal(size=(nlm, 2)) / (lsig[:, None] + 1.0)).astype(np.float32)
    q0[:lmax + 1, 1] = 0.0  # m = 0 coefficients are real
    qb = [tsphere.DeviceBuffer(2 * nlm), tsphere.DeviceBuffer(2 * nlm)]
    sp = tsphere.DeviceBuffer(cfg.npts)
    syn = [p.create_synth_binding(qb[i], sp, 1, 2 * nlm, cfg.npts) for i in (0, 1)]
    ana = [p.create_analys_binding(sp, qb[1 - i], 1, cfg.npts, 2 * nlm) for i in (0, 1)]
    cur = [0]

    def encode(n):
        for _ in range(n):
            p.encode_synth_into(s, syn[cur[0]])
            p.encode_analys_into(s, ana[cur[0]])
            cur[0] ^= 1

    qb[0].upload(q0.reshape(-1), s)
    encode(1)
    s.sync()
    back = qb[1].download(s).astype(np.float64)
    ref = q0.reshape(-1).astype(np.float64)
    rt = float(np.sqrt(np.sum((back - ref) ** 2) / np.sum(ref ** 2)))
    if not rt < 1e-3:
        raise SanityError(f"reference transform benchmark: round trip relL2 {rt:.3e} >= 1e-3")
    qb[0].upload(q0.reshape(-1), s)
    cur[0] = 0
    encode(warmup)
    s.sync()
    # one submission = one graph of `batch` round trips (an even count, so
    # every replay starts from the same buffer)
    import ctypes as C_

    _lib.call("tsphere_graph_begin", s.handle)
    try:
        encode(batch)
    finally:
        h = C_.c_void_p()
        rc = _lib.load().tsphere_graph_end(s.handle, C_.byref(h))
    _lib.check(rc, "graph capture")
    nb = max(1, -(-steps // batch))
    t0 = time.perf_counter()
    for _ in range(nb):
        _lib.call("tsphere_graph_launch", h.value, s.handle)
        s.sync()
    thr_ms = (time.perf_counter() - t0) * 1e3 / (nb * batch)
    lat = np.empty(latency_steps)
    for i in range(latency_steps):
        t1 = time.perf_counter()
        encode(1)
        s.sync()
        lat[i] = (time.perf_counter() - t1) * 1e3
    e0, e1 = Event(), Event()
    e0.record(s)
    for _ in range(nb):
        _lib.call("tsphere_graph_launch", h.value, s.handle)
    e1.record(s)
    gpu_ms = e0.elapsed_ms(e1) / (nb * batch)
    _lib.call("tsphere_graph_destroy", h.value)
    finite = bool(np.all(np.isfinite(qb[0].download(s))))
    p.destroy()
    return {"workload": f"1 field, lmax {lmax}, {nlat}x{nphi} (the reference's bench:sht default), "
                        f"round trip = synth + analys",
            "first_round_trip_rel_l2": rt,
            "round_trips": nb * batch, "batch": batch,
            "ms_per_round_trip": round(thr_ms, 5),
            "round_trips_per_s": round(1e3 / thr_ms, 1),
            "gpu_ms_per_round_trip": round(gpu_ms, 5),
            "latency": {"samples": latency_steps, "meanMs": round(float(lat.mean()), 4),
                        "medianMs": round(float(np.median(lat)), 4),
                        "minMs": round(float(lat.min()), 4)},
            "published": {"webgpu_rtx_pro_6000_ms": 0.084, "shtns_cuda_rtx_pro_6000_ms": 0.021},
            "speedup_vs_published": {"webgpu_rtx_pro_6000": round(0.084 / thr_ms, 2),
                                     "shtns_cuda_rtx_pro_6000": round(0.021 / thr_ms, 2)},
            "finite": finite}


def ensemble_bench(steps: int = 200, legendre: str = "fp32"):
    """256 Allen-Cahn runs, 64 each at lmax 63/127/255/511 on (L+1) x 2(L+1)
    Gauss grids, eps = U[0.01, 0.05] (mulberry32 seed 0, its own stream),
    dt = 1e-3, u0 = U(-0.1, 0.1) (mulberry32 seed = run index), all runs
    stepped by one HIP graph per 20 steps (groups as parallel branches)."""
    import tsphere
    from tsphere.ensemble import MixedEnsemble

    src = open(os.path.join(ROOT, "turing-sphere_amd", "mfiles", "allencahn.m")).read()
    groups = [(tsphere.ShtConfig(L, L, L + 1, 2 * (L + 1)), 64) for L in (63, 127, 255, 511)]
    ens = MixedEnsemble.create(src, ["eps2", "dt"], ["U"], ["u"], groups, legendre=legendre,
                               tune=True)
    eps = 0.01 + 0.04 * mulberry_uniform(0, 256)
    ens.set_params({"eps2": eps * eps, "dt": 1e-3})
    noise = []
    for g in ens.groups:
        for r in range(g.runs):
            noise.append((0.2 * mulberry_uniform(g.first_run + r, g.cfg.npts) - 0.1)
                         .astype(np.float32))
    ens.init(noise)
    ens.graph_steps = 20
    ens.step(20)
    ens.sync()
    t0 = time.perf_counter()
    ens.step(steps)
    ens.sync()
    dt = time.perf_counter() - t0
    U = ens.read("U")
    finite = all(bool(np.all(np.isfinite(u))) for u in U)
    tiles = ens.tuning
    ens.destroy()
    f_step = sum(64 * 2 * sum(flops_per_transform(L, L + 1, 2 * (L + 1)))
                 for L in (63, 127, 255, 511))
    return {"workload": "256 Allen-Cahn runs, 64 each at lmax 63/127/255/511, one graph",
            "steps": steps, "batched_steps_per_s": steps / dt,
            "run_steps_per_s": 256 * steps / dt, "ms_per_step": 1e3 * dt / steps,
            "sht_gflops": f_step * steps / dt / 1e9, "finite": finite,
            "tuned_tiles": [{k: v for k, v in t.items() if k in ("synth", "analys", "cached")}
                            for t in tiles]}


def adjoint_bench(steps: int = 200, every: int = 20):
    """Gradient of L = sum_grid (synth(U_200) - u*)^2 with respect to the
    initial state and a, b, D1, D2 of Schnakenberg at lmax 255 (256 x 512),
    dt = 0.01, 200 steps, checkpoints every 20; u* is the u field of a run
    with (a, b, D1, D2) x 1.05 from the same start."""
    import tsphere
    from tsphere.adjoint import AdjointModel

    L = 255
    cfg = tsphere.ShtConfig(L, L, L + 1, 2 * (L + 1))
    sht = tsphere.ShtPlan.create(cfg, max_fields=2)
    src = open(os.path.join(ROOT, "turing-sphere_amd", "mfiles", "schnakenberg.m")).read()
    names = ["a", "b", "D1", "D2", "dt"]
    p = {"a": 0.1, "b": 0.9, "D1": 6.5e-4, "D2": 0.026, "dt": 0.01}
    am = AdjointModel.create(sht, cfg, src, names, ["U", "V"], ["u", "v"], every)
    ru, rv = initial_grid(cfg.npts, 0)
    x0 = {"U": sht.analys((1.0 + ru).astype(np.float32)),
          "V": sht.analys((0.9 + rv).astype(np.float32))}
    am.set_params({k: (v * 1.05 if k != "dt" else v) for k, v in p.items()})
    target = sht.synth(am.forward(x0, steps)["U"])
    am.set_params(p)

    def gradient():
        uN = sht.synth(am.forward(x0, steps)["U"])
        cot = {"U": am.synth_adjoint(2.0 * (uN - target)),
               "V": np.zeros(2 * sht.nlm, np.float32)}
        return (uN,) + am.vjp(x0, steps, cot)

    t0 = time.perf_counter()
    gradient()          # first gradient: captures the step and segment graphs
    t1 = time.perf_counter()
    uN, grads, pg = gradient()
    dt = time.perf_counter() - t1
    out = {"workload": "Schnakenberg lmax 255 (256x512), 200 steps, loss on u_200, "
                       "checkpoints every 20",
           "ms_per_gradient": 1e3 * dt, "first_gradient_ms": 1e3 * (t1 - t0),
           "loss": float(np.sum((uN - target) ** 2)),
           "dL_dparams": {k: pg[k] for k in ("a", "b", "D1", "D2")},
           "finite": bool(np.all(np.isfinite(grads["U"])))}
    am.destroy()
    sht.destroy()
    return out


def default_run():
    """The reference's default benchmark run (runSpec.ts defaults: preset
    schnak-spots, lmax 63 on its dealiased 128 x 256 grid, seed 1, 100 warmup
    + 2000 steps in batches of 16 with one sync each, then 200 one-step
    latencies, then the state digest after exactly 2000 steps from the seed),
    through the session layer -- the number the reference's own benchmark
    prints for the same command."""
    from tsphere.session import parse_args, run

    rec = run(parse_args([]), batch=16, latency_steps=200)
    return {"command": rec["command"], "steps_per_s": rec["throughput"]["stepsPerSec"],
            "ms_per_step": rec["throughput"]["msPerStep"],
            "latency_median_ms": rec["latency"]["medianMs"], "digest": rec["digest"],
            "kernels": rec["compiled"]["kernels"], "ops_per_step": rec["compiled"]["opsPerStep"]}


def mulberry_uniform(seed: int, n: int) -> np.ndarray:
    """n mulberry32 deviates in [0, 1) (src/mgpu/noise.ts:15-24), vectorised."""
    k = np.arange(1, n + 1, dtype=np.uint64)
    t = ((np.uint64(seed & 0xFFFFFFFF) + k * np.uint64(0x6D2B79F5)) &
         np.uint64(0xFFFFFFFF)).astype(np.uint32)
    with np.errstate(over="ignore"):
        t = (t ^ (t >> np.uint32(15))) * (t | np.uint32(1))
        t ^= t + (t ^ (t >> np.uint32(7))) * (t | np.uint32(61))
        return (t ^ (t >> np.uint32(14))).astype(np.float64) / 4294967296.0


def hbm_calibration() -> int:
    """Three device-to-device copies of 256 MiB: each moves 262144 KiB in and
    262144 KiB out of HBM, which calibrates the FETCH_SIZE / WRITE_SIZE
    counters that the dominant kernel's traffic figure is read from."""
    import tsphere

    n = 64 * 1024 * 1024  # floats: 256 MiB
    a, b = tsphere.DeviceBuffer(n), tsphere.DeviceBuffer(n)
    s = tsphere.Stream()
    a.fill(1.0, s)
    s.sync()
    for _ in range(3):
        b.copy_from(a, s)
    s.sync()
    log("[bench] hbm calibration: 3 copies of 256 MiB done")
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--legendre", default="fp32", choices=["fp32", "f64rec", "bf16x3"],
                    help="the variant reported as `value` (both are measured)")
    ap.add_argument("--graph-steps", type=int, default=50)
    ap.add_argument("--cpu-rows", type=int, default=1024)
    ap.add_argument("--cpu-rows-1t", type=int, default=96,
                    help="latitude rows of the single-thread CPU sample")
    ap.add_argument("--latency-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="skip errors and batched extras")
    ap.add_argument("--ensemble-only", type=int, default=0, metavar="STEPS",
                    help="only run the 256-run mixed-lmax Allen-Cahn ensemble for STEPS steps "
                         "(a short workload to trace its graph's kernels) and print its figures")
    ap.add_argument("--sht-only", type=int, default=0, metavar="LMAX",
                    help="time only the batched round trips of 64 fields at LMAX")
    ap.add_argument("--hbm-calibration", action="store_true",
                    help="only copy 256 MiB device to device three times (the known-size "
                         "workload that FETCH_SIZE / WRITE_SIZE counters are scaled by)")
    args = ap.parse_args()
    if args.hbm_calibration:
        return hbm_calibration()
    if args.ensemble_only:
        print(json.dumps(ensemble_bench(args.ensemble_only), default=float), flush=True)
        return 0
    if args.sht_only:
        # the batched round trips alone (64 fields at this lmax), e.g. for a
        # kernel trace of the batched transform's kernels
        print(json.dumps({leg: batched_sht(args.sht_only, 64, leg)
                          for leg in ("fp32", "f64rec", "bf16x3")}, default=float), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): the step runs on a "
              f"single MI355X; across GPUs it would shard by spectral rows m and latitude rows.",
              file=sys.stderr)
        return 2

    variants = {}
    try:
        for leg in ("fp32", "f64rec", "bf16x3"):
            variants[leg] = run_variant(leg, args.steps, args.warmup, args.graph_steps,
                                        not args.quick,
                                        args.latency_steps if leg == args.legendre else 0)
        variants["fp32_valu"] = run_variant("fp32", args.steps, args.warmup, args.graph_steps,
                                            False, 0, plain=True)
    except SanityError as e:
        log(f"[bench] sanity check failed: {e}")
        return 3
    main_v = variants[args.legendre]
    kms = main_v["kernel_ms"]
    f_leg, f_fft = flops_per_transform(LMAX, NLAT, NPHI)
    # the step's Legendre launches: the spectral op (2 fields analysed + 2
    # synthesised per launch) where the step has one, else the two stages
    legk = ("leg_spectral",) if "leg_spectral" in kms else ("leg_synth", "leg_analys")
    dom = max(legk, key=lambda k: kms[k])
    nxf = 4 if dom == "leg_spectral" else 2
    achieved = nxf * f_leg / (kms[dom] * 1e-3)
    peak = PEAK_BF16 if args.legendre == "bf16x3" else PEAK_FP32
    steps_s = main_v["steps_per_s"]
    sht_flops_step = 4 * (f_leg + f_fft)
    # HBM bytes per launch of the dominant kernel, from the committed counter
    # summary (FETCH_SIZE / WRITE_SIZE runs, calibrated on a copy of known size)
    traffic = None
    prof = os.path.join(ROOT, "profiles", f"{dom}_brusselator_l1023_pmc.md")
    if args.legendre == "fp32" and os.path.exists(prof):
        import re

        m = re.search(r"^hbm_bytes_per_launch: (\d+)", open(prof).read(), re.M)
        traffic = int(m.group(1)) if m else None

    # achieved HBM rate of the Fourier-stage kernels on their algorithmic
    # bytes: a field's Fourier scratch is (M+1) nlat complex, its grid nlat
    # nphi reals; the grid kernel reads the scratch, writes the returned grids
    # and the analysed scratch
    fmb, spb = 2 * 8 * (LMAX + 1) * NLAT, 2 * 4 * NLAT * NPHI
    stage_bytes = {"fft_synth": fmb + spb, "fft_analys": spb + fmb, "grid_kernel": 2 * fmb + spb}
    hbm_gbs = {k: round(b / (kms[k] * 1e-3) / 1e9, 1) for k, b in stage_bytes.items() if k in kms}

    extras = {}
    if not args.quick:
        for leg in ("fp32", "f64rec", "bf16x3"):
            extras[f"sht_64fields_l511_{leg}"] = batched_sht(511, 64, leg)
            log(f"[bench] 64-field lmax 511 {leg}: {extras[f'sht_64fields_l511_{leg}']}")
        extras["reference_sht_bench"] = reference_sht_bench()
        log(f"[bench] reference transform benchmark: {extras['reference_sht_bench']}")
        # the same at the grids of src/sht/sht.ts:72-76's published WGSL round
        # trips (RTX PRO 6000; lmax 63 / 127 / 255 on 128x256 / 256x512 /
        # 512x1024: 37.8 / 66.6 / 131.9 us)
        extras["reference_sht_sizes"] = {}
        for lm, pub in ((63, 0.0378), (127, 0.0666), (255, 0.1319)):
            r = reference_sht_bench(lm, steps=400, warmup=20, latency_steps=20)
            extras["reference_sht_sizes"][f"lmax{lm}"] = {
                "ms_per_round_trip": r["ms_per_round_trip"],
                "gpu_ms_per_round_trip": r["gpu_ms_per_round_trip"],
                "published_webgpu_rtx_pro_6000_ms": pub,
                "speedup": round(pub / r["ms_per_round_trip"], 2)}
        log(f"[bench] reference transform sizes: {extras['reference_sht_sizes']}")
        extras["reference_default_run"] = default_run()
        log(f"[bench] default run: {extras['reference_default_run']}")
        extras["ensemble_allencahn_256runs"] = ensemble_bench()
        log(f"[bench] ensemble: {extras['ensemble_allencahn_256runs']}")
        extras["adjoint_schnakenberg_l255"] = adjoint_bench()
        log(f"[bench] adjoint: {extras['adjoint_schnakenberg_l255']}")

    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline on {args.cpu_rows} rows ...")
        cpu = cpu_baseline(args.cpu_rows, args.cpu_rows_1t)

    line = {
        "metric": "IMEX steps/s + SHT GFLOP/s at lmax=1023 (split-BF16 MFMA vs plain FP32)",
        "value": round(steps_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_v["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "f64rec": "f32 (f64 recurrence)", "bf16x3": "bf16x3"}[args.legendre],
        "data": "synthetic",
        "config": {"workload": "brusselator IMEX Euler, lmax 1023, Gauss 1024x2048, 2 species, "
                               "FP32 state, dt 0.002", "model": "brusselator", "lmax": LMAX,
                   "nlat": NLAT, "nphi": NPHI, "legendre": args.legendre,
                   "parallelism": "single GPU"},
        "sht_gflops": round(sht_flops_step * steps_s / 1e9, 2),
        "transforms_per_s": round(4 * steps_s, 2),
        "kernel_ms": {k: round(v, 5) for k, v in kms.items()},
        "latency": main_v.get("latency"),
        "hbm_gbs": hbm_gbs,
        "finite": main_v["finite"],
        "variants": {k: {kk: (round(vv, 8) if isinstance(vv, float) else vv)
                         for kk, vv in v.items() if kk != "kernel_ms"}
                     | {"kernel_ms": {kk: round(vv, 5) for kk, vv in v["kernel_ms"].items()}}
                     for k, v in variants.items()},
        "batched": extras,
        # the precision form's step against the plain FP32 step (vector-ALU
        # Legendre sums) and against the FP32 step with its sums on 4x4
        # matrix blocks (the default)
        "split_bf16_speedup": {
            "vs_plain_fp32_valu": round(variants["bf16x3"]["steps_per_s"] /
                                        variants["fp32_valu"]["steps_per_s"], 3),
            "vs_fp32_blocks": round(variants["bf16x3"]["steps_per_s"] /
                                    variants["fp32"]["steps_per_s"], 3)},
        "roofline": {"bound": "mfma", "kernel": dom,
                     "pipe": ("bf16 MFMA (split x3)" if args.legendre == "bf16x3"
                              else "fp32: recurrence on the vector ALU, sums on the fp32 matrix "
                                   "cores as 4x4 blocks (157.3 TF peak, matrix = packed vector)"),
                     "achieved": round(achieved / 1e12, 3),
                     "peak": round(peak / 1e12, 1), "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": traffic,
                     "flops_per_launch": nxf * f_leg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
