"""Legendre tile autotuning with an on-disk cache.

``tsphere_plan_tune`` (C) times every candidate workgroup shape of both
Legendre stages on the device and keeps the fastest in the plan.  This
module caches the winners per (device, form, lmax, mmax, nlat, nfield) in a
JSON file, so a plan of a known shape applies the cached tiles without
timing anything.  Cache file: ``$TSPHERE_TUNE_CACHE`` or
``~/.cache/tsphere/tiles.json``; unreadable or foreign entries are ignored.
"""
from __future__ import annotations

import ctypes as C
import json
import os

from . import _lib

_STAGES = ("synth", "analys")


def cache_path() -> str:
    return os.environ.get("TSPHERE_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "tsphere", "tiles.json"))


def _key(plan, nfield: int) -> str:
    c = plan.cfg
    return f"gfx950/v5/{plan.legendre}/L{c.lmax}/M{c.mmax}/nlat{c.nlat}/nphi{c.nphi}/nf{nfield}"


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(d: dict) -> None:
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + f".{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: tuning still applies to this plan


def tune_plan(plan, nfield: int, reps: int = 5, cache: bool = True) -> dict:
    """Apply cached tiles for this shape, or time the candidates and cache
    the winners.  Returns {"synth": tile, "analys": tile, "ms": [..],
    "cached": bool}."""
    key = _key(plan, nfield)
    if cache:
        hit = _load().get(key)
        if isinstance(hit, dict) and all(isinstance(hit.get(s), dict) for s in _STAGES):
            try:
                for s in _STAGES:
                    t = hit[s]
                    plan.set_tile(s, t.get("nf", 0), t.get("nwaves", 0), t.get("k", 0),
                                  t.get("ctc", 0), t.get("impl", 0), t.get("rows", 0))
                return {**hit, "cached": True}
            except (_lib.TsphereError, TypeError):
                pass  # stale entry: re-time
    best = (_lib.LegTileC * 2)()
    ms = (C.c_float * 2)()
    _lib.call("tsphere_plan_tune", plan.handle, nfield, reps, plan.stream.handle, best, ms)
    out = {s: {"nf": best[i].nf, "nwaves": best[i].nwaves, "k": best[i].k, "ctc": best[i].ctc,
               "impl": best[i].impl, "rows": best[i].rows} for i, s in enumerate(_STAGES)}
    out["ms"] = [float(ms[0]), float(ms[1])]
    if cache:
        d = _load()
        d[key] = out
        _store(d)
    return {**out, "cached": False}
