"""Launch autotuner of the batched master-pattern loop.

For a plan (or the plans of an energy series) it times complete runs of the path over a grid of
launch settings -- the multiple the padded matrix sizes are rounded to (which sets the bucket
edges) and the work space of one chunk (which sets how many matrices each GEMM launch holds) --
and keeps the fastest. The choice is cached per GPU (device name, compute units) and workload
signature (precision, strong-beam histogram by 32, depth count, direction count) in a JSON file,
so a later process on the same GPU applies it without timing:

    $CTEM_TUNE_CACHE, or ~/.cache/ctemsoft_amd/launch_tuning.json

The chunk size does not change any result. The bucket multiple pads some matrices further
(zero rows and columns, which leave the products unchanged), but a matrix padded past 224 takes
the mixed form's four-depth recursion instead of the on-chip one, so mixed results can change
within that form's error bound; FP64 results agree to rounding.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from . import ebsd

CANDIDATES = [(32, 24.0), (32, 64.0), (32, 8.0), (64, 24.0), (64, 64.0)]


def cache_path() -> str:
    return os.environ.get("CTEM_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "ctemsoft_amd", "launch_tuning.json")


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db: dict) -> None:
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: the choice is still applied in this process


def signature(ctx, plans, precision) -> str:
    name, cus = ctx.device_info()
    hist = np.zeros(64, np.int64)
    nz = set()
    for p in plans:
        ns = p.nns if getattr(p, "nns", None) is not None else p.select()[0]
        np.add.at(hist, np.minimum(np.asarray(ns) // 32, 63), 1)
        nz.add(len(p.lam))
    h = hashlib.sha1(hist.tobytes() + repr(sorted(nz)).encode()).hexdigest()[:16]
    return f"{name}|{cus}|{precision}|{len(plans)}|{h}"


def apply(plans, setting) -> None:
    for p in plans:
        p.set_tuning(*setting)


def _time(plans, precision, reps) -> float:
    for p in plans:  # warm: selection, buckets and squaring counts of this setting
        p.run(precision)
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        for p in plans:
            p.run(precision)
        best = min(best, time.perf_counter() - t0)
    return best


def tune(ctx, plans, precision="mixed", candidates=CANDIDATES, reps=2, use_cache=True, log=None):
    """Pick the fastest launch setting for these plans and apply it; returns
    (setting, {setting: seconds} or None when taken from the cache)."""
    key = signature(ctx, plans, precision)
    db = _load() if use_cache else {}
    if key in db:
        setting = tuple(db[key]["setting"])
        apply(plans, setting)
        if log:
            log(f"autotune: cached {setting} for {key}")
        return setting, None
    times = {}
    for cand in candidates:
        try:
            apply(plans, cand)
            times[cand] = _time(plans, precision, reps)
        except ebsd.CtemError as e:  # e.g. a chunk larger than the free device memory
            if log:
                log(f"autotune: {cand} not usable ({e})")
            continue
        if log:
            log(f"autotune: bucket multiple {cand[0]}, chunk {cand[1]:g} GiB: {times[cand] * 1e3:.1f} ms")
    if not times:
        raise RuntimeError("autotune: no candidate setting ran")
    setting = min(times, key=times.get)
    apply(plans, setting)
    db = _load()
    db[key] = {"setting": list(setting), "seconds": {f"{a},{b:g}": t for (a, b), t in times.items()}}
    _store(db)
    if log:
        log(f"autotune: chose bucket multiple {setting[0]}, chunk {setting[1]:g} GiB")
    return setting, times
