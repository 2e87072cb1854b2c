"""Autotuner for the launch shapes of the batched step.

The fused step of ``BatchedSimulation`` has six launch parameters (threads
per block and persistent blocks per compute unit of the fused operator
kernels, neighbour records in flight per lane, destination cells per block of
the list count pass, one persistent launch per evaluation with a grid
barrier or two launches, and the neighbour records staged in LDS or gathered).  None of them changes a result -- the per-particle sums
keep their order -- so the tuner may pick any of them by speed alone.

``autotune`` times real steps of the batch itself (the state is restored
afterwards) and walks the parameters one at a time (coordinate descent,
``rounds`` passes), keeping the fastest value of each.  The choice is cached
per device, dimension, precision and batch size in a JSON file, so a later
batch of the same shape starts from it without timing again.
"""
import json
import os
import time

import torch

from .batch import TUNE_FIELDS

SPACE = {
    "cells_per_block": (1, 2, 4, 8),
    "op_threads": (64, 128, 256),
    "op_blocks_per_cu": (0, 4, 8, 16),
    "nbr_tile": (1, 2, 4),
    "persistent": (0, 1),
    "staged": (0, 1),
}

DEFAULT = {"cells_per_block": 1, "op_threads": 256, "op_blocks_per_cu": 0, "nbr_tile": 2,
           "persistent": 0, "staged": 0}


def cache_path():
    base = os.environ.get("PYSPH_AMD_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "pysph_amd")
    return os.path.join(base, "autotune.json")


def signature(sim):
    """Cache key: device, dimension, precision, simulations and particles
    (rounded to 2 significant digits)."""
    name = torch.cuda.get_device_name(sim.device).replace(" ", "_")
    n = sim.n
    mag = 10 ** max(0, len(str(n)) - 2)
    return f"{name}/dim{sim.dim}/{sim.precision}/sims{sim.nsims}/n{round(n / mag) * mag}"


def _load_cache(path):
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _store_cache(path, key, entry):
    data = _load_cache(path)
    data[key] = entry
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + f".{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache


def time_steps(sim, steps, dt=1.0, warmup=1):
    """Mean ms per step of `steps` eager steps, timed with events on the
    batch's stream after `warmup` untimed steps."""
    for _ in range(warmup):
        sim.step(dt)
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(sim.stream)
    for _ in range(steps):
        sim.step(dt)
    end.record(sim.stream)
    end.synchronize()
    return start.elapsed_time(end) / steps


def autotune(sim, dt=1.0, steps=3, rounds=1, space=None, use_cache=True, log=None):
    """Pick the fastest launch shapes for `sim` and apply them.

    Returns {"tuning": {...}, "ms_per_step": best, "trials": [(config, ms)],
    "cached": bool}.  The simulation's state is the same afterwards as
    before (trial steps run on a snapshot that is restored)."""
    space = dict(SPACE if space is None else space)
    for k in space:
        if k not in TUNE_FIELDS:
            raise ValueError(f"unknown tuning field {k!r}")
    key = signature(sim)
    path = cache_path()
    if use_cache:
        hit = _load_cache(path).get(key)
        if hit and set(hit.get("tuning", {})) <= set(TUNE_FIELDS):
            sim.set_tuning(hit["tuning"])
            return {"tuning": sim.tuning(), "ms_per_step": hit.get("ms_per_step"),
                    "trials": [], "cached": True}
    snap = sim.snapshot()
    steps_done = sim.steps_done
    best = dict(DEFAULT)
    best.update({k: v for k, v in sim.tuning().items() if v})
    trials = []

    def trial(cfg):
        sim.restore(snap)
        sim.set_tuning(cfg)
        ms = time_steps(sim, steps, dt)
        sim.check()
        trials.append((dict(cfg), ms))
        if log:
            log(f"autotune {cfg} {ms:.3f} ms/step")
        return ms

    t0 = time.perf_counter()
    best_ms = trial(best)
    for _ in range(rounds):
        for field, values in space.items():
            for v in values:
                if v == best[field]:
                    continue
                cfg = dict(best, **{field: v})
                ms = trial(cfg)
                if ms < best_ms:
                    best, best_ms = cfg, ms
    sim.restore(snap)
    sim.set_tuning(best)
    sim.steps_done = steps_done
    sim.synchronize()
    if use_cache:
        _store_cache(path, key, {"tuning": best, "ms_per_step": best_ms,
                                 "tuned_s": time.perf_counter() - t0})
    return {"tuning": dict(best), "ms_per_step": best_ms, "trials": trials, "cached": False}
