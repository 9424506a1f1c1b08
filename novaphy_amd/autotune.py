"""Autotuner for the launch tunables of a device context.

The tunables are the block size of the neighbour passes (64 / 128 / 256), the
XCD-aware block order (1: each of the 8 XCDs sweeps one contiguous run of
particles so the neighbour rows its blocks share stay in its own L2, or, in
steps whose computing particles fall unevenly over those runs, chunks dealt
to the XCDs in turn; C >= 2: always chunks of C blocks; 0: the hardware's
round-robin order), the skin of the per-step candidate lists
(a fraction of h: larger keeps the lists valid through more motion per step,
smaller visits fewer candidates) and persistent passes (at most that many
per-lane workgroups per neighbour pass, each looping over particle blocks;
0: one workgroup per particle block).  The cell size is not tunable: the reference's
candidate order is defined on cells of size h, and the GPU reproduces it
exactly.  The sort is a counting sort with a per-cell rank, so there is no
radix to choose.

Results are bitwise independent of every tunable, so the tuner times real
steps of the workload it is tuning (typically the warm-up steps) and the
simulation state afterwards is exactly what it would have been untuned.
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict, Iterable, Optional, Tuple

import numpy as np

Candidate = Tuple[int, int, float, int]
# (the skin fraction matters only with skin lists on, npx_set_list_policy 1)
DEFAULT_CANDIDATES: Tuple[Candidate, ...] = (
    (128, 1, 0.1, 0), (64, 1, 0.1, 0), (256, 1, 0.1, 0), (128, 64, 0.1, 0), (64, 128, 0.1, 0),
    (128, 0, 0.1, 0), (128, 1, 0.1, 4096))


def _norm(c) -> Candidate:
    return (int(c[0]), int(c[1]), float(c[2]), int(c[3]) if len(c) > 3 else 0)


def _timed_steps(target, nsteps: int, step_args) -> float:
    """Median milliseconds per step over nsteps steps."""
    if hasattr(target, "step_timed"):
        return float(np.median(target.step_timed(nsteps)))
    dt, g, m = step_args
    times = []
    for _ in range(nsteps):
        target.synchronize()
        t0 = time.perf_counter()
        target.step(dt, g, m, 1)
        target.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(times))


def autotune(target, candidates: Iterable[Candidate] = DEFAULT_CANDIDATES, steps_per_candidate: int = 2,
             step_args=None, cache_path: Optional[str] = None, key: Optional[str] = None, rounds: int = 3):
    """Times each (block_size, xcd_remap, skin_fraction, persistent) candidate on ``target`` (a capi.Batch, or
    a capi.Context with ``step_args`` = (dt, gravity, mass)), applies the
    fastest and returns (best, {candidate: ms_per_step}).  The simulation
    moves on while it is timed, and the cost per step moves with it, so the
    candidates take turns: ``rounds`` rounds, each timing every candidate for
    ``steps_per_candidate`` steps (after one untimed step that captures its
    launch shape), and a candidate's time is its median over the rounds.
    With ``cache_path`` and ``key`` the choice is stored / reused."""
    if cache_path and key and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache:
            best = _norm(cache[key]["best"])
            target.set_tuning(*best)
            return best, {}
    cands = [_norm(c) for c in candidates]
    times: Dict[Candidate, list] = {c: [] for c in cands}
    for _ in range(max(1, rounds)):
        for c in cands:
            target.set_tuning(*c)
            _timed_steps(target, 1, step_args)           # graph capture / warm caches
            times[c].append(_timed_steps(target, steps_per_candidate, step_args))
    results: Dict[Candidate, float] = {c: float(np.median(v)) for c, v in times.items()}
    best = min(results, key=results.get)
    target.set_tuning(*best)
    if cache_path and key:
        cache = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                cache = json.load(f)
        cache[key] = {"best": list(best), "ms": {"/".join(str(x) for x in c): v for c, v in results.items()}}
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return best, results
