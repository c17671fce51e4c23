"""Launch-shape autotuner: workgroup size x group count per (solver, graph).

Counterpart of the reference's WG-size sweep (pannotia/scripts/tune_wg_size.py
and lonestar_gpu_opencl/scripts/tune_wg_size.py, which time each workgroup
size and keep the fastest, with results tabulated in wg_size_data.py).  Here
the sweep also covers the number of launched groups, because on a 256-CU
part the cost of an inter-workgroup barrier grows with the number of groups
taking part: barrier-bound solvers want few groups, bandwidth-bound ones
want the residency bound.  Results are cached in a JSON file keyed by device
and graph signature.
"""
from __future__ import annotations

import json
import os
import time
from pathlib import Path
from typing import Callable, Iterable

from . import api

CACHE = Path(os.environ.get("GDB_TUNE_CACHE", Path(__file__).resolve().parent / "tune_cache.json"))

DEFAULT_WGS = (256, 512, 1024)
DEFAULT_GROUPS = (8, 16, 32, 64, 128, 256, 512, 0)  # 0 = residency bound


def _load() -> dict:
    try:
        return json.loads(CACHE.read_text())
    except Exception:
        return {}


def _store(d: dict) -> None:
    try:
        CACHE.write_text(json.dumps(d, indent=1, sort_keys=True))
    except OSError:
        pass


def candidates(wgs_list: Iterable[int] = DEFAULT_WGS, groups_list: Iterable[int] = DEFAULT_GROUPS):
    for w in wgs_list:
        bound = api.max_resident_groups(w, 0)
        seen = set()
        for g in groups_list:
            gg = bound if g == 0 else min(g, bound)
            if gg in seen:
                continue
            seen.add(gg)
            yield w, gg


def tune(key: str, run: Callable[[int, int], float], repeats: int = 2,
         wgs_list: Iterable[int] = DEFAULT_WGS, groups_list: Iterable[int] = DEFAULT_GROUPS,
         use_cache: bool = True) -> dict:
    """Times run(wgs, groups) -> ms for every candidate; returns the fastest.

    `key` identifies the solver and graph; cached results are reused.
    """
    cache = _load() if use_cache else {}
    if use_cache and key in cache:
        return cache[key]
    best = None
    trials = []
    for w, g in candidates(wgs_list, groups_list):
        ms = min(run(w, g) for _ in range(repeats))
        trials.append((w, g, ms))
        if best is None or ms < best["ms"]:
            best = {"wgs": w, "groups": g, "ms": ms}
    best["trials"] = trials
    best["tuned_at"] = time.time()
    if use_cache:
        cache[key] = best
        _store(cache)
    return best
