"""Tile-shape autotuner for the split-K Biot-Savart kernel.

Candidates are (points per lane, work items per CU).  Each is timed with HIP
events over a few field evaluations (split-K Biot-Savart pass + fixed-order
reduction, the per-stage cost of tracking) on the real segment set and
points; the fastest wins.  Results are cached per (device, precision, design
sizes) in memory and in a JSON file (``$CCTAMD_TUNE_CACHE`` or
``~/.cache/cctamd/tune.json``).
"""
from __future__ import annotations

import json
import os
from typing import Dict, Optional, Sequence

import torch

from . import engine

PPL = (1, 2, 4)
ITEMS_PER_CU = (2, 4, 8, 16, 32)  # 32: 2-3 % faster on the 64-design population's uneven designs
_MEM: Dict[str, dict] = {}


def _cache_path() -> str:
    return os.environ.get("CCTAMD_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "cctamd",
                                                             "tune.json"))


def _load() -> dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key: str, val: dict) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = _load()
        data[key] = val
        with open(path, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
    except OSError:
        pass


# Bumped whenever the kernels or the candidate set change, so that cached choices of an
# older library are not reused.
KERNEL_SET = 10
TIE = 0.015


def _key(model: "engine.DeviceModel", npt: Sequence[int], precision: int) -> str:
    props = torch.cuda.get_device_properties(engine.device())
    return f"v{KERNEL_SET}|{props.name}|cu{props.multi_processor_count}|fp{precision}|" \
           f"seg{list(map(int, model.design_nseg))}|pt{list(map(int, npt))}"


def pick(cost: Dict[tuple, float], tie: float = TIE) -> tuple:
    """((ppl, items per CU), ms) from measured costs.  Shapes within ``tie`` of the fastest
    are a tie (timing noise is about 1 %): the one with the most work items per CU wins, its
    finer chunks balancing better as the particles move along the track."""
    fastest = min(cost.values())
    return max(((k, v) for k, v in cost.items() if v <= fastest * (1.0 + tie)), key=lambda kv: (kv[0][1], -kv[1]))


def tune(model: "engine.DeviceModel", npt: Sequence[int], precision: int, points,
         launches: int = 4, use_cache: bool = True, rounds: int = 2) -> dict:
    """Best {"ppl", "target_items"} for this model, point layout and precision.

    ``points`` is one device tensor (3, P) or a list of them (e.g. states sampled along a
    track: the FP32 kernel's cost depends on how close points come to the windings); a
    candidate's cost is its time summed over the point sets, best of ``rounds``."""
    key = _key(model, npt, precision)
    if use_cache:
        if key in _MEM:
            return _MEM[key]
        disk = _load().get(key)
        if disk:
            _MEM[key] = disk
            return disk
    sets = list(points) if isinstance(points, (list, tuple)) else [points]
    cus = engine.cu_count()
    ppls = PPL + ((-2, -4) if precision != 32 else ())
    cost: Dict[tuple, float] = {}
    for _ in range(max(1, rounds)):
        for ppl in ppls:
            for per_cu in ITEMS_PER_CU:
                plan = model.plan(npt, ppl, per_cu * cus)
                _, partial = model.workspace(plan)
                out = torch.empty_like(sets[0])
                model.field_tensors(plan, precision, sets[0], partial, out)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for pts in sets:
                    for _ in range(launches):
                        model.field_tensors(plan, precision, pts, partial, out)
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / (launches * len(sets))
                k = (ppl, per_cu)
                cost[k] = min(cost.get(k, ms), ms)
    (ppl, per_cu), ms = pick(cost)
    best = {"ppl": ppl, "target_items": per_cu * cus, "ms": ms}
    _MEM[key] = best
    if use_cache:
        _store(key, best)
    return best
