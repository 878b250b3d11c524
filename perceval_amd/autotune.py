"""Autotuner for the batched permanent kernel.

For a batch shape (order n, batch size) it times every kernel variant (row 1
of the Gray code in VGPRs or in LDS; row flips as a separate phase or fused
into the product) and lane-chunk length that the plan API
accepts, on the current GPU, and keeps the fastest.  Results are cached in a
JSON file keyed by device name, order and batch size, so tuning runs once per
machine.
"""
from __future__ import annotations

import json
import os

import numpy as np

from .permanent import PermanentPlan

_CACHE_PATH = os.environ.get("PCVL_AUTOTUNE_CACHE",
                             os.path.join(os.path.expanduser("~"), ".cache", "perceval_amd", "autotune.json"))
VARIANTS = (1, 2, 3, 4)
CHUNKS = (8, 10, 11, 12, 13, 14)


def _load() -> dict:
    try:
        with open(_CACHE_PATH) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(cache: dict):
    try:
        os.makedirs(os.path.dirname(_CACHE_PATH), exist_ok=True)
        with open(_CACHE_PATH, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    except OSError:
        pass


def _key(n: int, batch: int) -> str:
    import torch
    return f"{torch.cuda.get_device_name(0)}|perm|n={n}|batch={batch}"


def tune_permanent(n: int, batch: int = 4096, reps: int = 5, use_cache: bool = True) -> dict:
    """Fastest (variant, chunk_log2) for `batch` permanents of order n >= 9.
    Returns {"variant", "chunk_log2", "ms", "trials": [...]}."""
    import torch
    cache = _load() if use_cache else {}
    key = _key(n, batch)
    if key in cache:
        return cache[key]
    rng = np.random.RandomState(n)
    mats = (rng.randn(batch, n, n) + 1j * rng.randn(batch, n, n)) / np.sqrt(2 * n)
    d = torch.from_numpy(mats.reshape(-1)).cuda()
    out = torch.empty(batch, dtype=torch.complex128, device="cuda")
    trials = []
    for variant in VARIANTS:
        for c in CHUNKS:
            if c > n - 7:
                continue
            plan = PermanentPlan(np.full(batch, n, dtype=np.int32), max_chunk_log2=c, variant=variant)
            try:
                plan(d, out)
                torch.cuda.synchronize()
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record()
                for _ in range(reps):
                    plan(d, out)
                b.record()
                torch.cuda.synchronize()
                trials.append({"variant": variant, "chunk_log2": c, "ms": a.elapsed_time(b) / reps})
            finally:
                plan.close()
    best = min(trials, key=lambda t: t["ms"])
    res = {"variant": best["variant"], "chunk_log2": best["chunk_log2"], "ms": best["ms"], "trials": trials}
    cache[key] = res
    if use_cache:
        _save(cache)
    return res


def tuned_plan(orders, batch_hint: int | None = None) -> PermanentPlan:
    """A PermanentPlan for `orders`, tuned when every order is the same and >= 9."""
    orders = np.asarray(orders, dtype=np.int32).ravel()
    if orders.size and orders.min() == orders.max() and orders[0] >= 9:
        t = tune_permanent(int(orders[0]), batch_hint or int(orders.size))
        return PermanentPlan(orders, max_chunk_log2=t["chunk_log2"], variant=t["variant"])
    return PermanentPlan(orders)
