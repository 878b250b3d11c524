"""Autotuner for the batched permanent kernel.

For a batch shape (order n, batch size) it times every kernel variant (row 1
of the Gray code in VGPRs or in LDS; row flips as a separate phase or fused
into the product) and lane-chunk length that the plan API accepts, on the
current GPU, and keeps the fastest; for a mixed-order batch it tunes the
chunk length; for a SLOS plan it times every applicable layer kernel.
Results are cached in a JSON file keyed by device name and shape, so tuning
runs once per machine.
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
    """A PermanentPlan for `orders`: variant and chunk tuned when every order is
    the same and >= 9, chunk tuned for a mixed batch with orders >= 9."""
    orders = np.asarray(orders, dtype=np.int32).ravel()
    if orders.size and orders.min() == orders.max() and orders[0] >= 9:
        t = tune_permanent(int(orders[0]), batch_hint or int(orders.size))
        return PermanentPlan(orders, max_chunk_log2=t["chunk_log2"], variant=t["variant"])
    if orders.size and orders.max() >= 9:
        t = tune_permanent_mixed(orders)
        return PermanentPlan(orders, max_chunk_log2=t["chunk_log2"])
    return PermanentPlan(orders)


def _orders_key(orders) -> str:
    import hashlib
    o = np.asarray(orders, dtype=np.int32).ravel()
    return hashlib.sha1(o.tobytes()).hexdigest()[:16]


def tune_permanent_mixed(orders, reps: int = 3, use_cache: bool = True) -> dict:
    """Fastest chunk length for a mixed-order batch (one persistent launch).
    Returns {"chunk_log2", "ms", "trials": [...]}."""
    import torch
    orders = np.asarray(orders, dtype=np.int32).ravel()
    cache = _load() if use_cache else {}
    key = f"{torch.cuda.get_device_name(0)}|perm-mixed|{_orders_key(orders)}"
    if key in cache:
        return cache[key]
    rng = np.random.RandomState(int(orders.sum()) % (2 ** 31))
    flat = np.concatenate([((rng.randn(k, k) + 1j * rng.randn(k, k)) / np.sqrt(2 * max(k, 1))).ravel()
                           for k in orders]) if orders.size else np.zeros(1, complex)
    d = torch.from_numpy(flat).cuda()
    out = torch.empty(max(1, orders.size), dtype=torch.complex128, device="cuda")
    trials = []
    for c in CHUNKS:
        plan = PermanentPlan(orders, max_chunk_log2=c)
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
            trials.append({"chunk_log2": c, "ms": a.elapsed_time(b) / reps})
        finally:
            plan.close()
    best = min(trials, key=lambda t: t["ms"])
    res = {"chunk_log2": best["chunk_log2"], "ms": best["ms"], "trials": trials}
    cache[key] = res
    if use_cache:
        _save(cache)
    return res


SLOS_KERNELS = (1, 4, 6, 8)
SLOS_BLOCK_WIDTHS = (6, 8, 10, 12)


def tune_slos(in_occ, reps: int = 3, use_cache: bool = True) -> dict:
    """Fastest SLOS layer kernel for an input state (SlosPlan.set_kernel ids;
    kernel 8 at each suffix width of SLOS_BLOCK_WIDTHS below m), timed on a
    Haar unitary.  Returns {"kernel", "width", "ms", "trials": [...]}."""
    import torch
    from .slos import PROBS, SlosPlan
    from .utils.workloads import haar_unitary
    in_occ = [int(v) for v in in_occ]
    cache = _load() if use_cache else {}
    key = f"{torch.cuda.get_device_name(0)}|slos|{in_occ}"
    if key in cache and "width" in cache[key]:
        return cache[key]
    plan = SlosPlan(in_occ)
    dU = torch.from_numpy(haar_unitary(plan.m, seed=1)).cuda()
    out = torch.empty(max(1, plan.count), dtype=torch.float64, device="cuda")
    trials = []

    def timed():
        plan.run(dU, PROBS, out)
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            plan.run(dU, PROBS, out)
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    try:
        for k in SLOS_KERNELS:
            try:
                plan.set_kernel(k)
            except ValueError:
                continue   # not applicable to this plan's layer sizes
            widths = [0]
            if k == 8:
                widths = [w for w in SLOS_BLOCK_WIDTHS if w < plan.m] or [0]
            for w in widths:
                if k == 8:
                    plan.set_block_width(w)
                trials.append({"kernel": k, "width": plan.block_width if k == 8 else 0, "ms": timed()})
    finally:
        plan.close()
    best = min(trials, key=lambda t: t["ms"])
    res = {"kernel": best["kernel"], "width": best["width"], "ms": best["ms"], "trials": trials}
    cache[key] = res
    if use_cache:
        _save(cache)
    return res


def tuned_slos_plan(in_occ):
    """A SlosPlan running the layer kernel the tuner found fastest."""
    from .slos import SlosPlan
    plan = SlosPlan(in_occ)
    best = tune_slos(in_occ)
    plan.set_kernel(best["kernel"])
    if best["kernel"] == 8:
        plan.set_block_width(best["width"])
    return plan
