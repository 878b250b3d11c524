#!/usr/bin/env python3
"""Benchmark of the MI355X simulation core.

Headline (one JSON line on stdout): batched complex permanents at n = 20 --
4096 n x n submatrices of a seeded 64 x 64 Haar unitary, one persistent launch
per step, inputs resident in HBM -- in permanents per second, with the FP64
roofline of the permanent kernel and the CPU oracle timed on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one pass of the hot path over the batch.  Only one GPU is used;
--gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6      # MI355X dense FP64 vector, FMA = 2 flops
HBM_PEAK_GBS = 8000.0


def glynn_flops(n: int) -> float:
    """Algorithmic flops of one Glynn Gray-code permanent: 2^(n-1) terms, each
    n complex adds (column sums) + (n-1) complex multiplies + 1 complex add."""
    return float(2 ** (n - 1) * (8 * n - 4))


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def time_plan(plan, d_mats, out, steps: int, warmup: int):
    import torch
    for _ in range(warmup):
        plan(d_mats, out)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        plan(d_mats, out)
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps   # ms per step


def cpu_baseline_perm(mats, budget_s: float = 8.0):
    """Oracle (chunked double Glynn, OpenMP) on a growing sample of the same
    matrices until about budget_s of wall time has been spent."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    k, spent, done = 16, 0.0, 0
    while True:
        sample = mats[:k]
        t0 = time.perf_counter()
        oracle.perm_batch(sample, precise=False, threads=threads)
        dt = time.perf_counter() - t0
        spent += dt
        done = k
        if dt >= budget_s or k >= len(mats) or spent > 2 * budget_s:
            break
        k = min(len(mats), k * 2)
    return {"value": done / dt, "unit": "perms/s", "cores": threads, "kind": "port",
            "sample": f"{done} of the n=20 matrices, oracle chunked-double Glynn (oracle/pcvl_oracle.c), "
                      f"{threads} POSIX threads, {dt:.2f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--order", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-order batch")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE={world}); "
              f"the batch is embarrassingly parallel and would shard by matrices, see DESIGN.md",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from perceval_amd.permanent import PermanentPlan, pack
    from perceval_amd.utils.workloads import submatrix_batch

    n = args.order
    log(f"[bench] building {args.batch} x n={n} submatrices of a 64x64 Haar unitary")
    _, orders, mats = submatrix_batch(args.batch, n, n, seed=2024)
    o, flat = pack(mats)
    d_mats = torch.from_numpy(flat).cuda()
    out = torch.empty(len(o), dtype=torch.complex128, device="cuda")
    plan = PermanentPlan(o)
    log(f"[bench] plan {plan.info()}; warmup {args.warmup}, steps {args.steps}")
    ms = time_plan(plan, d_mats, out, args.steps, args.warmup)
    perms_per_s = len(o) / (ms * 1e-3)
    flops = len(o) * glynn_flops(n)
    achieved = flops / (ms * 1e-3) / 1e12
    log(f"[bench] {ms:.3f} ms/step, {perms_per_s:.1f} perms/s, {achieved:.2f} TFLOP/s")

    extras = {}
    if not args.no_mixed:
        _, morders, mmats = submatrix_batch(4096, 12, 24, seed=2025)
        mo, mflat = pack(mmats)
        md = torch.from_numpy(mflat).cuda()
        mout = torch.empty(len(mo), dtype=torch.complex128, device="cuda")
        mplan = PermanentPlan(mo)
        mms = time_plan(mplan, md, mout, max(3, args.steps // 4), 1)
        mflops = sum(glynn_flops(int(k)) for k in mo)
        extras["mixed_n12_24"] = {"perms_per_s": len(mo) / (mms * 1e-3), "ms_per_step": mms,
                                  "tflops": mflops / (mms * 1e-3) / 1e12,
                                  "frac_fp64_peak": mflops / (mms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
        log(f"[bench] mixed n=12..24: {mms:.3f} ms/step, {extras['mixed_n12_24']['tflops']:.2f} TFLOP/s")
        mplan.close()

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU oracle baseline")
        cpu = cpu_baseline_perm(mats)
        log(f"[bench] CPU {cpu['value']:.1f} perms/s on {cpu['cores']} threads")

    plan.close()
    line = {
        "metric": "perms/s at n=20 (%FP64 peak); samples/s m=60 n=20 learned-MIS vs plain CC",
        "value": perms_per_s,
        "unit": "perms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (fp64)",
        "data": "synthetic: seeded Haar 64x64 unitary, n x n submatrices (rows/cols without replacement)",
        "config": {"workload": f"batched Glynn permanents, {len(o)} matrices of order {n}",
                   "batch": len(o), "order": n, "parallelism": "single GPU"},
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None},
        "cpu_baseline": cpu,
    }
    line.update(extras)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
