#!/usr/bin/env python3
"""Benchmark of the MI355X simulation core.

Prints one JSON line.  Headline (`value`): batched complex permanents at
n = 20 -- 4096 n x n submatrices (rows and columns drawn without replacement)
of a seeded 64 x 64 Haar unitary, one persistent launch per step, inputs
resident in HBM -- in permanents per second, with the FP64 roofline of the
permanent kernel and the CPU oracle timed on this host.  The other measured
configurations are reported beside it:

  mixed_n12_24  4096 permanents, n uniform in 12..24, one launch
  slos_m24_n10  SLOS all_prob, Haar 24x24, 10 photons, 92,561,040 outputs
                (achieved algorithmic HBM GB/s); masked plans on the same input;
                m = 32, n = 10 (1,121,099,408 outputs)
  slos_tree     37 inputs (an 8-photon state and its 1- and 2-photon-loss
                variants, m = 20) through one shared SLOS tree vs one plan each
  gradient      adjoint phase gradients, Clements 12 modes (132 phases),
                6 photons, 256 parameter sets, 32 targets
  sampling      m = 60, n = 20: plain GPU Clifford-Clifford samples/s, the
                learned Metropolis independence sampler's samples/s, and the
                1- and 2-mode marginal TVD between the two sample sets
  sampling_multi  2,000 samples from each of 37 inputs (m = 20, an 8-photon
                state and its loss variants) in one batched call
  gurvits       NaiveApprox: Gurvits estimates, 64 order-30 matrices x 10^6
                samples (samples/s)
  copy          1 GiB device copy: the HBM rate a copy attains, and the
                reference that scales FETCH_SIZE / WRITE_SIZE readings
  cache         (only with --sections cache) read rate against working-set
                size, and the XCD of each workgroup of a 2,048-workgroup grid
  cc            (only with --sections cc) plain Clifford-Clifford alone,
                m = 60, n = 20
The perm section also times a pure-FMA FP64 probe, the vector rate the chip
sustains at the clock it holds under FP64 load.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--sections ...]

A step is one pass of the permanent hot path over the batch.  Only one GPU
is used; --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import statistics
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6      # MI355X dense FP64 vector, FMA = 2 flops
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the permanent kernel on the headline workload (4096
# matrices of order 20), measured with rocprofv3 --pmc on MI355X: FETCH_SIZE
# 30,798 KiB x 2.0 + WRITE_SIZE 1,313 KiB x 1.0, the factors that map a plain
# 1 GiB device copy onto 1 GiB read and 1 GiB written.  bench.py cannot read
# hardware counters itself, so the measured value is carried here.
PERM_N20_B4096_HBM_BYTES = (30798 * 2 + 1313) * 1024
SECTIONS = ("perm", "mixed", "slos", "slos_tree", "gradient", "sampling", "sampling_multi", "gurvits", "copy")


def glynn_flops(n: int) -> float:
    """Algorithmic flops of one Glynn Gray-code permanent: 2^(n-1) terms, each
    n complex adds (column sums) + (n-1) complex multiplies + 1 complex add."""
    return float(2 ** (n - 1) * (8 * n - 4))


def slos_bytes(m: int, n: int) -> float:
    """Algorithmic HBM bytes of fused SLOS all_prob: every layer k < n is
    written once and read once as parent (16 B complex), the last layer reads
    layer n-1 and writes 8-byte probabilities."""
    N = [math.comb(m + k - 1, k) for k in range(n + 1)]
    b = sum(16.0 * (N[k - 1] + N[k]) for k in range(1, n))
    return b + 16.0 * N[n - 1] + 8.0 * N[n]


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def cpu_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)


def cpu_model() -> str:
    """The host CPU's model name (first processor of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class Timer:
    """HIP events on torch's current stream, which is the stream every
    entry point of this library is launched on."""

    def __init__(self):
        import torch
        self.a = torch.cuda.Event(enable_timing=True)
        self.b = torch.cuda.Event(enable_timing=True)

    def __enter__(self):
        import torch
        torch.cuda.synchronize()
        self.a.record()
        return self

    def __exit__(self, *exc):
        import torch
        self.b.record()
        torch.cuda.synchronize()
        self.ms = self.a.elapsed_time(self.b)


def time_plan(plan, d_mats, out, steps: int, warmup: int):
    for _ in range(warmup):
        plan(d_mats, out)
    with Timer() as t:
        for _ in range(steps):
            plan(d_mats, out)
    return t.ms / steps


# ------------------------------------------------------------------ sections

def fp64_probe_tflops() -> float:
    """Attainable vector FP64 rate: pure FMA chains (pcvl_probe_fp64_fma),
    8 waves per SIMD, timed with HIP events on the launch stream."""
    import ctypes
    import torch
    from perceval_amd import _native
    lib = _native.load()
    blocks = torch.cuda.get_device_properties(0).multi_processor_count * 8
    out = torch.empty(blocks * 256, dtype=torch.float64, device="cuda")
    flops = ctypes.c_int64()
    iters = 100000

    def launch():
        _native.check(lib.pcvl_probe_fp64_fma(blocks, iters, _native.ptr(out), ctypes.byref(flops),
                                              _native.stream_handle()), "probe")
    launch()
    with Timer() as t:
        for _ in range(3):
            launch()
    if not bool(torch.isfinite(out).all()):
        raise RuntimeError("FP64 probe produced non-finite values")
    return flops.value / (t.ms / 3 * 1e-3) / 1e12


def fp64_mix_probe() -> dict:
    """FP64 matrix cores beside the vector ALU (pcvl_probe_fp64_mix): TFLOP/s
    of v_mfma_f64_16x16x4f64 alone, v_fma_f64 alone, and both in one loop."""
    import ctypes
    import torch
    from perceval_amd import _native
    lib = _native.load()
    blocks = torch.cuda.get_device_properties(0).multi_processor_count * 8
    out = torch.empty(blocks * 256, dtype=torch.float64, device="cuda")
    fm, fv = ctypes.c_int64(), ctypes.c_int64()
    iters = 20000
    res = {}
    for mode, name in ((0, "mfma"), (1, "valu"), (2, "both")):
        def launch():
            _native.check(lib.pcvl_probe_fp64_mix(blocks, iters, mode, _native.ptr(out), ctypes.byref(fm),
                                                  ctypes.byref(fv), _native.stream_handle()), "mix probe")
        launch()
        with Timer() as t:
            for _ in range(3):
                launch()
        sec = t.ms / 3 * 1e-3
        res[name] = {"ms": t.ms / 3, "mfma_tflops": fm.value / sec / 1e12, "valu_tflops": fv.value / sec / 1e12}
    if not bool(torch.isfinite(out).all()):
        raise RuntimeError("FP64 mix probe produced non-finite values")
    b = res["both"]
    res["combined_tflops"] = b["mfma_tflops"] + b["valu_tflops"]
    return res


def bench_copy(args, res):
    """Plain 1 GiB device-to-device copy, three times: the reference point that
    scales FETCH_SIZE / WRITE_SIZE counter readings to bytes, and the HBM rate
    a copy attains."""
    import torch
    nbytes = 1 << 30
    x = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    y.copy_(x)
    with Timer() as t:
        for _ in range(3):
            y.copy_(x)
    ms = t.ms / 3
    res["copy_1GiB"] = {"ms": ms, "GB_per_s": 2 * nbytes / (ms * 1e-3) / 1e9}
    log(f"[bench] copy 1 GiB: {ms:.3f} ms, {2 * nbytes / (ms * 1e-3) / 1e9:.0f} GB/s (read + write)")


def bench_cache(args, res):
    """Probes behind the SLOS kernel's design (not run by default): the read
    rate of a buffer read over and over against its size (where the last-level
    caches stop holding a working set), and the XCD that runs each workgroup
    of a 2,048-workgroup grid (the persistent launches assume blockIdx % 8)."""
    import torch
    from perceval_amd import _native
    lib = _native.load()
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    blocks = cus * 16
    sink = torch.empty(blocks * 256, dtype=torch.float64, device="cuda")
    rates = {}
    for mb in (2, 16, 32, 64, 80, 96, 128, 256, 512):
        nbytes = mb << 20
        buf = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        reps = max(1, (8 << 30) // nbytes)

        def run():
            _native.check(lib.pcvl_probe_read(_native.ptr(buf), nbytes, reps, blocks, _native.ptr(sink),
                                              _native.stream_handle()), "read probe")
        run()
        with Timer() as t:
            run()
        rates[str(mb)] = nbytes * reps / (t.ms * 1e-3) / 1e12
        del buf
    xcc = torch.empty(2048, dtype=torch.int32, device="cuda")
    _native.check(lib.pcvl_probe_xcc(2048, _native.ptr(xcc), _native.stream_handle()), "xcc probe")
    got = xcc.cpu().numpy()
    match = int((got == (np.arange(2048) % 8)).sum())
    res["cache_probe"] = {"read_TBps_by_MB": rates, "xcd_is_blockidx_mod_8": f"{match} of 2048 workgroups"}
    log("[bench] read rate by working set (MB: TB/s): " + ", ".join(f"{k}: {v:.1f}" for k, v in rates.items()) +
        f"; XCD = blockIdx % 8 for {match} of 2048 workgroups")


def bench_cc(args, res):
    """Plain GPU Clifford-Clifford alone (not run by default; the sampling
    section runs it too, with the learned sampler): m = 60, n = 20,
    --samples samples, for profiling the sampler by itself."""
    import torch
    from perceval_amd.sampling import cc_samples_device
    from perceval_amd.utils.workloads import haar_unitary
    m, n, N = 60, 20, args.samples
    dU = torch.from_numpy(haar_unitary(m, seed=60)).cuda()
    inp = [1] * n + [0] * (m - n)
    cc_samples_device(dU, inp, 256, seed=99)
    with Timer() as t:
        cc_samples_device(dU, inp, N, seed=1)
    res["cc_m60_n20"] = {"samples": N, "ms": t.ms, "samples_per_s": N / (t.ms * 1e-3)}
    log(f"[bench] plain CC alone: {N} samples in {t.ms:.1f} ms = {N / (t.ms * 1e-3):.0f} samples/s")


def bench_gurvits(args, res):
    """NaiveApprox: Gurvits estimates of 64 order-30 submatrices of a 60-mode
    Haar unitary (the size of the reference's NaiveApprox example), 10^6
    samples each, matrices resident in HBM."""
    import ctypes
    import torch
    from perceval_amd import _native
    from perceval_amd.permanent import pack
    from perceval_amd.utils.workloads import haar_unitary
    lib = _native.load()
    m, n, B, iters = 60, 30, 64, 1_000_000
    U = haar_unitary(m, seed=61)
    rng = np.random.RandomState(3)
    mats = [np.ascontiguousarray(U[np.sort(rng.choice(m, n, replace=False))][:, :n]) for _ in range(B)]
    orders, flat = pack(mats)
    offsets = (np.arange(B, dtype=np.int64) * n * n)
    d = torch.from_numpy(flat).cuda()
    est = torch.empty(B, dtype=torch.complex128, device="cuda")
    err = torch.empty(B, dtype=torch.float64, device="cuda")

    def run():
        _native.check(lib.pcvl_gurvits_batch(_native.ptr(d), orders.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                             offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), B, iters, 5, 0,
                                             _native.ptr(est), _native.ptr(err), _native.stream_handle()), "gurvits")
    run()
    with Timer() as t:
        run()
    rate = B * iters / (t.ms * 1e-3)
    flops = B * iters * (4 * n * n + 6 * (n - 1)) / (t.ms * 1e-3) / 1e12
    res["gurvits_m60_n30"] = {"ms": t.ms, "matrices": B, "samples_per_matrix": iters, "samples_per_s": rate,
                              "tflops": flops, "rel_err_bound_mean": float((err / est.abs()).mean().item())}
    log(f"[bench] Gurvits n=30: {B} x {iters} samples in {t.ms:.1f} ms = {rate:.3e} samples/s ({flops:.1f} TFLOP/s)")
    if not args.no_cpu:
        import oracle
        k = 20000
        t0 = time.perf_counter()
        oracle.gurvits(mats[0], k, seed=5)
        dt = time.perf_counter() - t0
        res["gurvits_cpu"] = {"value": k / dt, "unit": "samples/s", "cores": 1, "kind": "port", "cpu_model": cpu_model(),
                              "sample": f"{k} samples of one order-30 matrix, oracle (oracle/pcvl_oracle.c), 1 thread"}
        log(f"[bench] Gurvits CPU: {k / dt:.0f} samples/s on 1 thread")


def bench_perm(args, res):
    import torch
    from perceval_amd.permanent import PermanentPlan, pack
    from perceval_amd.utils.workloads import submatrix_batch
    n = args.order
    log(f"[bench] perm: {args.batch} x n={n} submatrices of a 64x64 Haar unitary")
    _, orders, mats = submatrix_batch(args.batch, n, n, seed=2024)
    o, flat = pack(mats)
    d_mats = torch.from_numpy(flat).cuda()
    out = torch.empty(len(o), dtype=torch.complex128, device="cuda")
    plan = PermanentPlan(o, max_chunk_log2=args.chunk_log2, variant=args.variant)
    log(f"[bench] plan {plan.info()} variant {args.variant}; warmup {args.warmup}, steps {args.steps}")
    ms = time_plan(plan, d_mats, out, args.steps, args.warmup)
    plan.close()
    achieved = len(o) * glynn_flops(n) / (ms * 1e-3) / 1e12
    res["perm"] = {"ms": ms, "perms_per_s": len(o) / (ms * 1e-3), "tflops": achieved}
    # the timed launch's own output, checked on a spread sample against the
    # oracle's long-double Glynn (outside the timed region)
    import oracle
    pick = np.unique(np.linspace(0, len(mats) - 1, 24).astype(int))
    got = out.cpu().numpy()[pick]
    ref = oracle.perm_batch([mats[i] for i in pick], precise=True, threads=cpu_threads())
    rel = float(np.max(np.abs(got - ref) / np.abs(ref)))
    res["perm"]["checked"] = int(len(pick))
    res["perm"]["max_rel_err"] = rel
    log(f"[bench] perm: {len(pick)} of the timed permanents against the oracle: max rel err {rel:.2e}")
    if not rel <= 1e-10:
        raise RuntimeError(f"timed permanents differ from the oracle (max rel err {rel:.3e})")
    log(f"[bench] perm: {ms:.3f} ms/step, {len(o) / (ms * 1e-3):.1f} perms/s, {achieved:.2f} TFLOP/s "
        f"({100 * achieved / FP64_PEAK_TFLOPS:.1f}% of FP64 peak)")
    probe = fp64_probe_tflops()
    res["perm"]["fp64_fma_probe_tflops"] = probe
    res["perm"]["frac_of_probe"] = achieved / probe
    log(f"[bench] FP64 FMA probe: {probe:.2f} TFLOP/s attainable; permanent kernel at {100 * achieved / probe:.1f}% "
        f"of it (instruction ceiling of Glynn at n={n}: {100 * (8 * n - 4) / (2 * (6 * n - 4)):.1f}%)")
    mix = fp64_mix_probe()
    res["perm"]["fp64_matrix_core_probe"] = mix
    log(f"[bench] FP64 matrix cores: alone {mix['mfma']['mfma_tflops']:.1f} TFLOP/s; vector FMA alone "
        f"{mix['valu']['valu_tflops']:.1f}; both in one loop {mix['both']['mfma_tflops']:.1f} + "
        f"{mix['both']['valu_tflops']:.1f} = {mix['combined_tflops']:.1f} TFLOP/s")
    if not args.no_cpu:
        import oracle
        th = cpu_threads()
        k, dt = 16, 0.0
        while True:
            t0 = time.perf_counter()
            oracle.perm_batch(mats[:k], precise=False, threads=th)
            dt = time.perf_counter() - t0
            if dt >= args.cpu_budget or k >= len(mats):
                break
            k = min(len(mats), k * 2)
        res["perm_cpu"] = {"value": k / dt, "unit": "perms/s", "cores": th, "kind": "port", "cpu_model": cpu_model(),
                           "sample": f"{k} of the {len(mats)} order-{n} matrices, oracle chunked-double Glynn "
                                     f"(oracle/pcvl_oracle.c), {th} POSIX threads, {dt:.2f} s wall"}
        log(f"[bench] perm CPU: {k / dt:.1f} perms/s on {th} threads")


def bench_mixed(args, res):
    import torch
    from perceval_amd.permanent import PermanentPlan, pack
    from perceval_amd.utils.workloads import submatrix_batch
    _, morders, mmats = submatrix_batch(4096, 12, 24, seed=2025)
    mo, mflat = pack(mmats)
    md = torch.from_numpy(mflat).cuda()
    mout = torch.empty(len(mo), dtype=torch.complex128, device="cuda")
    mplan = PermanentPlan(mo, max_chunk_log2=args.chunk_log2)
    mms = time_plan(mplan, md, mout, max(3, args.steps // 4), 1)
    # the same launch captured in a HIP graph and replayed
    from perceval_amd.graphs import graphed_permanents
    g = graphed_permanents(mplan, md, mout)
    with Timer() as t:
        for _ in range(3):
            g.replay()
    graph_ms = t.ms / 3
    g.close()
    mplan.close()
    # the project's autotuner on this batch shape (chunk length), cached per device
    from perceval_amd import autotune
    tuned = autotune.tune_permanent_mixed(mo, reps=2)
    mflops = sum(glynn_flops(int(k)) for k in mo)
    res["mixed_n12_24"] = {"perms_per_s": len(mo) / (mms * 1e-3), "ms_per_step": mms,
                           "tflops": mflops / (mms * 1e-3) / 1e12,
                           "frac_fp64_peak": mflops / (mms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                           "ms_graph_replay": graph_ms, "autotuned_chunk_log2": tuned["chunk_log2"],
                           "autotuned_ms": tuned["ms"]}
    log(f"[bench] mixed n=12..24: {mms:.3f} ms/step, {res['mixed_n12_24']['tflops']:.2f} TFLOP/s; HIP-graph replay "
        f"{graph_ms:.3f} ms; autotuned chunk 2^{tuned['chunk_log2']}: {tuned['ms']:.3f} ms")


def bench_slos(args, res):
    import torch
    from perceval_amd.slos import PROBS, SlosPlan
    from perceval_amd.utils.workloads import haar_unitary
    m, n = 24, 10
    U = haar_unitary(m, seed=24)
    inp = [1] * n + [0] * (m - n)
    plan = SlosPlan(inp)
    dU = torch.from_numpy(U).cuda()
    out = torch.empty(plan.count, dtype=torch.float64, device="cuda")

    def timed(reps=10, batches=3):
        # 3 untimed runs, then the median of 3 batches of 10
        for _ in range(3):
            plan.run(dU, PROBS, out)
        ms = []
        for _ in range(batches):
            with Timer() as t:
                for _ in range(reps):
                    plan.run(dU, PROBS, out)
            ms.append(t.ms / reps)
        return statistics.median(ms)

    per_kernel = {}
    # 1 per state, 4 per state in two phases, 6 successor runs of 3 children
    # per thread with a transposed gather, 8 prefix/suffix blocks, 0 the
    # library's own pick (the headline number)
    for kernel in args.slos_kernels:
        plan.set_kernel(kernel)
        per_kernel[kernel] = timed()
    # suffix width of the block kernel
    default_w = plan.block_width
    per_width = {}
    plan.set_kernel(8)
    for w in args.slos_widths:
        plan.set_block_width(w)
        per_width[w] = timed()
    plan.set_block_width(default_w)
    plan.set_kernel(0)
    if 0 not in per_kernel:
        per_kernel[0] = timed()
    # the default plan captured in a HIP graph (10 runs per replay)
    from perceval_amd.graphs import graphed_slos
    g = graphed_slos(plan, dU, PROBS, out, repeats=10)
    g.replay()
    with Timer() as t:
        for _ in range(3):
            g.replay()
    graph_ms = t.ms / 30
    g.close()
    plan.run(dU, PROBS, out)
    ms = per_kernel[0]
    total = float(out.sum().item())
    plan.close()
    # masked plans on the same input: heralds 0,1,0,1 on modes 20..23, and two
    # herald rows with a floor on mode 23 (every layer restricted to the
    # states that can still match)
    masked = {}
    for name, rows, at_least in (("heralds_0101", [[-1] * 20 + [0, 1, 0, 1]], []),
                                 ("rows_0101_1001_floor23", [[-1] * 20 + [0, 1, 0, 1], [-1] * 20 + [1, 0, 0, 1]], [23])):
        mplan = SlosPlan(inp, mask_cond=rows, at_least=at_least)
        mout = torch.empty(max(1, mplan.count), dtype=torch.float64, device="cuda")
        for _ in range(3):
            mplan.run(dU, PROBS, mout)
        with Timer() as t:
            for _ in range(10):
                mplan.run(dU, PROBS, mout)
        masked[name] = {"ms": t.ms / 10, "outputs": mplan.count}
        mplan.close()
    res["slos_m24_n10_masked"] = masked
    # a larger space: m=32, n=10, 1,121,099,408 outputs (9.0 GB) from a 273M-state parent layer
    U32 = torch.from_numpy(haar_unitary(32, seed=32)).cuda()
    big = SlosPlan([1] * 10 + [0] * 22)
    bout = torch.empty(big.count, dtype=torch.float64, device="cuda")
    for _ in range(2):
        big.run(U32, PROBS, bout)
    with Timer() as t:
        for _ in range(5):
            big.run(U32, PROBS, bout)
    big_ms = t.ms / 5
    big_gb = slos_bytes(32, 10) / 1e9
    res["slos_m32_n10"] = {"ms": big_ms, "outputs": big.count, "algorithmic_GB": big_gb,
                           "GB_per_s": big_gb / (big_ms * 1e-3), "block_width": big.block_width}
    log(f"[bench] SLOS m=32 n=10: {big.count} outputs in {big_ms:.2f} ms, {big_gb / (big_ms * 1e-3):.0f} GB/s algorithmic")
    del bout
    big.close()
    torch.cuda.empty_cache()
    log("[bench] SLOS m=24 n=10 masked: " + ", ".join(f"{k} {v['outputs']} outputs {v['ms']:.3f} ms"
                                                       for k, v in masked.items()))
    gbytes = slos_bytes(m, n) / 1e9
    res["slos_m24_n10"] = {"ms": ms, "outputs": plan.count, "algorithmic_GB": gbytes,
                           "GB_per_s": gbytes / (ms * 1e-3), "frac_hbm_peak": gbytes / (ms * 1e-3) / HBM_PEAK_GBS,
                           "sum_probs": total, "block_width": default_w, "ms_graph_replay": graph_ms,
                           "ms_by_kernel": {name: per_kernel[kid] for kid, name in
                                            ((1, "per_state"), (4, "two_phase"), (6, "successor_run3"),
                                             (8, "prefix_suffix_blocks")) if kid in per_kernel},
                           "ms_by_block_width": {str(w): v for w, v in per_width.items()}}
    log(f"[bench] SLOS m=24 n=10: {ms:.3f} ms, {gbytes / (ms * 1e-3):.0f} GB/s algorithmic, sum={total:.15f}; "
        f"HIP-graph replay {graph_ms:.3f} ms; by kernel " + ", ".join(f"{k}: {v:.3f}" for k, v in per_kernel.items()) +
        "; block kernel by width " + ", ".join(f"{w}: {v:.3f}" for w, v in per_width.items()))
    if not args.no_cpu:
        import oracle
        th = cpu_threads()
        t0 = time.perf_counter()
        oracle.slos_coefs(U, inp, "gather", threads=th)
        dt = time.perf_counter() - t0
        res["slos_cpu"] = {"value": dt * 1e3, "unit": "ms per all_prob (coefficients)", "cores": th, "kind": "port", "cpu_model": cpu_model(),
                           "sample": f"full m=24 n=10 layer recursion, oracle gather form, {th} POSIX threads"}
        log(f"[bench] SLOS CPU: {dt:.2f} s on {th} threads")


def bench_slos_tree(args, res):
    """Several inputs through one SLOS tree plan against one plan per input:
    m = 20, the 8-photon input |1^8 0^12> with its 8 one-photon-loss and 28
    two-photon-loss variants (37 inputs, the shape of a noisy-source
    simulation)."""
    import itertools
    import torch
    from perceval_amd.slos import PROBS, SlosPlan, SlosTree
    from perceval_amd.utils.workloads import haar_unitary
    m, n = 20, 8
    U = haar_unitary(m, seed=20)
    base = [1] * n + [0] * (m - n)
    inputs = [base]
    for lost in (1, 2):
        for drop in itertools.combinations(range(n), lost):
            inputs.append([0 if i in drop else v for i, v in enumerate(base)])
    dU = torch.from_numpy(U).cuda()
    tree = SlosTree(inputs)
    tree.run(dU, PROBS)
    with Timer() as t:
        for _ in range(3):
            outs = tree.run(dU, PROBS)
    tree_ms = t.ms / 3
    plans = [SlosPlan(s) for s in inputs]
    for p in plans:
        p.run(dU, PROBS)
    with Timer() as t:
        for _ in range(3):
            for p in plans:
                p.run(dU, PROBS)
    sep_ms = t.ms / 3
    sums = [float(o.sum().item()) for o in outs]
    for p in plans:
        p.close()
    res["slos_tree_m20_37inputs"] = {"ms": tree_ms, "ms_separate_plans": sep_ms, "speedup": sep_ms / tree_ms,
                                     "inputs": len(inputs), "layer_launches": tree.layer_launches,
                                     "outputs": int(tree.offsets[-1]),
                                     "max_abs_sum_err": max(abs(v - 1.0) for v in sums)}
    tree.close()
    log(f"[bench] SLOS tree m=20, {len(inputs)} inputs: {tree_ms:.3f} ms shared vs {sep_ms:.3f} ms one plan each "
        f"({sep_ms / tree_ms:.2f}x)")
    if not args.no_cpu:
        import oracle
        th = cpu_threads()
        t0 = time.perf_counter()
        for s in inputs:
            oracle.slos_coefs(U, s, "gather", threads=th)
        dt = time.perf_counter() - t0
        res["slos_tree_cpu"] = {"value": dt * 1e3, "unit": "ms for all inputs (coefficients)", "cores": th,
                                "kind": "port", "cpu_model": cpu_model(), "sample": f"{len(inputs)} inputs, oracle gather form, one run each"}
        log(f"[bench] SLOS tree CPU: {dt * 1e3:.1f} ms on {th} threads")


def bench_sampling_multi(args, res):
    """SamplesProvider-shaped batch: 2,000 samples from each of 37 inputs (an
    8-photon state in 20 modes and its 1- and 2-photon-loss variants) in one
    pcvl_cc_samples_multi call, against one pcvl_cc_samples call per input."""
    import itertools
    import torch
    from perceval_amd.sampling import cc_samples_device, cc_samples_multi_device
    from perceval_amd.utils.workloads import haar_unitary
    m, n, per = 20, 8, 2000
    U = haar_unitary(m, seed=21)
    base = [1] * n + [0] * (m - n)
    inputs = [base]
    for lost in (1, 2):
        for drop in itertools.combinations(range(n), lost):
            inputs.append([0 if i in drop else v for i, v in enumerate(base)])
    counts = [per] * len(inputs)
    dU = torch.from_numpy(U).cuda()
    cc_samples_multi_device(dU, inputs, counts, seed=1)
    with Timer() as t:
        cc_samples_multi_device(dU, inputs, counts, seed=2)
    multi_ms = t.ms
    for s in inputs[:2]:
        cc_samples_device(dU, s, per, seed=1)
    with Timer() as t:
        for k, s in enumerate(inputs):
            cc_samples_device(dU, s, per, seed=2, first_sample=k * per)
    sep_ms = t.ms
    tot = per * len(inputs)
    res["sampling_multi_m20_37inputs"] = {"samples": tot, "ms": multi_ms, "samples_per_s": tot / (multi_ms * 1e-3),
                                          "ms_one_call_per_input": sep_ms, "speedup": sep_ms / multi_ms}
    log(f"[bench] multi-input sampling: {tot} samples of {len(inputs)} inputs in {multi_ms:.1f} ms "
        f"({tot / (multi_ms * 1e-3):.0f} samples/s) vs {sep_ms:.1f} ms one call per input")


def bench_gradient(args, res):
    import torch
    from perceval_amd.components import clements_mesh
    from perceval_amd.gradient import MeshProgram, mesh_adjoint
    from perceval_amd.utils.states import fock_space
    prog = MeshProgram(clements_mesh(12))
    rng = np.random.RandomState(4)
    phases = torch.from_numpy(rng.uniform(0, 2 * math.pi, size=(256, prog.nparams))).cuda()
    states = fock_space(12, 6)
    targets = states[np.sort(rng.choice(len(states), 32, replace=False))]
    inp = [1] * 6 + [0] * 6
    mesh_adjoint(prog, phases, inp, targets)
    reps = 5
    with Timer() as t:
        for _ in range(reps):
            mesh_adjoint(prog, phases, inp, targets)
    ms = t.ms / reps
    res["gradient"] = {"ms_per_batch": ms, "sets_per_s": 256 / (ms * 1e-3)}
    log(f"[bench] gradient: {ms:.3f} ms per 256 sets, {256 / (ms * 1e-3):.0f} sets/s")
    if not args.no_cpu:
        import oracle
        ph = phases[:2].cpu().numpy()

        def loss(p):
            U = prog.unitary(p)
            return sum(abs(oracle.perm(oracle.submatrix(U, inp, s))) ** 2 / oracle.prodnfact(s) for s in targets)
        t0 = time.perf_counter()
        for b in range(len(ph)):
            for k in range(prog.nparams):
                a = ph[b].copy(); a[k] += 1e-6
                c = ph[b].copy(); c[k] -= 1e-6
                _ = (loss(a) - loss(c)) / 2e-6
        dt = time.perf_counter() - t0
        res["gradient_cpu"] = {"value": len(ph) / dt, "unit": "sets/s", "cores": 1, "kind": "port", "cpu_model": cpu_model(),
                               "sample": "2 parameter sets, central differences over 132 phases with oracle "
                                         "permanents, 1 thread"}
        log(f"[bench] gradient CPU (FD): {len(ph) / dt:.2f} sets/s")
    # the objective of benchmark/benchmark_QML-DE-solver.py on a phase mesh:
    # 6 photons in 6 modes, every one of the 462 outputs weighted by a
    # standardised lambda (x 200), Clements mesh of 6 modes (30 phases)
    prog6 = MeshProgram(clements_mesh(6))
    rng6 = np.random.RandomState(0)
    t6 = fock_space(6, 6)
    lam = rng6.rand(len(t6))
    lam = 200 * (lam - lam.mean()) / lam.std()
    ph6 = torch.from_numpy(rng6.uniform(0, 2 * math.pi, size=(256, prog6.nparams))).cuda()
    mesh_adjoint(prog6, ph6, [1] * 6, t6, weights=lam)
    with Timer() as t:
        for _ in range(reps):
            mesh_adjoint(prog6, ph6, [1] * 6, t6, weights=lam)
    ms6 = t.ms / reps
    res["gradient_qml_m6_n6"] = {"ms_per_batch": ms6, "sets_per_s": 256 / (ms6 * 1e-3), "targets": len(t6),
                                 "plans": (len(t6) + 63) // 64}
    log(f"[bench] gradient of sum_s lambda_s P(s) over all {len(t6)} outputs, m=n=6: {ms6:.3f} ms per 256 sets")
    if not args.no_cpu:
        import oracle
        x0 = ph6[0].cpu().numpy()
        t0 = time.perf_counter()
        for k in range(prog6.nparams):
            a = x0.copy(); a[k] += 1e-6
            c = x0.copy(); c[k] -= 1e-6
            _ = (np.dot(lam, oracle.slos_probs(prog6.unitary(a), [1] * 6)) -
                 np.dot(lam, oracle.slos_probs(prog6.unitary(c), [1] * 6))) / 2e-6
        dt = time.perf_counter() - t0
        res["gradient_qml_cpu"] = {"value": 1 / dt, "unit": "sets/s", "cores": 1, "kind": "port", "cpu_model": cpu_model(),
                                   "sample": "1 parameter set, central differences over 30 phases with the oracle's "
                                             "SLOS probabilities (the reference benchmark's numerical gradient), 1 thread"}
        log(f"[bench] gradient QML CPU (FD over SLOS): {1 / dt:.1f} sets/s")


def _tvd_curve(occ, exact, grid):
    from perceval_amd.marginals import marginal_tvds_exact
    return [(N, marginal_tvds_exact(occ[:N], exact)) for N in grid if N <= len(occ)]


def _samples_to_meet(curve, bar=0.01):
    """Smallest grid count from which on both the 1- and the 2-mode marginal
    TVD stay <= bar (None if the last point misses)."""
    ok = [t["tvd1_max"] <= bar and t["tvd2_max"] <= bar for _, t in curve]
    if not ok or not ok[-1]:
        return None
    i = len(ok) - 1
    while i > 0 and ok[i - 1]:
        i -= 1
    return curve[i][0]


def _batch_tau(x, batches=50):
    """Integrated autocorrelation time of a chain statistic by batch means:
    var(batch mean) * L / var(x)."""
    L = len(x) // batches
    if L < 2:
        return float("nan")
    bm = x[:L * batches].reshape(batches, L, -1).mean(axis=1)
    v = x[:L * batches].var(axis=0)
    keep = v > 0
    return float(np.mean(bm.var(axis=0, ddof=1)[keep] * L / v[keep]))


def bench_sampling(args, res):
    import torch
    from perceval_amd.marginals import exact_marginals, marginal_tvds_exact
    from perceval_amd.mis import LearnedMIS, ProposalNet, occ_to_seq, seq_to_occ
    from perceval_amd.sampling import cc_samples_device
    from perceval_amd.utils.workloads import haar_unitary
    m, n, N = 60, 20, args.samples
    U = haar_unitary(m, seed=60)
    inp = [1] * n + [0] * (m - n)
    dU = torch.from_numpy(U).cuda()
    cc_samples_device(dU, inp, 256, seed=99)                      # warm-up (code objects, allocator)
    with Timer() as t:
        train = cc_samples_device(dU, inp, N, seed=1)
    cc_ms = t.ms
    cc_rate = N / (cc_ms * 1e-3)
    log(f"[bench] plain CC: {N} samples in {cc_ms:.1f} ms = {cc_rate:.0f} samples/s")
    net = ProposalNet(m, n, hidden=256, seed=0)
    tseq = occ_to_seq(train, n)
    t0 = time.perf_counter()
    hist = net.train(tseq, epochs=args.epochs, batch=2048, lr=args.lr, lr_end=args.lr_end)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log(f"[bench] proposal training: {train_s:.2f} s, mean -log q per epoch {['%.3f' % h for h in hist]}")
    mis = LearnedMIS(U, inp, net, chunk=131072)
    x0 = occ_to_seq(cc_samples_device(dU, inp, 1, seed=7), n)[0]
    mis.run(4096, seed=3, x0_seq=x0)                              # warm-up
    with Timer() as t:
        out = mis.run(N, seed=5, x0_seq=x0)
    mis_ms = t.ms
    mis_rate = N / (mis_ms * 1e-3)
    acc = out["acceptance"]
    # a longer chain for the equal-TVD comparison (same seed: its first N
    # samples are the timed ones)
    Nl = max(N, args.mis_judge_samples)
    long = mis.run(Nl, seed=5, x0_seq=x0) if Nl > N else out
    mis.close()
    occ_mis = seq_to_occ(long["seq"], m).cpu().numpy()
    occ_cc = train.cpu().numpy()
    # judge: the exact 1- and 2-mode marginals (200 seeded pairs), so the
    # TVDs carry the sampler's error only
    ex = exact_marginals(U, inp, pairs=200, seed=0)
    tv = marginal_tvds_exact(occ_mis[:N], ex)
    tv_cc = marginal_tvds_exact(occ_cc, ex)
    grid = sorted({int(round(10000 * 2 ** (i / 2))) for i in range(0, 20)} | {N, Nl})
    curve_mis = _tvd_curve(occ_mis, ex, grid)
    curve_cc = _tvd_curve(occ_cc, ex, grid)
    n_mis, n_cc = _samples_to_meet(curve_mis), _samples_to_meet(curve_cc)
    # time for each sampler to reach both marginal TVDs <= 0.01
    t_mis = n_mis / mis_rate if n_mis else None
    t_cc = n_cc / cc_rate if n_cc else None
    equal = (t_cc / t_mis) if (t_mis and t_cc) else None
    tau = _batch_tau(occ_mis[:, :8].astype(np.float64))
    # break-even: MIS wall time including its CC training samples and the
    # training itself, against plain CC, at equal effective samples
    over_s = cc_ms * 1e-3 + train_s
    eff_mis = mis_rate / tau if tau == tau and tau > 0 else None
    breakeven = (over_s / (1.0 / cc_rate - 1.0 / eff_mis)) if (eff_mis and eff_mis > cc_rate) else None
    meets = bool(acc >= 0.01 and tv["tvd1_max"] <= 0.01 and tv["tvd2_max"] <= 0.01)
    res["sampling_m60_n20"] = {
        "samples": N, "cc_samples_per_s": cc_rate, "mis_samples_per_s": mis_rate,
        "mis_vs_cc": mis_rate / cc_rate, "mis_acceptance": acc,
        "mis_meets_tvd_0.01": meets, "tvd_mis_vs_exact": tv, "tvd_cc_vs_exact": tv_cc,
        "samples_to_tvd_0.01": {"mis": n_mis, "cc": n_cc},
        "mis_vs_cc_equal_tvd": equal, "mis_tau_int": tau,
        "mis_effective_samples_per_s": eff_mis, "breakeven_samples": breakeven,
        "train_seconds": train_s, "train_nll": [float(h) for h in hist],
        "tvd2_max_curve": {"mis": {str(k): v["tvd2_max"] for k, v in curve_mis},
                           "cc": {str(k): v["tvd2_max"] for k, v in curve_cc}}}
    if not meets:
        log(f"[bench] MIS marginals at {N} samples miss TVD <= 0.01 (acceptance {acc:.3f}, {tv})")
    log(f"[bench] MIS: {N} samples in {mis_ms:.1f} ms = {mis_rate:.0f} samples/s ({mis_rate / cc_rate:.2f}x plain CC "
        f"per sample), acceptance {acc:.3f}, tau_int {tau:.2f}; TVD vs exact {tv}; CC {tv_cc}")
    log(f"[bench] samples to both marginal TVDs <= 0.01: MIS {n_mis}, CC {n_cc}; MIS vs CC at equal TVD: "
        f"{equal if equal is None else round(equal, 2)}x; break-even after {breakeven} samples")
    if not args.no_cpu:
        import oracle
        th = cpu_threads()
        k = max(th, 64)
        t0 = time.perf_counter()
        oracle.cc_samples(U, inp, k, seed=1, threads=th)
        dt = time.perf_counter() - t0
        res["sampling_cpu"] = {"value": k / dt, "unit": "samples/s", "cores": th, "kind": "port", "cpu_model": cpu_model(),
                               "sample": f"{k} samples m=60 n=20, oracle Clifford-Clifford, {th} POSIX threads"}
        log(f"[bench] CC CPU: {k / dt:.1f} samples/s on {th} threads")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--order", type=int, default=20)
    ap.add_argument("--chunk-log2", type=int, default=0, help="permanent lane chunk (0 = default)")
    ap.add_argument("--variant", type=int, default=0, help="permanent kernel variant (0 = by order)")
    ap.add_argument("--samples", type=int, default=100000)
    ap.add_argument("--mis-judge-samples", type=int, default=400000,
                    help="chain length for the equal-TVD comparison of the learned sampler")
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--lr", type=float, default=2e-3, help="proposal training step size")
    ap.add_argument("--lr-end", type=float, default=None, help="cosine decay of the step size down to this value")
    ap.add_argument("--cpu-budget", type=float, default=5.0, help="seconds of CPU baseline per section")
    ap.add_argument("--sections", default=",".join(SECTIONS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baselines")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-order batch")
    ap.add_argument("--slos-kernels", type=lambda v: [int(x) for x in v.split(",") if x], default=[1, 4, 6, 8, 0],
                    help="SLOS layer kernels to time (0 = the plan's own pick)")
    ap.add_argument("--slos-widths", type=lambda v: [int(x) for x in v.split(",") if x], default=[6, 8, 10, 12],
                    help="suffix widths of the block SLOS kernel to time")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE={world}); "
              f"every workload here is a batch of independent items and would shard by items across GPUs",
              file=sys.stderr)
        sys.exit(2)

    sections = [s for s in args.sections.split(",") if s]
    if args.no_mixed and "mixed" in sections:
        sections.remove("mixed")
    res = {}
    for s in sections:
        {"perm": bench_perm, "mixed": bench_mixed, "slos": bench_slos, "gradient": bench_gradient, "copy": bench_copy,
         "gurvits": bench_gurvits, "slos_tree": bench_slos_tree, "sampling_multi": bench_sampling_multi,
         "sampling": bench_sampling, "cache": bench_cache, "cc": bench_cc}[s](args, res)

    perm = res.get("perm", {"ms": float("nan"), "perms_per_s": float("nan"), "tflops": float("nan")})
    line = {
        "metric": "perms/s at n=20 (%FP64 peak); samples/s m=60 n=20 learned-MIS vs plain CC",
        "value": perm["perms_per_s"],
        "unit": "perms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": perm["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (fp64)",
        "data": "synthetic: seeded Haar unitaries; n x n submatrices (rows/cols without replacement) of a 64x64 one",
        "config": {"workload": f"batched Glynn permanents, {args.batch} matrices of order {args.order}",
                   "batch": args.batch, "order": args.order, "parallelism": "single GPU"},
        "roofline": {"bound": "fp64-valu", "achieved": perm["tflops"], "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": perm["tflops"] / FP64_PEAK_TFLOPS,
                     "traffic": PERM_N20_B4096_HBM_BYTES if (args.order, args.batch) == (20, 4096) else None,
                     "compute": "FP64 vector ALU (no MFMA: Glynn's products are not a dense contraction); the peak "
                                "is MI355X's FP64 vector figure, equal to its dense FP64 MFMA figure",
                     "attainable_fp64_fma_tflops": perm.get("fp64_fma_probe_tflops"),
                     "frac_of_attainable": perm.get("frac_of_probe")},
        "cpu_baseline": res.get("perm_cpu"),
    }
    for k, v in res.items():
        if k not in ("perm", "perm_cpu"):
            line[k] = v
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
