"""HIP graphs of plan executions, through the library's C ABI.

`pcvl_perm_graph_create` / `pcvl_slos_graph_create` capture a plan's
execution on fixed device buffers into an instantiated hipGraph (the plans
enqueue kernels only: no allocation, no synchronisation, no memset), and
`pcvl_graph_launch` replays it with one call.  The tensors given here are
kept referenced for the graph's lifetime, since every replay reads and
writes them.
"""
from __future__ import annotations

import ctypes

from . import _native


class NativeGraph:
    """An instantiated graph; `replay(stream)` enqueues it (default: torch's
    current stream)."""

    def __init__(self, handle: ctypes.c_void_p, keep: tuple):
        self._lib = _native.load()
        self._g = handle
        self._keep = keep

    def replay(self, stream=None):
        _native.check(self._lib.pcvl_graph_launch(self._g, _native.stream_handle(stream)), "pcvl_graph_launch")

    def close(self):
        if getattr(self, "_g", None) is not None and self._g.value:
            _native.check(self._lib.pcvl_graph_destroy(self._g), "pcvl_graph_destroy")
            self._g = ctypes.c_void_p()
            self._keep = ()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def graphed_permanents(plan, d_mats, out, repeats: int = 1) -> NativeGraph:
    """Graph of `repeats` executions of a PermanentPlan on fixed buffers."""
    import torch
    if d_mats.dtype != torch.complex128 or not d_mats.is_cuda or not d_mats.is_contiguous():
        raise TypeError("packed matrices must be a contiguous complex128 CUDA tensor")
    if d_mats.numel() < plan.total_elems:
        raise ValueError("packed tensor too small for the plan")
    if out.dtype != torch.complex128 or out.numel() < len(plan.orders) or not out.is_contiguous():
        raise ValueError("out must be a contiguous complex128 tensor of the batch length")
    lib = _native.load()
    torch.cuda.synchronize()   # the buffers are ready before capture starts on its own stream
    g = ctypes.c_void_p()
    _native.check(lib.pcvl_perm_graph_create(plan._plan, _native.ptr(d_mats), _native.ptr(out), int(repeats),
                                             ctypes.byref(g)), "pcvl_perm_graph_create")
    return NativeGraph(g, (plan, d_mats, out))


def graphed_slos(plan, dU, mode, out, repeats: int = 1) -> NativeGraph:
    """Graph of `repeats` executions of a SlosPlan (mode COEFS or PROBS) for a
    fixed unitary tensor and output tensor."""
    import torch
    from .slos import COEFS
    if dU.dtype != torch.complex128 or not dU.is_cuda or not dU.is_contiguous() or tuple(dU.shape) != (plan.m, plan.m):
        raise ValueError(f"unitary must be a contiguous {plan.m} x {plan.m} complex128 CUDA tensor")
    dtype = torch.complex128 if mode == COEFS else torch.float64
    if out.dtype != dtype or out.numel() < plan.count or not out.is_contiguous():
        raise ValueError("out has the wrong dtype, size or layout")
    work = plan._workspace(dU.device)
    lib = _native.load()
    torch.cuda.synchronize()
    g = ctypes.c_void_p()
    _native.check(lib.pcvl_slos_graph_create(plan._plan, _native.ptr(dU), int(mode), _native.ptr(out),
                                             _native.ptr(work), work.numel(), int(repeats), ctypes.byref(g)),
                  "pcvl_slos_graph_create")
    return NativeGraph(g, (plan, dU, out, work))
