"""HIP-graph capture of launch plans.

Plan execution (`PermanentPlan.__call__`, `SlosPlan.run` with a preallocated
output and workspace) enqueues kernels only -- no allocation, no host
synchronisation, no memset (the permanent queue cursor is reset by a kernel)
-- so it can be captured once and replayed.  torch.cuda.CUDAGraph is the
capture/replay handle; under ROCm it wraps hipGraph.  The library launches
on torch's current stream, which during capture is the capture stream.
"""
from __future__ import annotations


class GraphedCall:
    """Capture `fn()` (a sequence of this library's plan executions writing into
    preallocated tensors) into a HIP graph; `replay()` re-runs it."""

    def __init__(self, fn, warmup: int = 1):
        import torch
        self._fn = fn
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                fn()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            fn()
        torch.cuda.synchronize()

    def replay(self):
        self.graph.replay()


def graphed_permanents(plan, d_mats, out, repeats: int = 1) -> GraphedCall:
    """Graph of `repeats` executions of a permanent plan on fixed buffers."""
    def fn():
        for _ in range(repeats):
            plan(d_mats, out)
    return GraphedCall(fn)
