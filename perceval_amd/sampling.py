"""Exact boson sampling on the GPU (host side of pcvl_cc_samples).

Clifford & Clifford 2017, algorithm A, as called by the CliffordClifford2017
backend (perceval/backends/_clifford2017.py:52-57).  Sample s of a call draws
its uniforms from the counter-based stream (seed, first_sample + s), so equal
arguments give equal samples.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native


def cc_samples_device(U, in_occ, count: int, seed: int, first_sample: int = 0, out=None, stream=None):
    """Samples as a (count, m) int32 CUDA tensor.  U: numpy or CUDA m x m complex128."""
    import torch
    lib = _native.load()
    in_occ = np.ascontiguousarray(np.asarray(list(in_occ), dtype=np.int32))
    m = int(in_occ.size)
    n = int(in_occ.sum())
    if isinstance(U, np.ndarray):
        dU = torch.from_numpy(np.ascontiguousarray(U, dtype=np.complex128)).cuda()
    else:
        dU = U.to(dtype=torch.complex128).contiguous()
    if tuple(dU.shape) != (m, m):
        raise ValueError(f"unitary must be {m} x {m}")
    if out is None:
        out = torch.empty((count, m), dtype=torch.int32, device=dU.device)
    need = ctypes.c_size_t()
    _native.check(lib.pcvl_cc_workspace_bytes(n, count, ctypes.byref(need)), "pcvl_cc_workspace_bytes")
    work = torch.empty(max(16, need.value), dtype=torch.uint8, device=dU.device)
    _native.check(lib.pcvl_cc_samples(_native.ptr(dU), m, in_occ.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                      int(count), int(seed) & (2 ** 64 - 1), int(first_sample), _native.ptr(out),
                                      _native.ptr(work), work.numel(), _native.stream_handle(stream)),
                  "pcvl_cc_samples")
    return out


def cc_samples(U, in_occ, count: int, seed: int, first_sample: int = 0) -> np.ndarray:
    """Samples as a (count, m) int32 numpy array of occupation vectors."""
    if count == 0:
        return np.zeros((0, len(list(in_occ))), dtype=np.int32)
    return cc_samples_device(U, in_occ, count, seed, first_sample).cpu().numpy()


def cc_samples_multi_device(U, inputs, counts, seed: int, first_sample: int = 0, stream=None):
    """Samples of several input states in one call (pcvl_cc_samples_multi):
    a (sum(counts), m) int32 CUDA tensor, rows grouped by input in order.
    Input i's rows are the samples pcvl_cc_samples would draw for it with
    first_sample + sum(counts[:i])."""
    import torch
    lib = _native.load()
    rows = np.ascontiguousarray(np.atleast_2d(np.asarray([list(s) for s in inputs], dtype=np.int32)))
    cnt = np.ascontiguousarray(np.asarray(list(counts), dtype=np.int64))
    if len(cnt) != rows.shape[0]:
        raise ValueError("one count per input state")
    m = rows.shape[1]
    if isinstance(U, np.ndarray):
        dU = torch.from_numpy(np.ascontiguousarray(U, dtype=np.complex128)).cuda()
    else:
        dU = U.to(dtype=torch.complex128).contiguous()
    if tuple(dU.shape) != (m, m):
        raise ValueError(f"unitary must be {m} x {m}")
    out = torch.empty((max(1, int(cnt.sum())), m), dtype=torch.int32, device=dU.device)
    _native.check(lib.pcvl_cc_samples_multi(_native.ptr(dU), m, rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                            cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), rows.shape[0],
                                            int(seed) & (2 ** 64 - 1), int(first_sample), _native.ptr(out),
                                            _native.stream_handle(stream)), "pcvl_cc_samples_multi")
    return out[:int(cnt.sum())]
