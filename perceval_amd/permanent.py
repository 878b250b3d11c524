"""Batched complex permanents on the GPU (host side of pcvl_perm_*).

Replaces ``xq.permanent_cx`` (perceval/backends/_naive.py:70-71).  A batch is a
packed complex128 device tensor holding row-major matrices back to back; the
orders may differ from matrix to matrix (0..32); orders 9..28 are handled in
one persistent launch, orders 29..32 in a second one.  ``PermanentPlan`` keeps the launch plan for a batch shape so
repeated batches (sampler acceptance, Naive all_prob over a fixed Fock space)
pay the host-side scheduling once.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native

MAX_ORDER = 32


class PermanentPlan:
    """Launch plan for a fixed list of matrix orders."""

    def __init__(self, orders, max_chunk_log2: int = 0, variant: int = 0):
        lib = _native.load()
        self.orders = np.ascontiguousarray(np.asarray(orders, dtype=np.int32).ravel())
        if self.orders.size and (self.orders.min() < 0 or self.orders.max() > MAX_ORDER):
            raise ValueError(f"matrix orders must lie in [0, {MAX_ORDER}]")
        sizes = self.orders.astype(np.int64) ** 2
        self.offsets = np.zeros(len(self.orders), dtype=np.int64)
        if len(sizes) > 1:
            self.offsets[1:] = np.cumsum(sizes)[:-1]
        self.total_elems = int(sizes.sum())
        self._plan = ctypes.c_void_p()
        _native.check(lib.pcvl_perm_plan_create_ex(
            self.orders.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            self.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            len(self.orders), int(max_chunk_log2), int(variant), ctypes.byref(self._plan)),
            "pcvl_perm_plan_create_ex")
        self._lib = lib

    def info(self) -> dict:
        nitems, grid, single = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        _native.check(self._lib.pcvl_perm_plan_info(self._plan, ctypes.byref(nitems), ctypes.byref(grid),
                                                    ctypes.byref(single)), "pcvl_perm_plan_info")
        return {"items": nitems.value, "grid": grid.value, "single_order": single.value}

    def __call__(self, packed, out=None, stream=None):
        """Permanents of the packed complex128 device tensor ``packed``."""
        import torch
        if packed.dtype != torch.complex128 or not packed.is_cuda:
            raise TypeError("packed matrices must be a complex128 CUDA tensor")
        if not packed.is_contiguous() or packed.numel() < self.total_elems:
            raise ValueError("packed tensor is not contiguous or too small for the plan")
        if out is None:
            out = torch.empty(len(self.orders), dtype=torch.complex128, device=packed.device)
        elif out.dtype != torch.complex128 or out.numel() < len(self.orders) or not out.is_contiguous():
            raise ValueError("out must be a contiguous complex128 tensor of the batch length")
        _native.check(self._lib.pcvl_perm_plan_execute(self._plan, _native.ptr(packed), _native.ptr(out),
                                                       _native.stream_handle(stream)), "pcvl_perm_plan_execute")
        return out

    def close(self):
        if getattr(self, "_plan", None) is not None and self._plan.value:
            _native.check(self._lib.pcvl_perm_plan_destroy(self._plan), "pcvl_perm_plan_destroy")
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pack(mats) -> tuple[np.ndarray, np.ndarray]:
    """Pack a list of square complex matrices: (orders int32, flat complex128)."""
    orders = np.array([np.shape(m)[0] for m in mats], dtype=np.int32)
    flat = np.concatenate([np.ascontiguousarray(m, dtype=np.complex128).ravel() for m in mats]) \
        if len(mats) else np.zeros(0, dtype=np.complex128)
    return orders, flat


def permanents(mats, device="cuda") -> np.ndarray:
    """Permanents of a list of square complex matrices, computed on the GPU."""
    import torch
    orders, flat = pack(mats)
    if len(orders) == 0:
        return np.zeros(0, dtype=np.complex128)
    plan = PermanentPlan(orders)
    try:
        d = torch.from_numpy(flat).to(device) if flat.size else torch.zeros(1, dtype=torch.complex128,
                                                                              device=device)
        out = plan(d)
        return out.cpu().numpy()
    finally:
        plan.close()


class SinglePermanents:
    """One permanent per call with no plan or device allocation per call: a
    plan and device buffers are kept per order (the per-output path of the
    Naive backend, _naive.py:46-49,70-71)."""

    def __init__(self, device: str = "cuda"):
        self.device = device
        self._by_order = {}

    def __call__(self, M) -> complex:
        import torch
        M = np.ascontiguousarray(np.asarray(M, dtype=np.complex128))
        n = M.shape[0]
        if M.shape != (n, n):
            raise ValueError("permanent of a non-square matrix")
        if n not in self._by_order:
            plan = PermanentPlan([n])
            d_in = torch.empty(max(1, n * n), dtype=torch.complex128, device=self.device)
            d_out = torch.empty(1, dtype=torch.complex128, device=self.device)
            self._by_order[n] = (plan, d_in, d_out)
        plan, d_in, d_out = self._by_order[n]
        if n:
            d_in[:n * n].copy_(torch.from_numpy(M.ravel()))
        plan(d_in, d_out)
        return complex(d_out.cpu().numpy()[0])

    def close(self):
        for plan, _, _ in self._by_order.values():
            plan.close()
        self._by_order = {}


def output_submatrices(U, occ, in_occ):
    """Naive submatrices of many output states (_naive.py:51-68): mats[k] =
    U[photon modes of occ[k], photon modes of in_occ], both ascending with
    repetition, and norms[k] = sqrt(prod occ[k]! prod in_occ!).  occ is
    (N, m) with n photons per row; mats is (N, n, n) complex128."""
    occ = np.asarray(occ, dtype=np.int64)
    in_occ = np.asarray(in_occ, dtype=np.int64)
    N, m = occ.shape
    n = int(in_occ.sum())
    # every row holds n photons, so repeating the mode indices by the flat
    # counts lists each state's photon modes in order, n per state
    rows = np.repeat(np.tile(np.arange(m), N), occ.reshape(-1)).reshape(N, n)
    cols = np.repeat(np.arange(in_occ.size), in_occ)
    mats = np.ascontiguousarray(np.asarray(U, dtype=np.complex128)[rows[:, :, None], cols[None, None, :]])
    fact = np.cumprod(np.concatenate([[1.0], np.arange(1, max(n, 1) + 1, dtype=float)]))
    norms = np.sqrt(np.prod(fact[occ], axis=1) * np.prod(fact[in_occ]))
    return mats, norms


class BatchPermanents:
    """Permanents of an (N, n, n) batch in one launch, the plan kept per
    (N, n): repeated all-output calls of the Naive backend allocate no plan."""

    def __init__(self):
        self._plans = {}

    def __call__(self, mats) -> np.ndarray:
        import torch
        mats = np.ascontiguousarray(np.asarray(mats, dtype=np.complex128))
        N, n = mats.shape[0], mats.shape[1]
        key = (N, n)
        if key not in self._plans:
            self._plans[key] = PermanentPlan(np.full(N, n, dtype=np.int32))
        return self._plans[key](torch.from_numpy(mats.reshape(-1)).cuda()).cpu().numpy()

    def close(self):
        for plan in self._plans.values():
            plan.close()
        self._plans = {}


_single = None


def permanent(M) -> complex:
    """Permanent of one square complex matrix, computed on the GPU (cached
    per-order plans, see SinglePermanents)."""
    global _single
    if _single is None:
        _single = SinglePermanents()
    return _single(M)


def gurvits(mats, iterations: int, seed: int = 0, stream0: int = 0):
    """Gurvits estimates of a batch of matrices (list of n x n complex arrays,
    n <= 48): returns (estimates complex128 (B,), 99% error bounds float64 (B,))
    as numpy arrays.  Matrix b draws its samples from the counter-based stream
    (seed, stream0 + b), so results are reproducible and batches can be split."""
    import torch
    lib = _native.load()
    orders, flat = pack(mats)
    offsets = np.zeros(len(orders), dtype=np.int64)
    if len(orders):
        offsets[1:] = np.cumsum(orders.astype(np.int64) ** 2)[:-1]
    d = torch.from_numpy(flat if flat.size else np.zeros(1, dtype=np.complex128)).cuda()
    est = torch.empty(max(1, len(orders)), dtype=torch.complex128, device="cuda")
    err = torch.empty(max(1, len(orders)), dtype=torch.float64, device="cuda")
    _native.check(lib.pcvl_gurvits_batch(_native.ptr(d), orders.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                         offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(orders),
                                         int(iterations), int(seed), int(stream0), _native.ptr(est), _native.ptr(err),
                                         _native.stream_handle()), "pcvl_gurvits_batch")
    return est.cpu().numpy()[:len(orders)], err.cpu().numpy()[:len(orders)]
