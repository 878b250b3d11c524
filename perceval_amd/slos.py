"""SLOS strong simulation and Fock-space ranking on the GPU (host side of
pcvl_slos_* and pcvl_fock_*).

A ``SlosPlan`` corresponds to one deployed SLOS path of the reference
(_SLOSBackend._deploy/_Path, perceval/backends/_slos.py:61-102,151-185): it
fixes the input state (hence the photon-addition order) and the Fock layers;
``run`` recomputes the final layer for a new unitary, which is what the
reference's set_circuit fast path does (_slos.py:130-141).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native

COEFS = 0
PROBS = 1


class SlosPlan:
    """mask_cond: optional conditions, one row per mask (-1 any count, c >= 0
    c photons; a 1-D row is one mask; see mask_conditions); at_least: the
    modes whose digit is a floor (at least c photons) instead of a count;
    mask_n: the photon count the masks were set for (None: the input's; else
    at least the input's).  A masked plan computes only the layers' states
    that can still match some row (the masked FSArray of _slos.py:161-164)
    and returns the kept output states only, in FSArray order: those matching
    a row, or with mask_n above the input's photon count, those that the
    missing photons can still complete (FSMask.keeps)."""

    def __init__(self, in_occ, mask_cond=None, at_least=None, mask_n=None):
        lib = _native.load()
        self.in_occ = np.ascontiguousarray(np.asarray(list(in_occ), dtype=np.int32))
        self.m = int(self.in_occ.size)
        self._plan = ctypes.c_void_p()
        self.mask_cond = None
        self.at_least = None
        i32p = ctypes.POINTER(ctypes.c_int32)
        if mask_cond is None:
            if at_least or mask_n:
                raise ValueError("at-least modes and a mask photon count need mask conditions")
            _native.check(lib.pcvl_slos_plan_create(self.m, self.in_occ.ctypes.data_as(i32p), ctypes.byref(self._plan)),
                          "pcvl_slos_plan_create")
        else:
            cond = np.asarray(mask_cond, dtype=np.int32)
            if cond.ndim == 1:
                cond = cond[None, :]
            cond = np.ascontiguousarray(cond)
            if cond.ndim != 2 or cond.shape[1] != self.m or cond.shape[0] < 1:
                raise ValueError("mask conditions must have one entry per mode")
            at = np.zeros(self.m, dtype=np.int32)
            for q in (at_least or []):
                if not 0 <= int(q) < self.m:
                    raise ValueError("at-least mode outside the circuit")
                at[int(q)] = 1
            self.mask_cond = cond
            self.at_least = at
            _native.check(lib.pcvl_slos_plan_create_masks(self.m, self.in_occ.ctypes.data_as(i32p), cond.shape[0],
                                                          cond.ctypes.data_as(i32p), at.ctypes.data_as(i32p),
                                                          int(mask_n or 0), ctypes.byref(self._plan)),
                          "pcvl_slos_plan_create_masks")
        self._lib = lib
        n, count, work = ctypes.c_int(), ctypes.c_int64(), ctypes.c_size_t()
        path = np.zeros(max(1, int(self.in_occ.sum())), dtype=np.int32)
        _native.check(lib.pcvl_slos_plan_info(self._plan, ctypes.byref(n), ctypes.byref(count), ctypes.byref(work),
                                              path.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
                      "pcvl_slos_plan_info")
        self.n = n.value
        self.count = count.value
        self.work_bytes = work.value
        self.path = path[:self.n].tolist()
        self._work = None

    # layer kernels accepted by set_kernel
    KERNELS = (0, 1, 4, 6, 8)

    def set_kernel(self, kernel: int):
        """0 = pick (8 where its layout fits, else 6, else 4, else 1),
        1 = one thread per output state, 4 = one thread per output state, walk
        and gather in two phases, 6 = runs of 3 consecutive states per thread
        stepped by successor and gathered lane-consecutively, 8 = blocks of
        states sharing their first m - w modes: one scalar walk per block,
        the last w modes from a table.  All give bit-identical results."""
        _native.check(self._lib.pcvl_slos_plan_set_kernel(self._plan, int(kernel)), "pcvl_slos_plan_set_kernel")

    def set_block_width(self, w: int):
        """Suffix width of kernel 8 (0 = default).  Results do not depend on it."""
        _native.check(self._lib.pcvl_slos_plan_set_block_width(self._plan, int(w)), "pcvl_slos_plan_set_block_width")

    @property
    def block_width(self) -> int:
        return int(self._lib.pcvl_slos_plan_block_width(self._plan))

    def _workspace(self, device):
        import torch
        if self._work is None or self._work.device != torch.device(device):
            self._work = torch.empty(max(16, self.work_bytes), dtype=torch.uint8, device=device)
        return self._work

    def run(self, U, mode: int = PROBS, out=None, stream=None):
        """Final layer for unitary U (m x m complex128, numpy or CUDA tensor):
        complex coefficients (mode COEFS) or probabilities (mode PROBS), as a
        CUDA tensor in FSArray order."""
        import torch
        if isinstance(U, np.ndarray):
            dU = torch.from_numpy(np.ascontiguousarray(U, dtype=np.complex128)).cuda()
        else:
            dU = U.to(dtype=torch.complex128).contiguous()
        if tuple(dU.shape) != (self.m, self.m):
            raise ValueError(f"unitary must be {self.m} x {self.m}")
        dtype = torch.complex128 if mode == COEFS else torch.float64
        if out is None:
            out = torch.empty(max(1, self.count), dtype=dtype, device=dU.device)[:self.count]
        elif out.dtype != dtype or out.numel() < self.count or not out.is_contiguous():
            raise ValueError("out has the wrong dtype, size or layout")
        if self.count == 0:
            return out
        work = self._workspace(dU.device)
        _native.check(self._lib.pcvl_slos_plan_execute(self._plan, _native.ptr(dU), mode, _native.ptr(out),
                                                       _native.ptr(work), work.numel(),
                                                       _native.stream_handle(stream)), "pcvl_slos_plan_execute")
        return out

    def close(self):
        if getattr(self, "_plan", None) is not None and self._plan.value:
            _native.check(self._lib.pcvl_slos_plan_destroy(self._plan), "pcvl_slos_plan_destroy")
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SlosTree:
    """Several input states in one plan: the _Path tree that
    SLOSBackend.preprocess([inputs]) builds (perceval/backends/_slos.py:37-86,
    170-185), so layers on a shared photon-addition prefix are computed once.
    ``run`` returns one CUDA tensor per input (views of one output buffer)."""

    def __init__(self, inputs):
        lib = _native.load()
        rows = np.ascontiguousarray(np.atleast_2d(np.asarray([list(s) for s in inputs], dtype=np.int32)))
        self.inputs = rows
        self.count, self.m = rows.shape
        self._plan = ctypes.c_void_p()
        _native.check(lib.pcvl_slos_tree_create(self.m, rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                self.count, ctypes.byref(self._plan)), "pcvl_slos_tree_create")
        self._lib = lib
        nmax, work, launches = ctypes.c_int(), ctypes.c_size_t(), ctypes.c_int64()
        offs = np.zeros(self.count + 1, dtype=np.int64)
        _native.check(lib.pcvl_slos_tree_info(self._plan, ctypes.byref(nmax),
                                              offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                              ctypes.byref(work), ctypes.byref(launches)), "pcvl_slos_tree_info")
        self.nmax = nmax.value
        self.offsets = offs
        self.work_bytes = work.value
        self.layer_launches = launches.value
        self._work = None

    def run(self, U, mode: int = PROBS, out=None, stream=None):
        import torch
        if isinstance(U, np.ndarray):
            dU = torch.from_numpy(np.ascontiguousarray(U, dtype=np.complex128)).cuda()
        else:
            dU = U.to(dtype=torch.complex128).contiguous()
        if tuple(dU.shape) != (self.m, self.m):
            raise ValueError(f"unitary must be {self.m} x {self.m}")
        dtype = torch.complex128 if mode == COEFS else torch.float64
        total = int(self.offsets[-1])
        if out is None:
            out = torch.empty(total, dtype=dtype, device=dU.device)
        elif out.dtype != dtype or out.numel() < total or not out.is_contiguous():
            raise ValueError("out has the wrong dtype, size or layout")
        if self._work is None or self._work.device != dU.device:
            self._work = torch.empty(max(16, self.work_bytes), dtype=torch.uint8, device=dU.device)
        _native.check(self._lib.pcvl_slos_tree_execute(self._plan, _native.ptr(dU), mode, _native.ptr(out),
                                                       _native.ptr(self._work), self._work.numel(),
                                                       _native.stream_handle(stream)), "pcvl_slos_tree_execute")
        return [out[int(self.offsets[i]):int(self.offsets[i + 1])] for i in range(self.count)]

    def close(self):
        if getattr(self, "_plan", None) is not None and self._plan.value:
            _native.check(self._lib.pcvl_slos_tree_destroy(self._plan), "pcvl_slos_tree_destroy")
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fock_count(m: int, n: int) -> int:
    return int(_native.load(require_gpu=False).pcvl_fock_count(m, n))


def fock_rank(occ, nmax: int | None = None) -> np.ndarray:
    """Ranks (FSArray.find) of occupation rows, computed on the GPU."""
    import torch
    occ = np.ascontiguousarray(np.atleast_2d(np.asarray(occ, dtype=np.int32)))
    count, m = occ.shape
    nmax = int(occ.sum(axis=1).max()) if nmax is None and count else (nmax or 0)
    lib = _native.load()
    d_occ = torch.from_numpy(occ).cuda()
    d_out = torch.empty(count, dtype=torch.int64, device="cuda")
    _native.check(lib.pcvl_fock_rank(_native.ptr(d_occ), count, m, nmax, _native.ptr(d_out),
                                     _native.stream_handle()), "pcvl_fock_rank")
    return d_out.cpu().numpy()


def fock_unrank(ranks, m: int, n: int) -> np.ndarray:
    """Occupation rows of the given ranks in the n-photon, m-mode space (GPU)."""
    import torch
    ranks = np.ascontiguousarray(np.asarray(ranks, dtype=np.int64).ravel())
    lib = _native.load()
    d_r = torch.from_numpy(ranks).cuda()
    d_out = torch.empty((len(ranks), m), dtype=torch.int32, device="cuda")
    _native.check(lib.pcvl_fock_unrank(_native.ptr(d_r), len(ranks), m, n, _native.ptr(d_out),
                                       _native.stream_handle()), "pcvl_fock_unrank")
    return d_out.cpu().numpy()


def mask_conditions(masks, m: int) -> np.ndarray:
    """FSMask strings -> (nmask, m) int32 conditions (-1 = any count).  A
    count c is the character chr(0x30 + c), so '0'..'9' and on up to 32
    photons (perceval/simulators/simulator.py:650-662 writes heralds so)."""
    if isinstance(masks, str):
        masks = [masks]
    rows = []
    for mk in masks:
        if len(mk) != m:
            raise ValueError("mask length must equal the number of modes")
        rows.append([mask_char_count(ch) for ch in mk])
    return np.ascontiguousarray(np.asarray(rows, dtype=np.int32))


def mask_char_count(ch: str) -> int:
    """One mask character: ' ' or '*' any count (-1), else chr(0x30 + c) for c in 0..32."""
    if ch in " *":
        return -1
    c = ord(ch) - 0x30
    if not 0 <= c <= 32:
        raise ValueError(f"mask character {ch!r} is neither a count nor a wildcard")
    return c


def fock_mask_select(m: int, n: int, masks, at_least_modes=None, mask_n=None):
    """Ranks (FSArray order) of the n-photon states the masks keep (matching a
    row; with mask_n > n, completable by the mask_n - n photons still to
    come), as a CUDA int64 tensor, selected on the GPU (pcvl_fock_mask_select)."""
    import torch
    lib = _native.load()
    cond = mask_conditions(masks, m)
    at = np.zeros(m, dtype=np.int32)
    for q in (at_least_modes or []):
        at[int(q)] = 1
    total = fock_count(m, n)
    out = torch.empty(max(1, total), dtype=torch.int64, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    _native.check(lib.pcvl_fock_mask_select(m, n, int(mask_n or 0), cond.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                            cond.shape[0],
                                            at.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _native.ptr(out),
                                            _native.ptr(cnt), _native.stream_handle()), "pcvl_fock_mask_select")
    return out[:int(cnt.item())]


def gather_f64(src, idx):
    """src[idx] for a CUDA float64 tensor and CUDA int64 indices (pcvl_gather_f64)."""
    import torch
    out = torch.empty(max(1, idx.numel()), dtype=torch.float64, device=src.device)
    _native.check(_native.load().pcvl_gather_f64(_native.ptr(src), _native.ptr(idx), idx.numel(), _native.ptr(out),
                                                 _native.stream_handle()), "pcvl_gather_f64")
    return out[:idx.numel()]
