"""ctypes binding of libpcvl_hip.so, the C ABI declared in include/pcvl_hip.h.

This is the only way Python reaches the HIP kernels.  torch supplies device
memory and the current stream; nothing here computes on the CPU.  If the
library or a GPU is missing, every entry point raises ``NativeUnavailable``:
there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpcvl_hip.so")

_lock = threading.Lock()
_lib = None


class NativeUnavailable(RuntimeError):
    """The HIP library is not built or no GPU is visible."""


class NativeError(RuntimeError):
    """A C-ABI entry point returned a negative status."""


c_int = ctypes.c_int
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_u64 = ctypes.c_uint64
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
c_size_t = ctypes.c_size_t
P = ctypes.POINTER

# name -> (restype, argtypes); must match include/pcvl_hip.h
SIGNATURES = {
    "pcvl_last_error": (ctypes.c_char_p, []),
    "pcvl_abi_version": (c_int, []),
    "pcvl_perm_plan_create": (c_int, [P(c_i32), P(c_i64), c_i64, c_int, P(c_void_p)]),
    "pcvl_perm_plan_create_ex": (c_int, [P(c_i32), P(c_i64), c_i64, c_int, c_int, P(c_void_p)]),
    "pcvl_perm_plan_execute": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "pcvl_perm_plan_destroy": (c_int, [c_void_p]),
    "pcvl_perm_plan_info": (c_int, [c_void_p, P(c_i64), P(c_int), P(c_int)]),
    "pcvl_perm_batch": (c_int, [c_void_p, P(c_i32), P(c_i64), c_i64, c_void_p, c_void_p]),
    "pcvl_perm_batch_noalloc": (c_int, [c_void_p, P(c_i32), P(c_i64), c_i64, c_void_p, c_void_p]),
    "pcvl_perm_noalloc_max_bytes": (c_size_t, []),
    "pcvl_fock_count": (c_i64, [c_int, c_int]),
    "pcvl_fock_rank": (c_int, [c_void_p, c_i64, c_int, c_int, c_void_p, c_void_p]),
    "pcvl_fock_unrank": (c_int, [c_void_p, c_i64, c_int, c_int, c_void_p, c_void_p]),
    "pcvl_fock_mask_select": (c_int, [c_int, c_int, c_int, P(c_i32), c_int, P(c_i32), c_void_p, c_void_p, c_void_p]),
    "pcvl_gather_f64": (c_int, [c_void_p, c_void_p, c_i64, c_void_p, c_void_p]),
    "pcvl_slos_plan_create": (c_int, [c_int, P(c_i32), P(c_void_p)]),
    "pcvl_slos_plan_create_masked": (c_int, [c_int, P(c_i32), P(c_i32), P(c_void_p)]),
    "pcvl_slos_plan_create_masks": (c_int, [c_int, P(c_i32), c_int, P(c_i32), P(c_i32), c_int, ctypes.POINTER(c_void_p)]),
    "pcvl_slos_plan_info": (c_int, [c_void_p, P(c_int), P(c_i64), P(c_size_t), P(c_i32)]),
    "pcvl_slos_plan_execute": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "pcvl_slos_plan_destroy": (c_int, [c_void_p]),
    "pcvl_slos_plan_set_kernel": (c_int, [c_void_p, c_int]),
    "pcvl_slos_plan_set_block_width": (c_int, [c_void_p, c_int]),
    "pcvl_slos_plan_block_width": (c_int, [c_void_p]),
    "pcvl_slos_tree_create": (c_int, [c_int, P(c_i32), c_int, ctypes.POINTER(c_void_p)]),
    "pcvl_slos_tree_info": (c_int, [c_void_p, P(c_int), P(c_i64), P(ctypes.c_size_t), P(c_i64)]),
    "pcvl_slos_tree_execute": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, ctypes.c_size_t, c_void_p]),
    "pcvl_slos_tree_destroy": (c_int, [c_void_p]),
    "pcvl_cc_workspace_bytes": (c_int, [c_int, c_i64, P(c_size_t)]),
    "pcvl_mlp_gemm": (c_int, [c_void_p, c_i64, c_i64, c_void_p, c_int, c_i64, c_i64, c_void_p, c_i64, c_int, c_int,
                              c_int, c_void_p, c_int, c_int, c_void_p]),
    "pcvl_mlp_relu_bwd": (c_int, [c_void_p, c_void_p, c_i64, c_void_p]),
    "pcvl_mlp_colsum": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "pcvl_mlp_wgrad_workspace": (c_int, [c_int, c_int, c_int, ctypes.POINTER(c_i64)]),
    "pcvl_mlp_wgrad": (c_int, [c_void_p, c_i64, c_void_p, c_i64, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                               c_i64, c_void_p]),
    "pcvl_mlp_ce": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, ctypes.c_float, c_void_p, c_void_p,
                            c_void_p]),
    "pcvl_mlp_adam": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_i64, ctypes.c_float,
                              ctypes.c_float, ctypes.c_float, ctypes.c_float, c_int, c_void_p]),
    "pcvl_mlp_to_bf16": (c_int, [c_void_p, c_void_p, c_i64, c_void_p]),
    "pcvl_mis_features": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "pcvl_mis_features_all": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pcvl_mis_sample_step": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_u64, c_u64, c_void_p,
                                     c_void_p, c_void_p]),
    "pcvl_mis_sample_fused": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_u64, c_u64,
                                      c_void_p, c_void_p, c_void_p]),
    "pcvl_mis_submatrices": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_i64, c_void_p, c_void_p]),
    "pcvl_mis_permanents": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_i64, c_void_p, c_void_p]),
    "pcvl_mis_logp": (c_int, [c_void_p, c_void_p, c_int, c_i64, c_double, c_void_p, c_void_p]),
    "pcvl_mis_chain": (c_int, [c_void_p, c_i64, c_double, c_u64, c_u64, c_void_p, c_void_p, c_void_p]),
    "pcvl_mis_seq_to_occ": (c_int, [c_void_p, c_i64, c_int, c_int, c_void_p, c_void_p]),
    "pcvl_mis_occ_to_seq": (c_int, [c_void_p, c_i64, c_int, c_int, c_void_p, c_void_p]),
    "pcvl_mis_gather": (c_int, [c_void_p, c_void_p, c_i64, c_int, c_void_p, c_void_p, c_void_p]),
    "pcvl_probe_fp64_fma": (c_int, [c_int, c_i64, c_void_p, ctypes.POINTER(c_i64), c_void_p]),
    "pcvl_probe_read": (c_int, [c_void_p, c_u64, c_int, c_int, c_void_p, c_void_p]),
    "pcvl_probe_xcc": (c_int, [c_int, c_void_p, c_void_p]),
    "pcvl_perm_graph_create": (c_int, [c_void_p, c_void_p, c_void_p, c_int, ctypes.POINTER(c_void_p)]),
    "pcvl_slos_graph_create": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_size_t, c_int,
                                       ctypes.POINTER(c_void_p)]),
    "pcvl_graph_launch": (c_int, [c_void_p, c_void_p]),
    "pcvl_graph_destroy": (c_int, [c_void_p]),
    "pcvl_probe_fp64_mix": (c_int, [c_int, c_i64, c_int, c_void_p, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64),
                                    c_void_p]),
    "pcvl_gurvits_batch": (c_int, [c_void_p, P(c_i32), P(c_i64), c_i64, c_i64, c_u64, c_u64, c_void_p, c_void_p,
                                   c_void_p]),
    "pcvl_mesh_adjoint": (c_int, [P(c_i32), P(c_i32), P(c_i32), P(c_double), c_int, c_int, c_void_p, c_int, c_int,
                                  P(c_i32), P(c_i32), c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pcvl_mesh_plan_create": (c_int, [P(c_i32), P(c_i32), P(c_i32), P(c_double), c_int, c_int, c_int, P(c_i32),
                                      P(c_i32), c_int, ctypes.POINTER(c_void_p)]),
    "pcvl_mesh_plan_execute": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pcvl_mesh_plan_destroy": (c_int, [c_void_p]),
    "pcvl_mesh_plan_set_weights": (c_int, [c_void_p, P(ctypes.c_double)]),
    "pcvl_cc_samples": (c_int, [c_void_p, c_int, P(c_i32), c_i64, c_u64, c_u64, c_void_p, c_void_p, c_size_t,
                                c_void_p]),
    "pcvl_cc_samples_multi": (c_int, [c_void_p, c_int, P(c_i32), P(c_i64), c_int, c_u64, c_u64, c_void_p, c_void_p]),
}


def load(require_gpu: bool = True):
    """Load the library (once).  Raises NativeUnavailable if it is missing."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise NativeUnavailable(
                    f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                    f"or `make -C perceval_amd/csrc`")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    if require_gpu:
        import torch
        if not torch.cuda.is_available():
            raise NativeUnavailable("no GPU visible: the simulation core runs only on the GPU")
    return _lib


def check(rc: int, what: str):
    if rc != 0:
        msg = _lib.pcvl_last_error().decode(errors="replace") if _lib is not None else ""
        if rc == -1:
            raise ValueError(f"{what}: {msg}")
        raise NativeError(f"{what} failed ({rc}): {msg}")


def stream_handle(stream=None) -> int:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def ptr(t) -> int:
    return int(t.data_ptr())
