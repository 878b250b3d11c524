"""pcvl_perm_batch_noalloc (no device allocation; plan block in the library's
per-GPU scratch array) against pcvl_perm_batch, through the C ABI: two calls
in a row, the second with a smaller batch, bit-identical results (the same
plan, kernels and summation order), and the size limit refused cleanly."""
import ctypes

import numpy as np
import pytest

import oracle
from perceval_amd.permanent import pack
from perceval_amd.utils.workloads import submatrix_batch

pytestmark = pytest.mark.gpu

P = ctypes.POINTER


def _lib():
    from perceval_amd import _native
    return _native.load()


def _run(lib, name, mats):
    import torch
    from perceval_amd import _native
    orders, flat = pack(mats)
    offsets = np.zeros(len(orders), dtype=np.int64)
    offsets[1:] = np.cumsum(orders.astype(np.int64) ** 2)[:-1]
    d = torch.from_numpy(flat).cuda()
    out = torch.full((len(orders),), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda")
    rc = getattr(lib, name)(_native.ptr(d), orders.ctypes.data_as(P(ctypes.c_int32)),
                            offsets.ctypes.data_as(P(ctypes.c_int64)), len(orders), _native.ptr(out),
                            _native.stream_handle())
    return rc, out.cpu().numpy()


def test_noalloc_matches_perm_batch_two_calls(gpu):
    lib = _lib()
    print(f"[no-alloc] scratch limit {lib.pcvl_perm_noalloc_max_bytes()} bytes", flush=True)
    _, _, big = submatrix_batch(600, 4, 22, seed=31)          # mixed orders, small (< 9) and wave kernels
    _, _, small = submatrix_batch(40, 12, 16, seed=32)
    for mats in (big, small):
        print(f"[no-alloc] batch of {len(mats)} matrices", flush=True)
        rc_a, ref = _run(lib, "pcvl_perm_batch", mats)
        rc_b, got = _run(lib, "pcvl_perm_batch_noalloc", mats)
        assert rc_a == 0 and rc_b == 0
        assert np.array_equal(got, ref)
        pick = list(range(0, len(mats), max(1, len(mats) // 16)))
        exact = oracle.perm_batch([mats[i] for i in pick], precise=True)
        assert np.max(np.abs(got[pick] - exact) / np.abs(exact)) < 1e-10


def test_noalloc_refuses_a_batch_beyond_its_scratch(gpu):
    lib = _lib()
    limit = lib.pcvl_perm_noalloc_max_bytes()
    # order-28 matrices at the default chunk need 2^27 / 2^11 / 64 = 1024 work
    # items (24 bytes) and partial sums (16 bytes) each: enough of them exceed
    # the limit; the call is refused before anything runs (no device reads)
    count = limit // (1024 * 40) + 1
    orders = np.full(count, 28, dtype=np.int32)
    offsets = np.zeros(count, dtype=np.int64)
    rc = lib.pcvl_perm_batch_noalloc(None, orders.ctypes.data_as(P(ctypes.c_int32)),
                                     offsets.ctypes.data_as(P(ctypes.c_int64)), count, None, None)
    assert rc == -4
    from perceval_amd import _native
    assert b"scratch" in lib.pcvl_last_error()
    del _native
