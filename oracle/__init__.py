"""CPU oracle for the Perceval simulation-core hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or as the timed
CPU baseline).  The product package ``perceval_amd`` never imports it.

Two layers:

* ``liboracle.so`` (``pcvl_oracle.c``, built by ``make -C oracle``): Glynn and
  Ryser permanents (long double and chunked double), descending-lex Fock
  rank/unrank/enumeration, the SLOS layer recursion (scatter and gather forms)
  and the Clifford & Clifford 2017 sampler.  Each C function cites the
  reference call site it restates.
* Pure-Python brute force for tiny cases (permutation-sum permanent, sorted
  enumeration of compositions, exact output distributions by permanents), used
  to pin the C layer.

Every function here is pinned in ``tests/test_oracle.py``; see the header of
``pcvl_oracle.c`` for the mapping to reference files.
"""
from __future__ import annotations

import ctypes
import itertools
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PCVL_ORACLE_LIB selects another build of the same source (the sanitizer
# build of `make -C oracle asan-test`)
_LIB_PATH = os.environ.get("PCVL_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build(force: bool = False) -> str:
    """Compile liboracle.so in place (``make -C oracle``)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_perm_glynn_ld.argtypes = [_dp, ctypes.c_int, _dp]
        L.oracle_perm_ryser_ld.argtypes = [_dp, ctypes.c_int, _dp]
        L.oracle_perm_glynn_d.argtypes = [_dp, ctypes.c_int, _dp]
        L.oracle_perm_glynn_ld_mt.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int]
        L.oracle_perm_batch.argtypes = [_dp, _i32p, _i64p, ctypes.c_int64, _dp, ctypes.c_int, ctypes.c_int]
        L.oracle_binom.argtypes = [ctypes.c_int64, ctypes.c_int64]
        L.oracle_binom.restype = ctypes.c_uint64
        L.oracle_fock_count.argtypes = [ctypes.c_int, ctypes.c_int]
        L.oracle_fock_count.restype = ctypes.c_uint64
        L.oracle_fock_rank.argtypes = [_i32p, ctypes.c_int]
        L.oracle_fock_rank.restype = ctypes.c_int64
        L.oracle_fock_unrank.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, _i32p]
        L.oracle_fock_enumerate.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, ctypes.c_int64]
        L.oracle_slos_path.argtypes = [_i32p, ctypes.c_int, _i32p]
        L.oracle_slos_scatter.argtypes = [_dp, ctypes.c_int, _i32p, _dp]
        L.oracle_slos_gather.argtypes = [_dp, ctypes.c_int, _i32p, _dp, ctypes.c_int]
        L.oracle_slos_probs.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _i32p, _dp]
        L.oracle_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_uniform.restype = ctypes.c_double
        L.oracle_cc_samples.argtypes = [_dp, ctypes.c_int, _i32p, ctypes.c_int64, ctypes.c_uint64,
                                        ctypes.c_uint64, _i32p, ctypes.c_int]
        L.oracle_gurvits.argtypes = [_dp, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint64, _dp]
        L.oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _cd(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _ci32(a: np.ndarray):
    return a.ctypes.data_as(_i32p)


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


# ---------------------------------------------------------------- permanents

def perm(a, method: str = "glynn_ld") -> complex:
    """Permanent of a square complex matrix: 'glynn_ld', 'ryser_ld' or 'glynn_d'."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    n = a.shape[0]
    assert a.shape == (n, n)
    out = np.zeros(2)
    fn = {"glynn_ld": lib().oracle_perm_glynn_ld, "ryser_ld": lib().oracle_perm_ryser_ld,
          "glynn_d": lib().oracle_perm_glynn_d}[method]
    _check(fn(_cd(a.view(np.float64)), n, _cd(out)), "perm")
    return complex(out[0], out[1])


def perm_large(a, threads: int = 0) -> complex:
    """Long-double Glynn of one matrix with its Gray-code chunks split over
    threads (oracle_perm_glynn_ld_mt): for orders 25..28."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    n = a.shape[0]
    assert a.shape == (n, n)
    out = np.zeros(2)
    _check(lib().oracle_perm_glynn_ld_mt(_cd(a.view(np.float64)), n, _cd(out), threads), "perm_large")
    return complex(out[0], out[1])


def perm_batch(mats: list, precise: bool = True, threads: int = 0) -> np.ndarray:
    """Permanents of a list of square matrices (POSIX threads over matrices)."""
    ns = np.array([m.shape[0] for m in mats], dtype=np.int32)
    offs = np.zeros(len(mats), dtype=np.int64)
    if len(mats) > 1:
        offs[1:] = np.cumsum(ns.astype(np.int64) ** 2)[:-1]
    packed = np.concatenate([np.ascontiguousarray(m, dtype=np.complex128).ravel() for m in mats]) \
        if mats else np.zeros(0, np.complex128)
    out = np.zeros(2 * len(mats))
    _check(lib().oracle_perm_batch(_cd(packed.view(np.float64)), _ci32(ns), offs.ctypes.data_as(_i64p),
                                   len(mats), _cd(out), 1 if precise else 0, threads), "perm_batch")
    return out.view(np.complex128)


def perm_brute(a) -> complex:
    """Permutation-sum permanent (n <= 8), pure Python."""
    a = np.asarray(a, dtype=np.complex128)
    n = a.shape[0]
    assert n <= 8
    tot = 0j
    for p in itertools.permutations(range(n)):
        t = 1 + 0j
        for i in range(n):
            t *= a[i, p[i]]
        tot += t
    return complex(tot) if n else 1 + 0j


# --------------------------------------------------------------------- Fock

def fock_count(m: int, n: int) -> int:
    return int(lib().oracle_fock_count(m, n))


def fock_rank(occ) -> int:
    occ = np.ascontiguousarray(occ, dtype=np.int32)
    return int(lib().oracle_fock_rank(_ci32(occ), len(occ)))


def fock_unrank(rank: int, m: int, n: int) -> list:
    out = np.zeros(m, dtype=np.int32)
    _check(lib().oracle_fock_unrank(rank, m, n, _ci32(out)), "unrank")
    return out.tolist()


def fock_enumerate(m: int, n: int) -> np.ndarray:
    cnt = fock_count(m, n)
    out = np.zeros((cnt, m), dtype=np.int32)
    _check(lib().oracle_fock_enumerate(m, n, _ci32(out), cnt), "enumerate")
    return out


def fock_enumerate_brute(m: int, n: int) -> list:
    """All n-photon states over m modes sorted in descending lexicographic order."""
    states = [s for s in itertools.product(range(n + 1), repeat=m) if sum(s) == n]
    return sorted(states, reverse=True)


# --------------------------------------------------------------------- SLOS

def slos_path(in_occ) -> list:
    in_occ = np.ascontiguousarray(in_occ, dtype=np.int32)
    out = np.zeros(max(1, int(in_occ.sum())), dtype=np.int32)
    n = lib().oracle_slos_path(_ci32(in_occ), len(in_occ), _ci32(out))
    return out[:n].tolist()


def slos_coefs(U, in_occ, form: str = "scatter", threads: int = 0) -> np.ndarray:
    """Final-layer SLOS coefficients (complex, FSArray order)."""
    U = np.ascontiguousarray(U, dtype=np.complex128)
    in_occ = np.ascontiguousarray(in_occ, dtype=np.int32)
    m = len(in_occ)
    n = int(in_occ.sum())
    out = np.zeros(fock_count(m, n), dtype=np.complex128)
    if form == "scatter":
        rc = lib().oracle_slos_scatter(_cd(U.view(np.float64)), m, _ci32(in_occ), _cd(out.view(np.float64)))
    else:
        rc = lib().oracle_slos_gather(_cd(U.view(np.float64)), m, _ci32(in_occ), _cd(out.view(np.float64)),
                                      threads)
    _check(rc, "slos")
    return out


def slos_probs(U, in_occ, form: str = "scatter", threads: int = 0) -> np.ndarray:
    c = slos_coefs(U, in_occ, form, threads)
    in_occ = np.ascontiguousarray(in_occ, dtype=np.int32)
    m = len(in_occ)
    n = int(in_occ.sum())
    out = np.zeros(len(c))
    _check(lib().oracle_slos_probs(_cd(c.view(np.float64)), m, n, _ci32(in_occ), _cd(out)), "slos_probs")
    return out


def prodnfact(s) -> int:
    p = 1
    for v in s:
        p *= math.factorial(int(v))
    return p


def submatrix(U, in_occ, out_occ) -> np.ndarray:
    """U[out photons, in photons] with repeated rows/cols (perceval/backends/_naive.py:51-68)."""
    rows = [j for j, v in enumerate(out_occ) for _ in range(v)]
    cols = [i for i, v in enumerate(in_occ) for _ in range(v)]
    return np.asarray(U, dtype=np.complex128)[np.ix_(rows, cols)]


def amplitude_brute(U, in_occ, out_occ) -> complex:
    """perm(U_{out,in}) / sqrt(prod s! prod t!) by permutation sum."""
    if sum(in_occ) != sum(out_occ):
        return 0j
    M = submatrix(U, in_occ, out_occ)
    return perm_brute(M) / math.sqrt(prodnfact(in_occ) * prodnfact(out_occ))


def distribution_exact(U, in_occ) -> dict:
    """Exact output distribution {state tuple: probability} from long-double permanents."""
    m = len(in_occ)
    n = int(sum(in_occ))
    out = {}
    for s in fock_enumerate(m, n):
        M = submatrix(U, in_occ, s)
        a = perm(M) if n else 1 + 0j
        out[tuple(int(x) for x in s)] = abs(a) ** 2 / (prodnfact(in_occ) * prodnfact(s))
    return out


# ----------------------------------------------------------------- sampling

def uniform(seed: int, stream: int, ctr: int) -> float:
    return float(lib().oracle_uniform(seed, stream, ctr))


def cc_samples(U, in_occ, count: int, seed: int, first_sample: int = 0, threads: int = 0) -> np.ndarray:
    """Clifford & Clifford 2017 samples as a (count, m) int32 array of occupation vectors."""
    U = np.ascontiguousarray(U, dtype=np.complex128)
    in_occ = np.ascontiguousarray(in_occ, dtype=np.int32)
    m = len(in_occ)
    out = np.zeros((count, m), dtype=np.int32)
    _check(lib().oracle_cc_samples(_cd(U.view(np.float64)), m, _ci32(in_occ), count, seed, first_sample,
                                   _ci32(out), threads), "cc_samples")
    return out


def gurvits(a, iters: int, seed: int, stream: int = 0):
    """Gurvits estimate of perm(a): (complex mean, 99% error bound), samples
    drawn from the counter-based stream (seed, stream, s)."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    n = a.shape[0] if a.ndim == 2 else 0
    out = np.zeros(3, dtype=np.float64)
    _check(lib().oracle_gurvits(_cd(a), n, int(iters), seed, stream, _cd(out)), "oracle_gurvits")
    return complex(out[0], out[1]), float(out[2])


def max_threads() -> int:
    return int(lib().oracle_max_threads())
