"""Host-side logic that needs no GPU: Fock types, circuit unitaries, masks,
the backend registry and error conventions, and the C-ABI library surface."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import perceval_amd as pcvl
from perceval_amd import (BS, BSDistribution, BackendFactory, BasicState, Circuit, FSMask, PERM, PS, StateVector,
                          Unitary, clements_mesh)
from perceval_amd.utils.states import fock_space

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fockstate_basics():
    s = BasicState("|2,0,1>")
    assert s.m == 3 and s.n == 3 and list(s) == [2, 0, 1] and str(s) == "|2,0,1>"
    assert s.prodnfact() == 2
    assert BasicState([2, 0, 1]) == s and hash(BasicState([2, 0, 1])) == hash(s)
    assert (BasicState([1]) * BasicState([0, 2])) == BasicState([1, 0, 2])
    with pytest.raises(ValueError):
        BasicState([1, -1])


def test_statevector_and_distribution():
    sv = 0.5 * StateVector([1, 0]) + 0.5j * StateVector([0, 1])
    assert sv[BasicState([1, 0])] == 0.5 and sv[[0, 1]] == 0.5j
    sv2 = StateVector(sv).normalize()
    assert sv2.norm() == pytest.approx(1.0)
    d = BSDistribution()
    d.add(BasicState([1, 0]), 0.25)
    d.add([1, 0], 0.25)
    assert d[BasicState([1, 0])] == 0.5 and d[[0, 1]] == 0.0


def test_fock_space_order_pinned():
    # tests/utils/test_statevector.py:430-438
    got = [list(r) for n in range(3) for r in fock_space(3, n).tolist()]
    assert got == [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [2, 0, 0], [1, 1, 0], [1, 0, 1], [0, 2, 0],
                   [0, 1, 1], [0, 0, 2]]
    assert len(list(pcvl.allstate_iterator(BasicState([1, 0, 0, 0])))) == 4


def test_mask_semantics():
    # tests/utils/test_mask.py:33-45
    mask = FSMask(6, 4, ["    11"])
    assert mask.match(BasicState("|0,0,1,1,1,1>"))
    assert not mask.match(BasicState("|0,0,1,1,1,0>"), False)
    assert mask.match(BasicState("|0,0,1,1,1,0>"), True)
    mask = FSMask(6, 4, ["   011", "   110"])
    assert mask.match(BasicState("|0,0,1,0,1,1>"))
    assert mask.match(BasicState("|0,0,1,1,1,0>"))
    assert not mask.match(BasicState("|0,0,1,1,1,1>"))


@pytest.mark.parametrize("masks,at_least,m,n", [
    (["  01 1"], [], 6, 4), (["1 2   ", "0 0  1"], [0, 2], 6, 5), (["   011", "   110"], [], 6, 4),
    (["1111"], [], 4, 5), (["21  "], [1], 4, 5), (["0101"], [0, 2], 4, 3)])
def test_mask_keeps_is_the_closure_of_the_matches(masks, at_least, m, n):
    """FSMask.keeps on k-photon states (mask set for n photons) is the set of
    states that some matching n-photon state contains: the matches
    (tests/utils/test_mask.py:33-45), closed under photon removal, layer by
    layer down to the vacuum.  Includes a mask no n-photon state matches
    (4 exact counts summing to 4 at n = 5): nothing is kept at any k."""
    from perceval_amd.utils.states import fock_space
    fm = FSMask(m, n, masks, at_least)
    layer = {tuple(r) for r in fock_space(m, n).tolist() if fm.match(list(r))}
    for k in range(n, -1, -1):
        got = {tuple(r) for r in fock_space(m, k).tolist() if fm.keeps(list(r))}
        assert got == layer, (masks, k)
        layer = {s[:j] + (s[j] - 1,) + s[j + 1:] for s in layer for j in range(m) if s[j] > 0}


def test_mask_counts_above_nine():
    """Herald masks write a count c as chr(0x30 + c), up to 32 photons
    (perceval/simulators/simulator.py:650-662): ':' is 10, 'P' is 32."""
    from perceval_amd.slos import mask_conditions
    assert mask_conditions([" :*0P"], 5).tolist() == [[-1, 10, -1, 0, 32]]
    mask = FSMask(3, 12, [" : "])
    assert mask.match(BasicState([1, 10, 1])) and not mask.match(BasicState([0, 9, 3]))
    with pytest.raises(ValueError):
        mask_conditions(["Q  "], 3)
    with pytest.raises(ValueError):
        FSMask(3, 2, ["/  "])


def test_bs_conventions_match_reference_formulas():
    th, tl, bl, tr, br = 1.1, 0.3, -0.4, 0.7, 0.2
    c, s = math.cos(th / 2), math.sin(th / 2)
    U = BS(th, tl, bl, tr, br).compute_unitary()
    exp = np.array([[np.exp(1j * (tl + tr)) * c, 1j * np.exp(1j * (tr + bl)) * s],
                    [1j * np.exp(1j * (tl + br)) * s, np.exp(1j * (bl + br)) * c]])
    assert np.allclose(U, exp)
    assert np.allclose(BS.H().compute_unitary(), np.array([[1, 1], [1, -1]]) / math.sqrt(2))
    assert np.allclose(BS.Ry().compute_unitary(), np.array([[1, -1], [1, 1]]) / math.sqrt(2))
    assert np.allclose(PS(0.5).compute_unitary(), [[np.exp(0.5j)]])
    assert np.allclose(PERM([2, 0, 1]).compute_unitary()[:, 0], [0, 0, 1])


def test_circuit_composition_left_multiplies():
    a, b = BS.H(), BS(0.3)
    c = Circuit(2) // a // b
    assert np.allclose(c.compute_unitary(), b.compute_unitary() @ a.compute_unitary())
    c3 = Circuit(3) // (1, PS(0.2)) // (1, BS.H())
    E1 = np.eye(3, dtype=complex)
    E1[1, 1] = np.exp(0.2j)
    E2 = np.eye(3, dtype=complex)
    E2[1:, 1:] = BS.H().compute_unitary()
    assert np.allclose(c3.compute_unitary(), E2 @ E1)


def test_clements_mesh_shape_and_unitarity():
    mesh = clements_mesh(12)
    params = mesh.get_parameters()
    assert len(params) == 132
    rng = np.random.RandomState(0)
    for p in params:
        p.set_value(rng.uniform(0, 2 * math.pi))
    U = mesh.compute_unitary()
    assert np.allclose(U @ U.conj().T, np.eye(12), atol=1e-12)


def test_backend_registry_names():
    assert sorted(BackendFactory.list()) == ["CliffordClifford2017", "Naive", "NaiveApprox", "SLOS"]
    with pytest.warns(UserWarning):
        b = BackendFactory.get_backend("NoSuchBackend")
    assert b.name == "SLOS"


@pytest.mark.parametrize("backend_name", ["SLOS", "Naive", "CliffordClifford2017"])
def test_backend_wrong_size(backend_name):
    # tests/backends/test_backends.py:106-113
    backend = BackendFactory.get_backend(backend_name)
    with pytest.raises(AssertionError):
        backend.set_circuit(Circuit(2))
        backend.set_input_state(BasicState([1, 1, 1]))


@pytest.mark.parametrize("backend_name", ["SLOS", "Naive", "CliffordClifford2017"])
def test_input_before_circuit_asserts(backend_name):
    with pytest.raises(AssertionError):
        BackendFactory.get_backend(backend_name).set_input_state(BasicState([1, 0]))


def _declared_symbols():
    with open(os.path.join(ROOT, "include", "pcvl_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(pcvl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from perceval_amd import _native
    lib = _native.load(require_gpu=False)
    declared = _declared_symbols()
    assert len(declared) >= 10
    for name in declared:
        assert hasattr(lib, name), name
        assert name in _native.SIGNATURES, f"{name} has no ctypes signature"
    assert lib.pcvl_abi_version() >= 100


def test_no_gpu_means_loud_failure():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from perceval_amd import _native
    from perceval_amd.permanent import permanents
    with pytest.raises(_native.NativeUnavailable):
        permanents([np.eye(3)])
    b = BackendFactory.get_backend("SLOS")
    b.set_circuit(BS.H())
    with pytest.raises(_native.NativeUnavailable):
        b.set_input_state(BasicState([1, 1]))


def test_entry_points_reject_bad_arguments_before_touching_the_device():
    """Argument checks return -1 with a message and never reach HIP, so they run
    without a GPU (the header's error convention)."""
    import ctypes
    from perceval_amd import _native
    lib = _native.load(require_gpu=False)

    def err(rc):
        assert rc == -1, rc
        return lib.pcvl_last_error().decode()

    cond = (ctypes.c_int32 * 4)(-1, 0, 1, -1)
    out = ctypes.c_void_p(1)
    assert "masks" in err(lib.pcvl_fock_mask_select(4, 2, 0, cond, 0, None, out, out, None))
    assert "mode count" in err(lib.pcvl_fock_mask_select(257, 2, 0, cond, 1, None, out, out, None))
    assert "photon count" in err(lib.pcvl_fock_mask_select(4, 64, 0, cond, 1, None, out, out, None))
    assert "mask photon count" in err(lib.pcvl_fock_mask_select(4, 2, -1, cond, 1, None, out, out, None))
    occ = (ctypes.c_int32 * 4)(1, 1, 0, 0)
    plan = ctypes.c_void_p()
    assert "mask photon count" in err(lib.pcvl_slos_plan_create_masks(4, occ, 1, cond, None, 1, ctypes.byref(plan)))
    n = (ctypes.c_int32 * 1)(49)
    offs = (ctypes.c_int64 * 1)(0)
    assert "order" in err(lib.pcvl_gurvits_batch(out, n, offs, 1, 100, 1, 0, out, out, None))
    assert "iterations" in err(lib.pcvl_gurvits_batch(out, n, offs, 1, 0, 1, 0, out, out, None))
    nbytes = ctypes.c_size_t()
    assert "n outside" in err(lib.pcvl_cc_workspace_bytes(31, 10, ctypes.byref(nbytes)))
    assert lib.pcvl_cc_workspace_bytes(30, 10, ctypes.byref(nbytes)) == 0 and nbytes.value > 0
    many = (ctypes.c_int32 * 2)(16, 15)
    assert "more than 30 photons" in err(lib.pcvl_cc_samples(out, 2, many, 1, 1, 0, out, out, 1 << 20, None))
    inputs = (ctypes.c_int32 * 3)(1, 1, 0)
    counts = (ctypes.c_int64 * 1)(-1)
    assert "negative" in err(lib.pcvl_cc_samples_multi(out, 3, inputs, counts, 1, 1, 0, out, None))
    assert "mode count" in err(lib.pcvl_cc_samples_multi(out, 0, inputs, counts, 1, 1, 0, out, None))
    plan = ctypes.c_void_p()
    assert "no input" in err(lib.pcvl_slos_tree_create(3, inputs, 0, ctypes.byref(plan)))
    bad = (ctypes.c_int32 * 3)(1, -1, 0)
    assert "negative" in err(lib.pcvl_slos_tree_create(3, bad, 1, ctypes.byref(plan)))
    assert "fused sampler" in err(lib.pcvl_mis_sample_fused(out, out, out, out, 10, 65, 3, 64, 1, 0, out, out, None))
    assert "fused sampler" in err(lib.pcvl_mis_sample_fused(out, out, out, out, 10, 8, 3, 12, 1, 0, out, out, None))
    f = ctypes.c_int64()
    assert "negative" in err(lib.pcvl_mlp_wgrad_workspace(-1, 3, 3, ctypes.byref(f)))
    assert "workspace" in err(lib.pcvl_mlp_wgrad(out, 3, out, 3, 100, 3, 3, out, out, out, 0, None))
    assert "probe" in err(lib.pcvl_probe_fp64_fma(0, 10, out, ctypes.byref(f), None))
    assert "1 <= n <= 28" in err(lib.pcvl_mis_permanents(out, 10, out, 29, out, 5, out, None))
    assert "2^36" in err(lib.pcvl_mis_chain(out, (1 << 36) + 1, 0.0, 1, 0, out, out, None))
    assert "2^36" in err(lib.pcvl_mis_chain(out, -1, 0.0, 1, 0, out, out, None))
    assert "null" in err(lib.pcvl_mis_chain(None, 10, 0.0, 1, 0, out, out, None))
    # adjoint plans: at most 512 targets; tables beyond the kernel's LDS give -4 (the caller splits the targets)
    one = (ctypes.c_int32 * 4)(1, 0, 0, 0)
    tgt = (ctypes.c_int32 * (4 * 513))(*([1, 0, 0, 0] * 513))
    plan = ctypes.c_void_p()
    assert "[1, 512]" in err(lib.pcvl_mesh_plan_create(None, None, None, None, 0, 4, 0, one, tgt, 513,
                                                       ctypes.byref(plan)))
    rc = lib.pcvl_mesh_plan_create(None, None, None, None, 0, 4, 20000, one, tgt, 1, ctypes.byref(plan))
    assert rc == -4 and "LDS" in lib.pcvl_last_error().decode() and not plan.value


def test_output_submatrices_match_the_per_state_construction():
    """The batched gather of the Naive submatrices (rows: output photons'
    modes, columns: input photons' modes, ascending with repetition,
    _naive.py:51-68) equals np.ix_ per state, norms sqrt(prod s! prod t!)."""
    from perceval_amd.permanent import output_submatrices
    from perceval_amd.utils.states import fock_space
    rng = np.random.default_rng(5)
    U = rng.normal(size=(5, 5)) + 1j * rng.normal(size=(5, 5))
    inp = [2, 0, 1, 0, 1]
    occ = fock_space(5, 4)
    mats, norms = output_submatrices(U, occ, inp)
    cols = [0, 0, 2, 4]
    for k, s in enumerate(occ.tolist()):
        rows = [i for i, v in enumerate(s) for _ in range(v)]
        assert np.array_equal(mats[k], U[np.ix_(rows, cols)])
        assert norms[k] == math.sqrt(np.prod([math.factorial(v) for v in s]) * 2)
