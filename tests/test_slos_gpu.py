"""GPU parity of the SLOS kernels and the Fock ranking against the oracle.

Bars: Fock ranks/states bit-exact; SLOS probabilities within 1e-12 absolute of
the CPU oracle; at the full m=24, n=10 size (92,561,040 outputs), the sum rule
and permanents of sampled outputs."""
import math

import numpy as np
import pytest

import oracle
from perceval_amd.utils.workloads import haar_unitary

pytestmark = pytest.mark.gpu

ABS_TOL = 1e-12


def _say(msg):
    print(f"[slos-gpu] {msg}", flush=True)


@pytest.mark.parametrize("m,n", [(1, 0), (1, 3), (2, 5), (3, 2), (5, 4), (8, 4), (12, 6), (24, 3)])
def test_fock_rank_unrank_exact(gpu, m, n):
    from perceval_amd.slos import fock_rank, fock_unrank
    states = oracle.fock_enumerate(m, n)
    ranks = fock_rank(states, nmax=n)
    assert np.array_equal(ranks, np.arange(len(states)))
    back = fock_unrank(np.arange(len(states)), m, n)
    assert np.array_equal(back, states)


def test_fock_rank_unrank_m24_n10_sampled(gpu):
    from perceval_amd.slos import fock_rank, fock_unrank
    total = oracle.fock_count(24, 10)
    rng = np.random.RandomState(0)
    ranks = np.concatenate([rng.randint(0, total, 4096), [0, 1, total - 2, total - 1]]).astype(np.int64)
    states = fock_unrank(ranks, 24, 10)
    ref = np.array([oracle.fock_unrank(int(r), 24, 10) for r in ranks], dtype=np.int32)
    assert np.array_equal(states, ref)
    assert np.array_equal(fock_rank(states, nmax=10), ranks)
    # invalid inputs are flagged, not wrapped
    bad = fock_unrank(np.array([-1, total], dtype=np.int64), 24, 10)
    assert (bad == -1).all()


KERNELS = (1, 2, 3, 4, 5, 6, 7)   # per state; incremental runs; 32-bit ranks; two-phase gather (4 / 8 loads
#                                   in flight); successor runs of 3 (compact slots) / 2 (full slots) with a
#                                   transposed gather


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("inp", [[1, 1, 1, 1, 0, 0, 0, 0], [2, 0, 1, 0, 0, 1, 0, 0], [0, 0, 0, 0, 0, 0, 0, 4]])
def test_slos_config1_vs_oracle(gpu, inp, kernel):
    from perceval_amd.slos import COEFS, PROBS, SlosPlan
    U = haar_unitary(8, seed=42)
    plan = SlosPlan(inp)
    plan.set_kernel(kernel)
    assert plan.path == oracle.slos_path(inp)
    probs = plan.run(U, PROBS).cpu().numpy()
    coefs = plan.run(U, COEFS).cpu().numpy()
    ref_c = oracle.slos_coefs(U, inp)
    ref_p = oracle.slos_probs(U, inp)
    assert len(probs) == 330
    assert np.max(np.abs(probs - ref_p)) < ABS_TOL
    assert np.max(np.abs(coefs - ref_c)) < ABS_TOL
    plan.close()


def test_slos_reference_unitary_u_random_8(gpu):
    import os
    from perceval_amd.slos import PROBS, SlosPlan
    U = np.load(os.path.join(os.path.dirname(__file__), "golden", "u_random_8.npy"))
    inp = [1, 1, 1, 1, 0, 0, 0, 0]
    p = SlosPlan(inp).run(U, PROBS).cpu().numpy()
    exact = np.array(list(oracle.distribution_exact(U, inp).values()))
    assert np.max(np.abs(p - exact)) < ABS_TOL


@pytest.mark.parametrize("kernel", KERNELS)
def test_slos_m12_n6_full_vs_oracle(gpu, kernel):
    from perceval_amd.slos import PROBS, SlosPlan
    U = haar_unitary(12, seed=7)
    inp = [1] * 6 + [0] * 6
    plan = SlosPlan(inp)
    plan.set_kernel(kernel)
    p = plan.run(U, PROBS).cpu().numpy()
    ref = oracle.slos_probs(U, inp, "gather")
    assert len(p) == 12376
    assert np.max(np.abs(p - ref)) < ABS_TOL


@pytest.mark.parametrize("inp", [[1, 1, 0, 2, 0, 0, 1, 0, 3, 0, 0, 1], [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 9],
                                 [5, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0], [1] * 12])
def test_slos_kernels_agree(gpu, inp):
    """Every kernel on bunched inputs (long runs through the last mode)."""
    from perceval_amd.slos import COEFS, SlosPlan
    U = haar_unitary(12, seed=sum(inp))
    plan = SlosPlan(inp)
    got = {}
    for kernel in KERNELS:
        plan.set_kernel(kernel)
        got[kernel] = plan.run(U, COEFS).cpu().numpy()
    ref = oracle.slos_coefs(U, inp)
    for kernel in KERNELS:
        assert np.max(np.abs(got[kernel] - ref)) < ABS_TOL, kernel
        assert np.max(np.abs(got[kernel] - got[1])) < 1e-14, kernel
    assert np.array_equal(got[1], got[3])   # same walk, same summation order
    assert np.max(np.abs(got[1] - got[4])) == 0.0 and np.max(np.abs(got[1] - got[5])) == 0.0
    for kernel in (6, 7):
        assert np.array_equal(got[4], got[kernel]), kernel   # same slots and summation order
    plan.close()


@pytest.mark.parametrize("inp", [[1] * 8 + [0] * 8, [2, 0, 3, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0]])
def test_slos_successor_kernels_bit_identical(gpu, inp):
    """The successor kernels (6, 7) against the two-phase kernel, on every
    coefficient and probability of a 490,314-state layer (a last tile of 10
    states, successor runs cut by the end of the layer)."""
    from perceval_amd.slos import COEFS, PROBS, SlosPlan
    U = haar_unitary(16, seed=7)
    plan = SlosPlan(inp)
    assert plan.count == 490314 and plan.count % 64 == 10
    out = {}
    for kernel in (4, 6, 7):
        plan.set_kernel(kernel)
        out[kernel] = (plan.run(U, COEFS).cpu().numpy(), plan.run(U, PROBS).cpu().numpy())
    for kernel in (6, 7):
        assert np.array_equal(out[4][0], out[kernel][0]) and np.array_equal(out[4][1], out[kernel][1]), kernel
    plan.close()


def test_slos_set_kernel_rejects_bad_values(gpu):
    from perceval_amd.slos import SlosPlan
    plan = SlosPlan([1, 1, 0])
    with pytest.raises(ValueError):
        plan.set_kernel(8)
    plan.close()


@pytest.mark.parametrize("kernel", KERNELS)
def test_slos_m24_n10_full_size(gpu, kernel):
    import torch
    from perceval_amd.slos import PROBS, SlosPlan
    U = haar_unitary(24, seed=24)
    inp = [1] * 10 + [0] * 14
    plan = SlosPlan(inp)
    plan.set_kernel(kernel)
    _say(f"kernel {kernel} plan: {plan.count} outputs, workspace {plan.work_bytes / 2**20:.0f} MiB")
    p = plan.run(U, PROBS)
    torch.cuda.synchronize()
    _say("all_prob done")
    assert p.numel() == 92561040
    total = float(p.sum().item())
    assert abs(total - 1.0) < 1e-10
    assert float(p.min().item()) >= 0.0
    rng = np.random.RandomState(1)
    idx = np.concatenate([[0, 1, p.numel() - 1], rng.randint(0, p.numel(), 61)])
    got = p[torch.from_numpy(idx).cuda()].cpu().numpy()
    for i, g in zip(idx, got):
        s = oracle.fock_unrank(int(i), 24, 10)
        M = oracle.submatrix(U, inp, s)
        ref = abs(oracle.perm(M)) ** 2 / oracle.prodnfact(s)
        assert abs(g - ref) < ABS_TOL, (i, s, g, ref)
    # the largest output probabilities agree with permanents too
    top = torch.topk(p, 8).indices.cpu().numpy()
    for i in top:
        s = oracle.fock_unrank(int(i), 24, 10)
        ref = abs(oracle.perm(oracle.submatrix(U, inp, s))) ** 2 / oracle.prodnfact(s)
        assert abs(float(p[int(i)].item()) - ref) < ABS_TOL
    plan.close()


def test_slos_vacuum_and_one_photon(gpu):
    from perceval_amd.slos import COEFS, PROBS, SlosPlan
    U = haar_unitary(5, seed=3)
    p0 = SlosPlan([0] * 5).run(U, PROBS).cpu().numpy()
    assert p0.tolist() == [1.0]
    c1 = SlosPlan([0, 0, 1, 0, 0]).run(U, COEFS).cpu().numpy()
    assert np.allclose(c1, U[:, 2], atol=0, rtol=0)


def test_slos_rejects_bad_inputs(gpu):
    from perceval_amd.slos import SlosPlan
    with pytest.raises(ValueError):
        SlosPlan([-1, 2])
    with pytest.raises(ValueError):
        SlosPlan([1] * 65)


def _tree_layer_launches(inputs):
    """Layer launches of the _Path tree (perceval/backends/_slos.py:61-86),
    restated: one per node that has children, one per input that ends at a
    node of depth >= 1."""
    def walk(k, targets):
        here = [t for t in targets if sum(t) == 0]
        rest = [list(t) for t in targets if sum(t) > 0]
        n = (1 if (k > 0 and rest) else 0) + (len(here) if k > 0 else 0)
        while rest:
            counts = [sum(col) for col in zip(*rest)]
            best = counts.index(max(counts))
            child, nxt = [], []
            for t in rest:
                if t[best]:
                    t[best] -= 1
                    child.append(t)
                else:
                    nxt.append(t)
            n += walk(k + 1, child)
            rest = nxt
        return n
    return walk(0, [list(s) for s in inputs])


@pytest.mark.parametrize("m,inputs", [
    (4, [[1, 1, 0, 0], [1, 0, 1, 0], [0, 1, 1, 0], [1, 1, 1, 0], [2, 0, 0, 0], [0, 0, 0, 0], [1, 0, 0, 0]]),
    (12, [[1] * 6 + [0] * 6] + [[0 if j == i else 1 for j in range(6)] + [0] * 6 for i in range(6)]),
    (6, [[3, 0, 0, 0, 0, 0], [2, 1, 0, 0, 0, 0], [0, 0, 0, 0, 0, 4], [1, 1, 1, 1, 1, 1]]),
])
def test_slos_tree_vs_oracle_and_single_plans(gpu, m, inputs):
    from perceval_amd.slos import COEFS, PROBS, SlosPlan, SlosTree
    U = haar_unitary(m, seed=m + len(inputs))
    tree = SlosTree(inputs)
    assert tree.layer_launches == _tree_layer_launches(inputs)
    probs = [p.cpu().numpy() for p in tree.run(U, PROBS)]
    coefs = [c.cpu().numpy() for c in tree.run(U, COEFS)]
    for inp, p, c in zip(inputs, probs, coefs):
        assert np.max(np.abs(p - oracle.slos_probs(U, inp))) < ABS_TOL, inp
        assert np.max(np.abs(c - oracle.slos_coefs(U, inp))) < ABS_TOL, inp
        single = SlosPlan(inp).run(U, PROBS).cpu().numpy()
        assert np.max(np.abs(p - single)) < 1e-13, inp
    tree.close()


def test_slos_backend_preprocess_shares_a_tree(gpu):
    from perceval_amd.backends import SLOSBackend
    from perceval_amd.components import Unitary
    from perceval_amd.slos import SlosTree
    from perceval_amd.utils.states import BasicState
    m = 8
    U = haar_unitary(m, seed=81)
    inputs = [BasicState([1, 1, 1, 1, 0, 0, 0, 0]), BasicState([1, 1, 1, 0, 0, 0, 0, 0]),
              BasicState([0, 1, 1, 1, 0, 0, 0, 0]), BasicState([1, 1, 0, 1, 0, 0, 0, 0])]
    b = SLOSBackend()
    b.set_circuit(Unitary(U))
    assert b.preprocess(inputs) is True
    assert all(isinstance(b._state_mapping[s], tuple) and isinstance(b._state_mapping[s][0], SlosTree)
               for s in inputs)
    assert b.preprocess(inputs) is False
    for s in inputs:
        b.set_input_state(s)
        assert np.max(np.abs(np.array(b.all_prob()) - oracle.slos_probs(U, list(s)))) < ABS_TOL
    # a new circuit of the same size keeps the tree and recomputes
    U2 = haar_unitary(m, seed=82)
    b.set_circuit(Unitary(U2))
    b.set_input_state(inputs[2])
    assert np.max(np.abs(np.array(b.all_prob()) - oracle.slos_probs(U2, list(inputs[2])))) < ABS_TOL


@pytest.mark.parametrize("m,n,masks,at_least", [
    (6, 4, ["  0  1"], None),
    (6, 4, ["1     ", "  2   "], None),
    (6, 4, ["   1 1"], [3]),
    (5, 3, ["*****"], None),
    (7, 5, ["0000000"], None),
    (8, 6, ["1 1 1 1 ", "  00  0 ", " 2      "], [0, 2]),
])
def test_fock_mask_select_matches_host_fsmask(gpu, m, n, masks, at_least):
    from perceval_amd.slos import fock_mask_select
    from perceval_amd.utils.mask import FSMask
    states = oracle.fock_enumerate(m, n)
    fm = FSMask(m, n, masks, at_least)
    ref = np.array([i for i, s in enumerate(states) if fm.match(list(s))], dtype=np.int64)
    got = fock_mask_select(m, n, masks, at_least).cpu().numpy()
    assert np.array_equal(got, ref)


def test_fock_mask_select_m24_n10_heralds(gpu):
    """Heralds: modes 20..23 hold exactly 0, 1, 0, 1 photons; the kept states
    are the 8-photon states of the other 20 modes, C(27, 8) of them."""
    import torch
    from perceval_amd.slos import PROBS, SlosPlan, fock_mask_select, fock_unrank, gather_f64
    mask = " " * 20 + "0101"
    idx = fock_mask_select(24, 10, [mask])
    assert idx.numel() == math.comb(27, 8)
    assert bool((idx[1:] > idx[:-1]).all())
    rng = np.random.RandomState(3)
    pick = np.concatenate([[0, idx.numel() - 1], rng.randint(0, idx.numel(), 200)])
    states = fock_unrank(idx[torch.from_numpy(pick).cuda()].cpu().numpy(), 24, 10)
    assert (states[:, 20:] == np.array([0, 1, 0, 1])).all() and (states.sum(axis=1) == 10).all()
    # the kept probabilities against permanents
    U = haar_unitary(24, seed=24)
    inp = [1] * 10 + [0] * 14
    p = SlosPlan(inp).run(U, PROBS)
    kept = gather_f64(p, idx).cpu().numpy()
    for k in pick[:20]:
        s = oracle.fock_unrank(int(idx[int(k)].item()), 24, 10)
        ref = abs(oracle.perm(oracle.submatrix(U, inp, s))) ** 2 / oracle.prodnfact(s)
        assert abs(kept[k] - ref) < ABS_TOL
