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


KERNELS = (1, 4, 6, 8)   # per state; two-phase gather; successor runs of 3 with a transposed gather;
#                          prefix/suffix blocks


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
    for kernel in KERNELS:   # same parents, same summation order
        assert np.array_equal(got[4], got[kernel]), kernel
    plan.close()


@pytest.mark.parametrize("inp", [[1] * 8 + [0] * 8, [2, 0, 3, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0]])
def test_slos_successor_and_block_kernels_bit_identical(gpu, inp):
    """The successor (6) and block (8) kernels against the two-phase kernel,
    on every coefficient and probability of a 490,314-state layer (a last
    tile of 10 states, successor runs cut by the end of the layer)."""
    from perceval_amd.slos import COEFS, PROBS, SlosPlan
    U = haar_unitary(16, seed=7)
    plan = SlosPlan(inp)
    assert plan.count == 490314 and plan.count % 64 == 10
    out = {}
    for kernel in (4, 6, 8):
        plan.set_kernel(kernel)
        out[kernel] = (plan.run(U, COEFS).cpu().numpy(), plan.run(U, PROBS).cpu().numpy())
    for kernel in (6, 8):
        assert np.array_equal(out[4][0], out[kernel][0]) and np.array_equal(out[4][1], out[kernel][1]), kernel
    plan.close()


def test_slos_probs_thirteen_photons(gpu):
    """13 photons in 8 modes (77,520 states): the kernels' probability
    epilogues in the double-precision factorial form (products above 12! do
    not fit 32 bits), against the oracle at 1e-12 and bit-identical to the
    two-phase kernel."""
    from perceval_amd.slos import PROBS, SlosPlan
    U = haar_unitary(8, seed=13)
    inp = [3, 0, 2, 1, 0, 4, 1, 2]
    plan = SlosPlan(inp)
    assert plan.count == 77520
    got = {}
    for kernel in (4, 6, 8, 0):
        plan.set_kernel(kernel)
        got[kernel] = plan.run(U, PROBS).cpu().numpy()
    assert np.max(np.abs(got[0] - oracle.slos_probs(U, inp))) < ABS_TOL
    for kernel in (6, 8, 0):
        assert np.array_equal(got[4], got[kernel]), kernel
    plan.close()


@pytest.mark.parametrize("inp", [[1] * 6 + [0] * 6, [1, 1, 0, 2, 0, 0, 1, 0, 3, 0, 0, 1], [0] * 11 + [7],
                                 [3, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1], [9, 5, 0], [4, 3]])
def test_slos_block_kernel_every_width(gpu, inp):
    """Kernel 8 at every suffix width (1 .. min(m-1, 16)) is bit-identical to
    the two-phase kernel, coefficients and probabilities: regimes with one
    block per wave (uniform walk) and several blocks per wave (per-lane walk),
    chunks cut by the end of a block, groups of every suffix photon count."""
    from perceval_amd.slos import COEFS, PROBS, SlosPlan
    m = len(inp)
    U = haar_unitary(m, seed=3 * m + sum(inp))
    plan = SlosPlan(inp)
    plan.set_kernel(4)
    ref = (plan.run(U, COEFS).cpu().numpy(), plan.run(U, PROBS).cpu().numpy())
    assert np.max(np.abs(ref[1] - oracle.slos_probs(U, inp, "gather"))) < ABS_TOL
    plan.set_kernel(8)
    for w in range(1, min(m - 1, 16) + 1):
        plan.set_block_width(w)
        assert plan.block_width == w
        got = (plan.run(U, COEFS).cpu().numpy(), plan.run(U, PROBS).cpu().numpy())
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), w
    plan.close()


def test_slos_block_width_rejects_bad_values(gpu):
    from perceval_amd.slos import SlosPlan
    plan = SlosPlan([1, 1, 0, 0])
    for bad in (-1, 4, 17):
        with pytest.raises(ValueError):
            plan.set_block_width(bad)
    plan.set_block_width(0)
    assert 1 <= plan.block_width <= 3
    plan.close()
    one = SlosPlan([3])            # one mode: no block layout, auto falls back
    assert one.block_width == 0
    with pytest.raises(ValueError):
        one.set_kernel(8)
    assert one.run(np.eye(1, dtype=complex)).cpu().numpy().tolist() == [1.0]
    one.close()


def test_slos_set_kernel_rejects_bad_values(gpu):
    from perceval_amd.slos import SlosPlan
    plan = SlosPlan([1, 1, 0])
    for bad in (-1, 2, 3, 5, 7, 9):
        with pytest.raises(ValueError):
            plan.set_kernel(bad)
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


def test_slos_m32_n10_a_billion_outputs(gpu):
    """m=32, n=10: C(41, 10) = 1,121,099,408 probabilities (9.0 GB) from a
    273M-state parent layer (two 4.4 GB ping-pong buffers), the default
    kernel: the probabilities sum to 1 and 64 of them, the first and last
    included, equal |perm|^2 / prod s! from the oracle within 1e-12."""
    import torch
    from perceval_amd.slos import PROBS, SlosPlan
    U = haar_unitary(32, seed=32)
    inp = [1] * 10 + [0] * 22
    plan = SlosPlan(inp)
    assert plan.count == math.comb(41, 10)
    _say(f"plan: {plan.count} outputs, workspace {plan.work_bytes / 2**30:.1f} GiB")
    p = plan.run(U, PROBS)
    torch.cuda.synchronize()
    _say("all_prob done")
    assert abs(float(p.sum().item()) - 1.0) < 1e-10
    assert float(p.min().item()) >= 0.0
    rng = np.random.RandomState(2)
    idx = np.concatenate([[0, p.numel() - 1], rng.randint(0, p.numel(), 62)])
    got = p[torch.from_numpy(idx).cuda()].cpu().numpy()
    for i, g in zip(idx, got):
        s = oracle.fock_unrank(int(i), 32, 10)
        ref = abs(oracle.perm(oracle.submatrix(U, inp, s))) ** 2 / oracle.prodnfact(s)
        assert abs(g - ref) < ABS_TOL, (i, s, g, ref)
    del p
    plan.close()
    torch.cuda.empty_cache()


def _perm3(M):
    """Permanent of a 3 x 3 matrix, the six products written out."""
    return (M[0, 0] * (M[1, 1] * M[2, 2] + M[1, 2] * M[2, 1]) + M[0, 1] * (M[1, 0] * M[2, 2] + M[1, 2] * M[2, 0])
            + M[0, 2] * (M[1, 0] * M[2, 1] + M[1, 1] * M[2, 0]))


def test_slos_wide_circuits(gpu):
    """Circuits wider than 64 modes with few photons: m=128, n=3 (357,760
    outputs) against the oracle's full distribution; m=100, n=2 against the
    two-photon closed form, every output; m=256, n=3 (2,829,056 outputs)
    summing to 1 with 64 outputs against 3 x 3 permanents written out; a
    herald mask at m=100 (masked plan and selection)."""
    import torch
    from perceval_amd.slos import PROBS, SlosPlan, fock_mask_select, fock_unrank
    U = haar_unitary(128, seed=128)
    inp = [1, 0, 1, 1] + [0] * 124
    p = SlosPlan(inp).run(U, PROBS).cpu().numpy()
    assert p.size == math.comb(130, 3)
    assert np.max(np.abs(p - oracle.slos_probs(U, inp))) < ABS_TOL

    U = haar_unitary(100, seed=100)
    inp = [0] * 100
    inp[7] = inp[42] = 1
    p = SlosPlan(inp).run(U, PROBS).cpu().numpy()
    states = oracle.fock_enumerate(100, 2)
    ref = np.empty(len(states))
    for k, st in enumerate(states):
        modes = np.repeat(np.arange(100), st)
        a, b = modes
        amp = U[a, 7] * U[b, 42] + U[a, 42] * U[b, 7]
        ref[k] = abs(amp) ** 2 / (2 if a == b else 1)
    assert np.max(np.abs(p - ref)) < ABS_TOL
    # heralds: mode 99 empty, mode 98 one photon
    cond = [-1] * 98 + [1, 0]
    masked = SlosPlan(inp, mask_cond=cond).run(U, PROBS).cpu().numpy()
    keep = (states[:, 98] == 1) & (states[:, 99] == 0)
    assert np.max(np.abs(masked - ref[keep])) < ABS_TOL
    idx = fock_mask_select(100, 2, [" " * 98 + "10"]).cpu().numpy()
    assert np.array_equal(idx, np.flatnonzero(keep))
    # a tree over two inputs of the wide circuit equals its single plans
    from perceval_amd.slos import SlosTree
    other = [0] * 100
    other[7] = other[60] = other[99] = 1
    outs = SlosTree([inp, other]).run(U, PROBS)
    assert np.max(np.abs(outs[0].cpu().numpy() - p)) < 1e-13
    assert np.max(np.abs(outs[1].cpu().numpy() - SlosPlan(other).run(U, PROBS).cpu().numpy())) < 1e-13

    U = haar_unitary(256, seed=256)
    inp = [0] * 256
    inp[0] = inp[100] = inp[255] = 1
    plan = SlosPlan(inp)
    pt = plan.run(U, PROBS)
    assert pt.numel() == math.comb(258, 3)
    assert abs(float(pt.sum().item()) - 1.0) < 1e-10
    rng = np.random.RandomState(5)
    ranks = np.concatenate([[0, pt.numel() - 1], rng.randint(0, pt.numel(), 62)])
    got = pt[torch.from_numpy(ranks).cuda()].cpu().numpy()
    occ = fock_unrank(ranks, 256, 3)
    for g, st in zip(got, occ):
        rows = np.repeat(np.arange(256), st)
        M = U[np.ix_(rows, [0, 100, 255])]
        ref = abs(_perm3(M)) ** 2 / np.prod([math.factorial(v) for v in st])
        assert abs(g - ref) < ABS_TOL
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


def _closure_keep(m, k, rows, at_least, mask_n):
    """Flags over the k-photon states (FSArray order) of those contained in
    some mask_n-photon state that matches a row: the matches, closed under
    photon removal mask_n - k times."""
    full = oracle.fock_enumerate(m, mask_n)
    layer = {tuple(s) for s in full[_rows_keep(full, rows, at_least)].tolist()}
    for _ in range(mask_n - k):
        layer = {s[:j] + (s[j] - 1,) + s[j + 1:] for s in layer for j in range(m) if s[j] > 0}
    return np.array([tuple(s) in layer for s in oracle.fock_enumerate(m, k).tolist()], dtype=bool)


@pytest.mark.parametrize("m,k,masks,at_least,mask_n", [
    (6, 2, ["  0  1"], [], 4),
    (6, 3, ["1     ", "  2   "], [], 5),
    (6, 2, ["   1 1"], [3], 5),
    (5, 2, ["11111"], [], 5),                    # every mode exact: the surplus has nowhere to go
    (5, 2, ["11111"], [], 6),                    # no 6-photon state matches: nothing kept
    (8, 4, ["1 1 1 1 ", "  00  0 ", " 2      "], [0, 2], 6),
])
def test_fock_mask_select_parts_of_a_larger_state(gpu, m, k, masks, at_least, mask_n):
    """Masks set for mask_n photons (set_mask's n, the photons of the whole
    separated input), applied to a k-photon part: the kept states are those
    that the other mask_n - k photons can complete to a match."""
    from perceval_amd.slos import fock_mask_select, mask_conditions
    ref = np.flatnonzero(_closure_keep(m, k, mask_conditions(masks, m), at_least, mask_n))
    got = fock_mask_select(m, k, masks, at_least, mask_n).cpu().numpy()
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("inp,rows,at_least,mask_n", [
    ([1, 1, 0, 1, 0, 0], [[-1, -1, -1, 0, 1, 1]], [], 5),
    ([1, 0, 1, 1, 0, 0, 1], [[1, -1, -1, 2, -1, -1, -1], [0, -1, -1, 0, -1, -1, 1]], [0, 3], 6),
    ([2, 0, 1, 0, 0, 0, 0, 1], [[0, 1, -1, -1, -1, -1, -1, 1], [2, 2, -1, -1, -1, -1, -1, 0]], [7], 7),
    ([1] * 6 + [0] * 6, [[-1] * 10 + [1, 1]], [], 7),
])
def test_slos_masks_plan_parts_of_a_larger_state(gpu, inp, rows, at_least, mask_n):
    """A masked plan for masks set for more photons than the input holds:
    its outputs are the states the missing photons can complete, each within
    1e-12 of the full plan's value, every block width."""
    from perceval_amd.slos import COEFS, PROBS, SlosPlan
    m = len(inp)
    U = haar_unitary(m, seed=3 * m + mask_n)
    keep = _closure_keep(m, sum(inp), np.asarray(rows), at_least, mask_n)
    full = SlosPlan(inp)
    pf = full.run(U, PROBS).cpu().numpy()[keep]
    cf = full.run(U, COEFS).cpu().numpy()[keep]
    plan = SlosPlan(inp, mask_cond=rows, at_least=at_least, mask_n=mask_n)
    assert plan.count == int(keep.sum()) > 0
    free = int((np.asarray(rows) < 0).all(axis=0).sum())
    for w in range(1, min(free - 1, 16) + 1):
        plan.set_block_width(w)
        assert np.max(np.abs(plan.run(U, PROBS).cpu().numpy() - pf)) < ABS_TOL, w
        assert np.max(np.abs(plan.run(U, COEFS).cpu().numpy() - cf)) < ABS_TOL, w
    plan.close()
    full.close()


def test_slos_backend_mask_for_the_whole_separated_state(gpu):
    """set_mask(heralds, n) with n the photons of the whole input and a part
    holding fewer (the reference simulator's separated inputs,
    perceval/simulators/simulator.py:412-470): all_prob returns the states the
    other photons can complete, through a masked plan, and matches the
    selection path on the full output layer."""
    from perceval_amd.backends import SLOSBackend
    from perceval_amd.components import Unitary
    from perceval_amd.slos import mask_conditions
    from perceval_amd.utils.states import FockState
    U = haar_unitary(8, seed=81)
    part = FockState([1, 0, 1, 0, 0, 0, 0, 0])
    b = SLOSBackend()
    b.set_circuit(Unitary(U))
    b.set_input_state(part)
    full = np.asarray(b.all_prob())
    b.set_mask("      11", 4)
    b.set_input_state(part)
    got = np.asarray(b.all_prob())
    assert b._plans_masked and all(p is not None for p in b._plans_masked.values())
    assert b._state_mapping[part]._plan is None        # the full plan was never built
    keep = _closure_keep(8, 2, mask_conditions("      11", 8), [], 4)
    assert got.shape == (int(keep.sum()),) and np.max(np.abs(got - full[keep])) < ABS_TOL
    states = list(b._get_iterator(part))
    assert [list(s) for s in states] == oracle.fock_enumerate(8, 2)[keep].tolist()


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


@pytest.mark.parametrize("m,inp,cond", [
    (6, [1, 1, 1, 1, 0, 0], [-1, -1, -1, -1, 1, 1]),
    (6, [1, 0, 1, 0, 1, 1], [1, -1, 0, -1, -1, -1]),
    (7, [2, 0, 1, 0, 1, 0, 1], [-1, 2, -1, -1, 0, -1, 1]),
    (8, [1, 1, 1, 1, 1, 0, 0, 0], [0, -1, -1, 1, -1, -1, -1, -1]),
    (9, [1, 1, 1, 0, 1, 1, 0, 0, 1], [-1, 1, -1, 1, -1, 1, -1, -1, -1]),
    (6, [1, 1, 1, 0, 0, 0], [2, 2, -1, -1, -1, -1]),      # heralds need 4 of 3 photons: empty
    (12, [1] * 6 + [0] * 6, [-1] * 10 + [3, 3]),         # every photon heralded: one output
    (16, [1] * 8 + [0] * 8, [1, -1, -1, -1, -1, 1] + [-1] * 9 + [0]),
])
def test_slos_masked_plan_vs_full(gpu, m, inp, cond):
    """A plan restricted to a herald mask against the full plan's output
    filtered by the same mask, every kept state, 1e-12 absolute: layers that
    drop the states which cannot reach the mask, a different summation order
    (herald modes first), kept outputs in FSArray order."""
    from perceval_amd.slos import COEFS, PROBS, SlosPlan
    U = haar_unitary(m, seed=5 * m + sum(cond))
    full = SlosPlan(inp)
    states = oracle.fock_enumerate(m, sum(inp))
    keep = np.array([all(c < 0 or s[j] == c for j, c in enumerate(cond)) for s in states], dtype=bool)
    pf = full.run(U, PROBS).cpu().numpy()[keep]
    cf = full.run(U, COEFS).cpu().numpy()[keep]
    plan = SlosPlan(inp, mask_cond=cond)
    assert plan.count == int(keep.sum())
    if plan.count == 0:
        assert plan.run(U, PROBS).numel() == 0
        return
    free = [j for j in range(m) if cond[j] < 0]
    for w in range(1, min(len(free) - 1, 16) + 1):
        plan.set_block_width(w)
        p = plan.run(U, PROBS).cpu().numpy()
        c = plan.run(U, COEFS).cpu().numpy()
        assert np.max(np.abs(p - pf)) < ABS_TOL, w
        assert np.max(np.abs(c - cf)) < ABS_TOL, w
    assert np.max(np.abs(pf - oracle.slos_probs(U, inp)[keep])) < ABS_TOL
    plan.close()
    full.close()


def test_slos_masked_plan_m24_n10_heralds(gpu):
    """Heralds 0, 1, 0, 1 on modes 20..23 of the m=24, n=10 configuration:
    the masked plan computes C(27, 8) = 2,220,075 outputs, each within 1e-12
    of the full plan's probability of the same state."""
    import torch
    from perceval_amd.slos import PROBS, SlosPlan, fock_mask_select, gather_f64
    U = haar_unitary(24, seed=24)
    inp = [1] * 10 + [0] * 14
    cond = [-1] * 20 + [0, 1, 0, 1]
    plan = SlosPlan(inp, mask_cond=cond)
    assert plan.count == math.comb(27, 8)
    _say(f"masked plan: {plan.count} outputs, workspace {plan.work_bytes / 2**20:.0f} MiB")
    got = plan.run(U, PROBS)
    idx = fock_mask_select(24, 10, [" " * 20 + "0101"])
    ref = gather_f64(SlosPlan(inp).run(U, PROBS), idx)
    torch.cuda.synchronize()
    err = float((got - ref).abs().max().item())
    _say(f"masked vs full: max abs err {err:.3e}")
    assert err < ABS_TOL
    plan.close()


def _rows_keep(states, rows, at_least):
    """States matching some condition row: exact counts, floors on the
    at-least modes, -1 any count."""
    rows = np.atleast_2d(np.asarray(rows))
    floor = np.zeros(states.shape[1], dtype=bool)
    floor[list(at_least)] = True
    keep = np.zeros(len(states), dtype=bool)
    for row in rows:
        keep |= np.all((row < 0) | np.where(floor, states >= row, states == row), axis=1)
    return keep


@pytest.mark.parametrize("inp,rows,at_least", [
    ([1, 1, 1, 1, 0, 0], [[-1, -1, -1, 0, 1, 1], [-1, -1, -1, 1, 1, 0]], []),       # two rows (test_mask.py:42)
    ([1, 0, 1, 1, 0, 1], [[-1, -1, 1, -1, -1, 1]], [2]),                               # a floor
    ([2, 0, 1, 0, 1, 0, 1], [[1, -1, -1, -1, -1, -1, 0], [-1, -1, 2, -1, -1, -1, -1]], []),
    ([1, 1, 0, 1, 1, 0, 1], [[1, -1, -1, 2, -1, -1, -1], [0, -1, -1, 0, -1, -1, 1]], [0, 3]),
    ([1, 1, 0, 1, 1, 0, 1], [[1, -1, -1, 2, -1, -1, -1], [0, -1, -1, 0, -1, -1, -1]], [0, 3]),  # row 2 matches all
    ([1] * 5 + [0] * 3, [[0, 1, -1, -1, -1, -1, -1, 1], [1, 0, -1, -1, -1, -1, -1, 1],
                         [2, 2, -1, -1, -1, -1, -1, 0]], [7]),
    ([1, 1, 1, 0, 0, 0], [[1, 1, -1, -1, -1, 3], [2, 2, -1, -1, -1, -1]], []),         # no row reachable: empty
    ([1] * 7 + [0] * 7, [[1] + [-1] * 12 + [0], [0] + [-1] * 12 + [1]], []),
])
def test_slos_masks_plan_vs_full(gpu, inp, rows, at_least):
    """A plan restricted to several mask rows and at-least modes against the
    full plan's output filtered by the same masks, 1e-12 absolute, every
    block width: layers hold the states some matching output extends, the
    final layer several herald configurations, put back in FSArray order."""
    from perceval_amd.slos import COEFS, PROBS, SlosPlan
    m = len(inp)
    U = haar_unitary(m, seed=7 * m + len(rows))
    full = SlosPlan(inp)
    keep = _rows_keep(oracle.fock_enumerate(m, sum(inp)), rows, at_least)
    pf = full.run(U, PROBS).cpu().numpy()[keep]
    cf = full.run(U, COEFS).cpu().numpy()[keep]
    plan = SlosPlan(inp, mask_cond=rows, at_least=at_least)
    assert plan.count == int(keep.sum())
    if plan.count == 0:
        assert plan.run(U, PROBS).numel() == 0
        return
    free = int((np.asarray(rows) < 0).all(axis=0).sum())
    for w in range(1, min(free - 1, 16) + 1):
        plan.set_block_width(w)
        p = plan.run(U, PROBS).cpu().numpy()
        c = plan.run(U, COEFS).cpu().numpy()
        assert np.max(np.abs(p - pf)) < ABS_TOL, w
        assert np.max(np.abs(c - cf)) < ABS_TOL, w
    assert np.max(np.abs(pf - oracle.slos_probs(U, inp)[keep])) < ABS_TOL
    plan.close()
    full.close()


def test_slos_masks_plan_m24_n10(gpu):
    """Two herald patterns on modes 20..23 of the m=24, n=10 configuration,
    one of them a floor on mode 23: every output of the masked plan within
    1e-12 of the full plan's probability of the same state, in the order of
    the GPU mask selection (FSArray order)."""
    import torch
    from perceval_amd.slos import PROBS, SlosPlan, fock_mask_select, gather_f64
    U = haar_unitary(24, seed=25)
    inp = [1] * 10 + [0] * 14
    masks = [" " * 20 + "0101", " " * 20 + "1001"]
    plan = SlosPlan(inp, mask_cond=[[-1] * 20 + [0, 1, 0, 1], [-1] * 20 + [1, 0, 0, 1]], at_least=[23])
    idx = fock_mask_select(24, 10, masks, [23])
    assert plan.count == idx.numel()
    _say(f"masks plan: {plan.count} outputs, workspace {plan.work_bytes / 2**20:.0f} MiB")
    got = plan.run(U, PROBS)
    ref = gather_f64(SlosPlan(inp).run(U, PROBS), idx)
    torch.cuda.synchronize()
    err = float((got - ref).abs().max().item())
    _say(f"masks vs full: max abs err {err:.3e}")
    assert err < ABS_TOL
    plan.close()


def test_slos_backend_masks_use_a_masked_plan(gpu):
    """SLOSBackend.all_prob with two masks and an at-least mode: the masked
    plan's values equal the full distribution filtered by FSMask matching."""
    from perceval_amd.backends import SLOSBackend
    from perceval_amd.components import Unitary
    from perceval_amd.utils.states import FockState
    U = haar_unitary(7, seed=71)
    b = SLOSBackend()
    b.set_circuit(Unitary(U))
    b.set_input_state(FockState([1, 1, 0, 1, 1, 0, 1]))
    full = np.asarray(b.all_prob())
    b.set_mask(["1  2   ", "0  0  1"], at_least_modes=[0, 3])
    b.set_input_state(FockState([1, 1, 0, 1, 1, 0, 1]))
    got = np.asarray(b.all_prob())
    assert b._plans_masked and all(p is not None for p in b._plans_masked.values())
    keep = _rows_keep(oracle.fock_enumerate(7, 5), [[1, -1, -1, 2, -1, -1, -1], [0, -1, -1, 0, -1, -1, 1]], [0, 3])
    assert got.shape == (int(keep.sum()),)
    assert np.max(np.abs(got - full[keep])) < ABS_TOL


def test_slos_masked_plan_rejects_bad_masks(gpu):
    from perceval_amd.slos import SlosPlan
    with pytest.raises(ValueError):
        SlosPlan([1, 1, 0], mask_cond=[-1, 1])            # wrong length
    with pytest.raises(ValueError):
        SlosPlan([1, 1, 0], mask_cond=[1, 1, -1])         # one free mode
    with pytest.raises(ValueError):
        SlosPlan([1, 1, 0], mask_cond=[-2, -1, -1])       # bad condition
    plan = SlosPlan([1, 1, 0, 0], mask_cond=[1, -1, -1, -1])
    with pytest.raises(ValueError):
        plan.set_kernel(6)
    plan.set_kernel(0)
    plan.close()
