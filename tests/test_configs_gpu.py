"""GPU parity at the measured configurations, through the C ABI:

  * Naive all_prob on the Haar 8x8 / |1,1,1,1,0,0,0,0> case (330 outputs);
  * every one of the 92,561,040 SLOS probabilities at m=24, n=10;
  * permanents of orders 25..28 and closed forms at order 28;
  * SLOS probabilities with 13 photons (factorial products beyond 32 bits);
  * the learned independence sampler at m=60, n=20 with a trained 2 x 256
    proposal network, judged by exact 1- and 2-mode marginals;
  * seed plumbing and the BSSamples wire format on sampler output.

Bars: permanents 1e-10 relative; SLOS 1e-12 absolute per probability; Fock
indexing and sample streams bit-exact; marginal TVD <= 0.01."""
import math
import os

import numpy as np
import pytest

import oracle
from perceval_amd.utils.workloads import haar_unitary

pytestmark = pytest.mark.gpu

ABS_TOL = 1e-12
REL_TOL = 1e-10


def _say(msg):
    print(f"[configs-gpu] {msg}", flush=True)


def _threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


# ------------------------------------------------------------------ config 1

def test_config1_naive_all_prob_vs_oracle_and_slos(gpu):
    from perceval_amd.backends import BackendFactory
    from perceval_amd.components import Unitary
    from perceval_amd.utils.states import BasicState, fock_space
    U = haar_unitary(8, seed=42)
    inp = BasicState([1, 1, 1, 1, 0, 0, 0, 0])
    naive = BackendFactory.get_backend("Naive")
    naive.set_circuit(Unitary(U))
    naive.set_input_state(inp)
    p_naive = np.array(naive.all_prob())
    assert len(p_naive) == 330
    exact = oracle.distribution_exact(U, list(inp))          # long-double permanents per output
    states = fock_space(8, 4)
    ref = np.array([exact[tuple(s)] for s in states])
    assert np.max(np.abs(p_naive - ref)) < ABS_TOL
    assert np.max(np.abs(p_naive - oracle.slos_probs(U, list(inp)))) < ABS_TOL
    slos = BackendFactory.get_backend("SLOS")
    slos.set_circuit(Unitary(U))
    slos.set_input_state(inp)
    assert np.max(np.abs(p_naive - np.array(slos.all_prob()))) < ABS_TOL
    # the single-amplitude path (one permanent per call) on a few outputs
    for k in (0, 1, 77, 329):
        s = BasicState(states[k].tolist())
        amp = naive.prob_amplitude(s)
        ref_amp = oracle.amplitude_brute(U, list(inp), list(s))
        assert abs(amp - ref_amp) < ABS_TOL, s
        assert abs(amp - slos.prob_amplitude(s)) < ABS_TOL, s


# ------------------------------------------------------------------ config 3

def test_config3_every_probability_vs_oracle(gpu):
    """All 92,561,040 probabilities of the default kernel against the oracle's
    gather-form SLOS, 1e-12 absolute each."""
    import torch
    from perceval_amd.slos import PROBS, SlosPlan
    U = haar_unitary(24, seed=24)
    inp = [1] * 10 + [0] * 14
    plan = SlosPlan(inp)
    _say(f"GPU all_prob m=24 n=10 ({plan.count} outputs)")
    p = plan.run(U, PROBS)
    torch.cuda.synchronize()
    got = p.cpu().numpy()
    del p
    plan.close()
    _say(f"oracle gather on {_threads()} threads")
    ref = oracle.slos_probs(U, inp, "gather", threads=_threads())
    _say("oracle done; comparing")
    assert got.shape == ref.shape == (92561040,)
    err = np.abs(got - ref)
    worst = int(np.argmax(err))
    _say(f"max abs err {err[worst]:.3e} at rank {worst}; sum {got.sum():.15f}")
    assert err[worst] < ABS_TOL, (worst, got[worst], ref[worst])
    assert abs(got.sum() - 1.0) < 1e-10


# ------------------------------------------------------------------ permanents 25..28

def test_permanent_orders_25_to_28_vs_oracle(gpu):
    from perceval_amd.permanent import permanents
    U = haar_unitary(64, seed=2528)
    rng = np.random.RandomState(28)
    mats = [U[np.ix_(rng.choice(64, n, replace=False), rng.choice(64, n, replace=False))] for n in (25, 26, 27, 28)]
    _say("GPU permanents of orders 25..28")
    got = permanents(mats)
    for M, g in zip(mats, got):
        _say(f"oracle order {M.shape[0]} on {_threads()} threads")
        ref = oracle.perm_large(M, threads=_threads())
        rel = abs(g - ref) / abs(ref)
        _say(f"order {M.shape[0]}: rel err {rel:.3e}")
        assert rel < REL_TOL, (M.shape[0], g, ref)


def test_permanent_orders_29_to_32(gpu):
    """Orders 29..32 (their own launch beside the one of orders <= 28): the
    all-ones matrix J/2 (perm = n!/2^n, heavy cancellation) at every order,
    a block-diagonal order-30 matrix (perm = product of the blocks'), one
    Haar submatrix of order 29 against the threaded long-double oracle, and
    a mixed batch holding orders 12, 20, 29 and 32 in one plan."""
    from perceval_amd.permanent import permanents
    U = haar_unitary(64, seed=2932)
    rng = np.random.RandomState(29)
    Js = [np.ones((n, n), dtype=complex) / 2 for n in range(29, 33)]
    A = U[:20, 20:40]
    B = haar_unitary(10, seed=10)
    D = np.zeros((30, 30), dtype=complex)
    D[:20, :20] = A
    D[20:, 20:] = B
    M29 = U[np.ix_(rng.choice(64, 29, replace=False), rng.choice(64, 29, replace=False))]
    M12, M20 = U[30:42, 40:52], U[40:60, 2:22]
    _say("GPU permanents of orders 29..32")
    got = permanents(Js + [D, M29])
    for n, g in zip(range(29, 33), got[:4]):
        ref = math.factorial(n) / 2 ** n
        assert abs(g - ref) / ref < REL_TOL, (n, g, ref)
    pd = oracle.perm(A) * oracle.perm(B)
    assert abs(got[4] - pd) / abs(pd) < REL_TOL
    _say(f"oracle order 29 on {_threads()} threads")
    ref29 = oracle.perm_large(M29, threads=_threads())
    assert abs(got[5] - ref29) / abs(ref29) < REL_TOL, (got[5], ref29)
    mixed = permanents([M12, Js[0], M20, Js[3]])
    refs = [oracle.perm(M12), math.factorial(29) / 2 ** 29, oracle.perm(M20), math.factorial(32) / 2 ** 32]
    for g, r in zip(mixed, refs):
        assert abs(g - r) / abs(r) < REL_TOL, (g, r)


def test_permanent_order28_closed_forms(gpu):
    """Order 28 against closed forms: the all-ones matrix (perm = 28!, the
    sum with the most cancellation: sum |terms| / |perm| = 1.26e4) and a
    block-diagonal matrix (perm = product of the blocks' permanents)."""
    from perceval_amd.permanent import permanents
    n = 28
    J = np.ones((n, n), dtype=complex) / 2
    A = haar_unitary(64, seed=7)[:20, 10:30]
    B = haar_unitary(8, seed=8)
    D = np.zeros((n, n), dtype=complex)
    D[:20, :20] = A
    D[20:, 20:] = B
    rng = np.random.RandomState(1)
    P = np.eye(n)[rng.permutation(n)]
    got = permanents([J, D, P @ D @ P.T])
    assert abs(got[0] - math.factorial(n) / 2 ** n) / (math.factorial(n) / 2 ** n) < REL_TOL
    ref = oracle.perm(A) * oracle.perm_brute(B)
    assert abs(got[1] - ref) / abs(ref) < REL_TOL
    assert abs(got[2] - ref) / abs(ref) < REL_TOL


# ------------------------------------------------------------------ 13 photons

@pytest.mark.parametrize("m,inp", [(8, [2, 2, 2, 2, 2, 1, 1, 1]), (14, [1] * 13 + [0])])
def test_slos_probs_13_photons(gpu, m, inp):
    """More than 12 photons: the probability epilogue's factorial products
    leave 32-bit range.  The default pick and the successor kernels against
    the oracle, and bit for bit against the two-phase kernel."""
    from perceval_amd.slos import COEFS, PROBS, SlosPlan
    U = haar_unitary(m, seed=13 + m)
    plan = SlosPlan(inp)
    ref = oracle.slos_probs(U, inp, "gather", threads=_threads())
    out = {}
    for kernel in (4, 0, 6, 8):
        plan.set_kernel(kernel)
        out[kernel] = (plan.run(U, PROBS).cpu().numpy(), plan.run(U, COEFS).cpu().numpy())
        assert np.max(np.abs(out[kernel][0] - ref)) < ABS_TOL, kernel
    for kernel in (0, 6, 8):
        assert np.array_equal(out[kernel][0], out[4][0]) and np.array_equal(out[kernel][1], out[4][1]), kernel
    plan.close()


# ------------------------------------------------------------------ config 5

def _batch_se(x: np.ndarray, batches: int = 50) -> np.ndarray:
    """Standard error of a chain mean from batch means (autocorrelation-aware)."""
    L = len(x) // batches
    bm = x[:L * batches].reshape(batches, L, *x.shape[1:]).mean(axis=1)
    return bm.std(axis=0, ddof=1) / math.sqrt(batches)


def test_config5_learned_mis_m60_n20(gpu):
    """Train the 2 x 256 proposal on 1e5 plain Clifford-Clifford samples
    (10 epochs), run the independence chain, and judge it by the exact 1- and
    2-mode marginals (200 seeded pairs): TVD <= 0.01 at 4e5 chain samples,
    1-mode TVD <= 0.01 already at 1e5; occupation means and pair moments
    within 5 batch-mean standard errors of their closed forms."""
    import itertools
    import torch
    from perceval_amd.marginals import exact_marginals, marginal_tvds_exact
    from perceval_amd.mis import LearnedMIS, ProposalNet, occ_to_seq, seq_to_occ
    from perceval_amd.sampling import cc_samples_device
    m, n = 60, 20
    U = haar_unitary(m, seed=60)
    inp = [1] * n + [0] * (m - n)
    dU = torch.from_numpy(U).cuda()
    _say("plain CC: 1e5 training samples")
    train = cc_samples_device(dU, inp, 100000, seed=1)
    net = ProposalNet(m, n, hidden=256, seed=0)
    hist = net.train(occ_to_seq(train, n), epochs=10, batch=2048, lr=2e-3)
    _say(f"trained: mean -log q per epoch {['%.3f' % h for h in hist]}")
    assert hist[-1] < hist[0]
    mis = LearnedMIS(U, inp, net, chunk=16384)
    x0 = occ_to_seq(cc_samples_device(dU, inp, 1, seed=7), n)[0]
    res = mis.run(400000, seed=5, x0_seq=x0)
    mis.close()
    occ = seq_to_occ(res["seq"], m).cpu().numpy()
    _say(f"chain: acceptance {res['acceptance']:.3f}")
    assert res["acceptance"] > 0.2
    assert (occ.sum(axis=1) == n).all()
    ex = exact_marginals(U, inp, pairs=200, seed=0)
    t_full = marginal_tvds_exact(occ, ex)
    t_1e5 = marginal_tvds_exact(occ[:100000], ex)
    t_cc = marginal_tvds_exact(train.cpu().numpy(), ex)
    _say(f"MIS 4e5 {t_full}; MIS 1e5 {t_1e5}; CC 1e5 {t_cc}")
    assert t_full["tvd1_max"] <= 0.01 and t_full["tvd2_max"] <= 0.01
    assert t_1e5["tvd1_max"] <= 0.01
    assert t_cc["tvd1_max"] <= 0.01 and t_cc["tvd2_max"] <= 0.01
    # closed-form moments
    x = occ.astype(float)
    mean = (np.abs(U[:, :n]) ** 2).sum(axis=1)
    se = np.maximum(_batch_se(x), 1e-4)
    assert np.all(np.abs(x.mean(axis=0) - mean) < 5 * se), np.max(np.abs(x.mean(axis=0) - mean) / se)
    rng = np.random.RandomState(0)
    for j, k in rng.choice(m, size=(8, 2), replace=False):
        exp = sum(abs(U[j, i] * U[k, i2] + U[j, i2] * U[k, i]) ** 2 for i, i2 in itertools.combinations(range(n), 2))
        xy = x[:, j] * x[:, k]
        assert abs(xy.mean() - exp) < 5 * max(_batch_se(xy), 1e-4), (j, k)


# ------------------------------------------------------------------ seeds and formats

def test_random_seed_reproduces_backend_samples(gpu):
    from perceval_amd import random_seed
    from perceval_amd.backends import BackendFactory
    from perceval_amd.components import Unitary
    from perceval_amd.utils.states import BasicState
    b = BackendFactory.get_backend("CliffordClifford2017")
    b.set_circuit(Unitary(haar_unitary(6, seed=5)))
    b.set_input_state(BasicState([1, 1, 1, 0, 0, 0]))
    random_seed(77)
    a1, a2 = b.samples(300), b.samples(300)
    random_seed(77)
    assert b.samples(300) == a1 and b.samples(300) == a2
    assert a1 != a2                                 # successive calls draw new streams
    random_seed(78)
    assert b.samples(300) != a1


def test_bssamples_wire_format_round_trips_multi_input_samples(gpu):
    from perceval_amd import serialization
    from perceval_amd.sampling import cc_samples_multi_device
    from perceval_amd.utils.states import BSSamples
    m = 10
    U = haar_unitary(m, seed=11)
    inputs = [[1, 1, 1, 0, 0, 0, 0, 0, 0, 0], [1, 1, 0, 0, 0, 0, 0, 0, 0, 0], [2, 0, 0, 0, 0, 0, 0, 0, 0, 1]]
    occ = cc_samples_multi_device(U, inputs, [4000, 1500, 500], seed=9).cpu().numpy()
    text = serialization.serialize(occ)
    back = serialization.deserialize(text)
    assert back == BSSamples.from_array(occ)
    body = serialization.serialize_bssamples_array(occ)
    assert np.array_equal(serialization.deserialize_bssamples_array(body), occ)
    assert len(text) < len(body)                     # the compressed form is the smaller one
