"""GPU Gurvits estimates (NaiveApprox) against the oracle restatement (same
counter-based samples, so the means agree to rounding), exact cases, the 99%
coverage of the error bound, and the NaiveApprox backend API."""
import numpy as np
import pytest

import oracle
from perceval_amd.utils.workloads import haar_unitary

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gurvits-gpu] {msg}", flush=True)


@pytest.mark.parametrize("orders", [[0, 1, 2, 3], [5, 8, 12], [16, 20, 24], [30, 31, 32], [40, 48]])
def test_gurvits_batch_vs_oracle(gpu, orders):
    from perceval_amd.permanent import gurvits
    U = haar_unitary(64, seed=sum(orders) + 1)
    rng = np.random.RandomState(len(orders))
    mats = [U[np.ix_(rng.choice(64, n, replace=False), rng.choice(64, n, replace=False))] * 2.0 for n in orders]
    iters, seed, stream0 = 3000, 99, 5
    est, err = gurvits(mats, iters, seed=seed, stream0=stream0)
    _say(f"orders {orders}")
    for b, M in enumerate(mats):
        ref_est, ref_err = oracle.gurvits(M, iters, seed=seed, stream=stream0 + b)
        scale = max(ref_err * np.sqrt(iters) / 2.5758293035489004, abs(ref_est), 1e-300)
        assert abs(est[b] - ref_est) <= 1e-11 * scale, (orders[b], est[b], ref_est)
        # the bound comes from E|X|^2 - |mean|^2, whose rounding (summation
        # order) matters only when the variance is at rounding level
        assert abs(err[b] - ref_err) <= 1e-6 * ref_err + 1e-6 * scale / np.sqrt(iters), (orders[b], err[b], ref_err)


def test_gurvits_exact_cases(gpu):
    from perceval_amd.permanent import gurvits
    est, err = gurvits([np.eye(7), np.array([[3 - 2j]]), np.zeros((0, 0))], 100, seed=3)
    assert est[0] == 1.0 and est[1] == 3 - 2j and est[2] == 1.0


def test_gurvits_coverage_and_convergence(gpu):
    from perceval_amd.permanent import gurvits
    U = haar_unitary(12, seed=4)
    A = U[:8, :8]
    exact = oracle.perm(A)
    est, err = gurvits([A] * 256, 4000, seed=11)
    misses = int(np.sum(np.abs(est - exact) > err))
    _say(f"coverage misses {misses} / 256")
    assert misses <= 10
    big, berr = gurvits([A], 4_000_000, seed=12)
    assert abs(big[0] - exact) <= berr[0] and berr[0] < 0.1 * np.mean(err)


def test_naive_approx_backend(gpu):
    from perceval_amd.backends import BackendFactory
    from perceval_amd.components import BS, Unitary
    from perceval_amd.utils.states import BasicState
    from perceval_amd.components import Circuit
    # the reference wiring test: identity circuit on one mode
    b = BackendFactory.get_backend("NaiveApprox")
    b.set_circuit(Circuit(1))
    b.set_input_state(BasicState([1]))
    assert b.all_prob() == [1.0]
    assert b.probability(BasicState([1])) == 1.0
    # confidence intervals contain the exact probabilities (8 modes, 4 photons)
    U = haar_unitary(8, seed=21)
    approx = BackendFactory.get_backend("NaiveApprox", gurvits_iterations=200000)
    exact = BackendFactory.get_backend("Naive")
    for be in (approx, exact):
        be.set_circuit(Unitary(U))
        be.set_input_state(BasicState([1, 1, 1, 1, 0, 0, 0, 0]))
    # one estimate per output; each is a 99% statement (coverage itself is
    # tested over 256 trials above), so the truth is checked at 1.5x the bound
    for out in ([1, 1, 1, 1, 0, 0, 0, 0], [0, 0, 0, 0, 1, 1, 1, 1], [2, 0, 1, 0, 0, 1, 0, 0], [4, 0, 0, 0, 0, 0, 0, 0]):
        ref = exact.prob_amplitude(BasicState(out))
        a, e = approx.prob_amplitude_with_error(BasicState(out))
        assert abs(a - ref) <= 1.5 * e, (out, a, ref, e)
        lo, hi = approx.probability_confidence_interval(BasicState(out))
        assert 0.0 <= lo <= hi <= 1.0 and hi - lo > 0.0
    # photon-number mismatch: amplitude 0 with no error
    assert approx.prob_amplitude_with_error(BasicState([1, 0, 0, 0, 0, 0, 0, 0])) == (0j, 0.0)
    probs = np.array(approx.all_prob())
    assert len(probs) == 330 and abs(probs.sum() - 1.0) < 0.05
    assert approx.name == "NaiveApprox"
    _ = BS  # noqa: F841
