"""Pins the CPU oracle (oracle/pcvl_oracle.c) before it is used as the checker.

Each oracle function is checked against something that is not a restatement of
its own formula: the reference's own test and documentation values, mpmath
goldens (Ryser at 50 digits), brute force on tiny inputs, closed forms and
invariants, and an independent algorithm (Ryser vs Glynn, scatter vs gather,
SLOS vs permanents).
"""
import itertools
import json
import math
import os
from collections import Counter

import numpy as np
import pytest

import oracle
from perceval_amd.components import BS, PS, Circuit, catalog
from perceval_amd.utils.workloads import haar_unitary

HERE = os.path.dirname(os.path.abspath(__file__))


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


# ------------------------------------------------------------------ permanents

def _golden():
    with open(os.path.join(HERE, "golden", "perm_mpmath.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _golden(), ids=lambda c: f"n{c['n']}")
@pytest.mark.parametrize("method", ["glynn_ld", "ryser_ld", "glynn_d"])
def test_perm_vs_mpmath(case, method):
    M = np.array(case["re"]) + 1j * np.array(case["im"])
    ref = complex(float(case["perm"][0]), float(case["perm"][1]))
    tol = 1e-13 if method != "glynn_d" else 1e-12
    assert _rel(oracle.perm(M, method), ref) < tol


@pytest.mark.parametrize("n", range(0, 8))
def test_perm_vs_brute_force(n):
    rng = np.random.RandomState(40 + n)
    M = rng.randn(n, n) + 1j * rng.randn(n, n)
    ref = oracle.perm_brute(M)
    for method in ("glynn_ld", "ryser_ld", "glynn_d"):
        assert _rel(oracle.perm(M, method), ref) < 1e-12


def test_perm_closed_forms():
    assert oracle.perm(np.eye(9)) == pytest.approx(1.0)
    for n in (3, 6, 10):
        assert oracle.perm(np.ones((n, n))).real == pytest.approx(math.factorial(n), rel=1e-14)
    d = np.array([1 + 2j, -0.5j, 3.0, 0.25 + 0.25j, -1.5])
    assert _rel(oracle.perm(np.diag(d)), np.prod(d)) < 1e-15
    a, b, c, e = 0.3 + 1j, -2 + 0.5j, 1.5j, 0.7 - 0.2j
    assert _rel(oracle.perm(np.array([[a, b], [c, e]])), a * e + b * c) < 1e-15


def test_perm_invariants_n16():
    rng = np.random.RandomState(2)
    A = haar_unitary(32, seed=3)[:16, 16:]
    p = oracle.perm(A)
    P = np.eye(16)[rng.permutation(16)]
    Q = np.eye(16)[rng.permutation(16)]
    d1 = np.exp(1j * rng.uniform(0, 6.3, 16))
    d2 = rng.uniform(0.5, 2, 16)
    assert _rel(oracle.perm(A.T), p) < 1e-13
    assert _rel(oracle.perm(P @ A @ Q), p) < 1e-13
    assert _rel(oracle.perm(np.diag(d1) @ A @ np.diag(d2)), np.prod(d1) * np.prod(d2) * p) < 1e-13
    assert _rel(oracle.perm(A, "ryser_ld"), p) < 1e-13
    assert _rel(oracle.perm(A, "glynn_d"), p) < 1e-11


def test_perm_batch_threads_match_single():
    _, orders, mats = __import__("perceval_amd.utils.workloads", fromlist=["x"]).submatrix_batch(24, 6, 14, seed=5)
    batch = oracle.perm_batch(mats, precise=True, threads=4)
    for M, b in zip(mats, batch):
        assert b == oracle.perm(M)


# ------------------------------------------------------------------------ Fock

def test_fock_order_pinned_by_reference():
    # tests/utils/test_statevector.py:430-438 (max_photon_state_iterator(3, 2))
    expected = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1],
                [2, 0, 0], [1, 1, 0], [1, 0, 1], [0, 2, 0], [0, 1, 1], [0, 0, 2]]
    got = [list(s) for n in range(3) for s in oracle.fock_enumerate(3, n).tolist()]
    assert got == expected


def test_fock_counts_pinned_by_reference():
    # tests/utils/test_statevector.py:441-452
    assert oracle.fock_count(4, 1) == 4
    assert sum(oracle.fock_count(6, n) for n in range(3)) == 28


@pytest.mark.parametrize("m,n", [(m, n) for m in range(1, 7) for n in range(0, 6)])
def test_fock_rank_unrank_brute(m, n):
    brute = oracle.fock_enumerate_brute(m, n)
    assert [tuple(x) for x in oracle.fock_enumerate(m, n).tolist()] == brute
    assert oracle.fock_count(m, n) == len(brute) == math.comb(m + n - 1, n)
    for i, s in enumerate(brute):
        assert oracle.fock_rank(s) == i
        assert tuple(oracle.fock_unrank(i, m, n)) == s


def test_fock_rank_large_space_roundtrip():
    m, n = 24, 10
    total = oracle.fock_count(m, n)
    assert total == 92561040
    rng = np.random.RandomState(1)
    for r in rng.randint(0, total, size=200).tolist() + [0, total - 1]:
        s = oracle.fock_unrank(r, m, n)
        assert sum(s) == n and oracle.fock_rank(s) == r


# ------------------------------------------------------------------------ SLOS

def _exact_probs(U, inp):
    return np.array(list(oracle.distribution_exact(U, inp).values()))


@pytest.mark.parametrize("inp", [[1, 1, 0, 1, 0], [2, 0, 1, 0, 0], [0, 0, 0, 0, 3], [1, 1, 1, 1, 1]])
def test_slos_vs_permanents(inp):
    U = haar_unitary(len(inp), seed=sum(inp) * 7 + len(inp))
    scatter = oracle.slos_probs(U, inp, "scatter")
    gather = oracle.slos_probs(U, inp, "gather", threads=3)
    assert np.max(np.abs(scatter - _exact_probs(U, inp))) < 1e-14
    assert np.array_equal(scatter, gather)
    assert scatter.sum() == pytest.approx(1.0, abs=1e-13)


def test_slos_single_photon_closed_form():
    U = haar_unitary(7, seed=4)
    probs = oracle.slos_probs(U, [0, 0, 1, 0, 0, 0, 0])
    # FSArray order for one photon: |1,0,..>, |0,1,..>, ... -> p_j = |U[j, 2]|^2
    assert np.allclose(probs, np.abs(U[:, 2]) ** 2, atol=1e-15)


def test_slos_path_rule():
    # _Path._decompose (perceval/backends/_slos.py:61-86): the mode with the most
    # remaining photons first, the lowest such mode on ties.
    assert oracle.slos_path([1, 1, 1, 1, 0, 0, 0, 0]) == [0, 1, 2, 3]
    assert oracle.slos_path([2, 3]) == [1, 0, 1, 0, 1]
    assert oracle.slos_path([0, 1, 1]) == [1, 2]


def _dist(U, inp):
    states = [tuple(s) for s in oracle.fock_enumerate(len(inp), sum(inp)).tolist()]
    return dict(zip(states, oracle.slos_probs(U, inp)))


def test_reference_backend_distributions():
    # tests/backends/test_backends.py:116-142 (BS.H), :221-249 (3-mode circuit, BS on |2,3>)
    d = _dist(BS.H().compute_unitary(), [2, 0])
    assert [d[(2, 0)], d[(1, 1)], d[(0, 2)]] == pytest.approx([0.25, 0.5, 0.25])
    d = _dist(BS.H().compute_unitary(), [1, 1])
    assert [d[(2, 0)], d[(1, 1)], d[(0, 2)]] == pytest.approx([0.5, 0.0, 0.5], abs=1e-15)
    d = _dist(BS.H(theta=2 * math.pi / 3).compute_unitary(), [2, 0])
    assert [d[(2, 0)], d[(1, 1)], d[(0, 2)]] == pytest.approx([0.0625, 0.375, 0.5625])
    d = _dist(BS.H(theta=2 * math.pi / 3).compute_unitary(), [1, 0])
    assert [d[(1, 0)], d[(0, 1)]] == pytest.approx([0.25, 0.75])
    c = Circuit(3) // BS.H() // (1, PS(math.pi / 4)) // (1, BS.H())
    d = _dist(c.compute_unitary(), [0, 1, 1])
    exp = {(0, 1, 1): 0, (1, 1, 0): 0.25, (1, 0, 1): 0.25, (2, 0, 0): 0, (0, 2, 0): 0.25, (0, 0, 2): 0.25}
    for s, p in exp.items():
        assert d[s] == pytest.approx(p, abs=1e-15)
    d = _dist(BS().compute_unitary(), [2, 3])
    exp = {(5, 0): 0.3125, (4, 1): 0.0625, (3, 2): 0.125, (2, 3): 0.125, (1, 4): 0.0625, (0, 5): 0.3125}
    for s, p in exp.items():
        assert d[s] == pytest.approx(p, abs=1e-15)


def test_reference_cnot_values():
    # tests/backends/test_backends.py:42-57,170-185: post-processed CNOT 1/9 and 7/9
    U = catalog["postprocessed cnot"].build_circuit().compute_unitary()
    d = _dist(U, [1, 0, 1, 0, 0, 0])
    assert d[(1, 0, 1, 0, 0, 0)] == pytest.approx(1 / 9)
    assert d[(1, 0, 0, 1, 0, 0)] == pytest.approx(0, abs=1e-15)
    assert sum(p for s, p in d.items() if s[4] or s[5]) == pytest.approx(7 / 9)
    d = _dist(U, [0, 1, 0, 1, 0, 0])
    assert d[(0, 1, 1, 0, 0, 0)] == pytest.approx(1 / 9)


def test_reference_docs_examples():
    # docs/source/reference/backends/naive.rst:16-22 -> 0.5j;
    # docs/source/reference/backends/slos.rst:18-29 -> four states at 0.25
    U = (Circuit(4) // BS() // (2, BS())).compute_unitary()
    amp = oracle.amplitude_brute(U, [1, 0, 1, 0], [1, 0, 0, 1])
    assert amp == pytest.approx(0.5j)
    d = _dist(U, [1, 0, 1, 0])
    nz = {s: p for s, p in d.items() if p > 1e-12}
    assert nz == pytest.approx({(1, 0, 1, 0): 0.25, (1, 0, 0, 1): 0.25, (0, 1, 1, 0): 0.25, (0, 1, 0, 1): 0.25})


def test_reference_evolve_signs():
    # tests/backends/test_backends.py:280-289: BS.H|1,1> -> (|2,0> - |0,2>)/sqrt(2)
    U = BS.H().compute_unitary()
    c = oracle.slos_coefs(U, [1, 1])
    amps = {s: c[i] * math.sqrt(oracle.prodnfact(s)) for i, s in
            enumerate(tuple(x) for x in oracle.fock_enumerate(2, 2).tolist())}
    assert amps[(2, 0)] == pytest.approx(math.sqrt(2) / 2)
    assert amps[(0, 2)] == pytest.approx(-math.sqrt(2) / 2)
    assert abs(amps[(1, 1)]) < 1e-15


def test_reference_data_unitary_slos():
    # tests/data/u_random_8 is a unitary the reference's tests load as input;
    # its values are copied into tests/golden/u_random_8.npy by make_golden.
    path = os.path.join(HERE, "golden", "u_random_8.npy")
    U = np.load(path)
    assert np.allclose(U @ U.conj().T, np.eye(8), atol=1e-8)
    p = oracle.slos_probs(U, [1, 1, 1, 1, 0, 0, 0, 0])
    assert len(p) == 330
    assert np.max(np.abs(p - _exact_probs(U, [1, 1, 1, 1, 0, 0, 0, 0]))) < 1e-14


# --------------------------------------------------------------------- sampler

def test_cc_band_pinned_by_reference():
    # tests/backends/test_backends.py:58-67: BS.H, |0,1>, 10000 samples, each
    # outcome in [47.5%, 52.5%]
    s = oracle.cc_samples(BS.H().compute_unitary(), [0, 1], 10000, seed=1)
    c = Counter(map(tuple, s.tolist()))
    assert 4750 < c[(0, 1)] < 5250 and 4750 < c[(1, 0)] < 5250


@pytest.mark.parametrize("inp", [[1, 1, 0, 0], [1, 0, 1, 1, 0], [2, 1, 0, 0]])
def test_cc_matches_exact_distribution(inp):
    U = haar_unitary(len(inp), seed=11)
    exact = oracle.distribution_exact(U, inp)
    N = 40000
    s = oracle.cc_samples(U, inp, N, seed=5)
    c = Counter(map(tuple, s.tolist()))
    # chi-square over outcomes with expected count >= 5
    chi2, dof = 0.0, -1
    for st, p in exact.items():
        if p * N >= 5:
            chi2 += (c.get(st, 0) - p * N) ** 2 / (p * N)
            dof += 1
    # p < 1e-4 threshold for the chi-square statistic (Wilson-Hilferty)
    z = ((chi2 / dof) ** (1 / 3) - (1 - 2 / (9 * dof))) / math.sqrt(2 / (9 * dof))
    assert z < 3.7, (chi2, dof)


def _moments(U, cols, m):
    """<n_j> = sum_i |U_ji|^2 and, for j != k, <n_j n_k> =
    sum_{i<i'} |U_ji U_ki' + U_ji' U_ki|^2 (single photons in modes cols)."""
    mean = np.array([sum(abs(U[j, i]) ** 2 for i in cols) for j in range(m)])
    pair = np.zeros((m, m))
    for j in range(m):
        for k in range(m):
            if j != k:
                pair[j, k] = sum(abs(U[j, i] * U[k, i2] + U[j, i2] * U[k, i]) ** 2
                                 for i, i2 in itertools.combinations(cols, 2))
    return mean, pair


def test_moment_identities_vs_brute_force():
    U = haar_unitary(5, seed=21)
    cols = [0, 2, 3]
    inp = [1, 0, 1, 1, 0]
    mean, pair = _moments(U, cols, 5)
    exact = oracle.distribution_exact(U, inp)
    bm = np.zeros(5)
    bp = np.zeros((5, 5))
    for s, p in exact.items():
        s = np.array(s)
        bm += p * s
        bp += p * np.outer(s, s)
    assert np.allclose(mean, bm, atol=1e-14)
    off = ~np.eye(5, dtype=bool)
    assert np.allclose(pair[off], bp[off], atol=1e-14)


def test_cc_moments_m16_n6():
    m, cols = 16, [0, 1, 2, 3, 4, 5]
    U = haar_unitary(m, seed=8)
    inp = [1] * 6 + [0] * 10
    s = oracle.cc_samples(U, inp, 20000, seed=3).astype(float)
    mean, pair = _moments(U, cols, m)
    se = np.sqrt(np.maximum(s.var(axis=0), 1e-3) / len(s))
    assert np.all(np.abs(s.mean(axis=0) - mean) < 5 * se)


def test_uniform_stream_is_deterministic_and_uniform():
    u = np.array([oracle.uniform(7, 3, i) for i in range(20000)])
    assert u.min() >= 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.01
    assert oracle.uniform(7, 3, 5) == oracle.uniform(7, 3, 5)
    assert oracle.uniform(7, 3, 5) != oracle.uniform(7, 4, 5)


# ------------------------------------------------------------- Gurvits estimate

_M64 = (1 << 64) - 1


def _smix(z):
    z = (z + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def _rand64(seed, stream, ctr):
    key = _smix(seed ^ _smix(stream))
    return _smix((key + ctr * 0x9E3779B97F4A7C15) & _M64)


def _glynn_sample(A, signs):
    x = np.array(signs, dtype=float)
    return np.prod(x) * np.prod(A @ x)


def test_gurvits_estimator_is_unbiased_over_all_sign_vectors():
    """Averaged over every x in {+-1}^n the estimator is exactly perm(A)."""
    import itertools
    rng = np.random.RandomState(5)
    for n in (1, 2, 3, 6, 9):
        A = rng.randn(n, n) + 1j * rng.randn(n, n)
        tot = sum(_glynn_sample(A, s) for s in itertools.product((1, -1), repeat=n))
        assert abs(tot / 2 ** n - oracle.perm(A)) <= 1e-12 * max(1.0, abs(oracle.perm(A)))


def test_gurvits_oracle_matches_python_restatement():
    """Sample s uses x_j = -1 where bit j of rand64(seed, stream, s) is set."""
    rng = np.random.RandomState(6)
    A = (rng.randn(7, 7) + 1j * rng.randn(7, 7)) / 3
    seed, stream, N = 1234, 17, 300
    xs = [_glynn_sample(A, [-1 if (_rand64(seed, stream, s) >> j) & 1 else 1 for j in range(7)]) for s in range(N)]
    mean = np.mean(xs)
    var = np.mean(np.abs(xs) ** 2) - abs(mean) ** 2
    est, err = oracle.gurvits(A, N, seed=seed, stream=stream)
    assert abs(est - mean) <= 1e-12 * np.sqrt(np.mean(np.abs(xs) ** 2))
    assert abs(err - 2.5758293035489004 * np.sqrt(var / N)) <= 1e-9 * err


def test_gurvits_oracle_exact_cases_and_coverage():
    est, err = oracle.gurvits(np.eye(6), 50, seed=1)
    assert est == 1.0
    est, err = oracle.gurvits(np.array([[2 + 1j]]), 5, seed=1)
    assert est == 2 + 1j
    # 99% intervals cover the exact permanent (Haar submatrix, 200 independent estimates)
    from perceval_amd.utils.workloads import haar_unitary
    U = haar_unitary(10, seed=9)
    A = U[:6, :6]
    exact = oracle.perm(A)
    misses = sum(abs(oracle.gurvits(A, 2000, seed=7, stream=k)[0] - exact) > oracle.gurvits(A, 2000, seed=7, stream=k)[1]
                 for k in range(200))
    assert misses <= 8   # expectation 2 at 99%


def test_perm_large_threaded_equals_single_thread_glynn():
    """oracle.perm_large (chunked over threads, for orders 25..28) against the
    single-threaded long-double Glynn, the mpmath goldens and a closed form;
    its result does not depend on the thread count."""
    import json
    import os
    U = haar_unitary(64, seed=3)
    rng = np.random.RandomState(0)
    for n in (0, 1, 2, 5, 12, 17, 20):
        M = U[np.ix_(rng.choice(64, n, replace=False), rng.choice(64, n, replace=False))]
        a = oracle.perm(M)
        b = oracle.perm_large(M, threads=4)
        assert abs(a - b) <= 1e-15 * max(abs(a), 1e-300), n
        assert b == oracle.perm_large(M, threads=1)
    for c in _golden():
        A = np.array(c["re"]) + 1j * np.array(c["im"])
        want = complex(float(c["perm"][0]), float(c["perm"][1]))
        assert _rel(oracle.perm_large(A, threads=3), want) < 1e-13, c["n"]
    J = np.ones((22, 22), dtype=complex) / 2
    assert abs(oracle.perm_large(J, threads=8) / (math.factorial(22) / 2 ** 22) - 1) < 1e-14
