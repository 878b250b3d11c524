"""GPU parity of the Clifford & Clifford sampler.

The GPU and the oracle draw from the same counter-based uniforms, so for the
same seed they produce the same samples except where a uniform falls within
rounding distance of a CDF boundary; the statistical checks then compare with
exact distributions (small sizes) and closed-form moments (m=60, n=20)."""
import itertools
import math
from collections import Counter

import numpy as np
import pytest

import oracle
from perceval_amd.utils.workloads import haar_unitary

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[cc-gpu] {msg}", flush=True)


@pytest.mark.parametrize("m,inp", [(8, [1, 1, 1, 1, 0, 0, 0, 0]), (6, [2, 0, 1, 0, 0, 1]), (3, [0, 0, 1])])
def test_same_seed_same_samples(gpu, m, inp):
    from perceval_amd.sampling import cc_samples
    U = haar_unitary(m, seed=m)
    g = cc_samples(U, inp, 2000, seed=99)
    c = oracle.cc_samples(U, inp, 2000, seed=99)
    assert (g.sum(axis=1) == sum(inp)).all()
    same = np.mean(np.all(g == c, axis=1))
    assert same >= 0.999, same


def test_same_seed_same_samples_m60_n20(gpu):
    from perceval_amd.sampling import cc_samples
    U = haar_unitary(60, seed=60)
    inp = [1] * 20 + [0] * 40
    _say("GPU 48 samples")
    g = cc_samples(U, inp, 48, seed=5, first_sample=1000)
    _say("oracle 48 samples")
    c = oracle.cc_samples(U, inp, 48, seed=5, first_sample=1000)
    _say("done")
    assert np.mean(np.all(g == c, axis=1)) >= 0.97


@pytest.mark.parametrize("n,count", [(26, 8), (28, 1)])
def test_same_seed_same_samples_up_to_28_photons(gpu, n, count):
    """Steps above 20 photons run split over many waves per sample (2^{k-20}
    parts of the 2^{k-2} Gray-code terms): the oracle's samples, seed for
    seed, at 26 and 28 photons (one sample at 28: a single sample fills the
    device)."""
    from perceval_amd.sampling import cc_samples
    m = n + 10
    U = haar_unitary(m, seed=m)
    inp = [1] * n + [0] * 10
    _say(f"GPU {count} samples, n={n}")
    g = cc_samples(U, inp, count, seed=7)
    _say(f"oracle {count} samples, n={n}")
    c = oracle.cc_samples(U, inp, count, seed=7)
    _say("done")
    assert (g.sum(axis=1) == n).all()
    assert np.array_equal(g, c)


@pytest.mark.parametrize("n,count", [(22, 6), (30, 2)])
def test_split_steps_do_not_depend_on_the_call(gpu, n, count):
    """A sample's draws depend on (seed, sample index) only: one call of
    `count` samples equals `count` calls of one sample each (first_sample
    set), with the split steps cut the same way in both, up to 30 photons."""
    from perceval_amd.sampling import cc_samples
    m = n + 8
    U = haar_unitary(m, seed=3 * m)
    inp = [1] * n + [0] * 8
    _say(f"n={n}: one call of {count}")
    together = cc_samples(U, inp, count, seed=11)
    _say(f"n={n}: {count} calls of one")
    apart = np.concatenate([cc_samples(U, inp, 1, seed=11, first_sample=i) for i in range(count)])
    assert (together.sum(axis=1) == n).all()
    assert np.array_equal(together, apart)


@pytest.mark.parametrize("inp", [[1, 1, 0, 1, 0], [2, 1, 0, 0]])
def test_distribution_chi_square(gpu, inp):
    from perceval_amd.sampling import cc_samples
    U = haar_unitary(len(inp), seed=13)
    exact = oracle.distribution_exact(U, inp)
    N = 50000
    s = cc_samples(U, inp, N, seed=17)
    cnt = Counter(map(tuple, s.tolist()))
    chi2, dof = 0.0, -1
    for st, p in exact.items():
        if p * N >= 5:
            chi2 += (cnt.get(st, 0) - p * N) ** 2 / (p * N)
            dof += 1
    z = ((chi2 / dof) ** (1 / 3) - (1 - 2 / (9 * dof))) / math.sqrt(2 / (9 * dof))
    assert z < 3.7, (chi2, dof)


def test_moments_m60_n20(gpu):
    from perceval_amd.sampling import cc_samples
    m = 60
    U = haar_unitary(m, seed=600)
    inp = [1] * 20 + [0] * 40
    N = 4000
    _say(f"GPU {N} samples m=60 n=20")
    s = cc_samples(U, inp, N, seed=11).astype(float)
    _say("done")
    assert (s.sum(axis=1) == 20).all()
    mean = (np.abs(U[:, :20]) ** 2).sum(axis=1)
    se = np.sqrt(np.maximum(s.var(axis=0), 1e-4) / N)
    assert np.all(np.abs(s.mean(axis=0) - mean) < 5 * se)
    # pair moments on a few mode pairs
    rng = np.random.RandomState(0)
    for j, k in rng.choice(m, size=(8, 2), replace=False):
        exp = sum(abs(U[j, i] * U[k, i2] + U[j, i2] * U[k, i]) ** 2 for i, i2 in itertools.combinations(range(20), 2))
        x = s[:, j] * s[:, k]
        assert abs(x.mean() - exp) < 5 * max(x.std(), 1e-3) / math.sqrt(N)


def test_deterministic_replay(gpu):
    from perceval_amd.sampling import cc_samples
    U = haar_unitary(10, seed=1)
    a = cc_samples(U, [1] * 5 + [0] * 5, 500, seed=3)
    b = cc_samples(U, [1] * 5 + [0] * 5, 500, seed=3)
    c = cc_samples(U, [1] * 5 + [0] * 5, 500, seed=4)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)


def test_cc_samples_multi_equals_single_input_calls(gpu):
    """Rows of input i equal pcvl_cc_samples on input i with first_sample +
    row0_i, bit for bit (same counter streams), across photon numbers."""
    from perceval_amd.sampling import cc_samples, cc_samples_multi_device
    m = 10
    U = haar_unitary(m, seed=101)
    inputs = [[1, 1, 1, 1, 0, 0, 0, 0, 0, 0], [1, 1, 1, 0, 0, 0, 0, 0, 0, 0], [0] * m,
              [2, 0, 1, 0, 0, 0, 0, 0, 0, 1], [1, 0, 1, 1, 0, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, 0, 0, 0, 0, 5]]
    counts = [300, 257, 5, 0, 64, 100]
    seed, first = 4242, 1000
    out = cc_samples_multi_device(U, inputs, counts, seed=seed, first_sample=first).cpu().numpy()
    assert out.shape == (sum(counts), m)
    row = 0
    for inp, c in zip(inputs, counts):
        ref = cc_samples(U, inp, c, seed=seed, first_sample=first + row)
        assert np.array_equal(out[row:row + c], ref), inp
        assert (out[row:row + c].sum(axis=1) == sum(inp)).all()
        row += c
    # and seed for seed with the oracle on one of them
    row0 = counts[0] + counts[1] + counts[2] + counts[3]
    ref = oracle.cc_samples(U, inputs[4], counts[4], seed=seed, first_sample=first + row0)
    assert np.array_equal(out[row0:row0 + counts[4]], ref)


def test_backend_samples_for_inputs(gpu):
    from perceval_amd.backends import BackendFactory
    from perceval_amd.components import Unitary
    from perceval_amd.utils.states import BasicState
    U = haar_unitary(6, seed=7)
    b = BackendFactory.get_backend("CliffordClifford2017")
    b.set_circuit(Unitary(U))
    want = {BasicState([1, 1, 0, 0, 0, 0]): 50, BasicState([1, 0, 1, 0, 0, 0]): 20, BasicState([0, 0, 0, 0, 0, 0]): 3}
    got = b.samples_for_inputs(want)
    assert list(got) == list(want)
    for s, c in want.items():
        assert len(got[s]) == c
        assert all(o.n == s.n and o.m == 6 for o in got[s])
    with pytest.raises(AssertionError):
        b.samples_for_inputs({BasicState([1, 1, 0]): 2})
