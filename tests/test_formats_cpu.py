"""CPU tests of the host-side formats and metrics: the BSSamples wire format
(perceval/serialization/_state_serialization.py:68-92), tvd_dist
(perceval/utils/dist_metrics.py:34-52), the marginal TVDs that judge the
learned sampler, and the seed plumbing of random_seed
(perceval/utils/_random.py:36-58)."""
import base64
import zlib

import numpy as np
import pytest

from perceval_amd import BSSamples, BasicState, random_seed, serialization, tvd_dist
from perceval_amd.mis import marginal_tvds
from perceval_amd.utils import _random


def _reference_sample_list():
    # tests/serialization/test_serialization.py:206-216 of the reference
    samples = BSSamples()
    for _ in range(50):
        samples.extend([BasicState([0, 1, 0])] * 11)
        samples.extend([BasicState([1, 0, 0])] * 13)
        samples.extend([BasicState([0, 0, 1])] * 17)
    return samples


def test_bssamples_round_trip_reference_case():
    samples = _reference_sample_list()
    assert serialization.deserialize(serialization.serialize(samples)) == samples
    assert serialization.deserialize(serialization.serialize(samples, compress=False)) == samples
    empty = BSSamples()
    assert serialization.deserialize(serialization.serialize(empty)) == empty


def test_bssamples_body_layout():
    """Distinct states in order of first appearance, then one index per
    sample (_state_serialization.py:68-79)."""
    s = [BasicState("|0,1>"), BasicState("|1,0>"), BasicState("|0,1>"), BasicState("|2,0>"), BasicState("|1,0>")]
    assert serialization.serialize_bssamples(s) == "|0,1>;|1,0>;|2,0>/0;1;0;2;1"
    assert serialization.serialize(s, compress=False) == ":PCVL:BSSamples:|0,1>;|1,0>;|2,0>/0;1;0;2;1"
    z = serialization.serialize(s)
    assert z.startswith(":PCVL:zip:")
    assert zlib.decompress(base64.b64decode(z[len(":PCVL:zip:"):])).decode() == \
        ":PCVL:BSSamples:|0,1>;|1,0>;|2,0>/0;1;0;2;1"
    assert serialization.serialize_bssamples([]) == "/"


def test_bssamples_array_form_equals_list_form():
    rng = np.random.RandomState(0)
    occ = rng.randint(0, 3, size=(2000, 5)).astype(np.int32)
    occ[rng.rand(2000) < 0.5] = occ[0]          # many repeats of one row
    body = serialization.serialize_bssamples_array(occ)
    assert body == serialization.serialize_bssamples(BSSamples.from_array(occ))
    assert np.array_equal(serialization.deserialize_bssamples_array(body), occ)
    assert serialization.serialize(occ, compress=False) == serialization.serialize(BSSamples.from_array(occ),
                                                                                   compress=False)


def test_bssamples_rejects_other_tags():
    with pytest.raises(NotImplementedError):
        serialization.deserialize(":PCVL:BSCount:{}")
    with pytest.raises(ValueError):
        serialization.deserialize("|0,1>/0")


def test_tvd_dist_cases():
    a = {BasicState("|1,0>"): 0.25, BasicState("|0,1>"): 0.75}
    b = {BasicState("|1,0>"): 0.5, BasicState("|0,1>"): 0.5}
    assert tvd_dist(a, a) == 0.0
    assert tvd_dist(a, b) == pytest.approx(0.25)
    assert tvd_dist(a, b) == tvd_dist(b, a)
    # disjoint supports: a missing key counts as probability 0
    c = {BasicState("|2,0>"): 1.0}
    assert tvd_dist(a, c) == pytest.approx(1.0)
    assert tvd_dist({}, {}) == 0.0


def test_marginal_tvds_known_answer():
    """Set A: half |1,0,0>, half |0,1,0>; set B: all |1,0,0>.  Modes 0 and 1
    differ by 0.5, mode 2 not at all; every mode pair's joint law differs by
    0.5."""
    a = np.array([[1, 0, 0]] * 50 + [[0, 1, 0]] * 50, dtype=np.int32)
    b = np.array([[1, 0, 0]] * 100, dtype=np.int32)
    t = marginal_tvds(a, b, pairs=30, seed=1)
    assert t["tvd1_max"] == pytest.approx(0.5)
    assert t["tvd1_mean"] == pytest.approx(1.0 / 3.0)
    assert t["tvd2_max"] == pytest.approx(0.5) and t["tvd2_mean"] == pytest.approx(0.5)
    same = marginal_tvds(a, a.copy(), pairs=30)
    assert same["tvd1_max"] == 0.0 and same["tvd2_max"] == 0.0


def test_marginal_tvds_match_tvd_dist():
    """The 1-mode marginal TVD equals tvd_dist of the two per-mode count laws."""
    rng = np.random.RandomState(2)
    a = rng.randint(0, 3, size=(500, 4)).astype(np.int32)
    b = rng.randint(0, 2, size=(700, 4)).astype(np.int32)
    t = marginal_tvds(a, b, pairs=10)
    per_mode = []
    for j in range(4):
        da = {v: c / len(a) for v, c in zip(*np.unique(a[:, j], return_counts=True))}
        db = {v: c / len(b) for v, c in zip(*np.unique(b[:, j], return_counts=True))}
        per_mode.append(tvd_dist(da, db))
    assert t["tvd1_max"] == pytest.approx(max(per_mode), abs=1e-15)
    assert t["tvd1_mean"] == pytest.approx(np.mean(per_mode), abs=1e-15)


def test_random_seed_makes_sampler_seeds_reproducible():
    random_seed(1234)
    first = [_random.next_sampler_seed() for _ in range(4)]
    py_draw = np.random.rand()
    random_seed(1234)
    assert [_random.next_sampler_seed() for _ in range(4)] == first
    assert np.random.rand() == py_draw           # numpy is seeded too
    random_seed(1235)
    assert [_random.next_sampler_seed() for _ in range(4)] != first
    assert len(set(first)) == 4                  # successive calls get distinct streams


# ---------------------------------------------------------------- exact marginals

@pytest.mark.parametrize("m,inp_modes", [(5, (0, 1, 2)), (6, (0, 2, 3, 5)), (7, (1, 2, 3, 4, 6))])
def test_exact_pair_marginals_vs_brute_force_distribution(m, inp_modes):
    """Every 2-mode law of perceval_amd.marginals against the marginal of the
    full output distribution, each probability a long-double permanent of the
    oracle."""
    import oracle
    from perceval_amd.marginals import mode_marginal, pair_marginal
    from perceval_amd.utils.workloads import haar_unitary
    U = haar_unitary(m, seed=m)
    inp = [1 if i in inp_modes else 0 for i in range(m)]
    n = sum(inp)
    dist = oracle.distribution_exact(U, inp)
    for j in range(m):
        ref1 = np.zeros(n + 1)
        for s, p in dist.items():
            ref1[s[j]] += p
        assert np.max(np.abs(mode_marginal(U, inp, j) - ref1)) < 1e-13
        for k in range(m):
            if k == j:
                continue
            ref = np.zeros((n + 1, n + 1))
            for s, p in dist.items():
                ref[s[j], s[k]] += p
            assert np.max(np.abs(pair_marginal(U, inp, j, k) - ref)) < 1e-13, (j, k)


def test_exact_marginals_m60_n20_sum_rule_and_moments():
    import itertools
    from perceval_amd.marginals import exact_marginals
    from perceval_amd.utils.workloads import haar_unitary
    U = haar_unitary(60, seed=60)
    inp = [1] * 20 + [0] * 40
    ex = exact_marginals(U, inp, pairs=12, seed=3)
    for j, law in ex["one"].items():
        assert abs(law.sum() - 1) < 1e-12 and law.min() > -1e-14
        assert abs(np.dot(np.arange(21), law) - (np.abs(U[j, :20]) ** 2).sum()) < 1e-12
    for (j, k), law in ex["two"].items():
        assert abs(law.sum() - 1) < 1e-12 and law.min() > -1e-14
        a = np.arange(21)
        pair = sum(abs(U[j, i] * U[k, i2] + U[j, i2] * U[k, i]) ** 2 for i, i2 in itertools.combinations(range(20), 2))
        assert abs(a @ law @ a - pair) < 1e-12
        assert np.max(np.abs(law.sum(axis=1) - ex["one"][j])) < 1e-12


def test_marginal_tvds_exact_of_the_law_itself_is_zero():
    from perceval_amd.marginals import exact_marginals, marginal_tvds_exact
    # a deterministic sampler: identity circuit, all photons stay put
    U = np.eye(4, dtype=complex)
    inp = [1, 0, 1, 0]
    ex = exact_marginals(U, inp, pairs=6)
    occ = np.array([inp] * 10, dtype=np.int32)
    t = marginal_tvds_exact(occ, ex)
    assert t["tvd1_max"] < 1e-15 and t["tvd2_max"] < 1e-15
    t2 = marginal_tvds_exact(np.array([[0, 1, 1, 0]] * 10, dtype=np.int32), ex)
    assert t2["tvd1_max"] == pytest.approx(1.0)
