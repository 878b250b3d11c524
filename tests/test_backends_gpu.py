"""Backend API on the GPU against the expected values the reference's own
tests and documentation hold.

The values are data (tests/golden/backend_cases.json, each case with the
reference file:line it comes from); this module only builds each case's
circuit, sets up the named backends through BackendFactory and applies one
checker per kind of case.  Strong-simulation values are held to 1e-12
absolute, the SLOS bar."""
import json
import math
import os

import numpy as np
import pytest

from perceval_amd import BS, PERM, PS, BackendFactory, BasicState, Circuit, SLOSBackend, catalog

pytestmark = pytest.mark.gpu

TOL = 1e-12
_CASES = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "backend_cases.json")))["cases"]


def _component(spec):
    if "bs" in spec:
        kw = {k: v for k, v in spec["bs"].items() if k != "convention"}
        return {"H": BS.H, "Rx": BS.Rx}[spec["bs"]["convention"]](**kw)
    if "ps" in spec:
        return PS(spec["ps"])
    if "perm" in spec:
        return PERM(spec["perm"])
    raise ValueError(spec)


def build_circuit(spec):
    if "identity" in spec:
        return Circuit(spec["identity"])
    if "catalog" in spec:
        return catalog[spec["catalog"]].build_circuit()
    if "bs" in spec:
        return _component(spec)
    c = Circuit(spec["m"])
    for modes, comp in spec["ops"]:
        c.add(tuple(modes) if len(modes) > 1 else modes[0], _component(comp))
    return c


def make_backend(name, case, circuit_spec=None):
    b = BackendFactory.get_backend() if name == "<default>" else BackendFactory.get_backend(name)
    if "mask" in case:
        b.set_mask(case["mask"])
    b.set_circuit(build_circuit(circuit_spec or case["circuit"]))
    return b


def _check_distribution(b, inp, expected):
    b.set_input_state(BasicState(inp))
    got = b.prob_distribution()
    want = {BasicState(k): v for k, v in expected.items()}
    for s, p in got.items():
        assert abs(p - want.get(s, 0.0)) < TOL, (s, p, want.get(s))
    for s, p in want.items():
        assert abs(got.get(s, 0.0) - p) < TOL, (s, p)
    assert abs(sum(got.values()) - 1.0) < TOL


def _distribution(case, name):
    _check_distribution(make_backend(name, case), case["input"], case["expected"])


def _probabilities(case, name):
    b = make_backend(name, case)
    for inp, out, p in case["table"]:
        b.set_input_state(BasicState(inp))
        assert abs(b.probability(BasicState(out)) - p) < TOL, (inp, out)


def _mass_where_any(case, name):
    b = make_backend(name, case)
    b.set_input_state(BasicState(case["input"]))
    mass = sum(p for s, p in b.prob_distribution().items() if any(s[j] for j in case["modes"]))
    assert abs(mass - case["expected"]) < TOL


def _exact_support(case, name):
    b = make_backend(name, case)
    b.set_input_state(BasicState(case["input"]))
    got = b.prob_distribution()
    want = {BasicState(k): v for k, v in case["expected"].items()}
    assert set(got.keys()) == set(want.keys())
    for s, p in want.items():
        assert abs(got[s] - p) < TOL


def _refresh(case, name):
    b = BackendFactory.get_backend(name)
    for step in case["steps"]:
        b.set_circuit(build_circuit(step["circuit"]))
        for inp in step.get("inputs", []):
            b.set_input_state(BasicState(inp))
        _check_distribution(b, step["check"], step["expected"])


def _evolve(case, name):
    b = make_backend(name, case)
    b.set_input_state(BasicState(case["input"]))
    sv = b.evolve()
    want = {BasicState(k): complex(*v) for k, v in case["expected"].items()}
    for s in set(sv.keys()) | set(want.keys()):
        assert abs(sv[s] - want.get(s, 0)) < TOL, s


def _sample_band(case, name):
    b = make_backend(name, case)
    b.set_input_state(BasicState(case["input"]))
    counts = b.samples(case["samples"]).to_counts()
    for s, (lo, hi) in case["band"].items():
        assert lo * case["samples"] < counts[BasicState(s)] < hi * case["samples"], s


def _amplitude(case, name):
    b = make_backend(name, case)
    b.set_input_state(BasicState(case["input"]))
    assert abs(b.prob_amplitude(BasicState(case["output"])) - complex(*case["expected"])) < TOL


CHECKS = {"distribution": _distribution, "probabilities": _probabilities, "mass_where_any": _mass_where_any,
          "exact_support": _exact_support, "refresh": _refresh, "evolve": _evolve, "sample_band": _sample_band,
          "amplitude": _amplitude}


@pytest.mark.parametrize("case,name", [(c, n) for c in _CASES for n in c["backends"]],
                         ids=[f"{c['id']}-{n}" for c in _CASES for n in c["backends"]])
def test_reference_case(gpu, case, name):
    CHECKS[case["kind"]](case, name)


def test_every_case_kind_has_a_checker():
    assert {c["kind"] for c in _CASES} <= set(CHECKS)


def test_slos_layers_are_shared_across_inputs(gpu):
    """_fsms grows one layer per photon number and keeps its layer objects
    when a new input reuses them (the deployed path of _slos.py:151-168)."""
    slos = SLOSBackend()
    slos.set_circuit(Circuit(2))
    slos.set_input_state(BasicState([0, 1]))
    one_photon = slos._fsms[-1]
    assert len(slos._fsms) == 2
    slos.set_input_state(BasicState([1, 0]))
    assert len(slos._fsms) == 2 and slos._fsms[-1] is one_photon
    slos.set_input_state(BasicState([1, 1]))
    assert len(slos._fsms) == 3 and slos._fsms[1] is one_photon


@pytest.mark.parametrize("name", ["Naive", "SLOS"])
def test_iterator_cache_cleared_by_a_wider_circuit(gpu, name):
    b = BackendFactory.get_backend(name)
    b.set_circuit(Circuit(5).add(0, BS.H()))
    b.set_input_state(BasicState([1, 1, 0, 0, 0]))
    b.evolve()
    if name == "Naive":
        assert len(b._cache_iterator) != 0
    b.set_circuit(Circuit(7).add(3, BS.H()))
    assert len(b._cache_iterator) == 0


@pytest.mark.parametrize("name", ["Naive", "SLOS"])
def test_prob_iterator_drops_states_below_min_p(gpu, name):
    """prob_iterator(min_p) (_abstract_backends.py:197-201) yields the states
    whose probability exceeds min_p, in all_prob order."""
    b = BackendFactory.get_backend(name)
    b.set_circuit(build_circuit({"m": 3, "ops": [[[0, 1], {"bs": {"convention": "H"}}], [[1], {"ps": math.pi / 4}],
                                                 [[1, 2], {"bs": {"convention": "H"}}]]}))
    b.set_input_state(BasicState("|0,1,1>"))
    probs = b.all_prob()
    states = list(b._get_iterator(BasicState("|0,1,1>")))
    for min_p in (1e-16, 0.2, 0.3):
        it = list(b.prob_iterator(min_p))
        assert [s for s, _ in it] == [s for s, p in zip(states, probs) if p > min_p]
        assert all(p > min_p for _, p in it)
    assert len(list(b.prob_iterator(1e-16))) == 4     # |0,1,1> and |2,0,0> have probability 0
    assert np.isclose(sum(p for _, p in b.prob_iterator()), 1.0, atol=TOL)
