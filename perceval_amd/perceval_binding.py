"""Reference-side binding: route Perceval's Naive, SLOS and
CliffordClifford2017 backends to this library.

`install(perceval.backends)` subclasses the module's own backend classes and
assigns the subclasses into its BACKEND_LIST, the one dict that
BackendFactory, Processor, SimulatorFactory and Stepper look names up in
(perceval/backends/__init__.py:39-58).  Every method of the parents that
reaches exqalibur's arithmetic is overridden:

  SLOS    set_circuit, set_input_state, preprocess (no _Path.compute /
          compute_slos_layer, _slos.py:130-185), prob_amplitude, all_prob,
          prob_distribution, evolve (no all_prob_normalize_output or
          exqalibur coefficients, :187-223), and the output-state iterator
          (GPU ranking instead of FSArray / FSMask enumeration);
  Naive   _compute_permanent (a cached per-order plan instead of
          xq.permanent_cx, _naive.py:70-71), all_prob / prob_distribution /
          evolve batched into one permanent launch, the iterator as above;
  CliffordClifford2017
          __init__, set_circuit, set_input_state, sample, samples (no
          xq.Clifford2017, _clifford2017.py:36-61).

The parents keep only bookkeeping: ABackend.set_circuit's compute_unitary
(numpy) and the input-state and mask checks.  The container types the
methods return (FockState, BSDistribution, StateVector, BSSamples) are taken
from the `types` argument, Perceval's own (perceval.utils) by default.

The same function accepts perceval_amd.backends as the parent module, which
is how the binding is tested where Perceval is not installed.
"""
from __future__ import annotations

import math
import types as _types

import numpy as np

from . import _native
from . import permanent as _permanent
from . import sampling as _sampling
from . import slos as _slos
from .utils._random import next_sampler_seed


def _rank(state) -> int:
    """FSArray rank of an occupation vector (descending lexicographic order):
    sum over q < m-1 of C(R_{q+1} + m-q-2, m-q-1), R_q = photons in modes >= q."""
    s = [int(v) for v in state]
    m, R, r = len(s), sum(s), 0
    for q in range(m - 1):
        R -= s[q]
        r += math.comb(R + m - q - 2, m - q - 1)
    return r


class _LazyPlan:
    """A single-input SLOS plan built on first use: under a mask all_prob
    usually runs a masked plan, and the full plan is then never built."""

    def __init__(self, occ):
        self._occ = list(occ)
        self._plan = None

    def run(self, *args, **kwargs):
        if self._plan is None:
            self._plan = _slos.SlosPlan(self._occ)
        return self._plan.run(*args, **kwargs)


def _default_types():
    import perceval.utils as pu
    return _types.SimpleNamespace(FockState=pu.FockState, BSDistribution=pu.BSDistribution,
                                  StateVector=pu.StateVector, BSSamples=pu.BSSamples)


def make_backends(backends_module, types=None) -> dict:
    """{name: class} of GPU-routed subclasses of the module's Naive, SLOS and
    Clifford2017 backends."""
    t = types or _default_types()
    FockState, BSDistribution, StateVector, BSSamples = t.FockState, t.BSDistribution, t.StateVector, t.BSSamples
    SLOSBase = backends_module.SLOSBackend
    NaiveBase = backends_module.NaiveBackend
    CCBase = backends_module.Clifford2017Backend

    class _GpuStates:
        """Output states in FSArray order (masked: the kept ones), from the
        GPU ranking, cached per photon number as the parents cache their
        iterators (_abstract_backends.py:152-158)."""

        def _ranks(self, istate):
            key = ("ranks", istate.m, istate.n)
            cache = self._gpu_cache
            if key not in cache:
                if self._masks_str is not None:
                    cache[key] = _slos.fock_mask_select(istate.m, istate.n, self._masks_str, self._no_limit_modes,
                                                        self._mask_n)
                else:
                    cache[key] = None
            return cache[key]

        def _get_iterator(self, input_state):
            n = input_state.n
            if n not in self._cache_iterator:
                m = input_state.m
                idx = self._ranks(input_state)
                ranks = np.arange(_slos.fock_count(m, n), dtype=np.int64) if idx is None else idx.cpu().numpy()
                rows = _slos.fock_unrank(ranks, m, n) if len(ranks) else np.zeros((0, m), dtype=np.int32)
                self._cache_iterator[n] = tuple(FockState(r) for r in rows.tolist())
            return self._cache_iterator[n]

    class MI355XSLOSBackend(_GpuStates, SLOSBase):
        def __init__(self, mask=None, use_symbolic=False):
            if use_symbolic:
                raise NotImplementedError("symbolic SLOS is not supported by the MI355X backend")
            self._gpu_cache = {}
            self._plans = {}
            SLOSBase.__init__(self, mask=mask, use_symbolic=False)

        def _reset(self):
            self._fsms = [[]]
            self._fsas = {}
            self._mk_l = [1]
            self._path_roots = []
            self._state_mapping = {}
            self._plans = {}
            self._gpu_cache = {}
            self._mask = None
            self.clear_iterator_cache()

        def set_circuit(self, circuit):
            previous = self._circuit
            assert not circuit.requires_polarization, "Circuit must not contain polarized components"
            self._input_state = None
            self._circuit = circuit
            self._umat = np.asarray(circuit.compute_unitary(), dtype=np.complex128)
            self._results = {}
            if not (self._plans and previous is not None and previous.m == circuit.m):
                self._reset()

        def set_input_state(self, input_state):
            super(SLOSBase, self).set_input_state(input_state)   # checks + mask, no SLOS arithmetic
            self.preprocess([input_state])

        def preprocess(self, input_list) -> bool:
            new = [s for s in input_list if s not in self._plans]
            if not new:
                return False
            m = self._circuit.m
            for s in new:
                for k in range(len(self._fsms), s.n + 1):
                    self._fsms.append(("layer", m, k))
            if len(new) == 1:
                self._plans[new[0]] = (_slos.SlosPlan(list(new[0])) if self._masks_str is None
                                       else _LazyPlan(new[0]))
            else:
                tree = _slos.SlosTree([list(s) for s in new])
                for i, s in enumerate(new):
                    self._plans[s] = (tree, i)
            for s in new:
                self._state_mapping[s] = self._plans[s]
            return True

        def _device(self, istate, mode):
            key = (istate, mode)
            if key not in self._results:
                plan = self._plans[istate]
                if isinstance(plan, tuple):
                    self._results[key] = plan[0].run(self._umat, mode)[plan[1]]
                else:
                    self._results[key] = plan.run(self._umat, mode)
            return self._results[key]

        def _host(self, istate, mode):
            key = ("host", istate, mode)
            if key not in self._results:
                self._results[key] = self._device(istate, mode).cpu().numpy()
            return self._results[key]

        def prob_amplitude(self, output_state) -> complex:
            if self._input_state.n != output_state.n:
                return complex(0)
            idx = _rank(output_state)
            coefs = self._host(self._input_state, _slos.COEFS)
            assert 0 <= idx < len(coefs)
            return complex(coefs[idx] * math.sqrt(output_state.prodnfact() / self._input_state.prodnfact()))

        def _mask_plan_args(self, istate):
            """(condition rows, at-least modes, mask photon count) for a
            masked plan, or None when the mask goes through selection from the
            full output layer (no mode constrained, fewer than two modes free
            in every row, more than 64 rows, a photon count outside 1..30, or
            a mask set for fewer photons than the input's)."""
            if self._masks_str is None or len(self._masks_str) > 64:
                return None
            cond = _slos.mask_conditions(self._masks_str, istate.m)
            free = (cond < 0).all(axis=0)
            if free.sum() < 2 or free.all() or istate.n < 1 or istate.n > 30:
                return None
            mask_n = self._mask_n or istate.n
            if mask_n < istate.n:
                return None
            return cond, tuple(int(q) for q in (self._no_limit_modes or ())), int(mask_n)

        def all_prob(self, input_state=None):
            if input_state is not None:
                self.set_input_state(input_state)
            istate = self._input_state
            args = self._mask_plan_args(istate)
            if args is not None:
                # every layer restricted to the states that can still match
                # some mask row (the masked FSArray of _slos.py:161-164), the
                # matching outputs in FSArray order
                cond, at_least, mask_n = args
                key = ("masked", istate, cond.tobytes(), at_least, mask_n)
                if key not in self._results:
                    try:
                        plan = _slos.SlosPlan(list(istate), mask_cond=cond, at_least=at_least, mask_n=mask_n)
                        self._results[key] = plan.run(self._umat, _slos.PROBS).cpu().numpy()
                    except _native.NativeError as e:
                        if "failed (-4)" not in str(e):   # out of range for a masked plan: select instead
                            raise
                        self._results[key] = None   # too many constrained configurations: select instead
                if self._results[key] is not None:
                    return self._results[key].tolist()
            idx = self._ranks(istate)
            if idx is None:
                return self._host(istate, _slos.PROBS).tolist()
            return _slos.gather_f64(self._device(istate, _slos.PROBS), idx).cpu().numpy().tolist()

        def prob_distribution(self):
            bsd = BSDistribution()
            for s, p in zip(self._get_iterator(self._input_state), self.all_prob()):
                bsd.add(s, p)
            return bsd

        def evolve(self):
            istate = self._input_state
            coefs = self._host(istate, _slos.COEFS)
            idx = self._ranks(istate)
            sel = np.arange(len(coefs)) if idx is None else idx.cpu().numpy()
            res = StateVector()
            pin = istate.prodnfact()
            for s, c in zip(self._get_iterator(istate), coefs[sel]):
                res += s * complex(c * math.sqrt(s.prodnfact() / pin))
            return res

        def clear_mask(self):
            super().clear_mask()
            self._gpu_cache = {}

    class MI355XNaiveBackend(_GpuStates, NaiveBase):
        def __init__(self, *args, **kwargs):
            self._gpu_cache = {}
            self._single = _permanent.SinglePermanents()
            self._batch = _permanent.BatchPermanents()
            NaiveBase.__init__(self, *args, **kwargs)

        def _compute_permanent(self, M):
            return self._single(M)

        def clear_mask(self):
            super().clear_mask()
            self._gpu_cache = {}

        def set_circuit(self, circuit):
            super().set_circuit(circuit)
            self._umat = np.asarray(self._umat, dtype=np.complex128)

        def _all_amplitudes(self):
            istate = self._input_state
            states = self._get_iterator(istate)
            if istate.n <= 1 or not states:
                return np.array([self.prob_amplitude(s) for s in states], dtype=complex)
            mats, norms = _permanent.output_submatrices(self._umat, [list(s) for s in states], list(istate))
            return self._batch(mats) / norms

        def all_prob(self, input_state=None):
            if input_state is not None:
                self.set_input_state(input_state)
            return (np.abs(self._all_amplitudes()) ** 2).tolist()

        def prob_distribution(self):
            bsd = BSDistribution()
            for s, p in zip(self._get_iterator(self._input_state), self.all_prob()):
                bsd.add(s, p)
            return bsd

        def evolve(self):
            res = StateVector()
            for s, a in zip(self._get_iterator(self._input_state), self._all_amplitudes()):
                res += s * complex(a)
            return res

    class MI355XClifford2017Backend(CCBase):
        def __init__(self):
            super(CCBase, self).__init__()        # no xq.Clifford2017 object

        def set_circuit(self, circuit):
            super(CCBase, self).set_circuit(circuit)
            self._umat = np.asarray(self._umat, dtype=np.complex128)

        def set_input_state(self, input_state):
            super(CCBase, self).set_input_state(input_state)

        def sample(self):
            return self.samples(1)[0]

        def samples(self, count: int):
            occ = _sampling.cc_samples(self._umat, list(self._input_state), int(count), seed=next_sampler_seed())
            out = BSSamples()
            for row in occ.tolist():
                out.append(FockState(row))
            return out

    return {"SLOS": MI355XSLOSBackend, "Naive": MI355XNaiveBackend, "CliffordClifford2017": MI355XClifford2017Backend}


def install(backends_module=None, types=None) -> dict:
    """Assign the GPU-routed classes into backends_module.BACKEND_LIST
    (perceval.backends by default); returns the classes they replaced."""
    if backends_module is None:
        import perceval.backends as backends_module
    classes = make_backends(backends_module, types)
    previous = {k: backends_module.BACKEND_LIST.get(k) for k in classes}
    backends_module.BACKEND_LIST.update(classes)
    return previous
