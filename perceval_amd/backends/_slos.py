"""SLOS backend (perceval/backends/_slos.py:105-223) on the GPU.

Each input state deployed on its own owns a path (a SlosPlan: photon-addition
order from _Path._decompose, :61-86, and the Fock layers); several inputs
deployed by one preprocess call share a tree (a SlosTree, one _Path root,
:37-86,170-185), whose shared prefixes are computed once.  A new circuit with
the same number of modes keeps the deployed paths and only recomputes the
coefficients (:130-141).  all_prob returns the final layer's probabilities in
FSArray order (:205-214); prob_amplitude looks a state up by its rank (:187-193).
"""
from __future__ import annotations

import math

import numpy as np

from .._native import NativeError
from ..slos import COEFS, PROBS, SlosPlan, SlosTree, fock_mask_select, gather_f64, mask_conditions
from ..utils.states import BSDistribution, FockState, StateVector, fock_space
from ._abstract_backends import AStrongSimulationBackend, _as_circuit


def fock_rank_host(state) -> int:
    """rank(s) = sum_{q<m-1} C(R_{q+1} + m-q-2, m-q-1), R_q = photons in modes >= q."""
    s = list(state)
    m = len(s)
    R = sum(s)
    r = 0
    for q in range(m - 1):
        R -= s[q]
        r += math.comb(R + m - q - 2, m - q - 1)
    return r


class _Layer:
    """Marker for one deployed Fock layer (m modes, k photons)."""

    __slots__ = ("m", "k", "count")

    def __init__(self, m: int, k: int):
        self.m = m
        self.k = k
        self.count = math.comb(m + k - 1, k)


class _LazyPlan:
    """A single-input SlosPlan built on first use.  Under a mask all_prob
    usually runs a masked plan, and the full plan's tables (tens of ms of
    host work and tens of MB of device memory at m=24, n=10) are then never
    built."""

    def __init__(self, occ):
        self._occ = list(occ)
        self._plan = None

    def _get(self):
        if self._plan is None:
            self._plan = SlosPlan(self._occ)
        return self._plan

    def run(self, *args, **kwargs):
        return self._get().run(*args, **kwargs)

    def __getattr__(self, name):
        return getattr(self._get(), name)


class SLOSBackend(AStrongSimulationBackend):
    def __init__(self, mask=None, use_symbolic=False):
        super().__init__()
        if use_symbolic:
            raise NotImplementedError("symbolic SLOS is not supported by this simulation core")
        self._symb = False
        self._reset()
        if mask is not None:
            self.set_mask(mask)

    @property
    def name(self) -> str:
        return "SLOS"

    def _reset(self):
        self._fsms: list = [[]]
        self._state_mapping: dict = {}
        self._results: dict = {}
        self._plans_masked: dict = {}
        self.clear_iterator_cache()

    def set_circuit(self, circuit):
        circuit = _as_circuit(circuit)
        previous = self._circuit
        assert not circuit.requires_polarization, "Circuit must not contain polarized components"
        self._input_state = None
        self._circuit = circuit
        self._umat = circuit.compute_unitary()
        self._results = {}
        if not (self._state_mapping and previous is not None and previous.m == circuit.m):
            self._reset()

    def set_input_state(self, input_state):
        super().set_input_state(input_state)
        self.preprocess([self._input_state])

    def clear_mask(self):
        super().clear_mask()
        self._reset()

    def _deploy(self, input_list):
        m = self._circuit.m
        for input_state in input_list:
            for k in range(len(self._fsms), input_state.n + 1):
                self._fsms.append(_Layer(m, k))

    def preprocess(self, input_list) -> bool:
        """Deploy paths for new input states (_slos.py:170-185).  Several new
        inputs share one tree plan (one _Path root), so their common
        photon-addition prefixes are computed once per circuit."""
        input_list = [FockState(s) for s in input_list]
        new = []
        for s in input_list:
            if s not in self._state_mapping and s not in new:
                new.append(s)
        if not new:
            return False
        self._deploy(new)
        if len(new) == 1:
            self._state_mapping[new[0]] = SlosPlan(list(new[0])) if self._masks_str is None else _LazyPlan(new[0])
        else:
            tree = SlosTree([list(s) for s in new])
            for i, s in enumerate(new):
                self._state_mapping[s] = (tree, i)
        return True

    # ------------------------------------------------------------ results
    def _compute(self, istate, mode):
        key = (istate, mode)
        if key not in self._results:
            plan = self._state_mapping[istate]
            if isinstance(plan, tuple):   # a tree: one execution gives every input of it
                tree, _ = plan
                outs = tree.run(self._umat, mode)
                for s, (t2, i) in [(s, v) for s, v in self._state_mapping.items() if isinstance(v, tuple)]:
                    if t2 is tree:
                        self._results[(s, mode)] = outs[i].cpu().numpy()
            else:
                self._results[key] = plan.run(self._umat, mode).cpu().numpy()
        return self._results[key]

    def _masked_indices_device(self, istate):
        """Ranks of the output states the mask keeps (CUDA int64), selected on
        the GPU and cached per photon number, or None without a mask."""
        if self._mask is None:
            return None
        key = ("mask_idx", istate.m, istate.n)
        if key not in self._results:
            self._results[key] = fock_mask_select(istate.m, istate.n, self._masks_str, self._no_limit_modes,
                                                  self._mask_n)
        return self._results[key]

    def _masked_indices(self, istate):
        idx = self._masked_indices_device(istate)
        return None if idx is None else idx.cpu().numpy()

    def prob_amplitude(self, output_state) -> complex:
        output_state = FockState(output_state)
        if self._input_state.n != output_state.n:
            return complex(0)
        idx = fock_rank_host(output_state)
        coefs = self._compute(self._input_state, COEFS)
        assert 0 <= idx < len(coefs)
        return complex(coefs[idx] * math.sqrt(output_state.prodnfact() / self._input_state.prodnfact()))

    def _mask_plan_args(self, istate):
        """(condition rows, at-least modes, mask photon count) of the mask for
        a masked plan, or None when the mask goes through selection from the
        full output layer (no mode constrained, fewer than two modes free in
        every row, more than 64 rows, a photon count outside 1..30, or a mask
        set for fewer photons than the input's)."""
        if self._mask is None or len(self._masks_str) > 64:
            return None
        cond = mask_conditions(self._masks_str, istate.m)
        free = (cond < 0).all(axis=0)
        if free.sum() < 2 or free.all() or istate.n < 1 or istate.n > 30:
            return None
        mask_n = self._mask_n or istate.n
        if mask_n < istate.n:
            return None
        return cond, tuple(int(q) for q in (self._no_limit_modes or ())), int(mask_n)

    def _masked_plan(self, istate, args):
        cond, at_least, mask_n = args
        key = ("masked_plan", istate, cond.tobytes(), at_least, mask_n)
        if key not in self._plans_masked:
            try:
                self._plans_masked[key] = SlosPlan(list(istate), mask_cond=cond, at_least=at_least, mask_n=mask_n)
            except NativeError as e:
                if "failed (-4)" not in str(e):   # out of range for a masked plan: select instead
                    raise
                self._plans_masked[key] = None   # too many constrained configurations: select instead
        return self._plans_masked[key]

    def all_prob(self, input_state=None):
        """SLOS specific signature, to enhance optimization in some computations"""
        if input_state is not None:
            self.set_input_state(input_state)
        if self._mask is None:
            return self._compute(self._input_state, PROBS).tolist()
        args = self._mask_plan_args(self._input_state)
        plan = None if args is None else self._masked_plan(self._input_state, args)
        if plan is not None:
            # every layer restricted to the states that can still match some
            # mask row, the matching outputs in FSArray order
            key = ("masked_probs", self._input_state)
            if key not in self._results:
                self._results[key] = plan.run(self._umat, PROBS).cpu().numpy()
            return self._results[key].tolist()
        # otherwise: select on the GPU and copy back only the kept probabilities
        idx = self._masked_indices_device(self._input_state)
        return gather_f64(self.all_prob_tensor(), idx).cpu().numpy().tolist()

    def all_prob_tensor(self, input_state=None):
        """Probabilities of every output state as a CUDA float64 tensor (no host copy)."""
        if input_state is not None:
            self.set_input_state(input_state)
        plan = self._state_mapping[self._input_state]
        if isinstance(plan, tuple):
            return plan[0].run(self._umat, PROBS)[plan[1]]
        return plan.run(self._umat, PROBS)

    def prob_distribution(self) -> BSDistribution:
        bsd = BSDistribution()
        for s, p in zip(self._get_iterator(self._input_state), self.all_prob()):
            bsd.add(s, p)
        return bsd

    def evolve(self) -> StateVector:
        istate = self._input_state
        coefs = self._compute(istate, COEFS)
        rows = fock_space(istate.m, istate.n)
        res = StateVector()
        pin = istate.prodnfact()
        for r, c in zip(rows, coefs):
            s = FockState(r)
            if self._mask is not None and not self._mask.keeps(s):
                continue
            res += s * complex(c * math.sqrt(s.prodnfact() / pin))
        return res
