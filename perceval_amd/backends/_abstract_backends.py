"""Backend plugin API, mirroring perceval/backends/_abstract_backends.py:39-208.

Same class names, methods, argument meanings and error behaviour:
AssertionError for a missing circuit or a circuit/state size mismatch
(:63-65), RuntimeError for a polarized circuit (:50-51), amplitude 0 for a
photon-number mismatch.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from collections.abc import Iterable

import numpy as np

from ..components import ACircuit, Unitary
from ..utils.globals import global_params
from ..utils.mask import FSMask
from ..utils.states import BSDistribution, BSSamples, FockState, StateVector, allstate_iterator


def _as_circuit(circuit) -> ACircuit:
    if isinstance(circuit, ACircuit):
        return circuit
    return Unitary(np.asarray(circuit, dtype=complex))


class ABackend(ABC):
    def __init__(self):
        self._circuit = None
        self._umat = None
        self._input_state = None

    def set_circuit(self, circuit):
        """Sets the circuit to simulate (a circuit object or an m x m unitary)."""
        circuit = _as_circuit(circuit)
        if circuit.requires_polarization:
            raise RuntimeError("Circuit must not contain polarized components")
        self._input_state = None
        self._circuit = circuit
        self._umat = circuit.compute_unitary()

    def set_input_state(self, input_state):
        """Sets an input Fock state (no annotations)."""
        input_state = FockState(input_state)
        self._check_state(input_state)
        self._input_state = input_state

    def _check_state(self, state):
        assert self._circuit is not None, 'Circuit must be set before the input state'
        assert self._circuit.m == state.m, f'Circuit({self._circuit.m}) and state({state.m}) size mismatch'

    @property
    @abstractmethod
    def name(self) -> str:
        """Returns the back-end name as a string"""


class ASamplingBackend(ABackend):
    @abstractmethod
    def sample(self) -> FockState:
        """Request one sample from the circuit given an input state"""

    @abstractmethod
    def samples(self, count: int) -> BSSamples:
        """Request samples from the circuit given an input state"""


class _StateProbIterator(Iterable):
    def __init__(self, states, probs):
        self.states = states
        self.probs = probs

    def __iter__(self):
        return zip(self.states, self.probs)


class AStrongSimulationBackend(ABackend):
    def __init__(self):
        super().__init__()
        self._mask_n = None
        self._cache_iterator: dict = dict()
        self._masks_str = None
        self._mask = None
        self._no_limit_modes = None

    def set_mask(self, masks, n=None, at_least_modes=None):
        self.clear_mask()
        if isinstance(masks, str):
            masks = [masks]
        mask_length = len(masks[0])
        for m in masks:
            m = m.replace("*", " ")
            assert len(m) == mask_length, "Inconsistent mask lengths"
        self._masks_str = masks
        self._mask_n = n
        self._no_limit_modes = at_least_modes
        self._init_mask()

    def _init_mask(self):
        if self._masks_str is not None and self._input_state is not None:
            instate = self._input_state
            assert len(self._masks_str[0]) == instate.m, "Mask and input state lengths have to be the same"
            self._mask = FSMask(instate.m, self._mask_n or instate.n, self._masks_str, self._no_limit_modes)

    def clear_mask(self):
        self._masks_str = None
        self._mask = None
        self._mask_n = None
        self.clear_iterator_cache()

    def set_input_state(self, input_state):
        super().set_input_state(input_state)
        self._init_mask()

    def _get_iterator(self, input_state):
        n_photons = input_state.n
        if n_photons not in self._cache_iterator:
            self._cache_iterator[n_photons] = tuple(allstate_iterator(input_state, self._mask))
        return self._cache_iterator[n_photons]

    def clear_iterator_cache(self):
        self._cache_iterator = dict()

    def set_circuit(self, circuit):
        circuit = _as_circuit(circuit)
        if self._circuit and circuit.m != self._circuit.m:
            self.clear_iterator_cache()
        super().set_circuit(circuit)

    @abstractmethod
    def prob_amplitude(self, output_state) -> complex:
        """Probability amplitude of one output state (circuit and input set)."""

    def probability(self, output_state) -> float:
        return abs(self.prob_amplitude(output_state)) ** 2

    def all_prob(self, input_state=None) -> list:
        """Probabilities of all output states (respecting the mask), in
        allstate_iterator order."""
        if input_state is not None:
            self.set_input_state(input_state)
        return [self.probability(s) for s in self._get_iterator(self._input_state)]

    def prob_distribution(self) -> BSDistribution:
        bsd = BSDistribution()
        for output_state, p in zip(self._get_iterator(self._input_state), self.all_prob()):
            bsd.add(output_state, p)
        return bsd

    def prob_iterator(self, min_p: float = global_params["min_p"]):
        probs = self.all_prob(self._input_state)
        states = [state for i, state in enumerate(self._get_iterator(self._input_state)) if probs[i] > min_p]
        return _StateProbIterator(states, [p for p in probs if p > min_p])

    def evolve(self) -> StateVector:
        res = StateVector()
        for output_state in self._get_iterator(self._input_state):
            res += output_state * self.prob_amplitude(output_state)
        return res
