"""Naive backend: one permanent per output amplitude
(perceval/backends/_naive.py:38-71).

amplitude(s) = perm(U[out photons, in photons]) / sqrt(prod s! prod t!)
(_naive.py:46-49), with the submatrix built as in _compute_submatrix
(:51-68): row r = output mode of the r-th output photon, column c = input mode
of the c-th input photon, both ascending with repetition.  The permanents run
on the GPU; all_prob / prob_distribution / evolve batch every output state of
the Fock space into one launch instead of looping.
"""
from __future__ import annotations

import math

import numpy as np

from ..permanent import BatchPermanents, output_submatrices
from ..utils.states import BSDistribution, FockState, StateVector
from ._abstract_backends import AStrongSimulationBackend


def _photon_modes(occ) -> list:
    return [i for i, v in enumerate(occ) for _ in range(int(v))]


class NaiveBackend(AStrongSimulationBackend):
    """Naive algorithm, no clever calculation path, does not cache anything,
    recompute all states on the fly"""

    @property
    def name(self) -> str:
        return "Naive"

    def _compute_submatrix(self, output_state) -> np.ndarray:
        n = self._input_state.n
        if n != FockState(output_state).n:
            return np.zeros((1, 1), dtype=complex)
        if n == 0:
            return np.ones((1, 1), dtype=complex)
        rows = _photon_modes(output_state)
        cols = _photon_modes(self._input_state)
        return np.ascontiguousarray(self._umat[np.ix_(rows, cols)])

    def _compute_permanent(self, M) -> complex:
        # a plan and device buffers kept per order: no allocation per call
        if getattr(self, "_single", None) is None:
            from ..permanent import SinglePermanents
            self._single = SinglePermanents()
        return self._single(M)

    def prob_amplitude(self, output_state) -> complex:
        output_state = FockState(output_state)
        p = output_state.prodnfact() * self._input_state.prodnfact()
        M = self._compute_submatrix(output_state)
        return self._compute_permanent(M) / math.sqrt(p) if M.size > 1 else complex(M[0, 0])

    def _all_amplitudes(self) -> np.ndarray:
        """Amplitudes of every state of the iterator, one batched GPU launch
        (the batch plan kept per size)."""
        states = self._get_iterator(self._input_state)
        if self._input_state.n <= 1 or not states:
            return np.array([self.prob_amplitude(s) for s in states], dtype=complex)
        mats, norms = output_submatrices(self._umat, [tuple(s) for s in states], tuple(self._input_state))
        if getattr(self, "_batch", None) is None:
            self._batch = BatchPermanents()
        return self._batch(mats) / norms

    def all_prob(self, input_state=None) -> list:
        if input_state is not None:
            self.set_input_state(input_state)
        return (np.abs(self._all_amplitudes()) ** 2).tolist()

    def prob_distribution(self) -> BSDistribution:
        bsd = BSDistribution()
        for s, p in zip(self._get_iterator(self._input_state), self.all_prob()):
            bsd.add(s, p)
        return bsd

    def evolve(self) -> StateVector:
        res = StateVector()
        for s, a in zip(self._get_iterator(self._input_state), self._all_amplitudes()):
            res += s * complex(a)
        return res
