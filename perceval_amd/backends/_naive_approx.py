"""NaiveApprox backend: the Naive backend with Gurvits estimates of the
permanents (perceval/backends/_naive_approx.py:36-73).

_compute_permanent returns the estimate of xq.estimate_permanent_cx(M,
iterations) (:52-54); prob_amplitude_with_error returns (amplitude, 99% error
bound on it) with both divided by sqrt(prod s! prod t!) (:56-62), and a 1 x 1
submatrix (n <= 1, or a photon-number mismatch) returns (M[0, 0], 0);
probability_confidence_interval clips [(|a| - e)^2, (|a| + e)^2] to [0, 1]
(:64-69).  The estimates run on the GPU (pcvl_gurvits_batch); all_prob
batches every output state of the space into one launch.
"""
from __future__ import annotations

import math

import numpy as np

from ..utils.states import FockState
from ._naive import NaiveBackend, _photon_modes


class NaiveApproxBackend(NaiveBackend):
    """Naive algorithm with Gurvits computations of permanents

    :param gurvits_iterations: Number of iterations to use for Gurvits estimation algorithm (default 10 000)
    """

    def __init__(self, gurvits_iterations: int = 10000, seed: int = 0):
        self._gurvits_iterations = gurvits_iterations
        self._seed = int(seed)
        self._calls = 0
        NaiveBackend.__init__(self)

    @property
    def name(self) -> str:
        return "NaiveApprox"

    def _estimate(self, mats) -> tuple:
        from ..permanent import gurvits
        # each call takes fresh streams, so repeated estimates are independent
        est, err = gurvits(mats, self._gurvits_iterations, seed=self._seed, stream0=self._calls)
        self._calls += len(mats)
        return est, err

    def _compute_permanent(self, M) -> complex:
        est, _ = self._estimate([np.asarray(M, dtype=complex)])
        return complex(est[0])

    def prob_amplitude_with_error(self, output_state) -> tuple:
        """Estimate of the probability amplitude and a 99% error bound on it."""
        output_state = FockState(output_state)
        M = self._compute_submatrix(output_state)
        if M.size <= 1:
            return complex(M[0, 0]), 0.0
        est, err = self._estimate([M])
        norm = math.sqrt(output_state.prodnfact() * self._input_state.prodnfact())
        return complex(est[0]) / norm, float(err[0]) / norm

    def probability_confidence_interval(self, output_state) -> list:
        """99% confidence interval for the true value of the probability."""
        mean, err = self.prob_amplitude_with_error(output_state)
        return [max((abs(mean) - err) ** 2, 0), min((abs(mean) + err) ** 2, 1)]

    def _all_amplitudes(self) -> np.ndarray:
        states = self._get_iterator(self._input_state)
        n = self._input_state.n
        if n <= 1:
            return np.array([self.prob_amplitude(s) for s in states], dtype=complex)
        cols = _photon_modes(self._input_state)
        pin = self._input_state.prodnfact()
        mats = [np.ascontiguousarray(self._umat[np.ix_(_photon_modes(s), cols)]) for s in states]
        norm = np.array([math.sqrt(s.prodnfact() * pin) for s in states])
        est, _ = self._estimate(mats)
        return est / norm
