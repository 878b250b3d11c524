"""CliffordClifford2017 sampling backend (perceval/backends/_clifford2017.py:36-61)
on the GPU: exact boson sampling by algorithm A of Clifford & Clifford 2017,
with the per-photon marginals of many samples computed in one launch."""
from __future__ import annotations

from ..utils._random import next_sampler_seed
from ..utils.states import BSSamples, FockState
from ._abstract_backends import ASamplingBackend


class Clifford2017Backend(ASamplingBackend):
    def __init__(self):
        super().__init__()

    def set_circuit(self, circuit):
        super().set_circuit(circuit)  # computes the circuit unitary as _umat

    def set_input_state(self, input_state):
        super().set_input_state(input_state)

    def sample(self) -> FockState:
        return self.samples(1)[0]

    def samples(self, count: int) -> BSSamples:
        """count samples of the output distribution, drawn in parallel on the GPU."""
        from ..sampling import cc_samples
        occ = cc_samples(self._umat, list(self._input_state), int(count), seed=next_sampler_seed())
        return BSSamples.from_array(occ)

    @property
    def name(self) -> str:
        return "CliffordClifford2017"
