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

    def samples_for_inputs(self, counts: dict) -> dict:
        """Samples of several input states at once ({input: count} ->
        {input: BSSamples}), one batched GPU call: what
        SamplesProvider.prepare (noisy_sampling_simulator.py:55-74) obtains by
        one set_input_state + samples(count) per input."""
        from ..sampling import cc_samples_multi_device
        inputs = [FockState(s) for s in counts]
        for s in inputs:
            self._check_state(s)
        occ = cc_samples_multi_device(self._umat, [list(s) for s in inputs], [int(counts[s]) for s in counts],
                                      seed=next_sampler_seed()).cpu().numpy()
        res, row = {}, 0
        for s, c in zip(inputs, counts.values()):
            res[s] = BSSamples.from_array(occ[row:row + int(c)])
            row += int(c)
        return res

    @property
    def name(self) -> str:
        return "CliffordClifford2017"
