"""Fock states and result containers.

Light, integer-exact stand-ins for the exqalibur types the reference aliases
(perceval/utils/states.py:42-49): FockState/BasicState, StateVector,
BSDistribution, BSCount, BSSamples, plus allstate_iterator
(perceval/utils/states.py:268-285) in FSArray order (descending lexicographic,
tests/utils/test_statevector.py:430-438).
"""
from __future__ import annotations

import math
import re
from collections.abc import Iterable

import numpy as np


class FockState:
    """Occupation vector of an m-mode Fock state (no annotations)."""

    __slots__ = ("_occ", "_hash")

    def __init__(self, state=None):
        if state is None:
            occ = ()
        elif isinstance(state, FockState):
            occ = state._occ
        elif isinstance(state, str):
            occ = _parse_ket(state)
        elif isinstance(state, int):
            occ = (0,) * state
        else:
            occ = tuple(int(v) for v in state)
        if any(v < 0 for v in occ):
            raise ValueError("photon counts must be non-negative")
        self._occ = occ
        self._hash = hash(occ)

    @property
    def m(self) -> int:
        return len(self._occ)

    @property
    def n(self) -> int:
        return sum(self._occ)

    def prodnfact(self) -> int:
        """prod_i s_i! (exqalibur FockState.prodnfact, used at _naive.py:47, _slos.py:192)."""
        p = 1
        for v in self._occ:
            p *= math.factorial(v)
        return p

    def __getitem__(self, i):
        if isinstance(i, slice):
            return FockState(self._occ[i])
        return self._occ[i]

    def __iter__(self):
        return iter(self._occ)

    def __len__(self):
        return len(self._occ)

    def __eq__(self, other):
        if isinstance(other, FockState):
            return self._occ == other._occ
        if isinstance(other, (list, tuple)):
            return self._occ == tuple(other)
        return NotImplemented

    def __hash__(self):
        return self._hash

    def __lt__(self, other):
        return self._occ < FockState(other)._occ

    def __str__(self):
        return "|" + ",".join(str(v) for v in self._occ) + ">"

    __repr__ = __str__

    def __mul__(self, other):
        if isinstance(other, FockState):   # tensor product = concatenation of modes
            return FockState(self._occ + other._occ)
        if isinstance(other, (int, float, complex, np.number)):
            return StateVector({self: complex(other)})
        return NotImplemented

    def __rmul__(self, other):
        if isinstance(other, (int, float, complex, np.number)):
            return StateVector({self: complex(other)})
        return NotImplemented

    def __add__(self, other):
        return StateVector(self) + other

    def tolist(self) -> list:
        return list(self._occ)


BasicState = FockState


def _parse_ket(s: str) -> tuple:
    s = s.strip()
    m = re.fullmatch(r"\|([0-9,\s]*)>", s)
    if not m:
        raise ValueError(f"cannot parse Fock state {s!r}")
    body = m.group(1).strip()
    if not body:
        return ()
    return tuple(int(v) for v in body.split(","))


class StateVector:
    """Superposition of Fock states: FockState -> complex amplitude."""

    def __init__(self, init=None):
        self._amp: dict = {}
        if init is None:
            return
        if isinstance(init, StateVector):
            self._amp = dict(init._amp)
        elif isinstance(init, dict):
            for k, v in init.items():
                self._amp[FockState(k)] = complex(v)
        else:
            self._amp[FockState(init)] = 1 + 0j

    def __iadd__(self, other):
        other = _as_sv(other)
        for k, v in other._amp.items():
            self._amp[k] = self._amp.get(k, 0j) + v
        return self

    def __add__(self, other):
        r = StateVector(self)
        r += other
        return r

    __radd__ = __add__

    def __sub__(self, other):
        return self + (-1) * _as_sv(other)

    def __mul__(self, a):
        if isinstance(a, (int, float, complex, np.number)):
            return StateVector({k: v * complex(a) for k, v in self._amp.items()})
        return NotImplemented

    __rmul__ = __mul__

    def __getitem__(self, state):
        return self._amp.get(FockState(state), 0j)

    def __len__(self):
        return len(self._amp)

    def __iter__(self):
        return iter(self._amp)

    def keys(self):
        return self._amp.keys()

    def items(self):
        return self._amp.items()

    def norm(self) -> float:
        return math.sqrt(sum(abs(v) ** 2 for v in self._amp.values()))

    def normalize(self):
        nrm = self.norm()
        if nrm > 0:
            for k in self._amp:
                self._amp[k] /= nrm
        return self

    @property
    def m(self) -> int:
        return next(iter(self._amp)).m if self._amp else 0

    @property
    def n(self):
        return sorted({k.n for k in self._amp})

    def __str__(self):
        return "+".join(f"({v:.6g}){k}" for k, v in self._amp.items())


def _as_sv(x) -> StateVector:
    return x if isinstance(x, StateVector) else StateVector(x)


class BSDistribution(dict):
    """FockState -> probability (exqalibur BSDistribution).

    add() ignores probabilities below 1e-16 (global_params["min_p"]): the
    reference's masked CNOT test (tests/backends/test_backends.py:205-218)
    expects a 2-entry distribution where the other 8 masked states have
    probability 0 up to rounding."""

    MIN_P = 1e-16

    def __init__(self, d=None):
        super().__init__()
        if d is not None:
            if isinstance(d, dict):
                for k, v in d.items():
                    self.add(k, v)
            else:
                self.add(d, 1.0)

    def add(self, state, p: float):
        p = float(p)
        if p < self.MIN_P:
            return
        k = FockState(state)
        super().__setitem__(k, super().get(k, 0.0) + p)

    def __getitem__(self, state):
        return super().get(FockState(state), 0.0)

    def __setitem__(self, state, p):
        super().__setitem__(FockState(state), float(p))

    def normalize(self):
        tot = sum(self.values())
        if tot > 0:
            for k in list(self.keys()):
                super().__setitem__(k, super().get(k) / tot)
        return self


class BSCount(dict):
    """FockState -> integer count (exqalibur BSCount)."""

    def __getitem__(self, state):
        return super().get(FockState(state), 0)

    def __setitem__(self, state, v):
        super().__setitem__(FockState(state), int(v))

    def total(self) -> int:
        return sum(self.values())


class BSSamples(list):
    """List of sampled FockStates (exqalibur BSSamples)."""

    @classmethod
    def from_array(cls, occ: np.ndarray) -> "BSSamples":
        occ = np.asarray(occ)
        return cls(FockState(row) for row in occ.tolist())

    def to_counts(self) -> BSCount:
        c = BSCount()
        for s in self:
            c[s] = c[s] + 1
        return c


def fock_space(m: int, n: int) -> np.ndarray:
    """All n-photon states over m modes, (C(m+n-1, n), m) int32, FSArray order."""
    count = math.comb(m + n - 1, n) if m > 0 else (1 if n == 0 else 0)
    out = np.zeros((count, m), dtype=np.int32)
    if count == 0:
        return out
    s = [0] * m
    s[0] = n
    for k in range(count):
        out[k] = s
        i = m - 2
        while i >= 0 and s[i] == 0:
            i -= 1
        if i < 0:
            break
        tail = sum(s[i + 1:])
        for j in range(i + 1, m):
            s[j] = 0
        s[i] -= 1
        s[i + 1] = tail + 1
    return out


def allstate_iterator(input_state, mask=None) -> Iterable:
    """All output states with the input's photon number(s), FSArray order
    (perceval/utils/states.py:268-285); with a mask, the states it keeps
    (FSMask.keeps)."""
    m = input_state.m
    ns = input_state.n
    ns = [ns] if isinstance(ns, int) else list(ns)
    for n in ns:
        for row in fock_space(m, n):
            st = FockState(row)
            if mask is None or mask.keeps(st):
                yield st


def max_photon_state_iterator(m: int, n_max: int):
    for n in range(n_max + 1):
        for row in fock_space(m, n):
            yield FockState(row)
