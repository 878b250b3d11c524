"""Fock-space masks (exqalibur FSMask semantics as exercised by
tests/utils/test_mask.py:33-45 and AStrongSimulationBackend.set_mask,
perceval/backends/_abstract_backends.py:103-146).

A mask is a string with one character per mode: chr(0x30 + c) (a digit for
c <= 9, up to 32 photons: perceval/simulators/simulator.py:650-662) fixes the
photon count of that mode, a space or '*' accepts any count.  Several masks are OR-ed.  On
`at_least_modes` a digit is a lower bound instead of an exact count.  With
allow_partial=True a state may hold fewer photons than a digit requires (it
can still be completed by photons added later).
"""
from __future__ import annotations


def _count(ch: str) -> int:
    c = ord(ch) - 0x30
    if not 0 <= c <= 32:
        raise ValueError(f"mask character {ch!r} is neither a count nor a wildcard")
    return c


class FSMask:
    def __init__(self, m: int, n: int, masks, at_least_modes=None):
        if isinstance(masks, str):
            masks = [masks]
        self.m = m
        self.n = n
        self.at_least = set(at_least_modes or [])
        self.conditions = []
        for mk in masks:
            if len(mk) != m:
                raise ValueError("mask length must equal the number of modes")
            self.conditions.append([None if ch in " *" else _count(ch) for ch in mk])

    def _match_one(self, cond, state, allow_partial: bool) -> bool:
        for i, c in enumerate(cond):
            if c is None:
                continue
            v = state[i]
            if i in self.at_least:
                if v < c and not allow_partial:
                    return False
            elif allow_partial:
                if v > c:
                    return False
            elif v != c:
                return False
        return True

    def match(self, state, allow_partial: bool = False) -> bool:
        return any(self._match_one(c, state, allow_partial) for c in self.conditions)

    def _missing(self, cond, state):
        """Photons that state lacks for cond (None: an exact count exceeded)."""
        d = 0
        for i, c in enumerate(cond):
            if c is None:
                continue
            v = state[i]
            if i in self.at_least:
                d += max(0, c - v)
            elif v > c:
                return None
            else:
                d += c - v
        return d

    def keeps(self, state) -> bool:
        """Membership of the masked Fock space of state.n photons (the masked
        FSArray a SLOS layer holds, perceval/backends/_slos.py:161-164): the
        self.n - state.n photons still to come can complete some row (they
        cover its missing photons, and a free or at-least mode takes the
        rest).  At state.n >= self.n this is match(state)."""
        spare = self.n - sum(state)
        if spare <= 0:
            return self.match(state)
        for cond in self.conditions:
            d = self._missing(cond, state)
            if d is None or d > spare:
                continue
            if d == spare or any(c is None or i in self.at_least for i, c in enumerate(cond)):
                return True
        return False
