"""Seed plumbing (perceval/utils/_random.py:36-58): seeds Python's random,
numpy, and the counter-based generator of the GPU samplers."""
from __future__ import annotations

import random

import numpy as np

_state = {"seed": None}


def random_seed(seed: int | None = None):
    random.seed(seed)
    np.random.seed(seed)
    _state["seed"] = random.randint(0, 2 ** 32 - 1) if seed is None else int(seed) & (2 ** 64 - 1)


def next_sampler_seed() -> int:
    """Seed for the next GPU sampling call: derived from the last random_seed(),
    or drawn from Python's random module if none was set."""
    if _state["seed"] is None:
        _state["seed"] = random.randint(0, 2 ** 32 - 1)
    s = _state["seed"]
    _state["seed"] = (s * 6364136223846793005 + 1442695040888963407) & (2 ** 64 - 1)
    return s
