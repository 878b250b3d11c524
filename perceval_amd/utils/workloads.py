"""Seeded synthetic inputs with the shapes of the measured configurations.

Haar unitaries follow Matrix.random_unitary / _unitarize_matrix
(perceval/utils/matrix.py:141-149,166-173): Z = randn + 1j*randn drawn from a
numpy RandomState, Q, R = qr(Z), U = Q * sign(Re diag R).
"""
from __future__ import annotations

import numpy as np


def haar_unitary(n: int, seed=None, rng: np.random.RandomState | None = None) -> np.ndarray:
    rng = rng if rng is not None else np.random.RandomState(seed)
    z = rng.randn(n, n) + 1j * rng.randn(n, n)
    q, r = np.linalg.qr(z)
    d = np.sign(np.diagonal(np.real(r)))
    return q @ np.diag(d)


def submatrix_batch(count: int, nmin: int, nmax: int, m: int = 64, seed: int = 2024):
    """``count`` n x n submatrices (rows and columns drawn without replacement)
    of one seeded m x m Haar unitary, n uniform in [nmin, nmax]."""
    rng = np.random.RandomState(seed)
    U = haar_unitary(m, rng=rng)
    orders = rng.randint(nmin, nmax + 1, size=count).astype(np.int32)
    mats = []
    for n in orders:
        rows = rng.choice(m, size=n, replace=False)
        cols = rng.choice(m, size=n, replace=False)
        mats.append(np.ascontiguousarray(U[np.ix_(rows, cols)]))
    return U, orders, mats
