"""Haar-random unitaries with the reference's convention
(perceval/utils/matrix.py:141-149,166-173): Z = randn(n, n) + 1j randn(n, n)
from numpy's global RNG, Q, R = qr(Z), U = Q diag(sign(Re diag R))."""
from __future__ import annotations

import numpy as np


class Matrix:
    @staticmethod
    def random_unitary(n: int) -> np.ndarray:
        u = np.random.randn(n, n) + 1j * np.random.randn(n, n)
        return Matrix._unitarize_matrix(n, u)

    @staticmethod
    def parametrized_unitary(n: int, parameters) -> np.ndarray:
        assert len(parameters) == 2 * n ** 2, "parameters do not have the right size"
        a = np.reshape(parameters[:n ** 2], (n, n))
        b = np.reshape(parameters[n ** 2:], (n, n))
        return Matrix._unitarize_matrix(n, a + 1j * b)

    @staticmethod
    def _unitarize_matrix(n: int, u: np.ndarray) -> np.ndarray:
        q, r = np.linalg.qr(u)
        r_diag = np.sign(np.diagonal(np.real(r)))
        n_u = np.zeros((n, n))
        np.fill_diagonal(n_u, val=r_diag)
        return np.matmul(q, n_u)

    @staticmethod
    def eye(n: int) -> np.ndarray:
        return np.eye(n, dtype=complex)


def is_unitary(u: np.ndarray, tol: float = 1e-8) -> bool:
    u = np.asarray(u)
    return u.shape[0] == u.shape[1] and np.allclose(u @ u.conj().T, np.eye(u.shape[0]), atol=tol)
