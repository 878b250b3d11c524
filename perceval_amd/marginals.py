"""Exact 1- and 2-mode output marginals of single-photon boson sampling.

A diagnostic, not a simulation path: it judges samplers (plain
Clifford-Clifford and the learned independence sampler) against the exact
law of their marginals, so that a marginal TVD carries the sampler's own
error only, not that of a second finite sample set.

For single photons in input modes T (A = U[:, T], m x n, orthonormal
columns), the probability generating function of the output counts is a
permanent:

    E[prod_l x_l^(n_l)] = perm(A^H diag(x) A),

expanding the output state prod_t (sum_l A_lt a_l^+)|0> and pairing the
creation and annihilation operators (Wick).  For the modes j and k only
(x_l = 1 elsewhere) this is

    G(x, y) = perm(I + (x-1) conj(u) u^T + (y-1) conj(v) v^T),   u = A[j], v = A[k],

the permanent of the identity plus a rank-2 matrix.  perm(I + X) is the sum
of perm(X[S, S]) over row sets S; for a rank-2 X each of those splits into
the rows R and columns C taking the first term (|R| = |C| = r):
r! (|S|-r)! (x-1)^r (y-1)^(|S|-r) times products of u, v entries.  One pass
over the n photons accumulates these products by (|S|, |R|, |C|), O(n^4)
work in all, giving the factorial moments F[r, l] = E[C(n_j, r) C(n_k, l)];
the marginal law follows by inclusion-exclusion,

    P(n_j = a, n_k = b) = sum_{r>=a, l>=b} (-1)^(r-a+l-b) C(r, a) C(l, b) F[r, l].

F[1, 0] and F[1, 1] are the closed-form moments sum_t |u_t|^2 and
sum_{t<t'} |u_t v_t' + u_t' v_t|^2.
"""
from __future__ import annotations

import math

import numpy as np


def _input_columns(U, in_occ) -> np.ndarray:
    in_occ = [int(v) for v in in_occ]
    if any(v > 1 for v in in_occ):
        raise ValueError("exact marginals are implemented for single-photon inputs (occupations 0 or 1)")
    U = np.asarray(U, dtype=np.complex128)
    return U[:, [i for i, v in enumerate(in_occ) if v]]


def _factorial_moments(u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """F[r, l] = E[C(n_j, r) C(n_k, l)] for the rows u = A[j], v = A[k]."""
    n = len(u)
    # T[s, a, c]: sum over photon sets S (|S| = s) with R, C subsets of S
    # (|R| = a, |C| = c) of prod_{R} conj(u) prod_{S\R} conj(v) prod_{C} u prod_{S\C} v
    T = np.zeros((n + 1, n + 1, n + 1), dtype=np.complex128)
    T[0, 0, 0] = 1.0
    for t in range(n):
        uu = abs(u[t]) ** 2
        vv = abs(v[t]) ** 2
        uv = np.conj(u[t]) * v[t]      # t in R (conj u) and not in C (v)
        vu = np.conj(v[t]) * u[t]      # t in C (u) and not in R (conj v)
        new = T.copy()
        new[1:, 1:, 1:] += uu * T[:-1, :-1, :-1]
        new[1:, 1:, :] += uv * T[:-1, :-1, :]
        new[1:, :, 1:] += vu * T[:-1, :, :-1]
        new[1:, :, :] += vv * T[:-1, :, :]
        T = new
    F = np.zeros((n + 1, n + 1))
    for r in range(n + 1):
        for l in range(n + 1 - r):
            F[r, l] = (math.factorial(r) * math.factorial(l) * T[r + l, r, r]).real
    return F


def _inclusion_exclusion(F: np.ndarray) -> np.ndarray:
    n = F.shape[0] - 1
    W = np.array([[(-1) ** (r - a) * math.comb(r, a) for r in range(n + 1)] for a in range(n + 1)], dtype=float)
    return W @ F @ W.T


def pair_marginal(U, in_occ, j: int, k: int) -> np.ndarray:
    """(n+1) x (n+1) array P[a, b] = P(n_j = a, n_k = b), j != k."""
    if j == k:
        raise ValueError("pair_marginal needs two distinct modes")
    A = _input_columns(U, in_occ)
    return _inclusion_exclusion(_factorial_moments(A[j], A[k]))


def mode_marginal(U, in_occ, j: int) -> np.ndarray:
    """(n+1,) array P[a] = P(n_j = a)."""
    A = _input_columns(U, in_occ)
    other = (j + 1) % A.shape[0]
    return pair_marginal(U, in_occ, j, other).sum(axis=1) if A.shape[0] > 1 else np.array([0.0] * A.shape[1] + [1.0])


def sample_pairs(m: int, pairs: int, seed: int = 0) -> list:
    """The seeded mode pairs the marginal TVDs are taken over."""
    rng = np.random.RandomState(seed)
    return [tuple(int(x) for x in rng.choice(m, 2, replace=False)) for _ in range(pairs)]


def exact_marginals(U, in_occ, pairs: int = 200, seed: int = 0) -> dict:
    """Exact 1-mode laws of every mode and 2-mode laws of the seeded pairs."""
    m = np.asarray(U).shape[0]
    pl = sample_pairs(m, pairs, seed)
    two = {p: pair_marginal(U, in_occ, *p) for p in pl}
    one = {}
    for j in range(m):
        src = next(((p, 0) for p in pl if p[0] == j), None) or next(((p, 1) for p in pl if p[1] == j), None)
        if src is not None:
            one[j] = two[src[0]].sum(axis=1 - src[1])
        else:
            one[j] = mode_marginal(U, in_occ, j)
    return {"one": one, "two": two, "pairs": pl}


def marginal_tvds_exact(occ: np.ndarray, exact: dict) -> dict:
    """TVD of a sample set's 1-mode marginals (every mode) and 2-mode
    marginals (the pairs of `exact`) against the exact laws."""
    occ = np.asarray(occ)
    N = len(occ)
    n = len(next(iter(exact["one"].values()))) - 1
    one = []
    for j, law in exact["one"].items():
        emp = np.bincount(occ[:, j], minlength=n + 1)[:n + 1] / N
        one.append(0.5 * np.abs(emp - law).sum())
    two = []
    for (j, k), law in exact["two"].items():
        idx = occ[:, j].astype(np.int64) * (n + 1) + occ[:, k]
        emp = np.bincount(idx, minlength=(n + 1) ** 2)[:(n + 1) ** 2].reshape(n + 1, n + 1) / N
        two.append(0.5 * np.abs(emp - law).sum())
    return {"tvd1_max": float(max(one)), "tvd1_mean": float(np.mean(one)),
            "tvd2_max": float(max(two)), "tvd2_mean": float(np.mean(two))}
