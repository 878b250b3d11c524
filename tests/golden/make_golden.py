"""Writes the golden fixtures under tests/golden/:

* perm_mpmath.json: permanents of seeded complex matrices computed with mpmath
  at 50 significant digits by Ryser's formula (independent of the oracle's
  Glynn code and of its long-double arithmetic);
* u_random_8.npy: the 8 x 8 unitary of the reference's test data file
  tests/data/u_random_8 (one matrix row of Python complex literals per line),
  read from /root/reference when that tree is present.

Run from the repository root:  python tests/golden/make_golden.py
"""
import itertools
import json
import os

import mpmath
import numpy as np

mpmath.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))


def haar(n, seed):
    rng = np.random.RandomState(seed)
    z = rng.randn(n, n) + 1j * rng.randn(n, n)
    q, r = np.linalg.qr(z)
    return q @ np.diag(np.sign(np.diagonal(np.real(r))))


def ryser_mp(a):
    n = len(a)
    A = [[mpmath.mpc(complex(a[i][j])) for j in range(n)] for i in range(n)]
    total = mpmath.mpc(0)
    for size in range(1, n + 1):
        for cols in itertools.combinations(range(n), size):
            prod = mpmath.mpc(1)
            for i in range(n):
                prod *= mpmath.fsum(A[i][j] for j in cols)
            total += (-1) ** size * prod
    return (-1) ** n * total


def main():
    cases = []
    for n, seed in [(1, 1), (2, 2), (3, 3), (4, 4), (5, 5), (6, 6), (8, 8), (10, 10), (12, 12)]:
        # an n x n block of a 2n x 2n Haar unitary (Haar submatrix statistics)
        U = haar(2 * n, seed)
        M = U[:n, n:2 * n]
        p = ryser_mp(M.tolist())
        cases.append({"n": n, "seed": seed, "re": [[float(x.real) for x in row] for row in M],
                      "im": [[float(x.imag) for x in row] for row in M],
                      "perm": [mpmath.nstr(p.real, 30), mpmath.nstr(p.imag, 30)]})
        print(n, mpmath.nstr(p, 20), flush=True)
    with open(os.path.join(HERE, "perm_mpmath.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "digits": 50, "cases": cases}, f, indent=1)
    src = "/root/reference/tests/data/u_random_8"
    if os.path.exists(src):
        with open(src) as f:
            rows = [[complex(tok) for tok in line.split()] for line in f if line.strip()]
        np.save(os.path.join(HERE, "u_random_8.npy"), np.array(rows, dtype=np.complex128))


if __name__ == "__main__":
    main()
