"""Writes the golden fixtures under tests/golden/ (run from the repo root:
``python tests/golden/make_golden.py``).

The reference ships no stored vectors, only known-answer tests inside its
test code; those KATs are restated inline in the tests.  These fixtures add
independent expected values for small seeded inputs, none of them produced by
this repository's code:

* LAPACK getrf (scipy.linalg.lu_factor): packed LU and 0-based ipiv; LAPACK's
  idamax picks the first maximum, as rawPivotSearch does
  (src/Numeric/LinearAlgebra/Massiv/Internal/Kernel.hs:1541-1556), so the
  pivots are pinned exactly and the factors to rounding;
* LAPACK potrf (scipy.linalg.cholesky, lower);
* 50-digit mpmath solutions of A x = b (for n <= 12) -- the exact answer the
  FP64 solves approximate.

Inputs: A, B iid U[-1, 1] (numpy default_rng(seed)); SPD S = G G^T / n + I
with G iid N(0, 1), the SPD family of the n=8192 solve workload.
"""
import os

import mpmath
import numpy as np
import scipy.linalg as sl

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_small.npz")
SIZES = [1, 2, 3, 5, 10, 12, 33, 63, 64, 65, 100, 130]
NRHS = 3


def mp_solve(A, b):
    mpmath.mp.dps = 50
    M = mpmath.matrix(A.tolist())
    v = mpmath.matrix(b.tolist())
    x = mpmath.lu_solve(M, v)
    return np.array([float(x[i]) for i in range(A.shape[0])])


def main():
    data = {"sizes": np.array(SIZES)}
    for n in SIZES:
        rng = np.random.default_rng(1000 + n)
        A = rng.uniform(-1.0, 1.0, (n, n))
        B = rng.uniform(-1.0, 1.0, (n, NRHS))
        G = rng.standard_normal((n, n))
        S = G @ G.T / n + np.eye(n)
        lu, piv = sl.lu_factor(A)
        L = sl.cholesky(S, lower=True)
        data[f"A_{n}"] = A
        data[f"B_{n}"] = B
        data[f"S_{n}"] = S
        data[f"lapack_lu_{n}"] = lu
        data[f"lapack_ipiv_{n}"] = piv.astype(np.int64)
        data[f"lapack_chol_{n}"] = L
        if n <= 12:
            data[f"mp_x_{n}"] = mp_solve(A, B[:, 0])
            data[f"mp_xs_{n}"] = mp_solve(S, B[:, 0])
    np.savez_compressed(OUT, **data)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
