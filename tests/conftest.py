import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "linear-massiv_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden_small.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def lmx_gpu():
    """The product library on a GPU; fails (does not skip) when the HIP
    library or the device is missing, so a GPU run can never pass silently."""
    import lmx
    lmx.init(0)
    return lmx


def known_pivot_matrix(n, seed):
    """-> (A, L, U, sigma) with A[sigma[j]] = row j of L U, where L is unit
    lower with |l_ij| <= 1/2 and 1 <= |u_jj| <= 2.  At step j of partial
    pivoting the candidates are l_ij u_jj, so row j of L U wins by a factor
    of two in exact arithmetic; the multipliers and the off-diagonal part of
    U shrink like 1/sqrt(i), which keeps L and U well conditioned, so FP32
    rounding cannot close that margin."""
    rng = np.random.default_rng(seed)
    scale = 0.5 / np.sqrt(np.maximum(np.arange(n), 1))
    L = np.tril(rng.uniform(-1, 1, (n, n)) * scale[:, None], -1) + np.eye(n)
    U = np.triu(rng.uniform(-1, 1, (n, n)) * scale[None, :], 1)
    U += np.diag(rng.choice([-1.0, 1.0], n) * rng.uniform(1, 2, n))
    sigma = rng.permutation(n)
    A = np.empty((n, n))
    A[sigma] = L @ U
    return A, L, U, sigma


def swaps_to_perm(ipiv):
    """LAPACK swap list (0-based) -> permutation: row i of P A is row perm[i] of A."""
    perm = np.arange(len(ipiv))
    for k, p in enumerate(ipiv):
        perm[[k, p]] = perm[[p, k]]
    return perm


@pytest.fixture(scope="session")
def known_pivots():
    return known_pivot_matrix


def tie_block_matrix(n, seed, zero_cols=0):
    """-> A (n x n) whose nonzeros form 2 x 2 blocks [[a, b], [c, d]] with
    |a| = |c| and |b| = |d| (so each pivot search in a block column is a tie
    between the block's two rows) and 1 x 1 blocks, scattered over the matrix
    by independent row and column permutations, plus `zero_cols` empty
    row/column pairs (a zero column is an n-way tie at 0, and the matrix is
    singular).  The signs make every 2 x 2 block nonsingular.  Every
    multiplier of partial pivoting is 0 or +-1 and every product and partial
    sum is a small integer, so any order of the same operations gives the
    same bits: the pivot choices alone decide the factors."""
    rng = np.random.default_rng(seed)
    vals = np.array([1.0, 2.0, 3.0, 4.0])
    m = n - zero_cols
    nb2 = m // 3  # 2 x 2 blocks; the rest are 1 x 1
    B = np.zeros((n, n))
    i = 0
    for _ in range(nb2):
        s, t = rng.choice(vals, 2)
        g1, g2, g3 = rng.choice([-1.0, 1.0], 3)
        B[i:i + 2, i:i + 2] = [[g1 * s, g3 * t], [g2 * s, -g1 * g2 * g3 * t]]
        i += 2
    while i < m:
        B[i, i] = rng.choice([-1.0, 1.0]) * rng.choice(vals)
        i += 1
    rows, cols = rng.permutation(n), rng.permutation(n)
    return np.ascontiguousarray(B[rows][:, cols])


@pytest.fixture(scope="session")
def tie_blocks():
    return tie_block_matrix
