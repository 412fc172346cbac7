"""The known-pivot matrices of the FP32 LU tests, pinned on the CPU: LAPACK's
sgetrf and dgetrf (scipy) choose the planted pivot order and recover L and U
at working precision."""
import numpy as np
import pytest
import scipy.linalg as sl
from conftest import swaps_to_perm


@pytest.mark.parametrize("n", [5, 33, 200, 1100])
def test_known_pivot_matrix_vs_lapack(known_pivots, n):
    A, L, U, sigma = known_pivots(n, n)
    for getrf, dtype, tol in ((sl.lapack.sgetrf, np.float32, 1e-5), (sl.lapack.dgetrf, np.float64, 1e-12)):
        lu, piv, info = getrf(A.astype(dtype))
        assert info == 0
        np.testing.assert_array_equal(swaps_to_perm(piv), sigma)
        assert np.abs(np.tril(lu.astype(float), -1) + np.eye(n) - L).max() <= tol
        assert np.abs(np.triu(lu.astype(float)) - U).max() <= 2 * tol
