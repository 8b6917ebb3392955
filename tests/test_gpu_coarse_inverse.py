"""The coarse correction's level-2 inverse (k_coarse_invert: Gauss-Jordan
with partial pivoting, a 4x4 register tile per thread) on dense matrices
through its test hook: A @ Ainv = I to round-off for sizes 1..128, on
matrices whose elimination needs row interchanges at most steps, on an exact
permutation matrix, and a zero pivot reported as TDYHIP_ERR_SINGULAR."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inverse(A):
    from tdycore_amd import _lib
    L = _lib.load()
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    out = np.zeros_like(A)
    rc = L.tdyhip_test_dense_inverse(A.ctypes.data_as(C.c_void_p), n, out.ctypes.data_as(C.c_void_p))
    return rc, out


@pytest.mark.parametrize("n", [1, 2, 5, 37, 64, 100, 127, 128])
def test_inverse_with_pivoting(n):
    import torch
    assert torch.cuda.is_available()
    print(f"[case] dense inverse n={n}", flush=True)
    rng = np.random.default_rng(100 + n)
    # a well-conditioned matrix whose rows are shuffled: the largest entry of
    # each column is away from the diagonal, so partial pivoting swaps rows
    B = rng.standard_normal((n, n)) + n * np.eye(n)
    A = B[rng.permutation(n)]
    rc, X = _inverse(A)
    assert rc == 0
    assert np.abs(A @ X - np.eye(n)).max() <= 1e-11
    assert np.allclose(X, np.linalg.inv(A), rtol=0, atol=1e-12 * np.abs(X).max())


def test_permutation_inverse_is_exact():
    rng = np.random.default_rng(7)
    n = 96
    P = np.eye(n)[rng.permutation(n)]
    rc, X = _inverse(P)
    assert rc == 0
    assert np.array_equal(X, P.T)


def test_zero_pivot_is_reported():
    from tdycore_amd import _lib
    A = np.eye(40)
    A[:, 17] = 0.0  # column 17 has no nonzero at or below the diagonal once it is reached
    rc, _ = _inverse(A)
    assert rc == _lib.ERR_SINGULAR
