"""The coarse correction's dense inverses (gj_invert_tiles: Gauss-Jordan with
partial pivoting, a 4x4 register tile per thread) on dense matrices through
their test hook, at the level-2 inverse's width (128) and the group
inverses' (64): A @ Ainv = I to round-off for sizes 1..128, on matrices
whose elimination needs row interchanges at most steps, the two widths equal
bit for bit where both apply (same pivots, same operations per entry), an
exact permutation matrix, and a zero pivot reported as TDYHIP_ERR_SINGULAR."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inverse(A, width=128):
    from tdycore_amd import _lib
    L = _lib.load()
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    out = np.zeros_like(A)
    rc = L.tdyhip_test_dense_inverse(A.ctypes.data_as(C.c_void_p), n, width, out.ctypes.data_as(C.c_void_p))
    return rc, out


def _shuffled(n):
    # a well-conditioned matrix whose rows are shuffled: the largest entry of
    # each column is away from the diagonal, so partial pivoting swaps rows
    rng = np.random.default_rng(100 + n)
    B = rng.standard_normal((n, n)) + n * np.eye(n)
    return B[rng.permutation(n)]


@pytest.mark.parametrize("n", [1, 2, 5, 37, 64, 100, 127, 128])
def test_inverse_with_pivoting(n):
    import torch
    assert torch.cuda.is_available()
    print(f"[case] dense inverse n={n} width=128", flush=True)
    A = _shuffled(n)
    rc, X = _inverse(A)
    assert rc == 0
    assert np.abs(A @ X - np.eye(n)).max() <= 1e-11
    assert np.allclose(X, np.linalg.inv(A), rtol=0, atol=1e-12 * np.abs(X).max())


@pytest.mark.parametrize("n", [1, 3, 5, 32, 37, 63, 64])
def test_group_width_equals_level2_width(n):
    import torch
    assert torch.cuda.is_available()
    print(f"[case] dense inverse n={n} width=64", flush=True)
    A = _shuffled(n)
    rc64, X64 = _inverse(A, 64)
    rc128, X128 = _inverse(A, 128)
    assert rc64 == 0 and rc128 == 0
    assert np.abs(A @ X64 - np.eye(n)).max() <= 1e-11
    assert np.array_equal(X64, X128)


def test_width_limits_are_checked():
    from tdycore_amd import _lib
    rc, _ = _inverse(np.eye(65), 64)
    assert rc == _lib.ERR_ARG
    rc, _ = _inverse(np.eye(4), 96)
    assert rc == _lib.ERR_ARG


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
