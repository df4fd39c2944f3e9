"""The GPU shape fit's LU schedule against the reference's Crout LU (no GPU).

shape_lu (csrc/ps_shape.hip) factors each shape's normal matrix with one
wave: lanes hold rows, and column j's entries
a[i][j] - sum_{k < min(i, j)} a[i][k] a[k][j] are accumulated right-looking
(at step k the final a[k][j] is broadcast and every row below k subtracts
its term), the pivot is the last row of largest scaled magnitude (NaN never
wins), rows are swapped whole, the zero pivot becomes 1e-20 and the column
below the pivot is scaled by its reciprocal.  This file restates that
schedule in numpy -- one rounding per operation, as in the kernel -- and
requires the factors and the pivot rows to be bitwise the oracle's ludcmp
(period_search/ludcmp.c:9-69; the oracle's restatement is bitwise the
reference's own compiled source, tests/test_oracle_ref.py) for the sizes the
shape fit takes (ncoef up to 121) on normal matrices like the shape fit's,
on tie-heavy integer matrices and on matrices with a zero pivot.  The GPU
kernel is held to the host's LU through whole shape tables
(test_gpu_shape.py)."""
import numpy as np
import pytest

import oracle


def lu_schedule(A):
    """shape_lu's arithmetic and pivot order: (factors, pivot rows)."""
    a = np.array(A, dtype=np.float64)
    n = a.shape[0]
    vv = 1.0 / np.abs(a).max(axis=1)
    indx = np.zeros(n, int)
    for j in range(n):
        acc = a[:, j].copy()
        for k in range(j):
            u = acc[k]
            rows = np.arange(n) > k
            acc[rows] = acc[rows] - a[rows, k] * u
        a[:, j] = acc
        best, imax = -1.0, -1
        for r in range(j, n):
            dum = vv[r] * abs(acc[r])
            if dum >= best:
                best, imax = dum, r
        if imax < 0:
            imax = j
        if imax != j:
            a[[imax, j], :] = a[[j, imax], :]
            vv[imax] = vv[j]
        indx[j] = imax
        if a[j, j] == 0.0:
            a[j, j] = 1.0e-20
        if j != n - 1:
            dum = 1.0 / a[j, j]
            a[j + 1:, j] = a[j + 1:, j] * dum
    return a, indx


def reference(A):
    """The oracle's ludcmp on 1-indexed storage of leading dimension n + 1."""
    n = A.shape[0]
    s = n + 1
    a = np.zeros((s, s))
    a[1:, 1:] = A
    a = np.ascontiguousarray(a).ravel()
    indx = np.zeros(s, dtype=np.int32)
    d = np.zeros(1)
    oracle.olib.orc_ludcmp(oracle._d(a), n, indx.ctypes.data_as(oracle._ip), oracle._d(d))
    return a.reshape(s, s)[1:, 1:].copy(), indx[1:].astype(int) - 1


def _normal(rng, n):
    # like the shape fit's: a basis sampled at many directions, squared
    X = rng.normal(size=(4 * n + 8, n)) * np.exp(rng.normal(0, 1.5, size=n))[None, :]
    return X.T @ X


@pytest.mark.parametrize("n", [1, 2, 5, 16, 28, 49, 64, 65, 81, 100, 121])
def test_lu_schedule_is_bitwise_ludcmp(n):
    rng = np.random.default_rng(1000 + n)
    cases = [_normal(rng, n), rng.normal(size=(n, n))]
    ints = rng.integers(-3, 4, size=(n, n)).astype(float)  # many equal scaled magnitudes
    ints[np.diag_indices(n)] += 1.0
    cases.append(ints)
    for A in cases:
        fa, ia = lu_schedule(A)
        fr, ir = reference(A)
        assert np.array_equal(ia, ir)
        assert np.array_equal(fa, fr), np.abs(fa - fr).max()


def test_lu_schedule_zero_pivot():
    # a zero column below the diagonal after elimination: the pivot becomes 1e-20
    A = np.array([[2.0, 4.0, 1.0], [1.0, 2.0, 3.0], [4.0, 8.0, 5.0]])
    fa, ia = lu_schedule(A)
    fr, ir = reference(A)
    assert np.array_equal(ia, ir) and np.array_equal(fa, fr)
    assert 1e-20 in np.diag(fr)
