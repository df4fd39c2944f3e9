"""The register solvers' schedule against the reference's Gauss-Jordan (no GPU).

Both register solvers of ps_engine.hip -- gauss_solve_wave (one wave, up to
63 unknowns: lane c holds column c) and gauss_solve_block (four waves, 64-127
unknowns: lane l holds columns l and l + 64, wave w the rows w + 4r) -- run
the same schedule, which this file restates.  They keep [A|b] in registers
and never move a row: the reference's row swaps (period_search/gauss_errc.c:47-53)
become a logical-row permutation, pivots are found while eliminating, and
ties are judged by the logical row-major index (the reference's `>=` scan
keeps the last).  This file restates that schedule in numpy -- physical rows
fixed, perm/inv bookkeeping, the candidate rule, the element update
a - pr * dum with one rounding per operation as in the kernel (FMA
contraction off) -- and requires the solution to be bitwise the oracle's
gauss_errc (itself bitwise the reference's, tests/test_oracle_ref.py) on
random SPD matrices like LM normal matrices, tie-heavy integer matrices and
singular ones, for the sizes the two solvers take (3-63 and 64-127
unknowns).  The GPU kernels are held to the oracle through whole fits
(test_gpu_parity.py::test_fits_match_oracle_across_solver_sizes)."""
import numpy as np
import pytest

import oracle


def block_schedule(A, b):
    """The register solvers' arithmetic and pivot order.  Returns (code, x)."""
    n = len(b)
    M = np.concatenate([A, b[:, None]], axis=1).astype(np.float64)
    perm = np.arange(n)  # logical row -> physical row
    inv = np.arange(n)   # physical row -> logical row
    done = np.zeros(n, bool)

    def candidate(M):
        j = inv  # logical row of each physical row
        mask = (~done[j])[:, None] & (~done)[None, :]
        vals = np.abs(M[:, :n])
        ok = mask & ~np.isnan(vals)
        if not ok.any():
            return -1
        idx = j[:, None] * n + np.arange(n)[None, :]
        best = vals[ok].max()
        return int(idx[ok & (vals == best)].max())

    be = candidate(M)
    prev = None
    for _ in range(n):
        if be < 0:
            if prev is None:
                return 1, None
            irow, icol = prev
        else:
            irow, icol = divmod(be, n)
        if done[icol]:
            return 1, None
        prev = (irow, icol)
        phr, phc = perm[irow], perm[icol]
        if irow != icol:
            perm[irow], perm[icol] = phc, phr
            inv[phr], inv[phc] = icol, irow
        pp = phr
        piv = M[pp, icol]
        if piv == 0.0:
            return 2, None
        pivinv = 1.0 / piv
        done[icol] = True
        pr = M[pp] * pivinv
        dum = M[:, icol].copy()
        prod = pr[None, :] * dum[:, None]
        M = M - prod
        M[pp] = pr
        be = candidate(M)
    x = np.empty(n)
    x[inv] = M[:, n]
    return 0, x


def reference(A, b):
    """The oracle's gauss_errc on 1-indexed [201][201] storage."""
    n = len(b)
    a = np.zeros((oracle.NP, oracle.NP))
    a[1:n + 1, 1:n + 1] = A
    x = np.zeros(oracle.NP)
    x[1:n + 1] = b
    code = oracle.olib.orc_gauss_errc(oracle._d(a), n, oracle._d(x))
    return code, x[1:n + 1].copy()


def _spd(rng, n):
    # like an LM normal matrix: J^T J of a Jacobian with uneven column scales, damped diagonal
    J = rng.normal(size=(3 * n, n)) * np.exp(rng.normal(0, 2, size=n))[None, :]
    A = J.T @ J
    A[np.diag_indices(n)] *= 1.1
    return A


@pytest.mark.parametrize("n", [3, 17, 40, 63, 64, 65, 87, 100, 127])
def test_block_schedule_is_bitwise_gauss_errc(n):
    rng = np.random.default_rng(n)
    cases = [(_spd(rng, n), rng.normal(size=n)) for _ in range(3)]
    # tie-heavy: small integers, many equal magnitudes
    cases.append((rng.integers(-3, 4, size=(n, n)).astype(float) + 7 * np.eye(n), rng.integers(-5, 6, size=n) * 1.0))
    for A, b in cases:
        code_r, x_r = reference(A, b)
        code_b, x_b = block_schedule(A, b)
        assert code_b == code_r == 0
        assert np.array_equal(x_b, x_r), np.abs(x_b - x_r).max()


@pytest.mark.parametrize("n", [5, 70])
def test_block_schedule_singular_codes(n):
    rng = np.random.default_rng(100 + n)
    A = _spd(rng, n)
    A[:, 2] = A[:, 1]  # rank deficient
    A[2, :] = A[1, :]
    b = rng.normal(size=n)
    code_r, _ = reference(A, b)
    code_b, _ = block_schedule(A, b)
    assert code_b == code_r and code_r != 0
    Z = np.zeros((n, n))
    assert block_schedule(Z, b)[0] == reference(Z, b)[0]


def diag_schedule(A, b):
    """gauss_solve_wave_diag and gauss_solve_block_diag (the mixed forms'
    solves, one wave up to 63 unknowns, four waves from 64): the element updates of
    gauss_errc with the pivot the not-yet-pivoted row of largest |a_jj|, ties
    to the larger j, and no row swaps.  Returns (code, x)."""
    n = len(b)
    M = np.concatenate([A, b[:, None]], axis=1).astype(np.float64)
    done = np.zeros(n, bool)
    for _ in range(n):
        d = np.abs(np.diagonal(M[:, :n]))
        ok = ~done & ~np.isnan(d)
        if not ok.any():
            return 1, None
        icol = int(np.flatnonzero(ok & (d == d[ok].max())).max())
        piv = M[icol, icol]
        if piv == 0.0:
            return 2, None
        pivinv = 1.0 / piv
        done[icol] = True
        pr = M[icol] * pivinv
        dum = M[:, icol].copy()
        M = M - pr[None, :] * dum[:, None]
        M[icol] = pr
    return 0, M[:, n].copy()


@pytest.mark.parametrize("n", [3, 17, 40, 55, 63, 64, 65, 87, 100, 127])
def test_diagonal_pivots_are_gauss_errc_on_normal_matrices(n):
    # damped normal matrices are symmetric positive definite: complete
    # pivoting finds its maximum on the diagonal, so the pivot order and
    # every rounding are gauss_errc's
    rng = np.random.default_rng(1000 + n)
    for _ in range(4):
        A, b = _spd(rng, n), rng.normal(size=n)
        code_r, x_r = reference(A, b)
        code_d, x_d = diag_schedule(A, b)
        assert code_d == code_r == 0
        assert np.array_equal(x_d, x_r), np.abs(x_d - x_r).max()


def test_diagonal_pivots_solve_general_systems():
    # not positive definite: other pivots than gauss_errc's, the same solution
    rng = np.random.default_rng(7)
    for n in (64, 90):
        A = rng.integers(-3, 4, size=(n, n)).astype(float) + 7 * np.eye(n)
        b = rng.integers(-5, 6, size=n) * 1.0
        code_r, x_r = reference(A, b)
        code_d, x_d = diag_schedule(A, b)
        assert code_d == code_r == 0
        assert np.allclose(x_d, x_r, rtol=1e-12, atol=1e-12 * np.abs(x_r).max())
        assert np.allclose(A @ x_d, b, rtol=0, atol=1e-10 * np.abs(b).max())
    assert diag_schedule(np.zeros((70, 70)), np.ones(70))[0] == 2
