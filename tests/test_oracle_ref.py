"""The oracle against the reference's own numeric sources (oracle/_ref).

_ref/libpsref.so is the reference's period_search/*.c compiled unchanged
(g++ -O3, as period_search/Makefile) behind a harness that defines the
globals.  The oracle must reproduce it bit for bit: shape basis, one model
evaluation, and whole LM fits.
"""
import numpy as np
import pytest

import oracle

ref = oracle.load_ref()
needs_ref = pytest.mark.skipif(ref is None, reason="oracle/_ref not built (reference sources absent)")


def _refarr(name, shape):
    return np.ctypeslib.as_array(ref.ref_dptr(name.encode()), shape=(int(np.prod(shape)),)).reshape(shape)


@pytest.fixture(scope="module")
def loaded(small_input):
    from periodsearch_amd import Input

    inp = Input(small_input)
    o = oracle.Oracle(small_input)
    o.prepare(inp.conw_r0)
    oracle.ref_load_problem(ref, o)
    return inp, o


@needs_ref
def test_shape_basis_bitwise(loaded):
    inp, o = loaded
    F, nc = o.iget("Numfac"), o.iget("Ncoef")
    assert np.array_equal(_refarr("Nor", (1001, 4))[1:F + 1, 1:4], o.darr("Nor", (1001, 4))[1:F + 1, 1:4])
    assert np.array_equal(_refarr("Darea", (1001,))[1:F + 1], o.darr("Darea", (1001,))[1:F + 1])
    assert np.array_equal(_refarr("Dsph", (1001, 201))[1:F + 1, 1:nc + 1], o.darr("Dsph", (1001, 201))[1:F + 1, 1:nc + 1])
    assert np.array_equal(_refarr("cg_first", (201,))[1:nc + 1], o.darr("cg_first", (201,))[1:nc + 1])


@needs_ref
def test_mrqcof_bitwise(loaded):
    inp, o = loaded
    cg = o.initial_cg(inp.freq_start, 3)
    A, b, chi = o.mrqcof(cg)
    ma = o.iget("Ncoef") + 8
    alpha = np.zeros(201 * 201)
    beta = np.zeros(201)
    rchi = ref.ref_mrqcof(oracle._d(np.ascontiguousarray(cg)), ma, oracle._d(alpha), oracle._d(beta))
    mfit = A.shape[0]
    assert rchi == chi
    assert np.array_equal(alpha.reshape(201, 201)[1:mfit + 1, 1:mfit + 1], A)
    assert np.array_equal(beta[1:mfit + 1], b)


@needs_ref
@pytest.mark.parametrize("pole", [1, 6, 10])
def test_fit_bitwise(loaded, pole):
    inp, o = loaded
    cgf = o.darr("cg_first", (201,)).copy()
    ref.ref_set_cg_first(oracle._d(cgf))
    par = np.ascontiguousarray(o.darr("par", (4,)).copy())
    iap = np.ascontiguousarray(o.iarr("ia_par", (4,)).copy())
    freq = inp.freq_start - 2 * inp.freq_step
    r1 = o.fit(freq, pole)
    r2 = oracle.FitResult()
    b = [0, 0, 0, 0, 0, 60, 60, 60, -60, -60, -60][pole]
    lam = [0, 0, 90, 180, 270, 60, 180, 300, 60, 180, 300][pole]
    ref.ref_fit(freq, b, lam, oracle._d(par), oracle._i(iap), o.get("cl"), o.get("stop_condition"),
                o.iget("n_iter_min_in"), inp.conw_r0, r2)
    for k in ("dev", "chisq", "rchisq", "prd", "la", "be", "dark", "totarea", "alamda"):
        assert getattr(r1, k) == getattr(r2, k), k
    assert r1.niter == r2.niter == 50


@needs_ref
def test_gauss_errc_bitwise_and_singular():
    rng = np.random.default_rng(3)
    n = 12
    M = rng.normal(size=(n, n))
    A = M @ M.T + n * np.eye(n)
    b = rng.normal(size=n)
    a1 = np.zeros((201, 201))
    a1[1:n + 1, 1:n + 1] = A
    b1 = np.zeros(201)
    b1[1:n + 1] = b
    a2, b2 = a1.copy(), b1.copy()
    assert oracle.olib.orc_gauss_errc(oracle._d(a1), n, oracle._d(b1)) == 0
    assert ref.ref_gauss_errc(oracle._d(a2), n, oracle._d(b2)) == 0
    assert np.array_equal(b1, b2)
    assert np.allclose(b1[1:n + 1], np.linalg.solve(A, b), rtol=1e-12)
    # singular: a zero row gives a zero pivot (code 2)
    A[4, :] = 0
    A[:, 4] = 0
    a1[1:n + 1, 1:n + 1] = A
    a2 = a1.copy()
    assert oracle.olib.orc_gauss_errc(oracle._d(a1), n, oracle._d(b1.copy())) == ref.ref_gauss_errc(
        oracle._d(a2), n, oracle._d(b2.copy())) == 2
