"""Errors at the C ABI: calls out of order, invalid arguments and invalid
problems return their error codes from the host, before any kernel runs, and
leave the engine usable.

The reference's CUDA seam returns 1/0 and exits the process on fatal
conditions (period_search_cuda/period_search/start_CUDA.cu:774); here every
entry point returns 0 or a negative code (include/periodsearch.h) and the
message is kept for ps_last_error.  The shapes covered follow the reference's
own limits (period_search/constants.h): empty batches, pole indices outside
the ten-pole table, reversed or empty period ranges, light-curve point counts
that do not add up, and more shape coefficients than MAX_LM = 10 gives.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PS_OK, PS_ERR_ARG, PS_ERR_LIMIT, PS_ERR_STATE = 0, -1, -3, -5


def _rows(n):
    from periodsearch_amd import ROW_DTYPE

    return np.zeros(max(n, 1), dtype=ROW_DTYPE)


def test_calls_out_of_order_and_bad_arguments(small_incl_input):
    from periodsearch_amd import Engine, Input, PsError, lib

    inp = Input(small_incl_input)
    f = inp.freq_start - 2 * inp.freq_step
    eng = Engine(0)
    h = eng._h
    # nothing set yet
    nc, ma, mf = C.c_int(), C.c_int(), C.c_int()
    assert lib.ps_dims(h, C.byref(nc), C.byref(ma), C.byref(mf)) == PS_ERR_STATE
    rows = _rows(2)
    assert lib.ps_scan(h, inp.freq_start, inp.freq_step, 1, 2, inp.conw_r0, rows.ctypes.data) == PS_ERR_STATE
    assert "no problem set" in lib.ps_last_error(h).decode()
    with pytest.raises(PsError):
        eng.fit_batch([f], [0], inp.conw_r0)
    # settings out of range
    assert lib.ps_set_precision(h, 7) == PS_ERR_ARG
    assert lib.ps_set_groups(h, 0) == PS_ERR_ARG and lib.ps_set_groups(h, 5) == PS_ERR_ARG
    assert lib.ps_set_slots(h, 0) == PS_ERR_ARG
    eng.set_problem(inp.problem, keep=inp)
    # an empty batch is no work, not an error
    assert len(eng.fit_batch([], [], inp.conw_r0)) == 0
    # pole indices outside the table
    for bad in (10, -1):
        with pytest.raises(PsError, match="pole index"):
            eng.fit_batch([f, f], [0, bad], inp.conw_r0)
    # period ranges: reversed, starting at 0 (rows are 1-based), no output
    assert lib.ps_scan(h, inp.freq_start, inp.freq_step, 5, 4, inp.conw_r0, rows.ctypes.data) == PS_ERR_ARG
    assert lib.ps_scan(h, inp.freq_start, inp.freq_step, 0, 1, inp.conw_r0, rows.ctypes.data) == PS_ERR_ARG
    assert lib.ps_scan(h, inp.freq_start, inp.freq_step, 1, 2, inp.conw_r0, None) == PS_ERR_ARG
    # several-asteroid batch: problem index out of range
    with pytest.raises(PsError):
        eng.fit_batch_multi([1], [f], [0], [inp.conw_r0])
    # after all of that the engine gives the fits of a fresh engine, bit for bit
    got = eng.fit_batch([f] * 3, [0, 4, 9], inp.conw_r0)
    eng.close()
    fresh = Engine(0)
    fresh.set_problem(inp.problem, keep=inp)
    want = fresh.fit_batch([f] * 3, [0, 4, 9], inp.conw_r0)
    fresh.close()
    for key in ("chisq", "dev", "prd", "la", "be", "totarea", "niter", "err"):
        assert np.array_equal(got[key], want[key]), key


def test_invalid_problems_are_rejected(small_incl_input):
    from periodsearch_amd import Engine, Input, PsError, lib

    inp = Input(small_incl_input)
    eng = Engine(0)
    # light-curve point counts that do not add up to n_points
    p = inp.problem
    n0 = p.n_points
    p.n_points = n0 - 1
    try:
        with pytest.raises(PsError, match="do not add up"):
            eng.set_problem(p, keep=inp)
    finally:
        p.n_points = n0
    # mmax above lmax
    m0 = p.mmax
    p.mmax = p.lmax + 1
    try:
        assert lib.ps_set_problem(eng._h, C.byref(p)) == PS_ERR_ARG
    finally:
        p.mmax = m0
    # the problem as read is accepted afterwards
    eng.set_problem(inp.problem, keep=inp)
    nc, ma, mf = eng.dims()
    assert (nc, ma) == (49, 57)
    eng.close()
