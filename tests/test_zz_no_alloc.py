"""ps_fit_batch_noalloc (fits whose slot memory lives in the library's fixed
per-device array; inputs and outputs in caller-owned device memory) against
ps_fit_batch, through the C ABI: two calls in a row, the second smaller, each
bitwise equal to ps_fit_batch's fits; a pole index out of range is rejected
and leaves the engine usable."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the engine library itself loaded (found in this
    process's mappings, so that no second runtime is brought in)."""
    import periodsearch_amd  # noqa: F401  (loads libpsamd.so and its HIP runtime)

    path = next(line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64.so" in line)
    hip = C.CDLL(path)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


H2D, D2H = 1, 2


def _device_call(eng, freq, pole, conw_r):
    import periodsearch_amd as ps

    hip = _hip()
    freq = np.ascontiguousarray(freq, np.float64)
    pole = np.ascontiguousarray(pole, np.int32)
    n = len(freq)
    out = np.zeros(n, dtype=ps.FIT_DTYPE)
    ptrs = []
    for nbytes in (freq.nbytes, pole.nbytes, out.nbytes):
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), nbytes) == 0
        ptrs.append(p)
    d_freq, d_pole, d_out = ptrs
    try:
        assert hip.hipMemcpy(d_freq, freq.ctypes.data, freq.nbytes, H2D) == 0
        assert hip.hipMemcpy(d_pole, pole.ctypes.data, pole.nbytes, H2D) == 0
        rc = ps.lib.ps_fit_batch_noalloc(eng._h, d_freq, d_pole, n, conw_r, d_out)
        assert hip.hipDeviceSynchronize() == 0
        if rc == 0:
            assert hip.hipMemcpy(out.ctypes.data, d_out, out.nbytes, D2H) == 0
    finally:
        for p in ptrs:
            hip.hipFree(p)
    return rc, (out if rc == 0 else None)


def test_noalloc_fits_equal_fit_batch(small_incl_input):
    import periodsearch_amd as ps
    from periodsearch_amd import Engine, Input

    inp = Input(small_incl_input)
    eng = Engine(0)
    eng.set_problem(inp.problem, keep=inp)
    rng = np.random.default_rng(3)
    for n_periods in (40, 7):  # 400 fits (two refill rounds in the fixed array), then 70
        n = rng.integers(0, 60, size=n_periods)
        freq = np.repeat(inp.freq_start - n * inp.freq_step, 10)
        pole = np.tile(np.arange(10, dtype=np.int32), n_periods)
        print(f"noalloc call: {len(freq)} fits", flush=True)
        rc, got = _device_call(eng, freq, pole, inp.conw_r0)
        assert rc == 0, ps.lib.ps_last_error(eng._h)
        want = eng.fit_batch(freq, pole, inp.conw_r0)
        for k in ps.FIT_RESULT_FIELDS:
            assert np.array_equal(got[k], want[k]), k
    rc, _ = _device_call(eng, [inp.freq_start] * 3, [0, 10, 1], inp.conw_r0)
    assert rc == -1
    ok = eng.fit_batch([inp.freq_start], [0], inp.conw_r0)
    assert ok["niter"][0] == 50
    eng.close()
