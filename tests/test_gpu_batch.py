"""Several asteroids in one batch (ps_set_problems / ps_fit_batch_multi):
asteroids of different point counts, light curves, lmax (4, 6, 8) and stop
rules share the slots and the work queue, in one mixed order.  Every fit
must equal, bit for bit, the same fit run with its asteroid alone: a fit's
arithmetic does not depend on its neighbours in the batch (the kernel
instances for the largest coefficient count add only zero columns)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    dict(seed=11, lmax=6, mmax=6, nrows=6, n_dense=4, dense_pts=60, n_sparse=160, noise=0.02, span_years=8,
         period_h=7.3, incl_deg=10.0, stop=50),
    dict(seed=12, lmax=4, mmax=4, nrows=6, n_dense=3, dense_pts=50, n_sparse=90, noise=0.03, span_years=6,
         period_h=5.1, incl_deg=5.0, stop=50),
    dict(seed=13, lmax=8, mmax=8, nrows=6, n_dense=5, dense_pts=70, n_sparse=250, noise=0.02, span_years=9,
         period_h=11.0, incl_deg=12.0, stop=0.001),  # converge-or-100 rule: uneven fit lengths
]


# lmax 4 and 6 only: at most 54 free parameters, so the several-asteroid
# kernels run the one-wave register solver (CASES, with lmax 8, the LDS solver)
CASES_WAVE = [CASES[0], CASES[1], dict(CASES[0], seed=14, period_h=9.2, stop=0.001)]


def _inputs(tmp_path, cases=CASES):
    from periodsearch_amd import Input, synth

    inps = []
    for k, c in enumerate(cases):
        c = dict(c)
        stop = c.pop("stop")
        ast = synth.make_asteroid(**c)
        path = os.path.join(tmp_path, f"in{k}")
        synth.write_input(path, ast, c["period_h"] * 0.99, c["period_h"] * 1.01, stop=stop)
        inps.append(Input(path))
    return inps


@pytest.mark.parametrize("prec_name,cases", [("fp64", "lds"), ("bf16x3", "lds"), ("fp64", "wave"),
                                              ("bf16x3", "wave"), ("bf16x6", "wave")])
def test_multi_problem_batch_equals_single_runs(tmp_path, prec_name, cases):
    from periodsearch_amd import PS_PREC_BF16X3, PS_PREC_BF16X6, PS_PREC_FP64, Engine, PsError

    prec = {"fp64": PS_PREC_FP64, "bf16x3": PS_PREC_BF16X3, "bf16x6": PS_PREC_BF16X6}[prec_name]
    inps = _inputs(str(tmp_path), CASES if cases == "lds" else CASES_WAVE)
    items = []  # (problem, freq, pole)
    for k, inp in enumerate(inps):
        for n in (3, 9):
            f = inp.freq_start - (n - 1) * inp.freq_step
            items += [(k, f, m) for m in range(10)]
    single = {}
    for k, inp in enumerate(inps):
        eng = Engine(0)
        eng.set_precision(prec)
        eng.set_problem(inp.problem, keep=inp)
        sel = [it for it in items if it[0] == k]
        out = eng.fit_batch([it[1] for it in sel], [it[2] for it in sel], inp.conw_r0)
        for it, o in zip(sel, out):
            single[it] = o
        eng.close()
    order = np.random.default_rng(5).permutation(len(items))
    mixed = [items[i] for i in order]
    eng = Engine(0)
    eng.set_precision(prec)
    eng.set_problems([inp.problem for inp in inps], keep=inps)
    out = eng.fit_batch_multi([it[0] for it in mixed], [it[1] for it in mixed], [it[2] for it in mixed],
                              [inp.conw_r0 for inp in inps])
    niter = out["niter"]
    print(f"{prec_name}: {len(mixed)} fits of {len(inps)} asteroids, LM iterations {niter.min()}..{niter.max()}")
    assert niter.min() < niter.max()  # the stop rule < 1 gives uneven fits
    for it, o in zip(mixed, out):
        ref = single[it]
        for key in ("chisq", "dev", "prd", "la", "be", "totarea", "dark", "niter", "err"):
            assert o[key] == ref[key], (it, key, o[key], ref[key])
    with pytest.raises(PsError):
        eng.scan(inps[0].freq_start, inps[0].freq_step, 1, 2, inps[0].conw_r0)
    eng.close()
