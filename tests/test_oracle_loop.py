"""Pins of the oracle's fit loop and convexity-weight precalc (no GPU).

* The whole driver -- parse, precalc, scan, row text -- against the
  reference-held sample run: rows 1-3 of period_search/period_search_out,
  whose row 1 prints conw_r * escl^2 (period_search_BOINC.cpp:927-928) and so
  pins the precalc's outcome.
* The stop rule of period_search_BOINC.cpp:846-892 from its algebra:
  iter_diff is 1e40, then 1e30 - dev_1, then the last improvement of dev by
  more than 1e-10, so (a) a threshold below 1e-10 can never stop the loop (it
  runs to the 100-iteration cap), (b) a threshold of 0.99 stops it at the
  first improvement after iteration 1 (dev <= 1 here) but not before
  n_iter_min, and (c) for a real threshold the stop iteration follows from
  the dev trajectory, which is the same whatever the stop rule (the rule only
  decides when to stop): the trajectory of a 50-iteration run predicts where
  the convergence run stops.
* How far the reference arithmetic itself moves a 50-iteration fit when its
  inputs move by one ulp: the bar against which the GPU's FP64 fits are held
  (tests/test_gpu_parity.py).
* The oracle's FP32-derivative switch, which measures what an FP32 Jacobian
  alone does to the reference arithmetic.
"""
import os

import numpy as np
import pytest

import oracle
from periodsearch_amd import Input, synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def incl_path(tmp_path_factory):
    p = str(tmp_path_factory.mktemp("loop") / "period_search_in")
    synth.write_workload("small", p, incl_deg=10.0)
    return p


def _orc(path, stop=None, n_iter_min=None):
    inp = Input(path)
    o = oracle.Oracle(path)
    if stop is not None:
        o.set("stop_condition", stop)
    if n_iter_min is not None:
        o.set("n_iter_min_in", n_iter_min)
    o.prepare(inp.conw_r0)
    return inp, o


@pytest.mark.slow
def test_oracle_driver_reproduces_reference_rows_1_to_3(tmp_path):
    """Precalc + scan + text of the oracle's driver on the reference's own
    sample input give its sample output's first three rows byte for byte."""
    out = str(tmp_path / "out")
    oracle.run_oracle_main(os.path.join(GOLDEN, "ps_sample_in"), out, 1, 3, timeout=900)
    got = open(out).read().splitlines()
    ref = open(os.path.join(GOLDEN, "ps_sample_out")).read().splitlines()[:3]
    assert got == ref
    # row 1's fourth column is the final convexity weight times escl^2
    assert got[0].split()[3] == "0.1"


def test_stop_below_the_update_floor_runs_to_the_cap(incl_path):
    inp, o = _orc(incl_path, stop=1e-300, n_iter_min=0)
    for m in (1, 6):
        r = o.fit(inp.freq_start, m)
        assert r.niter == 100


def test_stop_at_first_improvement_and_minimum_iterations(incl_path):
    for nmin in (0, 10):
        inp, o = _orc(incl_path, stop=0.99, n_iter_min=nmin)
        for m in (1, 4, 9):
            r = o.fit(inp.freq_start - 5 * inp.freq_step, m)
            dev = o.darr("dev_trace", (102,))[1:r.niter + 1].copy()
            # dev_old holds dev_1 until the first later iteration whose dev is
            # below it by more than 1e-10
            first = None
            for k in range(2, len(dev) + 1):
                if dev[0] - dev[k - 1] > 1e-10:
                    first = k
                    break
            assert first is not None and first <= max(nmin, 2) + 50
            assert r.niter == max(first, nmin), (m, nmin, r.niter, first)


def _predicted_stop(dev, thr, nmin):
    """Stop iteration of period_search_BOINC.cpp:868-892 for a dev trajectory."""
    iter_diff, dev_old = 1e40, 1e30
    for k in range(1, len(dev) + 1):
        d = dev[k - 1]
        if dev_old - d > 1e-10:
            iter_diff = dev_old - d
            dev_old = d
        if not ((k < 100 and iter_diff > thr) or k < nmin):
            return k
    return None


def test_convergence_stop_follows_the_50_iteration_trajectory(incl_path):
    inp, o50 = _orc(incl_path)
    freq = inp.freq_start - 20 * inp.freq_step
    checked = 0
    for m in range(1, 11):
        r50 = o50.fit(freq, m)
        assert r50.niter == 50
        dev = o50.darr("dev_trace", (102,))[1:51].copy()
        for thr, nmin in ((1e-3, 10), (1e-4, 10), (1e-4, 0), (3e-6, 10)):
            k = _predicted_stop(dev, thr, nmin)
            if k is None:
                continue  # would run past 50: not predicted by this trajectory
            _, oc = _orc(incl_path, stop=thr, n_iter_min=nmin)
            rc = oc.fit(freq, m)
            assert rc.niter == k, (m, thr, nmin, rc.niter, k)
            # the same iterations, so the same numbers
            assert np.array_equal(oc.darr("dev_trace", (102,))[1:k + 1], dev[:k])
            checked += 1
    assert checked >= 20


def ulp_spread(path, fits, perturb="tim"):
    """Relative chi2 change of each (freq, pole) fit when every time (or
    brightness) moves up by one ulp."""
    inp, base = _orc(path)
    _, pert = _orc(path)
    n = pert.iget("ndata") - 3
    arr = pert.darr("tim" if perturb == "tim" else "brightness", (20004,))
    arr[1:n + 1] = np.nextafter(arr[1:n + 1], np.inf)
    out = []
    for f, m in fits:
        a, b = base.fit(f, m), pert.fit(f, m)
        out.append(abs(a.chisq - b.chisq) / a.chisq)
    return np.array(out)


def test_one_ulp_moves_50_iteration_fits_beyond_1e9(incl_path):
    """The reference arithmetic does not fix a 50-iteration fit's chi2 to
    1e-9: one ulp of the input times moves some fits by far more (LM accepts
    or rejects a trial by a hair, and the paths part).  A per-fit 1e-9 bar on
    another implementation is therefore only meaningful where the reference's
    own fit is stable under such a perturbation."""
    inp = Input(incl_path)
    fits = [(inp.freq_start - (n - 1) * inp.freq_step, m) for n in (1, 40) for m in range(1, 11)]
    spread = ulp_spread(incl_path, fits)
    print("1-ulp time perturbation, chi2 relative change per fit:", " ".join("%.1e" % v for v in spread))
    assert (spread > 1e-9).sum() >= 3
    assert np.median(spread) < 1e-6


def test_fp32_derivative_switch(incl_path):
    """The oracle's dyda_f32 switch (its derivatives rounded to float: what a
    Jacobian held in FP32 alone does to the reference arithmetic, used to
    explain the mixed forms' chi2 at the 5,000-point shape): off by default
    and when cleared; on, alpha and beta move by FP32 rounding, between 1e-12
    and 1e-6 of their largest entry, while chi2, which uses no derivative,
    stays bitwise the same."""
    inp, o = _orc(incl_path)
    cg = o.initial_cg(inp.freq_start, 3)
    A0, b0, c0 = o.mrqcof(cg)
    o.set("dyda_f32", 1)
    A1, b1, c1 = o.mrqcof(cg)
    o.set("dyda_f32", 0)
    A2, b2, c2 = o.mrqcof(cg)
    assert np.array_equal(A0, A2) and np.array_equal(b0, b2) and c0 == c2
    ea = np.abs(A1 - A0).max() / np.abs(A0).max()
    eb = np.abs(b1 - b0).max() / np.abs(b0).max()
    assert 1e-12 < ea < 1e-6 and 1e-12 < eb < 1e-6, (ea, eb)
    assert c1 == c0
