"""The command-line driver (period_search_amd) end to end on the GPU: output
rows against the oracle's driver, the validator rule, checkpoint/resume
(period_search_BOINC.cpp:85-101, 237-251) from each of the three reference
checkpoint formats, including an optimised-CPU-app checkpoint taken inside a
convexity-weight round, and the learned scan (--learned) against the
library's ps_scan_learned."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from periodsearch_amd import DRIVER_PATH, synth
from validator_rule import compare, read_rows

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def short_scan(tmp_path_factory):
    d = tmp_path_factory.mktemp("drv")
    path = str(d / "period_search_in")
    # 6 trial periods: 1/P_end = 1/7 - 5 * freq_step (8-yr span, step coef 0.5);
    # lmax=mmax=3 keeps the oracle's CPU run short
    cfg = {k: v for k, v in synth.WORKLOADS["small"].items() if k not in ("seed", "per_start", "per_end")}
    cfg.update(lmax=3, mmax=3, incl_deg=10.0)
    ast = synth.make_asteroid(0, **cfg)
    synth.write_input(path, ast, 7.0, 7.0009)
    return str(d), path


def _run_driver(wd, extra=()):
    return subprocess.run([DRIVER_PATH, "--batch", "2", *extra], cwd=wd, capture_output=True, text=True,
                          timeout=300)


def test_driver_matches_oracle_and_resumes(short_scan):
    wd, path = short_scan
    r = _run_driver(wd)
    assert r.returncode == 0, r.stderr
    gpu = open(os.path.join(wd, "period_search_out")).read()
    ref_out = os.path.join(wd, "oracle_out")
    oracle.run_oracle_main(path, ref_out, timeout=600)
    orc = open(ref_out).read()
    g_rows, o_rows = read_rows(gpu), read_rows(orc)
    assert len(g_rows) == len(o_rows) >= 5
    assert compare(gpu, orc)
    # Rows are held to the validator's quantities.  Summation order differs
    # between the GPU and the CPU in the last bits, and 50 LM iterations
    # amplify that on a few (period, pole) trajectories (at lmax 3, per-fit
    # chi2 differences of up to 4e-3 on 2-3 of 60 fits, the rest below 1e-8),
    # which can move a row's best pole to another minimum.  So: period within
    # one scan step (2.5e-5 relative here) in every row and within 1e-6 in all
    # but at most one; rms and chi2 within 1e-4 relative in all but at most
    # one row, and within 1e-2 there.
    close = 0
    for g, o in zip(g_rows, o_rows):
        assert abs(g[0] - o[0]) <= 2.5e-5 * o[0]       # period
        assert abs(g[1] - o[1]) <= 1e-2 * o[1] and abs(g[2] - o[2]) <= 1e-2 * o[2]
        close += (abs(g[0] - o[0]) <= 1e-6 * o[0] and abs(g[1] - o[1]) <= 1e-4 * o[1] + 1e-6
                  and abs(g[2] - o[2]) <= 1e-4 * o[2] + 1e-6)
    assert close >= len(g_rows) - 1
    state = open(os.path.join(wd, "period_search_state")).read().split()
    assert int(state[0]) == len(g_rows) and int(state[1]) == 1
    # resume: a base-format checkpoint after 3 rows and a truncated output
    lines = gpu.splitlines(keepends=True)
    with open(os.path.join(wd, "period_search_out"), "w") as f:
        f.writelines(lines[:3])
    with open(os.path.join(wd, "period_search_state"), "w") as f:
        f.write("3")
    r = _run_driver(wd)
    assert r.returncode == 0, r.stderr
    assert open(os.path.join(wd, "period_search_out")).read() == gpu


def _copy_input(src, d):
    os.makedirs(d)
    shutil.copy(src, os.path.join(d, "period_search_in"))
    return d


@pytest.fixture(scope="module")
def full_run(short_scan, tmp_path_factory):
    """One uninterrupted driver run: its output text and final checkpoint."""
    wd = _copy_input(short_scan[1], str(tmp_path_factory.mktemp("full") / "run"))
    r = _run_driver(wd)
    assert r.returncode == 0, r.stderr
    from periodsearch_amd import state_read

    return open(os.path.join(wd, "period_search_out")).read(), state_read(os.path.join(wd, "period_search_state"))


def _first_round_sum(path, n_done):
    """The first convexity-weight round's dark sum after n_done of its
    periods, as the round adds it (NaN dark counts 1.0,
    period_search_BOINC.cpp:698-703)."""
    from periodsearch_amd import Engine, Input

    inp = Input(path)
    eng = Engine(0)
    eng.set_problem(inp.problem, keep=inp)
    rows = eng.scan(inp.freq_start, inp.freq_step, 1, n_done, inp.conw_r0)
    eng.close()
    s = 0.0
    for d in rows["dark"]:
        s = s + (1.0 if np.isnan(d) else float(d))
    return inp.conw_r0, s


def test_driver_resumes_from_each_checkpoint_format(short_scan, full_run, tmp_path):
    from periodsearch_amd import PS_STATE_CUDA, PS_STATE_OC, state_read

    text, st = full_run
    assert st["format"] == PS_STATE_CUDA and st["new_conw"] == 1
    lines = text.splitlines(keepends=True)
    conw_r0, s4 = _first_round_sum(short_scan[1], 4)
    cases = {
        "cuda": (3, "3 1 %.17g" % st["conw_r"]),
        "oc_after_precalc": (3, "3 1 %.17g 0 0" % st["conw_r"]),
        # inside the first precalc round, 4 of its periods done
        "oc_inside_precalc": (0, "0 0 %.17g %.17g 4" % (conw_r0, s4)),
        # unreadable: a fresh start, the output truncated
        "garbage": (3, "garbage\n"),
    }
    for name, (keep, state_text) in cases.items():
        wd = _copy_input(short_scan[1], str(tmp_path / name))
        with open(os.path.join(wd, "period_search_out"), "w") as f:
            f.writelines(lines[:keep])
        with open(os.path.join(wd, "period_search_state"), "w") as f:
            f.write(state_text)
        r = _run_driver(wd)
        assert r.returncode == 0, (name, r.stderr)
        assert open(os.path.join(wd, "period_search_out")).read() == text, name
        end = state_read(os.path.join(wd, "period_search_state"))
        assert end["nlines"] == len(lines) and end["conw_r"] == st["conw_r"], (name, end)
    # the OC app's own format is written on request and resumed from
    wd = _copy_input(short_scan[1], str(tmp_path / "write_oc"))
    r = _run_driver(wd, ["--state-format", "oc"])
    assert r.returncode == 0, r.stderr
    assert open(os.path.join(wd, "period_search_state")).read() == "%d 1 %.17g 0 0" % (len(lines), st["conw_r"])
    # a dark sum that forces the round's weight up: the weight doubles at
    # least once from that round on, so the sum taken from the checkpoint is used
    wd = _copy_input(short_scan[1], str(tmp_path / "forced"))
    with open(os.path.join(wd, "period_search_state"), "w") as f:
        f.write("0 0 %.17g 1e6 4" % conw_r0)
    r = _run_driver(wd)
    assert r.returncode == 0, r.stderr
    ratio = state_read(os.path.join(wd, "period_search_state"))["conw_r"] / conw_r0
    assert ratio >= 2 and ratio == 2.0 ** round(np.log2(ratio))


def test_driver_learned_scan_matches_the_library(tmp_path):
    """--learned on a held-out asteroid (292 trial periods, three blocks of
    128): the driver's rows and judged flags are the ones ps_scan_learned
    gives through the Python binding with the same network, options and
    convexity weight, and its best judged row is the plain FP64 scan's best
    row."""
    from periodsearch_amd import PS_PREC_BF16X6, PS_PREC_FP64, Engine, Input, format_row, state_read, synth
    from periodsearch_amd import proposal as pr

    cfg = pr.draw_asteroid(5001)
    ast = synth.make_asteroid(**cfg)
    P = cfg["period_h"]
    wd = str(tmp_path)
    path = os.path.join(wd, "period_search_in")
    synth.write_input(path, ast, P * 0.99, P * 1.01)
    r = subprocess.run([DRIVER_PATH, "--learned", "--prec", "bf16x6", "--block", "128", "--candidates", "4",
                        "--halo", "1", "--judged-out", "judged"], cwd=wd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = open(os.path.join(wd, "period_search_out")).read().splitlines(keepends=True)
    got_judged = np.loadtxt(os.path.join(wd, "judged"), dtype=int)
    conw_r = state_read(os.path.join(wd, "period_search_state"))["conw_r"]

    inp = Input(path)
    n = inp.n_periods
    assert len(got) == n == 292
    eng, e64 = Engine(0), Engine(0)
    eng.set_precision(PS_PREC_BF16X6)
    e64.set_precision(PS_PREC_FP64)
    for e in (eng, e64):
        e.set_problem(inp.problem, keep=inp)
    prop = pr.Proposal.load(max_batch=128)
    rows, judged, nf, _ = pr.scan_learned(eng, inp, prop, conw_r, 1, n, n_candidates=4, halo=1, block=128,
                                          judge=e64, refine=8)
    assert np.array_equal(got_judged[:, 0], np.arange(1, n + 1))
    assert np.array_equal(got_judged[:, 1].astype(bool), judged)
    assert got[1:] == [format_row(rr) for rr in rows[1:]]
    assert got[0].split()[:3] == format_row(rows[0]).split()[:3]  # row 1 prints the weight as its 4th number
    plain = e64.scan(inp.freq_start, inp.freq_step, 1, n, conw_r)
    kp = int(np.nanargmin(plain["chisq"]))
    kl = int(np.flatnonzero(judged)[np.nanargmin(rows["chisq"][judged])])
    print(f"{n} periods, {int(judged.sum())} judged, {nf} LM fits (plain {10 * n}); best row plain {kp} learned {kl}")
    assert kl == kp and rows["chisq"][kl] == plain["chisq"][kp]
    prop.close()
    eng.close()
    e64.close()
