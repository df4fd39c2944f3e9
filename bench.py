"""Throughput of the trial-period fit loop on one MI355X.

Workload (the work-unit shape): a synthetic asteroid with 1080 relative
points (6 dense light curves of 80 points + 600 sparse points over 15 years,
5% noise), shape lmax=mmax=6 (49 coefficients) on nrows=8 (512 facets), trial
periods 2-100 h with step coefficient 0.5, 10 initial poles per period, stop
condition 50 (every fit runs 50 LM iterations).

A step is one batch of consecutive trial periods (``--periods`` per step, x10
poles = fits) run through the engine's scan (fits + best-pole reduction).
The problem is uploaded before the timed region; rows come back to the host
inside it.  The convexity-weight precalc is excluded (it runs identically in
any implementation).

Forms: --prec fp64 (plain), fp32 / bf16 / bf16x3 (mixed precision), and
--form learned (the default, with bf16x6: every trial period of the step
fitted once from the network's start, in each --block-period block the best
--candidates periods and two neighbours each refitted with the plain ten
poles; the value is then the plain-equivalent rate, 10 fits per trial period
covered, and lm_fits_per_s the LM fits actually run).

Prints one JSON line with fits/s, the roofline position of the dominant
kernel (the model kernel: facet stage + brightness-derivative contraction,
timed per launch with HIP events on the engine's streams), the reference's
CPU build timed on this host's cores on a bounded sample, and (after the
timed region) the matched-chi2 check: the best row over the trial periods
around the true period against the plain FP64 scan's.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector = matrix), dense
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 matrix, dense
BF16_PEAK_TFLOPS = 2500.0  # MI355X bf16 matrix, dense
FP32_VECTOR_PEAK_TFLOPS = 157.3  # MI355X FP32 vector, published (counts packed FP32)
FP32_VECTOR_UNPACKED_TFLOPS = 78.6  # the same without packed FP32, which this build does not emit
FACET_FLOP_PER_PAIR = 40.0  # facet stage per (point, facet) pair: mu, mu0, scattering law, its derivatives, sums
WORKLOAD = "wu"
# HBM bytes per model evaluation of the model kernel on this workload, from the
# FETCH_SIZE / WRITE_SIZE counter runs (copy-calibrated) summarised in
# profiles/model_mfma_kernel_fp64_wu1080.md and
# profiles/model_mixed_kernel_bf16x3_learned_wu1080.md; the mixed forms share
# the FP32 Jacobian layout.
MODEL_HBM_BYTES_PER_EVAL = {"fp64": 469.4e3, "fp32": 242.2e3, "bf16": 242.2e3, "bf16x3": 242.2e3, "bf16x6": 242.2e3}


def cpu_baseline_reference(path: str, conw_r: float, n_procs: int, periods_each: int = 1) -> dict | None:
    """The reference's own fit loop (period_search/*.c compiled unchanged into
    oracle/_ref, single-threaded like the reference) in n_procs processes, each
    a disjoint slice of the period index (the BOINC split).  The rate is the
    fits over the span from the first fit's start to the last fit's end (the
    processes' start-up and input parsing are outside it)."""
    import oracle

    if not os.path.exists(oracle.REF_LIB):
        return None
    procs = []
    for k in range(n_procs):
        n0 = 1 + k * periods_each
        procs.append(subprocess.Popen([sys.executable, "-m", "oracle.ref_bench", path, repr(conw_r), str(n0),
                                       str(n0 + periods_each - 1)], cwd=ROOT, stdout=subprocess.PIPE,
                                      stderr=subprocess.DEVNULL, text=True))
    res = []
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            return None
        res.append(json.loads(out.strip().splitlines()[-1]))
    fits = sum(r["fits"] for r in res)
    span = max(r["t1"] for r in res) - min(r["t0"] for r in res)
    return {"value": fits / span, "unit": "fits/s", "cores": n_procs, "kind": "reference",
            "sample": f"{fits} fits: periods 1..{n_procs * periods_each} of the same workload, {n_procs} "
                      f"single-threaded processes running the reference's period_search/*.c (g++ -O3, unchanged) "
                      f"through oracle/_ref, {span:.1f} s from the first fit's start to the last fit's end"}


def cpu_baseline(path: str, conw_r: float, n_procs: int, periods_each: int = 1) -> dict:
    """The reference build where oracle/_ref exists, else the oracle
    (single-threaded, like the reference) on n_procs host cores, each process
    a disjoint slice of the period index (the BOINC split)."""
    import oracle

    ref = cpu_baseline_reference(path, conw_r, n_procs, periods_each)
    if ref is not None:
        return ref
    exe = oracle.ORACLE_MAIN
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as d:
        t0 = time.perf_counter()
        procs = []
        for k in range(n_procs):
            n0 = 1 + k * periods_each
            procs.append(subprocess.Popen([exe, path, os.path.join(d, f"out{k}"), str(n0), str(n0 + periods_each - 1),
                                           repr(conw_r)], stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
        rc = [p.wait() for p in procs]
        wall = time.perf_counter() - t0
    if any(rc):
        return None
    fits = n_procs * periods_each * 10
    return {"value": fits / wall, "unit": "fits/s", "cores": n_procs, "kind": "port",
            "sample": f"{fits} fits: periods 1..{n_procs * periods_each} of the same workload, "
                      f"{n_procs} single-threaded oracle processes, {wall:.1f} s wall"}


def bench_batch(args, prec) -> int:
    """64 uneven asteroids in one batch: per step, the `--window` trial
    periods nearest each asteroid's true period x 10 poles, all asteroids'
    fits through the same slots and work queue (ps_fit_batch_multi)."""
    import numpy as np

    from periodsearch_amd import Engine, Input, synth

    tmp = tempfile.mkdtemp(prefix="psbatch_")
    paths, periods = [], []
    for i in range(64):
        cfg = synth.uneven_asteroid(1000 + i)
        ast = synth.make_asteroid(**cfg)
        paths.append(os.path.join(tmp, f"in{i}"))
        periods.append(cfg["period_h"])
        synth.write_input(paths[-1], ast, periods[-1] * 0.98, periods[-1] * 1.02)
    eng = Engine(0)
    # shape setup of the 64 asteroids: host (ps_input_read builds each basis)
    # against the GPU (deferred reads, then one ps_shape_basis_many call)
    t0 = time.perf_counter()
    host_inps = [Input(p) for p in paths]
    t_host = time.perf_counter() - t0
    eng.shape_basis_inputs([Input(paths[0], defer_shape=True)])  # module load, first allocation
    t0 = time.perf_counter()
    inps = [Input(p, defer_shape=True) for p in paths]
    t_read = time.perf_counter() - t0
    eng.shape_basis_inputs(inps)
    t_gpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(3):
        [Input(p, defer_shape=True) for p in paths]
    t_parse = (time.perf_counter() - t0) / 3
    del host_inps
    ns = [int(round((inp.freq_start - 1.0 / P) / inp.freq_step)) + 1 for inp, P in zip(inps, periods)]
    eng.set_precision(prec)
    eng.set_problems([inp.problem for inp in inps], keep=inps)
    conw = [inp.conw_r0 for inp in inps]
    W = args.window

    def items(step):
        prob, freq, pole = [], [], []
        for k, inp in enumerate(inps):
            for j in range(W):
                n = max(1, ns[k] - W // 2 + j + step * W)
                f = inp.freq_start - (n - 1) * inp.freq_step
                prob += [k] * 10
                freq += [f] * 10
                pole += list(range(10))
        return prob, freq, pole

    for s_ in range(args.warmup):
        eng.fit_batch_multi(*items(s_), conw)
    eng.set_profiling(True)
    lat, fits = [], 0
    t0 = time.perf_counter()
    for s_ in range(args.steps):
        out = eng.fit_batch_multi(*items(args.warmup + s_), conw)
        ms, nfit, iters = eng.last_batch()
        lat += list((out["niter"] + 2) * (ms / max(iters, 1)))
        fits += nfit
    wall = time.perf_counter() - t0
    st = eng.kernel_stats()
    npts = [inp.ndata for inp in inps]
    lmaxes = [inp.problem.lmax for inp in inps]
    res = {
        "metric": "LM trial-period fits/sec on 1 MI355X (mixed+learned vs FP64) at matched chi2",
        "value": round(fits / wall, 2), "unit": "fits/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": args.prec, "data": "synthetic",
        "config": {"workload": f"64 uneven asteroids in one batch: {min(npts)}-{max(npts)} points, lmax "
                               f"{sorted(set(lmaxes))}, periods 2-50 h, {W} trial periods x 10 poles each per step",
                   "fits_per_step": 64 * W * 10, "precision": args.prec},
        "fit_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                           "p99": round(float(np.percentile(lat, 99)), 2),
                           "note": "(LM iterations + 2) x GPU time per LM iteration of the batch"},
        "model_evals": st["evals"],
        "shape_setup_ms": {"host": round((t_host - t_parse) * 1e3, 2), "gpu": round((t_gpu - t_read) * 1e3, 2),
                           "asteroids": 64, "shapes": len({(i.problem.n_fac, i.problem.lmax) for i in inps}),
                           "note": "shape basis of all 64 asteroids beyond parsing their inputs: host ps_shape_basis"
                                   " per asteroid vs one ps_shape_basis_many call"},
    }
    print(json.dumps(res), flush=True)
    eng.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--periods", type=int, default=4096, help="trial periods per step")
    ap.add_argument("--slots", type=int, default=0, help="resident fits (0 = engine default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (counter calibration) and exit")
    # default: the fastest form whose best row matches the plain FP64 scan's
    # chi2 within 1e-4 on this workload (three-part split bf16, 1.2e-5; the
    # two-part split is faster but 1.6e-4 off)
    ap.add_argument("--prec", choices=("fp64", "fp32", "bf16", "bf16x3", "bf16x6"), default="bf16x6")
    ap.add_argument("--form", choices=("plain", "learned"), default="learned")
    ap.add_argument("--workload", choices=("wu", "batch64"), default="wu",
                    help="batch64: 64 uneven asteroids in one batch of fits (plain start)")
    ap.add_argument("--window", type=int, default=16, help="batch64: trial periods per asteroid per step")
    ap.add_argument("--autotune", action="store_true",
                    help="time slot counts and slot groups on this workload first and keep the fastest")
    ap.add_argument("--candidates", type=int, default=16, help="learned form: periods verified per block")
    ap.add_argument("--block", type=int, default=4096, help="learned form: trial periods per block")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the matched-chi2 check after the timed region (the form's best row against the "
                         "plain FP64 scan's around the true period)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_mib": args.calibrate_copy, "copies": 3}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the trial-period index would shard across GPUs in contiguous ranges)", file=sys.stderr)
        return 2

    from periodsearch_amd import (PS_PREC_BF16, PS_PREC_BF16X3, PS_PREC_BF16X6, PS_PREC_FP32, PS_PREC_FP64, Engine,
                                  Input, synth)

    prec = {"fp64": PS_PREC_FP64, "fp32": PS_PREC_FP32, "bf16": PS_PREC_BF16, "bf16x3": PS_PREC_BF16X3,
            "bf16x6": PS_PREC_BF16X6}[args.prec]
    if args.workload == "batch64":
        return bench_batch(args, prec)

    tmp = tempfile.mkdtemp(prefix="psbench_")
    path = os.path.join(tmp, "period_search_in")
    ast = synth.write_workload(WORKLOAD, path)
    inp = Input(path)
    eng = Engine(0)
    eng.set_precision(prec)
    # learned form: one batch per block holds the block's proposed fits and the
    # previous block's verification fits (candidates x 5 periods x 10 poles)
    learned_batch = args.block + 10 * args.candidates * 5
    if args.slots:
        eng.set_slots(args.slots)
    elif args.form == "learned":
        eng.set_slots(learned_batch)
    eng.set_problem(inp.problem, keep=inp)
    tuned = None
    if args.autotune:
        from periodsearch_amd.autotune import autotune

        # periods far from the timed ones (the same per-fit cost under stop 50)
        if args.form == "learned":
            tuned = autotune(eng, inp, inp.conw_r0, n_fits=learned_batch, slots=(2048, 3072, learned_batch),
                             n_from=inp.n_periods // 2)
        else:
            tuned = autotune(eng, inp, inp.conw_r0, n_from=inp.n_periods // 2)
    nc, ma, mfit = eng.dims()
    N, F = inp.ndata, inp.problem.n_fac
    conw_r = inp.conw_r0
    P = args.periods
    n_next = 1

    lm_fits = 0
    if args.form == "learned":
        from periodsearch_amd.proposal import Proposal, scan_learned

        prop = Proposal.load()

    def step():
        nonlocal n_next, lm_fits
        if args.form == "learned":
            rows, _, nf, _ = scan_learned(eng, inp, prop, conw_r, n_next, n_next + P - 1,
                                          n_candidates=args.candidates, halo=2, block=args.block)
        else:
            rows = eng.scan(inp.freq_start, inp.freq_step, n_next, n_next + P - 1, conw_r)
            nf = 10 * P
        lm_fits += nf
        n_next += P
        return rows

    for _ in range(args.warmup):
        step()
    eng.set_profiling(True)
    lm_fits = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    wall = time.perf_counter() - t0
    st = eng.kernel_stats()
    fits = args.steps * P * 10
    value = fits / wall

    def matched_check() -> dict:
        """Outside the timed region: over the P trial periods around the true
        period, the best row of this form (for the learned form, from its
        plain-verified rows) against the best row of the plain FP64 scan."""
        import numpy as np

        n_true = int(round((inp.freq_start - 1.0 / ast["period_h"]) / inp.freq_step)) + 1
        n0 = max(1, n_true - P // 2)
        n1 = min(inp.n_periods, n0 + P - 1)
        if args.form == "learned":
            rows, ver, _, _ = scan_learned(eng, inp, prop, conw_r, n0, n1, n_candidates=args.candidates, halo=2,
                                           block=args.block)
            k = int(np.flatnonzero(ver)[np.argmin(rows["chisq"][ver])])
        else:
            rows = eng.scan(inp.freq_start, inp.freq_step, n0, n1, conw_r)
            k = int(np.argmin(rows["chisq"]))
        if prec == PS_PREC_FP64 and args.form != "learned":
            r64 = rows
        else:
            ref = Engine(0)
            ref.set_precision(PS_PREC_FP64)
            ref.set_problem(inp.problem, keep=inp)
            r64 = ref.scan(inp.freq_start, inp.freq_step, n0, n1, conw_r)
            ref.close()
        k64 = int(np.argmin(r64["chisq"]))
        step_h = float(r64["period_h"][k64]) ** 2 * inp.freq_step
        return {"window": f"trial periods {n0}..{n1} around the true period {ast['period_h']} h",
                "best_period_h": round(float(rows["period_h"][k]), 8), "chi2": float(rows["chisq"][k]),
                "fp64_best_period_h": round(float(r64["period_h"][k64]), 8), "fp64_chi2": float(r64["chisq"][k64]),
                "period_steps_apart": round(abs(float(rows["period_h"][k]) - float(r64["period_h"][k64])) / step_h, 3),
                "chi2_rel": abs(float(rows["chisq"][k]) - float(r64["chisq"][k64])) / float(r64["chisq"][k64])}
    # algorithmic work of the dominant kernel: the contraction 2*N*F*Ncoef per
    # model evaluation (masked facets counted, as the reference's dense loop)
    flop_eval = 2.0 * N * F * nc
    achieved = (st["evals"] * flop_eval) / (st["model_ms"] * 1e-3) / 1e12 if st["model_ms"] > 0 else 0.0
    peak = {"fp64": FP64_PEAK_TFLOPS, "fp32": FP32_PEAK_TFLOPS, "bf16": BF16_PEAK_TFLOPS,
            "bf16x3": BF16_PEAK_TFLOPS, "bf16x6": BF16_PEAK_TFLOPS}[args.prec]
    # the model kernel's other half: the facet stage on the vector units (FP64
    # in the plain form, FP32 in the mixed forms), which bounds the mixed forms
    vec_eval = FACET_FLOP_PER_PAIR * N * F
    vec_achieved = (st["evals"] * vec_eval) / (st["model_ms"] * 1e-3) / 1e12 if st["model_ms"] > 0 else 0.0
    vec_peak = FP64_PEAK_TFLOPS if args.prec == "fp64" else FP32_VECTOR_PEAK_TFLOPS
    vector = {"stage": "facet stage, %d flop per point-facet pair" % FACET_FLOP_PER_PAIR,
              "dtype": "f64" if args.prec == "fp64" else "f32", "achieved": round(vec_achieved, 3),
              "peak": vec_peak, "unit": "TFLOP/s", "frac": round(vec_achieved / vec_peak, 4)}
    if args.prec != "fp64":
        vector["peak_without_packed_fp32"] = FP32_VECTOR_UNPACKED_TFLOPS
        vector["frac_without_packed_fp32"] = round(vec_achieved / FP32_VECTOR_UNPACKED_TFLOPS, 4)
    iters = 50
    flop_fit = (iters + 1) * (2.0 * N * F * nc + (N + 3) * mfit * (mfit + 1))
    out = {
        "metric": "LM trial-period fits/sec on 1 MI355X (mixed+learned vs FP64) at matched chi2",
        "value": round(value, 2),
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32 contraction and alpha (MFMA, f32 accumulate), f64 geometry/chi2/solve",
                  "bf16": "bf16 contraction (MFMA, f32 accumulate), f32 facet stage and alpha, f64 geometry/chi2/solve",
                  "bf16x6": "three-part split-bf16 contraction (6 MFMA products, f32 accumulate), f32 facet stage "
                            "and alpha, f64 geometry/chi2/solve",
                  "bf16x3": "split-bf16 contraction (3 MFMA products, f32 accumulate), f32 facet stage and alpha, "
                            "f64 geometry/chi2/solve"}[args.prec],
        "data": "synthetic",
        "config": {"workload": "1080 points (6x80 dense + 600 sparse, 15 yr, 5% noise), lmax=mmax=6, "
                               "nrows=8 (512 facets), 2-100 h, stop 50, 10 poles",
                   "periods_per_step": P, "fits_per_step": P * 10, "precision": args.prec,
                   "form": {"fp64": "plain FP64", "fp32": "mixed FP32", "bf16": "mixed BF16",
                            "bf16x3": "mixed split BF16", "bf16x6": "mixed three-part split BF16"}[args.prec] + (" + learned proposal" if args.form == "learned"
                                                                         else "")},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4),
                     "traffic": round(MODEL_HBM_BYTES_PER_EVAL[args.prec] * st["evals"] / max(st["model_launches"], 1)),
                     "kernel": "model_kernel", "model_ms_total": round(st["model_ms"], 3),
                     "model_launches": st["model_launches"], "evals": st["evals"],
                     "flop_per_eval": flop_eval, "vector": vector},
        "job_tflops": round(lm_fits * flop_fit / wall / 1e12, 3),
        "lm_fits_per_s": round(lm_fits / wall, 2),
    }
    if not args.no_check:
        out["matched_chi2"] = matched_check()
    if tuned:
        out["config"]["autotuned"] = {"slots": tuned["best"]["slots"], "groups": tuned["best"]["groups"]}
    else:
        out["config"]["slots"] = args.slots or (learned_batch if args.form == "learned" else 2048)
    if args.form == "learned":
        out["config"]["learned"] = (f"1 proposed start per trial period; in each {args.block}-period block the "
                                    f"{args.candidates} best periods and 2 neighbours each refitted from the 10 "
                                    f"plain poles (in the next block's batch); value = 10 x trial periods covered "
                                    f"per second")
    if not args.no_cpu:
        try:
            ncpu = min(len(os.sched_getaffinity(0)), 16)
        except AttributeError:
            ncpu = min(os.cpu_count() or 1, 16)
        out["cpu_baseline"] = cpu_baseline(path, conw_r, ncpu)
        if out["cpu_baseline"]:
            out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
