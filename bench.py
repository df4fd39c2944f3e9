"""Throughput of the trial-period fit loop on one MI355X.

Workload (default ``--workload c2``, the mixed-precision configuration's
shape): a synthetic asteroid with 5,000 relative points (20 dense light
curves of 150 points + 2,000 sparse points, 2% noise, two apparitions 420 days
apart), shape lmax = mmax = 8 (81 coefficients) on nrows = 10 (800 facets),
trial periods 4-12 h with step coefficient 0.5 (6,720 trial periods), 10
initial poles per period, stop condition 50 (every fit runs 50 LM
iterations, the reference's default and its real work units' setting).
``--workload wu`` is the work-unit shape (1,080 points, lmax 6, 512 facets,
2-100 h); ``--workload batch64`` the 64 uneven asteroids in one batch.

Three forms are timed in one run, each with its inputs resident in HBM
before its timed region and its rows returned inside it, the engine on its
default path (poll graphs):
  * plain FP64 -- ``--periods`` trial periods x 10 poles per step
    (engine.scan): the comparator, ``fp64_fits_per_s``;
  * the mixed form alone (``--prec``, default the three-part split bf16) on
    the same periods: ``mixed_fits_per_s``;
  * the headline form, mixed + learned (``--form learned``): per step one
    whole learned scan of the work unit (c2: all 6,720 trial periods, in
    blocks of ``--block``): one LM fit per period from the network's start,
    and the ``--candidates`` lowest-rms periods per block with ``--halo``
    neighbours refitted from the ten plain poles by an FP64 engine (the
    plain method judges).  ``value`` is the LM fits actually run per second
    (screening + FP64 judge fits); the trial periods covered per second and
    the plain-equivalent rate (10 fits per covered period) are reported
    beside it, and every ratio to the FP64 rate.
With ``--form plain`` the value is the mixed form's own rate.

After the timed regions: one profiled pass of the headline form (the model
kernel timed per launch with HIP events, the LM iterations as plain launches,
because events captured into a HIP graph cannot be timed here) for the
roofline figures; the matched-chi2 check (the headline's best judged row
against the plain FP64 scan's best row over the trial periods around the
true period); and the reference's own fit loop (period_search/*.c compiled
unchanged into oracle/_ref) timed on this host's cores on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
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
HOST_CPU_SHARE = 16  # CPU threads the GPU box allots to one GPU
# HBM bytes per model evaluation of the model kernel (copy-calibrated
# FETCH_SIZE + WRITE_SIZE counter runs; profiles/), by workload and precision
MODEL_HBM_BYTES_PER_EVAL = {
    ("c2", "fp64"): 4223.0e3, ("c2", "bf16"): 2333.0e3, ("c2", "bf16a"): 2333.0e3,  # bf16a: the bf16 model kernel
}
PREC_NAMES = ("fp64", "fp32", "bf16", "bf16x3", "bf16x6", "bf16a")
DTYPE = {"fp64": "f64",
         "fp32": "f32 contraction and alpha (MFMA, f32 accumulate), f64 geometry/chi2/solve",
         "bf16": "bf16 contraction (MFMA, f32 accumulate), f32 facet stage and alpha, f64 geometry/chi2/solve",
         "bf16x3": "split-bf16 contraction (3 MFMA products, f32 accumulate), f32 facet stage and alpha, "
                   "f64 geometry/chi2/solve",
         "bf16x6": "three-part split-bf16 contraction (6 MFMA products, f32 accumulate), f32 facet stage and alpha, "
                   "f64 geometry/chi2/solve",
         "bf16a": "bf16 contraction and alpha products (MFMA, f32 accumulate), f32 facet stage, f32/f64 beta, "
                  "f64 geometry/chi2/solve"}
FORM = {"fp64": "plain FP64", "fp32": "mixed FP32", "bf16": "mixed BF16", "bf16x3": "mixed split BF16",
        "bf16x6": "mixed three-part split BF16", "bf16a": "mixed BF16 with bf16 alpha (screening)"}
WORKLOAD_TEXT = {
    "c2": "5000 points (20x150 dense + 2000 sparse, 2 apparitions over 420 d, 2% noise), lmax=mmax=8 (81 coef), "
          "nrows=10 (800 facets), 4-12 h (6720 trial periods), stop 50, 10 poles",
    "wu": "1080 points (6x80 dense + 600 sparse, 15 yr, 5% noise), lmax=mmax=6, nrows=8 (512 facets), 2-100 h, "
          "stop 50, 10 poles",
}


def host_info() -> dict:
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                info["cpu"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        mhz = [float(l.split(":")[1]) for l in open("/proc/cpuinfo") if l.startswith("cpu MHz")]
        if mhz:
            info["mhz_mean"] = round(sum(mhz) / len(mhz))
    except (OSError, ValueError):
        pass
    try:
        info["compiler"] = subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    except (OSError, IndexError):
        pass
    info["flags"] = "-O3 (period_search/Makefile:15-18), no -march, single-threaded per process"
    info["python"] = platform.python_version()
    return info


def cpu_baseline(path: str, conw_r: float, n_procs: int, n_poles: int) -> dict | None:
    """The reference's own fit loop (period_search/*.c compiled unchanged into
    oracle/_ref, single-threaded like the reference) in n_procs processes,
    each fitting its own trial period (the BOINC split) from the first
    n_poles poles.  The rate counts the fits from the first fit's start to the
    last fit's end (process start-up and input parsing are outside it)."""
    import oracle

    if not os.path.exists(oracle.REF_LIB):
        return None
    procs = []
    for k in range(n_procs):
        procs.append(subprocess.Popen([sys.executable, "-m", "oracle.ref_bench", path, repr(conw_r), str(1 + k),
                                       str(1 + k), str(n_poles)], cwd=ROOT, stdout=subprocess.PIPE,
                                      stderr=subprocess.DEVNULL, text=True))
    res = []
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            return None
        res.append(json.loads(out.strip().splitlines()[-1]))
    fits = sum(r["fits"] for r in res)
    span = max(r["t1"] for r in res) - min(r["t0"] for r in res)
    return {"value": round(fits / span, 3), "unit": "fits/s", "cores": n_procs, "kind": "reference",
            "sample": f"{fits} fits: trial periods 1..{n_procs} of the same workload, {n_poles} pole(s) each, "
                      f"{n_procs} single-threaded processes running the reference's period_search/*.c (g++ -O3, "
                      f"unchanged) through oracle/_ref, {span:.1f} s from the first fit's start to the last fit's end",
            "host": host_info()}


def bench_batch(args, prec) -> int:
    """64 uneven asteroids in one batch: per step, the `--window` trial
    periods nearest each asteroid's true period x 10 poles, all asteroids'
    fits through the same slots and work queue (ps_fit_batch_multi).
    Fit latency is measured per fit on the device clock, from the fit
    taking its slot to its summary being written (ps_fit_out.latency_us)."""
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
    lat, fits = [], 0
    t0 = time.perf_counter()
    for s_ in range(args.steps):
        out = eng.fit_batch_multi(*items(args.warmup + s_), conw)
        lat += list(out["latency_us"] * 1e-3)
        fits += len(out)
    wall = time.perf_counter() - t0
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
                           "note": "device wall clock per fit, slot taken to summary written"},
        "shape_setup_ms": {"host": round((t_host - t_parse) * 1e3, 2), "gpu": round((t_gpu - t_read) * 1e3, 2),
                           "asteroids": 64, "shapes": len({(i.problem.n_fac, i.problem.lmax) for i in inps})},
    }
    print(json.dumps(res), flush=True)
    eng.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("c2", "wu", "batch64"), default="c2")
    ap.add_argument("--periods", type=int, default=0,
                    help="trial periods per step of the plain FP64 and mixed legs (0: 256 for c2, 2048 for wu)")
    ap.add_argument("--prec", choices=PREC_NAMES, default="bf16a",
                    help="precision of the mixed leg and of the learned form's screening fits (the judge is FP64)")
    ap.add_argument("--form", choices=("plain", "learned"), default="learned")
    ap.add_argument("--candidates", type=int, default=8, help="learned form: periods judged per block")
    ap.add_argument("--halo", type=int, default=1, help="learned form: neighbours judged on either side")
    ap.add_argument("--block", type=int, default=0, help="learned form: trial periods per block (0: 2240 / 4096)")
    ap.add_argument("--slots", type=int, default=0, help="resident fits (0 = engine default)")
    ap.add_argument("--window", type=int, default=16, help="batch64: trial periods per asteroid per step")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the matched-chi2 check")
    ap.add_argument("--no-profile", action="store_true", help="skip the profiled pass (no roofline)")
    ap.add_argument("--only-headline", action="store_true", help="time only the headline form (profiler runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (counter calibration) and exit")
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

    import numpy as np

    from periodsearch_amd import PS_PREC_FP64, Engine, Input, synth
    from periodsearch_amd.proposal import Proposal, scan_learned

    prec = PREC_NAMES.index(args.prec)
    if args.workload == "batch64":
        return bench_batch(args, prec)

    tmp = tempfile.mkdtemp(prefix="psbench_")
    path = os.path.join(tmp, "period_search_in")
    ast = synth.write_workload(args.workload, path)
    inp = Input(path)
    conw_r = inp.conw_r0
    P = args.periods or (256 if args.workload == "c2" else 2048)
    block = args.block or (2240 if args.workload == "c2" else 4096)
    learned = args.form == "learned"
    n_true = int(round((inp.freq_start - 1.0 / ast["period_h"]) / inp.freq_step)) + 1
    # learned steps: the whole work unit (c2), or one block around the true period (wu)
    span = inp.n_periods if args.workload == "c2" else block
    judge_cap = 10 * args.candidates * (2 * args.halo + 1)

    eng = Engine(0)
    eng.set_precision(prec)
    if args.slots:
        eng.set_slots(args.slots)
    elif learned:
        eng.set_slots(block)
    eng.set_problem(inp.problem, keep=inp)
    e64 = Engine(0)
    e64.set_precision(PS_PREC_FP64)
    e64.set_problem(inp.problem, keep=inp)
    nc, ma, mfit = eng.dims()
    N, F = inp.ndata, inp.problem.n_fac
    prop = Proposal.load() if learned else None

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    def time_leg(fn, steps, warmup, name):
        for k in range(warmup):
            fn(k)
        t0 = time.perf_counter()
        work = 0
        for k in range(steps):
            work += fn(warmup + k)
        dt = time.perf_counter() - t0
        log(f"{name}: {work} LM fits in {dt:.2f} s")
        return work, dt

    def plain_step(e):
        def fn(k):
            n0 = 1 + (k * P) % max(1, inp.n_periods - P)
            e.scan(inp.freq_start, inp.freq_step, n0, n0 + P - 1, conw_r)
            return 10 * P
        return fn

    last = {}

    def learned_step(k):
        # the whole range (c2), or the block of trial periods around the true
        # period (wu), the same in every step
        n0 = 1 if args.workload == "c2" else max(1, min(n_true - span // 2, inp.n_periods - span + 1))
        n1 = min(inp.n_periods, n0 + span - 1)
        rows, judged, nf, _ = scan_learned(eng, inp, prop, conw_r, n0, n1, n_candidates=args.candidates,
                                           halo=args.halo, block=block, judge=e64)
        last.update(rows=rows, judged=judged, n0=n0, n1=n1, nf=nf)
        return nf

    out = {}
    if not args.only_headline:
        fits64, wall64 = time_leg(plain_step(e64), args.steps, args.warmup, "plain fp64 leg")
        out["fp64_fits_per_s"] = round(fits64 / wall64, 2)
        if args.prec != "fp64":
            fitsm, wallm = time_leg(plain_step(eng), args.steps, args.warmup, f"plain {args.prec} leg")
            out["mixed_fits_per_s"] = round(fitsm / wallm, 2)
    if learned:
        lm_fits, wall = time_leg(learned_step, args.steps, args.warmup, f"learned {args.prec} leg (headline)")
        periods = args.steps * (last["n1"] - last["n0"] + 1)
        value = lm_fits / wall
    else:
        lm_fits, wall = time_leg(plain_step(eng), args.steps, args.warmup, f"plain {args.prec} leg (headline)")
        periods = args.steps * P
        value = lm_fits / wall
    out["periods_per_s"] = round(periods / wall, 2)
    out["plain_equivalent_fits_per_s"] = round(10 * periods / wall, 2)
    if "fp64_fits_per_s" in out:
        r = out["fp64_fits_per_s"]
        out["vs_fp64"] = {"lm_fits_per_s": round(value / r, 3),
                          "plain_equivalent_fits_per_s": round(10 * periods / wall / r, 3)}
        if "mixed_fits_per_s" in out:
            out["vs_fp64"]["mixed_alone"] = round(out["mixed_fits_per_s"] / r, 3)

    # profiled pass of the headline form (outside the timed region): the
    # model kernel per launch, timed with HIP events between plain launches
    # (this runtime cannot time events captured into a graph)
    st = None
    if not args.no_profile:
        eng.set_profiling(True)
        (learned_step if learned else plain_step(eng))(0)
        st = eng.kernel_stats()
        eng.set_profiling(False)
        log("profiled pass done")

    def matched_check() -> dict:
        """The best judged row of the headline scan against the plain FP64
        scan's best row over the trial periods around the true period (and
        wherever the judged best lies)."""
        W = 256
        n0 = max(1, n_true - W // 2)
        n1 = min(inp.n_periods, n0 + W - 1)
        if learned:
            rows, judged = last["rows"], last["judged"]
            kk = np.flatnonzero(judged)
            k = int(kk[np.nanargmin(rows["chisq"][kk])])
            n_best = last["n0"] + k
            row = rows[k]
        else:
            r = eng.scan(inp.freq_start, inp.freq_step, n0, n1, conw_r)
            k = int(np.nanargmin(r["chisq"]))
            n_best = n0 + k
            row = r[k]
        lo, hi = min(n0, n_best - 8), max(n1, n_best + 8)
        r64 = e64.scan(inp.freq_start, inp.freq_step, lo, hi, conw_r)
        k64 = int(np.nanargmin(r64["chisq"]))
        step_h = float(r64["period_h"][k64]) ** 2 * inp.freq_step
        return {"window": f"trial periods {lo}..{hi} around the true period {ast['period_h']} h",
                "best_trial_period": n_best, "best_period_h": round(float(row["period_h"]), 8),
                "chi2": float(row["chisq"]), "fp64_best_trial_period": lo + k64,
                "fp64_best_period_h": round(float(r64["period_h"][k64]), 8), "fp64_chi2": float(r64["chisq"][k64]),
                "period_steps_apart": round(abs(float(row["period_h"]) - float(r64["period_h"][k64])) / step_h, 3),
                "chi2_rel": abs(float(row["chisq"]) - float(r64["chisq"][k64])) / float(r64["chisq"][k64])}

    flop_eval = 2.0 * N * F * nc
    peak = {"fp64": FP64_PEAK_TFLOPS, "fp32": FP32_PEAK_TFLOPS}.get(args.prec, BF16_PEAK_TFLOPS)
    res = {
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
        "dtype": DTYPE[args.prec],
        "data": "synthetic",
        "config": {"workload": WORKLOAD_TEXT[args.workload], "precision": args.prec,
                   "form": FORM[args.prec] + (" + learned proposal, FP64 judge" if learned else ""),
                   "value": ("LM fits actually run per second (screening fits + FP64 judge fits)" if learned
                             else "LM fits per second")},
    }
    res.update(out)
    res["lm_fits_per_s"] = round(value, 2)
    if learned:
        res["config"]["learned"] = (f"per step one learned scan of trial periods {last['n0']}..{last['n1']} in "
                                    f"blocks of {block}: 1 proposed start per period; the {args.candidates} "
                                    f"lowest-rms periods per block and {args.halo} neighbours each refitted from the "
                                    f"10 plain poles in FP64 on a second engine, overlapped with the next block")
        res["config"]["lm_fits_per_step"] = int(lm_fits / args.steps)
        res["config"]["judged_rows_per_step"] = int(last["judged"].sum())
    else:
        res["config"]["periods_per_step"] = P
    res["config"]["plain_legs"] = f"{P} trial periods x 10 poles per step, {args.steps} steps after {args.warmup} warm-up"
    if st and st["model_ms"] > 0:
        ms_launch = st["model_ms"] / max(st["model_launches"], 1)
        evals_launch = st["evals"] / max(st["model_launches"], 1)
        achieved = evals_launch * flop_eval / (ms_launch * 1e-3) / 1e12
        vec = evals_launch * FACET_FLOP_PER_PAIR * N * F / (ms_launch * 1e-3) / 1e12
        vpeak = FP64_PEAK_TFLOPS if args.prec == "fp64" else FP32_VECTOR_PEAK_TFLOPS
        tb = MODEL_HBM_BYTES_PER_EVAL.get((args.workload, args.prec))
        res["roofline"] = {
            "bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": round(tb * evals_launch) if tb else None,
            "kernel": {"fp64": "model_f64_kernel", "fp32": "model_mixedw_kernel", "bf16": "model_mixedw_kernel",
                       "bf16a": "model_mixedw_kernel"}.get(
                args.prec, "model_mixed_kernel"),
            "ms_per_launch": round(ms_launch, 4), "evals_per_launch": round(evals_launch, 1),
            "launches": int(st["model_launches"]),
            "flop_per_eval": flop_eval,
            "timing": "HIP events around each model launch, in a profiled pass after the timed regions (plain "
                      "launches: events inside a HIP graph cannot be timed)",
            "vector": {"stage": "facet stage, %d flop per point-facet pair" % FACET_FLOP_PER_PAIR,
                       "achieved": round(vec, 3), "peak": vpeak, "unit": "TFLOP/s", "frac": round(vec / vpeak, 4),
                       "peak_without_packed_fp32": FP32_VECTOR_UNPACKED_TFLOPS if args.prec != "fp64" else None},
        }
    if not args.no_check and (learned or args.prec != "fp64"):
        res["matched_chi2"] = matched_check()
        log("matched-chi2 check done")
    if not args.no_cpu:
        n_cpu = min(HOST_CPU_SHARE, os.cpu_count() or 1)
        try:
            n_cpu = min(n_cpu, len(os.sched_getaffinity(0)))
        except AttributeError:
            pass
        log(f"cpu baseline on {n_cpu} cores")
        res["cpu_baseline"] = cpu_baseline(path, conw_r, n_cpu, 1 if args.workload == "c2" else 10)
        log("cpu baseline done")
        if res["cpu_baseline"]:
            cb = res["cpu_baseline"]["value"]
            res["vs_cpu"] = {"lm_fits_per_s": round(value / cb, 1),
                             "plain_equivalent_fits_per_s": round(10 * periods / wall / cb, 1)}
            if "fp64_fits_per_s" in out:
                res["vs_cpu"]["fp64"] = round(out["fp64_fits_per_s"] / cb, 1)
    print(json.dumps(res), flush=True)
    if prop:
        prop.close()
    eng.close()
    e64.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
