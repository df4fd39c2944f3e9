"""CPU baseline: the reference's own fit loop, timed on a slice of trial periods.

Test infrastructure, run only by bench.py's CPU leg.  The arithmetic is the
reference's period_search/*.c, compiled unchanged into oracle/_ref/libpsref.so
(g++ -O3, as period_search/Makefile:15-18).  ref_harness.cpp's ref_fit restates
one fit of the driver loop (period_search_BOINC.cpp:819-911) around the
reference's mrqmin/mrqcof/bright/curv.  The oracle parses the input, prepares
the round and builds the shape basis; its tables are copied into the
reference's globals (oracle.ref_load_problem).  tests/test_oracle_ref.py shows
that these tables and whole fits are bit-identical between the two.

usage: python -m oracle.ref_bench <period_search_in> <conw_r> <n_from> <n_to> [n_poles]

For trial periods n_from..n_to (1-based, the reference's index n,
period_search_BOINC.cpp:807-812), runs the ten-pole fits
(period_search_BOINC.cpp:197-206, 819-845) in the reference's order, or the
first n_poles of them (a bounded sample where one fit takes many seconds:
under a stop condition > 1 every fit of an asteroid costs the same).  Prints
one JSON line: fits run, and the wall clock (time.time()) at the first fit's
start and after the last fit's end.
"""
from __future__ import annotations

import json
import sys
import time

import numpy as np

import oracle

POLE_BETA = [0, 0, 0, 0, 60, 60, 60, -60, -60, -60]      # period_search_BOINC.cpp:197-206
POLE_LAMBDA = [0, 90, 180, 270, 60, 180, 300, 60, 180, 300]


def main(argv) -> int:
    path, conw_r, n_from, n_to = argv[0], float(argv[1]), int(argv[2]), int(argv[3])
    n_poles = int(argv[4]) if len(argv) > 4 else 10
    ref = oracle.load_ref()
    if ref is None:
        print(json.dumps({"error": "oracle/_ref not built"}), flush=True)
        return 2
    o = oracle.Oracle(path)
    o.prepare(conw_r)
    oracle.ref_load_problem(ref, o)
    cgf = np.ascontiguousarray(o.darr("cg_first", (oracle.NP,)).copy())
    ref.ref_set_cg_first(oracle._d(cgf))
    par = np.ascontiguousarray(o.darr("par", (4,)).copy())
    iap = np.ascontiguousarray(o.iarr("ia_par", (4,)).copy())
    cl, stop, nmin = o.get("cl"), o.get("stop_condition"), o.iget("n_iter_min_in")
    f0, df = o.get("freq_start"), o.get("freq_step")
    r = oracle.FitResult()
    fits = 0
    t0 = time.time()
    for n in range(n_from, n_to + 1):
        freq = f0 - (n - 1) * df
        for m in range(n_poles):
            ref.ref_fit(freq, POLE_BETA[m], POLE_LAMBDA[m], oracle._d(par), oracle._i(iap), cl, stop, nmin, conw_r, r)
            fits += 1
    t1 = time.time()
    print(json.dumps({"fits": fits, "t0": t0, "t1": t1, "chisq_last": r.chisq}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
