"""Autotuner for the engine's launch configuration.

For a given problem and precision it times one batch of real fits (the
reference's starts at consecutive trial periods, full LM runs) under each
candidate (resident slots per refill round, slot groups on their own
streams) and keeps the fastest.  The tile shapes of the kernels (64 points x
32 facets, 16-column matrix blocks) are fixed at compile time; what varies
with the problem is how many fits are resident and how the per-slot kernels
overlap the model kernels.
"""
from __future__ import annotations

import itertools
import time

import numpy as np

SLOTS = (1024, 2048, 3072)
GROUPS = (1, 2, 3)


def autotune(eng, inp, conw_r: float, n_fits: int = 6144, slots=SLOTS, groups=GROUPS, n_from: int = 1,
             log=None) -> dict:
    """Times `n_fits` fits (n_fits/10 trial periods from n_from, 10 poles)
    under every (slots, groups) candidate on `eng` (problem already set),
    leaves the engine on the fastest and returns the timings."""
    n_per = max(1, n_fits // 10)
    n = np.arange(n_from, n_from + n_per)
    freq = np.repeat(inp.freq_start - (n - 1) * inp.freq_step, 10)
    pole = np.tile(np.arange(10), n_per)
    eng.fit_batch(freq[:10], pole[:10], conw_r)  # first-use costs out of the timings
    res = []
    for sl, gr in itertools.product(slots, groups):
        eng.set_slots(sl)
        eng.set_groups(gr)
        t0 = time.perf_counter()
        eng.fit_batch(freq, pole, conw_r)
        dt = time.perf_counter() - t0
        res.append(dict(slots=sl, groups=gr, fits_per_s=len(freq) / dt))
        if log:
            log(f"[autotune] slots {sl} groups {gr}: {len(freq) / dt:.1f} fits/s")
    best = max(res, key=lambda r: r["fits_per_s"])
    eng.set_slots(best["slots"])
    eng.set_groups(best["groups"])
    return dict(best=best, trials=res)
