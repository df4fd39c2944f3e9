"""The oracle's fits in parallel CPU processes (test infrastructure).

Each worker is a fresh interpreter (spawn: no state of a GPU-using parent is
inherited) that reads one period_search_in into an oracle context and runs
the fits it is given.  Options per call: the stop condition and minimum
iteration count (as the input header would set them), a one-ulp upward
move of every time or brightness (the reference arithmetic's own
sensitivity, see test_oracle_loop.py), and the derivatives rounded to float
(what a Jacobian held in FP32 alone does to the reference arithmetic)."""
from __future__ import annotations

import multiprocessing as mp
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np

FIELDS = ("chisq", "dev", "prd", "la", "be", "niter", "dark", "totarea")


def _work(args):
    path, conw_r, stop, n_iter_min, perturb, chunk, dyda_f32 = args
    import oracle

    o = oracle.Oracle(path)
    if stop is not None:
        o.set("stop_condition", stop)
    if n_iter_min is not None:
        o.set("n_iter_min_in", n_iter_min)
    if dyda_f32:
        o.set("dyda_f32", 1)
    o.prepare(conw_r)
    if perturb:
        n = o.iget("ndata") - 3
        a = o.darr({"tim": "tim", "y": "brightness"}[perturb], (20004,))
        a[1:n + 1] = np.nextafter(a[1:n + 1], np.inf)
    out = []
    for freq, pole0 in chunk:
        r = o.fit(freq, pole0 + 1)
        out.append(tuple(float(getattr(r, k)) for k in FIELDS))
    return out


def workers() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def fits(path: str, conw_r: float, items, stop=None, n_iter_min=None, perturb=None, nproc=None,
         dyda_f32=False) -> np.ndarray:
    """items: (freq [1/h], 0-based pole) pairs.  Returns a structured array
    with FIELDS per item, in order.  dyda_f32: the oracle's derivatives
    rounded to float (a Jacobian held in FP32)."""
    items = list(items)
    nproc = nproc or workers()
    k = max(1, -(-len(items) // nproc))
    chunks = [items[i:i + k] for i in range(0, len(items), k)]
    ctx = mp.get_context("spawn")
    with ProcessPoolExecutor(max_workers=len(chunks), mp_context=ctx) as ex:
        res = list(ex.map(_work, [(path, conw_r, stop, n_iter_min, perturb, c, dyda_f32) for c in chunks]))
    flat = [r for part in res for r in part]
    dt = np.dtype([(k, "f8") for k in FIELDS])
    return np.array(flat, dtype=dt)


def ulp_envelope(path: str, conw_r: float, items, base=None, **kw) -> np.ndarray:
    """Per item: the largest relative chi2 change of the oracle's own fit when
    every time, or every brightness, moves up by one ulp."""
    if base is None:
        base = fits(path, conw_r, items, **kw)
    env = np.zeros(len(base))
    for p in ("tim", "y"):
        r = fits(path, conw_r, items, perturb=p, **kw)
        env = np.maximum(env, np.abs(r["chisq"] - base["chisq"]) / np.abs(base["chisq"]))
    return env
