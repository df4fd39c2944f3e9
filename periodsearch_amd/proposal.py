"""Learned starting points for the trial-period fits.

The reference starts every trial period from the same ten poles and the
ellipsoid shape (period_search/period_search_BOINC.cpp:819-844) and keeps the
best of the ten fits.  This module adds a second way to scan:

  * a network (the GPU Mlp: 71 -> 256 -> 256 -> 12, ReLU, FP32) maps the
    phase-folded light curve at a trial period (Engine.lc_features: 64 bins,
    mean solar phase angle, mean |Earth latitude|) and the asteroid's
    amplitude-aspect features (aspect_features) to a pole and the l <= 2
    shape coefficients;
  * scan_learned fits every trial period from the proposal (and, if asked,
    its mirror pole): 1 or 2 LM fits instead of 10, each a full LM run under
    the problem's stop rule; then it refits the best candidates and their
    neighbours from the plain ten-pole start, reduced to rows as Engine.scan
    does.  The plain fits judge the result: the best period and its chi2 are
    taken from plain rows only.

Pole encoding.  The light curves of a pole (lambda, beta) and of
(lambda + 180, beta) are nearly the same for orbits close to the ecliptic, so
the network predicts the doubled-longitude vector (cos b cos 2l, cos b sin 2l,
sin b) and both longitudes are fitted.  Training targets are taken with
lambda folded into [-90, 90); folding a fit by 180 deg also turns the shape by
180 deg about the spin axis, which flips the sign of the m = 1 coefficients.

Training (train()): synthetic asteroids drawn like the multi-asteroid batch
(point counts, lmax, periods, poles and noise vary), the plain ten-pole fits
at the nine trial periods nearest the true one, the best pole of each as the
target, Adam on the GPU.  The result is stored with its feature and target
scalings in an .npz next to this file (proposal_net.npz).
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import sys
import tempfile
import time

import numpy as np

from . import FIT_DTYPE, ROW_DTYPE, Engine, Input, Mlp, PsLearnedOpts, PsProposalNet, lib, n_coef, synth

NBINS = 64
N_ASPECT = 5
N_IN = NBINS + 2 + N_ASPECT
HIDDEN = 256
N_OUT = 12  # 3 pole + 9 shape coefficients with l <= 2
DEFAULT_NET = os.path.join(os.path.dirname(os.path.abspath(__file__)), "proposal_net.npz")


def low_order_index(lmax: int, mmax: int):
    """Indices in the coefficient vector of (l, m) with l <= 2, in the order
    c00, c10, c20, c11 (cos, sin), c21 (cos, sin), c22 (cos, sin), and a mask
    of the m = 1 entries.  Coefficient order: for m = 0..mmax, l = m..lmax,
    one entry for m = 0, a (cos, sin) pair otherwise."""
    pos = {}
    k = 0
    for m in range(mmax + 1):
        for l in range(m, lmax + 1):
            if m == 0:
                pos[(l, m, 0)] = k
                k += 1
            else:
                pos[(l, m, 0)] = k
                pos[(l, m, 1)] = k + 1
                k += 2
    want = [(0, 0, 0), (1, 0, 0), (2, 0, 0), (1, 1, 0), (1, 1, 1), (2, 1, 0), (2, 1, 1), (2, 2, 0), (2, 2, 1)]
    idx = np.array([pos[w] for w in want])
    m1 = np.array([w[1] == 1 for w in want])
    return idx, m1


def aspect_features(inp: Input) -> np.ndarray:
    """Amplitude-aspect features of one asteroid (the same at every trial
    period), computed by the library (ps_aspect_features).

    A light curve's amplitude grows with the aspect angle between the spin
    axis and the line of sight.  For views near the ecliptic it varies with
    the viewing longitude lv as a0 - a1 cos 2(lv - lambda_pole), so the second
    harmonic of the amplitude against the viewing longitude points at the
    pole longitude modulo 180 deg (the network's doubled-longitude output).
    From the dense light curves (at least 10 points within one day): amplitude
    A = (max - min) / mean brightness and lv = longitude of the mean Earth
    direction; features [c, s] = sum (A - mean A) (cos 2lv, sin 2lv) /
    (count * mean A), their length, the mean amplitude, and the count of dense
    light curves / 10.  Zeros with fewer than two dense light curves."""
    out = np.zeros(N_ASPECT, np.float32)
    rc = lib.ps_aspect_features(C.byref(inp.problem), out.ctypes.data_as(C.POINTER(C.c_float)))
    if rc != 0:
        raise RuntimeError(f"ps_aspect_features failed: {rc}")
    return out


def features(eng: Engine, inp: Input, freq: np.ndarray) -> np.ndarray:
    """Network inputs at trial frequencies freq: the folded light curve
    (Engine.lc_features) and the asteroid's amplitude-aspect features."""
    x = eng.lc_features(freq, NBINS)
    a = aspect_features(inp)
    return np.concatenate([x, np.broadcast_to(a, (len(x), N_ASPECT))], axis=1).astype(np.float32)


def reference_starts(inp: Input, freq: np.ndarray) -> np.ndarray:
    """The reference's starting parameters (period_search_BOINC.cpp:819-844)
    for every (trial frequency, pole): [len(freq) * n_poles][ma]."""
    p = inp.problem
    nc = n_coef(p.lmax, p.mmax)
    ma = nc + 8
    arr = lambda ptr, n: np.ctypeslib.as_array(ptr, shape=(n,)).copy()
    cg_first = arr(p.cg_first, nc)
    bet, lam = arr(p.beta_pole, p.n_poles), arr(p.lambda_pole, p.n_poles)
    out = np.zeros((len(freq), p.n_poles, ma))
    out[:, :, :nc] = cg_first
    out[:, :, nc] = (np.pi / 180) * (90 - bet)
    out[:, :, nc + 1] = (np.pi / 180) * lam
    out[:, :, nc + 2] = (24 * 2 * np.pi / (1 / np.asarray(freq)))[:, None]
    out[:, :, nc + 3:nc + 6] = np.array(list(p.par))
    out[:, :, nc + 6] = np.log(p.cl)
    out[:, :, nc + 7] = 1.0
    return out.reshape(-1, ma)


def encode_target(a: np.ndarray, lmax: int, mmax: int) -> np.ndarray:
    """Target vector of a fitted parameter set a[ma]: doubled-longitude pole
    and the l <= 2 coefficients, with lambda folded into [-90, 90)."""
    nc = n_coef(lmax, mmax)
    idx, m1 = low_order_index(lmax, mmax)
    beta = np.pi / 2 - a[nc]
    lam = np.mod(a[nc + 1], 2 * np.pi)
    c = a[idx].copy()
    if np.pi / 2 <= lam < 3 * np.pi / 2:
        lam -= np.pi
        c[m1] = -c[m1]
    cb = np.cos(beta)
    return np.concatenate([[cb * np.cos(2 * lam), cb * np.sin(2 * lam), np.sin(beta)], c])


class Proposal:
    """A trained network with its input and target scalings."""

    def __init__(self, params: np.ndarray, x_mean, x_std, t_mean, t_std, max_batch: int = 4096, device: int = 0):
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        self.x_mean, self.x_std = f32(x_mean), f32(x_std)
        self.t_mean, self.t_std = f32(t_mean), f32(t_std)
        self.net = Mlp(N_IN, HIDDEN, N_OUT, max_batch=max_batch, device=device)
        self.net.set_params(params)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        self.c_net = PsProposalNet(self.net._h, N_IN, N_OUT, NBINS, max_batch, fp(self.x_mean), fp(self.x_std),
                                   fp(self.t_mean), fp(self.t_std))

    @classmethod
    def load(cls, path: str = DEFAULT_NET, **kw) -> "Proposal":
        z = np.load(path, allow_pickle=False)
        return cls(z["params"], z["x_mean"], z["x_std"], z["t_mean"], z["t_std"], **kw)

    def predict(self, feats: np.ndarray) -> np.ndarray:
        x = (np.asarray(feats, np.float32) - self.x_mean) / self.x_std
        out = np.zeros((x.shape[0], N_OUT), np.float32)
        mb = self.net.max_batch
        for i in range(0, x.shape[0], mb):
            out[i:i + mb] = self.net.forward(x[i:i + mb])
        return out * self.t_std + self.t_mean

    def starts(self, inp: Input, freq: np.ndarray, feats: np.ndarray) -> np.ndarray:
        """Two starting parameter sets per trial frequency: the proposed pole
        and its mirror (lambda + 180 deg, m = 1 coefficients negated), the
        proposed l <= 2 coefficients (c00, the size, stays the ellipsoid's)
        and the reference's values for everything else."""
        p = inp.problem
        nc = n_coef(p.lmax, p.mmax)
        idx, m1 = low_order_index(p.lmax, p.mmax)
        y = self.predict(feats).astype(np.float64)
        base = reference_starts(inp, freq).reshape(len(freq), p.n_poles, -1)[:, 0, :]
        v = y[:, :3]
        beta = np.arcsin(np.clip(v[:, 2] / np.maximum(np.linalg.norm(v, axis=1), 1e-12), -1, 1))
        lam = 0.5 * np.arctan2(v[:, 1], v[:, 0])
        out = np.repeat(base[:, None, :], 2, axis=1)
        for k in (0, 1):
            c = y[:, 3:].copy()
            if k == 1:
                c[:, m1] = -c[:, m1]
            out[:, k, idx[1:]] = c[:, 1:]
            out[:, k, nc] = np.pi / 2 - beta
            out[:, k, nc + 1] = np.mod(lam + k * np.pi, 2 * np.pi)
        return out.reshape(2 * len(freq), -1)

    def close(self):
        self.net.close()


def _validator_ok(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Per row pair, the project validator's acceptance rule
    (validator/period_search_validator4.cpp:88-103): |(x - y) / (x + y)| / 2
    at most 0.1 for the period and the rms, 0.5 for chi2."""
    def ok(x, y, tol):
        return np.abs((x - y) / (x + y)) / 2 <= tol
    return ok(a["period_h"], b["period_h"], 0.1) & ok(a["rms"], b["rms"], 0.1) & ok(a["chisq"], b["chisq"], 0.5)


def scan_learned(eng: Engine, inp: Input, prop: Proposal, conw_r: float, n_from: int, n_to: int,
                 n_candidates: int = 4, halo: int = 2, mirror: bool = False, block: int = 1024,
                 judge: Engine | None = None, refine: int = 8):
    """Rows n_from..n_to of the learned scan (ps_scan_learned): every trial
    period fitted from the proposed start (and, if asked, its mirror pole);
    per block of `block` periods the n_candidates periods of lowest rms and
    halo periods on either side of each are refitted from the reference's
    ten poles by `judge` (another engine with the same problem, e.g. in FP64,
    working while the next block is screened) or by eng; then the best
    judged row's unjudged neighbours are judged until it is a local minimum
    of the judged rows (at most `refine` rounds).  Returns (rows,
    judged mask, number of LM fits run, rows of the proposed starts alone);
    only judged rows are results in the reference's sense."""
    n = n_to - n_from + 1
    rows = np.zeros(n, dtype=ROW_DTYPE)
    screened = np.zeros(n, dtype=ROW_DTYPE)
    judged = np.zeros(n, dtype=np.int32)
    nf = C.c_longlong()
    opts = PsLearnedOpts(block, n_candidates, halo, 1 if mirror else 0, refine,
                         judge._h if judge is not None else None)
    eng._check(lib.ps_scan_learned(eng._h, C.byref(inp.problem), C.byref(prop.c_net), C.byref(opts),
                                   inp.freq_start, inp.freq_step, n_from, n_to, conw_r, rows.ctypes.data,
                                   judged.ctypes.data_as(C.POINTER(C.c_int)), screened.ctypes.data, C.byref(nf)),
               "ps_scan_learned")
    return rows, judged.astype(bool), nf.value, screened


# ---------------------------------------------------------------- training
def draw_asteroid(seed: int) -> dict:
    """Parameters of one training asteroid (uneven sizes, shapes, periods)."""
    rng = np.random.default_rng(seed)
    lmax = int(rng.choice([4, 6, 8]))
    return dict(seed=seed, lmax=lmax, mmax=lmax, nrows=6, n_dense=int(rng.integers(2, 9)),
                dense_pts=int(rng.integers(40, 121)), n_sparse=int(rng.integers(100, 1500)),
                noise=float(rng.uniform(0.02, 0.05)), span_years=float(rng.uniform(4, 15)),
                period_h=float(np.exp(rng.uniform(np.log(2.0), np.log(50.0)))),
                pole=(float(rng.uniform(0, 360)), float(np.degrees(np.arcsin(rng.uniform(-1, 1))))),
                incl_deg=float(rng.uniform(0, 15)))


def samples_for(eng: Engine, cfg: dict, tmpdir: str, n_near: int = 4):
    """Features and targets at the 2 n_near + 1 trial periods of the
    reference grid nearest the true period: plain ten-pole fits, best pole."""
    ast = synth.make_asteroid(**{k: v for k, v in cfg.items()})
    P = cfg["period_h"]
    path = os.path.join(tmpdir, "period_search_in")
    synth.write_input(path, ast, P * 0.98, P * 1.02)
    inp = Input(path)
    eng.set_problem(inp.problem, keep=inp)
    n_true = int(round((inp.freq_start - 1.0 / P) / inp.freq_step)) + 1
    n = np.arange(n_true - n_near, n_true + n_near + 1)
    n = n[(n >= 1) & (n <= inp.n_periods)]
    freq = inp.freq_start - (n - 1) * inp.freq_step
    fits, params = eng.fit_batch_from(reference_starts(inp, freq), inp.conw_r0, want_params=True)
    npol = inp.problem.n_poles
    x = features(eng, inp, freq)
    t = np.zeros((len(n), N_OUT))
    keep = np.zeros(len(n), dtype=bool)
    for i in range(len(n)):
        f = fits[i * npol:(i + 1) * npol]
        # the first strictly smaller dev wins and a NaN never does
        # (period_search_BOINC.cpp:913-924); a period whose pole fits all
        # diverged gives no target
        dev = f["dev"]
        if not np.isfinite(dev).any():
            continue
        b = int(np.nanargmin(np.where(np.isfinite(dev), dev, np.nan)))
        t[i] = encode_target(params[i * npol + b], cfg["lmax"], cfg["mmax"])
        keep[i] = True
    return x[keep], t[keep], len(fits)


def make_training_set(n_asteroids: int, seed0: int = 1_000_000, log=print):
    eng = Engine(0)
    xs, ts, nfit = [], [], 0
    t0 = time.time()
    with tempfile.TemporaryDirectory() as d:
        for i in range(n_asteroids):
            x, t, nf = samples_for(eng, draw_asteroid(seed0 + i), d)
            xs.append(x)
            ts.append(t)
            nfit += nf
            if (i + 1) % 100 == 0:
                log(f"[proposal] {i + 1}/{n_asteroids} asteroids, {nfit} fits, {time.time() - t0:.0f} s")
    eng.close()
    return np.concatenate(xs).astype(np.float32), np.concatenate(ts).astype(np.float32), nfit


def train(x: np.ndarray, t: np.ndarray, epochs: int = 200, batch: int = 512, lr: float = 1e-3, seed: int = 0,
          log=print):
    """Adam on the GPU network; returns (params, scalings, final loss)."""
    rng = np.random.default_rng(seed)
    x_mean, x_std = x.mean(0), x.std(0) + 1e-6
    t_mean, t_std = t.mean(0), t.std(0) + 1e-6
    t_mean[:3], t_std[:3] = 0.0, 1.0  # the pole vector is already O(1)
    xn = ((x - x_mean) / x_std).astype(np.float32)
    tn = ((t - t_mean) / t_std).astype(np.float32)
    w = np.ones(N_OUT, np.float32)
    w[:3] = 3.0  # the pole matters most for where the fit ends
    sizes = [(N_IN, HIDDEN), (HIDDEN,), (HIDDEN, HIDDEN), (HIDDEN,), (HIDDEN, N_OUT), (N_OUT,)]
    params = np.concatenate([rng.normal(0, np.sqrt(2.0 / s[0]), s).ravel() if len(s) == 2 else np.zeros(s)
                             for s in sizes]).astype(np.float32)
    net = Mlp(N_IN, HIDDEN, N_OUT, max_batch=batch)
    net.set_params(params)
    loss = 0.0
    for ep in range(epochs):
        perm = rng.permutation(len(xn))
        tot, nb = 0.0, 0
        for i in range(0, len(perm) - batch + 1, batch):
            b = perm[i:i + batch]
            tot += net.train_step(xn[b], tn[b], w, lr * (0.5 * (1 + np.cos(np.pi * ep / epochs))))
            nb += 1
        loss = tot / max(nb, 1)
        if ep % 20 == 0 or ep == epochs - 1:
            log(f"[proposal] epoch {ep} loss {loss:.4f}")
    params = net.get_params()
    net.close()
    return params, dict(x_mean=x_mean, x_std=x_std, t_mean=t_mean, t_std=t_std), loss


def evaluate(prop: Proposal, n_asteroids: int = 16, seed0: int = 5000, window: int = 200, cfg_over=None,
             mirror: bool = False, n_candidates: int = 4, halo: int = 2, stop: float = 1e-4, prec: int = 0,
             log=print) -> dict:
    """Held-out asteroids (seeds seed0..), each on a window of trial periods
    around its true period:
      * the plain FP64 scan (ten poles per period) and scan_learned with
        screening fits in precision `prec` and the judge fits in FP64 (for
        prec = FP64 one engine does both); whether the best judged row is the
        plain scan's best row (same trial period, and its chi2), the LM fits
        each scan ran, and how every screened row compares with the plain row
        of its period;
      * at the true period under a convergence stop rule (stop < 1): mean LM
        iterations per fit and the best chi2 from the ten plain starts and
        from the proposed starts."""
    e64 = Engine(0)
    eng = Engine(0) if prec else e64
    if prec:
        eng.set_precision(prec)
    res = []
    with tempfile.TemporaryDirectory() as d:
        for i in range(n_asteroids):
            cfg = draw_asteroid(seed0 + i)
            cfg.update(cfg_over or {})
            ast = synth.make_asteroid(**cfg)
            P = cfg["period_h"]
            path = os.path.join(d, "period_search_in")
            synth.write_input(path, ast, P * 0.97, P * 1.03)
            inp = Input(path)
            e64.set_problem(inp.problem, keep=inp)
            if prec:
                eng.set_problem(inp.problem, keep=inp)
            n_true = int(round((inp.freq_start - 1.0 / P) / inp.freq_step)) + 1
            n0, n1 = max(1, n_true - window // 2), min(inp.n_periods, n_true + window // 2)
            t0 = time.perf_counter()
            plain = e64.scan(inp.freq_start, inp.freq_step, n0, n1, inp.conw_r0)
            t_plain = time.perf_counter() - t0
            t0 = time.perf_counter()
            rows, ver, nf, learned = scan_learned(eng, inp, prop, inp.conw_r0, n0, n1, n_candidates=n_candidates,
                                                  halo=halo, mirror=mirror, judge=e64 if prec else None)
            t_learn = time.perf_counter() - t0
            kp = int(np.nanargmin(plain["chisq"]))
            kl = int(np.flatnonzero(ver)[np.nanargmin(rows["chisq"][ver])])
            rel = np.abs(learned["chisq"] - plain["chisq"]) / plain["chisq"]
            r = dict(seed=seed0 + i, n_periods=n1 - n0 + 1, best_plain=kp, best_learned=kl,
                     p_plain=float(plain["period_h"][kp]), p_learned=float(rows["period_h"][kl]),
                     chi2_plain=float(plain["chisq"][kp]), chi2_learned=float(rows["chisq"][kl]),
                     step_h=float(P * P * inp.freq_step), fits_plain=10 * (n1 - n0 + 1), fits_learned=nf,
                     t_plain=t_plain, t_learned=t_learn,
                     screen_rank=int(np.sum(learned["rms"] < learned["rms"][kp])),
                     screened_within_1e4=float(np.mean(rel <= 1e-4)),
                     screened_within_validator=float(np.mean(_validator_ok(learned, plain))))
            # iterations to convergence at the true period
            synth.write_input(path, ast, P * 0.97, P * 1.03, stop=stop)
            inp2 = Input(path)
            e64.set_problem(inp2.problem, keep=inp2)
            f = np.array([inp2.freq_start - (n_true - 1) * inp2.freq_step])
            fp = e64.fit_batch_from(reference_starts(inp2, f), inp2.conw_r0)
            fl = e64.fit_batch_from(prop.starts(inp2, f, features(e64, inp2, f)), inp2.conw_r0)
            r.update(iter_plain=float(fp["niter"].mean()), iter_learned=float(fl["niter"].mean()),
                     conv_chi2_plain=float(fp["chisq"].min()), conv_chi2_learned=float(fl["chisq"].min()))
            res.append(r)
            log(f"[proposal] eval {i}: P {P:.4f} h best row plain {kp} learned {kl} (screen rank {r['screen_rank']}), "
                f"chi2 {r['chi2_plain']:.6g} / {r['chi2_learned']:.6g}, fits {r['fits_plain']} / {nf}; "
                f"converged: iterations {r['iter_plain']:.1f} / {r['iter_learned']:.1f}, "
                f"chi2 {r['conv_chi2_plain']:.6g} / {r['conv_chi2_learned']:.6g}")
    if prec:
        eng.close()
    e64.close()
    return dict(runs=res,
                same_period=float(np.mean([abs(r["p_plain"] - r["p_learned"]) <= r["step_h"] for r in res])),
                same_row=float(np.mean([r["best_plain"] == r["best_learned"] for r in res])),
                screened_within_1e4=float(np.mean([r["screened_within_1e4"] for r in res])),
                screened_within_validator=float(np.mean([r["screened_within_validator"] for r in res])),
                chi2_rel_max=float(max(abs(r["chi2_learned"] - r["chi2_plain"]) / r["chi2_plain"] for r in res)),
                fit_ratio=float(sum(r["fits_plain"] for r in res) / sum(r["fits_learned"] for r in res)),
                time_ratio=float(sum(r["t_plain"] for r in res) / sum(r["t_learned"] for r in res)),
                iter_plain=float(np.mean([r["iter_plain"] for r in res])),
                iter_learned=float(np.mean([r["iter_learned"] for r in res])),
                conv_chi2_ratio=float(np.median([r["conv_chi2_learned"] / r["conv_chi2_plain"] for r in res])))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="train the proposal network on synthetic asteroids")
    ap.add_argument("--asteroids", type=int, default=2200)
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--out", default=DEFAULT_NET)
    ap.add_argument("--data", default="", help="also save the training set (.npz)")
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    x, t, nfit = make_training_set(a.asteroids, log=log)
    log(f"[proposal] training set: {len(x)} samples from {nfit} plain fits")
    if a.data:
        np.savez_compressed(a.data, x=x, t=t)
    params, sc, loss = train(x, t, epochs=a.epochs, log=log)
    np.savez(a.out, params=params, n_fits=nfit, n_samples=len(x), loss=loss, **sc)
    log(f"[proposal] saved {a.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
