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
import os
import sys
import tempfile
import time

import numpy as np

from . import FIT_DTYPE, ROW_DTYPE, Engine, Input, Mlp, n_coef, synth

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
    """Amplitude-aspect features of one asteroid (the same at every trial period).

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
    p = inp.problem
    n = p.n_points
    t = np.ctypeslib.as_array(p.tim, shape=(n,))
    y = np.ctypeslib.as_array(p.brightness, shape=(n,))
    ee = np.ctypeslib.as_array(p.ee, shape=(3 * n,)).reshape(n, 3)
    counts = np.ctypeslib.as_array(p.lc_points, shape=(p.n_lc,))
    amp, lon = [], []
    q = 0
    for k in counts:
        sl = slice(q, q + int(k))
        q += int(k)
        if k < 10 or np.ptp(t[sl]) > 1.0:
            continue
        m = y[sl].mean()
        if not m > 0:
            continue
        amp.append(np.ptp(y[sl]) / m)
        e = ee[sl].mean(0)
        lon.append(np.arctan2(e[1], e[0]))
    out = np.zeros(N_ASPECT, np.float32)
    if len(amp) >= 2:
        a, lv = np.array(amp), np.array(lon)
        am = a.mean()
        c = float(np.sum((a - am) * np.cos(2 * lv)) / (len(a) * am))
        s = float(np.sum((a - am) * np.sin(2 * lv)) / (len(a) * am))
        out[:] = [c, s, np.hypot(c, s), am, len(a) / 10.0]
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
        self.x_mean, self.x_std = np.asarray(x_mean, np.float32), np.asarray(x_std, np.float32)
        self.t_mean, self.t_std = np.asarray(t_mean, np.float32), np.asarray(t_std, np.float32)
        self.net = Mlp(N_IN, HIDDEN, N_OUT, max_batch=max_batch, device=device)
        self.net.set_params(params)

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


def _rows_from_fits(fits: np.ndarray, per: int, n_points: int) -> np.ndarray:
    """Best-of-`per` rows (the pole reduction of period_search_BOINC.cpp:900-924:
    the first fit with the smallest rms, lambda wrapped to [0, 360),
    chi2 = rms^2 (ndata - 3) with ndata = n_points + the 3 convexity rows)."""
    f = fits.reshape(-1, per)
    # the reference keeps the first fit whose dev is strictly below the best
    # so far, starting from 1e40: a NaN dev (a diverged fit) never wins, and
    # if no fit qualifies the row keeps dev 1e40 and zeros
    # (period_search_BOINC.cpp:913-924)
    dev = f["dev"]
    ok = dev < 1e40  # False for NaN
    key = np.where(ok, dev, np.inf)
    k = np.argmin(key, axis=1)
    any_ok = ok.any(axis=1)
    b = f[np.arange(len(f)), k]
    rows = np.zeros(len(f), dtype=ROW_DTYPE)
    dev_b = np.where(any_ok, b["dev"], 1e40)
    rows["period_h"] = np.where(any_ok, 24 * b["prd"], 0.0)
    rows["rms"] = dev_b
    rows["chisq"] = dev_b ** 2 * n_points
    with np.errstate(invalid="ignore", divide="ignore"):
        rows["dark"] = np.where(any_ok, b["dark"] / b["totarea"] * 100, 0.0)
    la = np.where(any_ok, b["la"], 0.0)
    rows["lambda"] = np.where(la < 0, la + 360, la)
    rows["beta"] = np.where(any_ok, b["be"], 0.0)
    return rows


def _pick_candidates(rms: np.ndarray, n_candidates: int, halo: int) -> np.ndarray:
    """Mask of the n_candidates lowest-rms rows (no two within halo) and
    halo rows on either side of each."""
    verified = np.zeros(len(rms), bool)
    picked = []
    for k in np.argsort(rms, kind="stable"):
        if len(picked) >= n_candidates:
            break
        if any(abs(int(k) - q) <= halo for q in picked):
            continue
        picked.append(int(k))
    for k in picked:
        verified[max(0, k - halo):min(len(rms), k + halo + 1)] = True
    return verified


def scan_learned(eng: Engine, inp: Input, prop: Proposal, conw_r: float, n_from: int, n_to: int,
                 n_candidates: int = 4, halo: int = 2, mirror: bool = False, block: int = 1024):
    """Rows n_from..n_to, in blocks of `block` trial periods: every period
    fitted from the proposed start (and, if asked, its mirror pole); in each
    block the n_candidates periods with the lowest rms and halo periods on
    either side of each are refitted from the reference's ten poles.  Those
    refits run in the same batch as the next block's proposed fits (one
    batch of LM fits per block instead of two).  Returns (rows, verified
    mask, number of LM fits run, rows of the proposed starts alone); only
    verified rows are judged results."""
    n = np.arange(n_from, n_to + 1)
    freq = inp.freq_start - (n - 1) * inp.freq_step
    per = 2 if mirror else 1
    npol = inp.problem.n_poles
    learned = np.zeros(len(n), dtype=ROW_DTYPE)
    rows = np.zeros(len(n), dtype=ROW_DTYPE)
    verified = np.zeros(len(n), bool)
    n_fits = 0
    pending = np.zeros(0, int)  # positions of the previous block awaiting their plain refits
    blocks = list(range(0, len(n), block))
    for bi in range(len(blocks) + 1):
        parts = []
        if bi < len(blocks):
            b0, b1 = blocks[bi], min(len(n), blocks[bi] + block)
            starts = prop.starts(inp, freq[b0:b1], features(eng, inp, freq[b0:b1]))
            parts.append(starts if mirror else starts[0::2])
        if len(pending):
            parts.append(reference_starts(inp, freq[pending]))
        if not parts:
            break
        fits = eng.fit_batch_from(np.concatenate(parts), conw_r)
        n_fits += len(fits)
        m = 0
        new_pending = np.zeros(0, int)
        if bi < len(blocks):
            m = (b1 - b0) * per
            learned[b0:b1] = _rows_from_fits(fits[:m], per, inp.ndata)
            rows[b0:b1] = learned[b0:b1]
            new_pending = b0 + np.flatnonzero(_pick_candidates(learned["rms"][b0:b1], n_candidates, halo))
        if len(pending):
            rows[pending] = _rows_from_fits(fits[m:], npol, inp.ndata)
            verified[pending] = True
        pending = new_pending
    return rows, verified, n_fits, learned


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
    for i in range(len(n)):
        f = fits[i * npol:(i + 1) * npol]
        b = int(np.argmin(f["dev"]))
        t[i] = encode_target(params[i * npol + b], cfg["lmax"], cfg["mmax"])
    return x, t, len(fits)


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
             mirror: bool = False, n_candidates: int = 4, stop: float = 1e-4, log=print) -> dict:
    """Held-out asteroids (seeds seed0..), each on a window of trial periods
    around its true period:
      * the plain scan (ten poles per period) and scan_learned; whether the
        best period (plain rows only, from both scans) agrees within one scan
        step, the chi2 of the two minima, the LM fits each scan ran;
      * at the true period under a convergence stop rule (stop < 1): mean LM
        iterations per fit and the best chi2 from the ten plain starts and
        from the proposed starts."""
    eng = Engine(0)
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
            eng.set_problem(inp.problem, keep=inp)
            n_true = int(round((inp.freq_start - 1.0 / P) / inp.freq_step)) + 1
            n0, n1 = max(1, n_true - window // 2), min(inp.n_periods, n_true + window // 2)
            t0 = time.perf_counter()
            plain = eng.scan(inp.freq_start, inp.freq_step, n0, n1, inp.conw_r0)
            t_plain = time.perf_counter() - t0
            t0 = time.perf_counter()
            rows, ver, nf, learned = scan_learned(eng, inp, prop, inp.conw_r0, n0, n1, n_candidates=n_candidates,
                                                  mirror=mirror)
            t_learn = time.perf_counter() - t0
            kp = int(np.argmin(plain["chisq"]))
            kl = int(np.flatnonzero(ver)[np.argmin(rows["chisq"][ver])])
            r = dict(seed=seed0 + i, n_periods=n1 - n0 + 1, best_plain=kp, best_learned=kl,
                     p_plain=float(plain["period_h"][kp]), p_learned=float(rows["period_h"][kl]),
                     chi2_plain=float(plain["chisq"][kp]), chi2_learned=float(rows["chisq"][kl]),
                     step_h=float(P * P * inp.freq_step), fits_plain=10 * (n1 - n0 + 1), fits_learned=nf,
                     t_plain=t_plain, t_learned=t_learn,
                     screen_rank=int(np.sum(learned["rms"] < learned["rms"][kp])))
            # iterations to convergence at the true period
            synth.write_input(path, ast, P * 0.97, P * 1.03, stop=stop)
            inp2 = Input(path)
            eng.set_problem(inp2.problem, keep=inp2)
            f = np.array([inp2.freq_start - (n_true - 1) * inp2.freq_step])
            fp = eng.fit_batch_from(reference_starts(inp2, f), inp2.conw_r0)
            fl = eng.fit_batch_from(prop.starts(inp2, f, features(eng, inp2, f)), inp2.conw_r0)
            r.update(iter_plain=float(fp["niter"].mean()), iter_learned=float(fl["niter"].mean()),
                     conv_chi2_plain=float(fp["chisq"].min()), conv_chi2_learned=float(fl["chisq"].min()))
            res.append(r)
            log(f"[proposal] eval {i}: P {P:.4f} h best row plain {kp} learned {kl} (screen rank {r['screen_rank']}), "
                f"chi2 {r['chi2_plain']:.6g} / {r['chi2_learned']:.6g}, fits {r['fits_plain']} / {nf}; "
                f"converged: iterations {r['iter_plain']:.1f} / {r['iter_learned']:.1f}, "
                f"chi2 {r['conv_chi2_plain']:.6g} / {r['conv_chi2_learned']:.6g}")
    eng.close()
    return dict(runs=res,
                same_period=float(np.mean([abs(r["p_plain"] - r["p_learned"]) <= r["step_h"] for r in res])),
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
