"""CPU oracle for the proposal network and its features (test infrastructure).

Only the tests use this module: it is the checker for the GPU network of
periodsearch_amd/csrc/ps_mlp.hip and the feature kernel of
periodsearch_amd/csrc/ps_engine.hip (lc_features).  Neither has a
counterpart in the reference, which starts every fit from fixed poles
(period_search/period_search_BOINC.cpp:819-844).

Network: x -> relu(x W1 + b1) -> relu(. W2 + b2) -> . W3 + b3, parameters
flat as W1 [n_in][H], b1 [H], W2 [H][H], b2 [H], W3 [H][n_out], b3 [n_out];
loss (1/B) sum_b sum_k w_k (y_bk - t_bk)^2.  Gradients by the chain rule in
float64; the CPU tests pin them against central finite differences.
"""
from __future__ import annotations

import numpy as np


def unpack(p: np.ndarray, n_in: int, h: int, n_out: int):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    out = []
    for shape in ((n_in, h), (h,), (h, h), (h,), (h, n_out), (n_out,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    assert o == p.size
    return out


def forward(p, x, n_in, h, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, h, n_out)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ W1 + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ W3 + b3
    return y, (x, h1, h2)


def loss_and_grad(p, x, t, w, n_in, h, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, h, n_out)
    y, (x, h1, h2) = forward(p, x, n_in, h, n_out)
    B = x.shape[0]
    d = y - np.asarray(t, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    loss = float((w * d * d).sum() / B)
    dy = 2.0 * w * d / B
    gW3 = h2.T @ dy
    gb3 = dy.sum(0)
    d2 = (dy @ W3.T) * (h2 > 0)
    gW2 = h1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (h1 > 0)
    gW1 = x.T @ d1
    gb1 = d1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, g


def features(tim, ynorm, freq, nbins, c_alpha, c_aspect):
    """Phase-folded features at trial frequencies freq (1/h): per bin the mean
    of ynorm over the points whose phase frac(24 freq t) falls in it."""
    out = np.zeros((len(freq), nbins + 2))
    tim = np.asarray(tim, dtype=np.float64)
    for i, f in enumerate(freq):
        ph = tim * (24.0 * f)
        fr = ph - np.floor(ph)
        b = np.minimum((fr * nbins).astype(int), nbins - 1)
        s = np.bincount(b, weights=ynorm, minlength=nbins)
        c = np.bincount(b, minlength=nbins)
        out[i, :nbins] = np.where(c > 0, s / np.maximum(c, 1), 0.0)
        out[i, nbins] = c_alpha
        out[i, nbins + 1] = c_aspect
    return out
