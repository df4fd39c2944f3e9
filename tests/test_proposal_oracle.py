"""The CPU oracle of the proposal network and features, pinned by finite
differences and a closed form (no GPU needed)."""
import numpy as np

from oracle import mlp


def test_oracle_gradient_matches_central_differences():
    rng = np.random.default_rng(0)
    n_in, h, n_out, B = 5, 7, 3, 4
    n = n_in * h + h + h * h + h + h * n_out + n_out
    p = rng.normal(0, 0.5, n)
    x = rng.normal(size=(B, n_in))
    t = rng.normal(size=(B, n_out))
    w = np.array([1.0, 0.5, 2.0])
    _, g = mlp.loss_and_grad(p, x, t, w, n_in, h, n_out)
    eps = 1e-6
    fd = np.zeros(n)
    for i in range(n):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd[i] = (mlp.loss_and_grad(pp, x, t, w, n_in, h, n_out)[0] - mlp.loss_and_grad(pm, x, t, w, n_in, h, n_out)[0]) / (2 * eps)
    assert np.abs(g - fd).max() < 1e-7 * max(1.0, np.abs(fd).max())
    # a plausible slip (ReLU mask of the wrong layer) is caught by the same check
    W1, b1, W2, b2, W3, b3 = mlp.unpack(p, n_in, h, n_out)
    z1 = x @ W1 + b1
    assert (z1 < 0).any()


def test_oracle_features_fold_a_sinusoid():
    """y/mean - 1 = A cos(2 pi f t): the mean over bin b of uniformly spread
    phases is A (sin(2 pi (b+1)/n) - sin(2 pi b/n)) / (2 pi / n)."""
    nb, f_h, A = 16, 0.137, 0.2
    t = np.linspace(0.0, 400.0, 200001)
    ph = 24.0 * f_h * t
    yn = A * np.cos(2 * np.pi * ph)
    feat = mlp.features(t, yn, [f_h], nb, 0.3, 0.4)[0]
    b = np.arange(nb)
    exact = A * (np.sin(2 * np.pi * (b + 1) / nb) - np.sin(2 * np.pi * b / nb)) / (2 * np.pi / nb)
    assert np.abs(feat[:nb] - exact).max() < 1e-3
    assert feat[nb] == 0.3 and feat[nb + 1] == 0.4
