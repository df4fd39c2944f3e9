"""Host-side amplitude-aspect features of the learned proposal (CPU)."""
import os

import numpy as np


def _input(tmp_path, **kw):
    from periodsearch_amd import Input, synth

    cfg = dict(lmax=4, mmax=4, nrows=6, n_dense=8, dense_pts=60, n_sparse=100, noise=0.01, span_years=8,
               period_h=6.0, incl_deg=5.0)
    cfg.update(kw)
    ast = synth.make_asteroid(kw.pop("seed", 3), **{k: v for k, v in cfg.items() if k != "seed"})
    path = os.path.join(str(tmp_path), "period_search_in")
    synth.write_input(path, ast, 5.9, 6.1)
    return Input(path)


def test_amplitude_harmonic_points_at_the_pole_longitude(tmp_path):
    """A pole in the ecliptic: the light-curve amplitude is largest when the
    asteroid is seen along the equator, 90 deg from the pole longitude, so the
    doubled-angle estimate 0.5 atan2(-s, -c) lands near the pole longitude
    (mod 180 deg).  Checked at four pole longitudes."""
    from periodsearch_amd.proposal import N_ASPECT, aspect_features

    for k, lam in enumerate((20.0, 65.0, 110.0, 155.0)):
        f = aspect_features(_input(tmp_path, pole=(lam, 0.0), seed=30 + k))
        assert f.shape == (N_ASPECT,) and np.isfinite(f).all()
        c, s, r, am, nd = f
        assert abs(r - np.hypot(c, s)) < 1e-6 and am > 0 and abs(nd - 0.8) < 1e-6
        est = np.degrees(0.5 * np.arctan2(-s, -c))
        err = abs((est - lam + 90.0) % 180.0 - 90.0)
        assert err < 25.0, (lam, est)


def test_no_dense_light_curves_give_zeros(tmp_path):
    from periodsearch_amd.proposal import aspect_features

    f = aspect_features(_input(tmp_path, n_dense=1, seed=5))
    assert np.array_equal(f, np.zeros_like(f))
