"""Synthetic asteroids written as ordinary period_search_in files.

A convex asteroid with a known pole, period and shape is observed from the
Earth over several years; its relative brightness comes from the same
scattering model the fit uses (Lommel-Seeliger + Lambert, exponential-linear
phase function; period_search/bright.c, phasec.c) with Gaussian noise.  The
geometry is a circular 2.5 AU asteroid orbit and a 1 AU Earth orbit, both in
the ecliptic.  Dense light curves are runs of points at a fixed cadence;
sparse points are single epochs grouped into light curves of at most
``sparse_lc_max`` points.  All light curves are relative (flag 0).

The file header carries the reference's default fit settings
(period_search/period_search_in:1-14).  Named workloads:

  small  400 points (4x60 dense + 160 sparse), lmax=mmax=6, nrows=6, 7.0-7.6 h, 8 yr
  wu     1080 points (6x80 dense + 600 sparse), lmax=mmax=6, nrows=8, 2-100 h, 15 yr
  big    5000 points (20x150 dense + 2000 sparse), lmax=mmax=8, nrows=10, 4-12 h
"""
from __future__ import annotations

import numpy as np

YEAR = 365.25
JD_EPOCH = 2455000.5


def n_coef(lmax: int, mmax: int) -> int:
    return sum(2 if m else 1 for m in range(mmax + 1) for _ in range(m, lmax + 1))


def coef_degrees(lmax: int, mmax: int) -> np.ndarray:
    """Degree l of each Laplace-series coefficient in the reference order
    (m = 0..mmax, l = m..lmax, cos then sin; period_search/curv.c:21-33)."""
    out = []
    for m in range(mmax + 1):
        for l in range(m, lmax + 1):
            out.append(l)
            if m:
                out.append(l)
    return np.array(out)


def orbit_vectors(jd: np.ndarray, a_ast: float = 2.5, phase0: float = 0.7, incl_deg: float = 0.0):
    """Asteroid-centric vectors to the Sun and to the Earth (AU).  The
    asteroid's circular orbit is inclined by incl_deg to the ecliptic (0 =
    coplanar; then every aspect lies in the ecliptic and the poles lambda and
    lambda+180 deg fit the data equally well)."""
    t = jd - JD_EPOCH
    th_e = 2 * np.pi * t / YEAR
    th_a = 2 * np.pi * t / (YEAR * a_ast ** 1.5) + phase0
    earth = np.stack([np.cos(th_e), np.sin(th_e), np.zeros_like(t)], 1)
    ci, si = np.cos(np.radians(incl_deg)), np.sin(np.radians(incl_deg))
    ast = a_ast * np.stack([np.cos(th_a), np.sin(th_a) * ci, np.sin(th_a) * si], 1)
    return -ast, earth - ast


def _rotation(beta_colat: float, lam: float):
    cb, sb, cl, sl = np.cos(beta_colat), np.sin(beta_colat), np.cos(lam), np.sin(lam)
    return np.array([[cb * cl, cb * sl, -sb], [-sl, cl, 0.0], [sb * cl, sb * sl, cb]])


def model_brightness(cg: np.ndarray, basis: dict, ee: np.ndarray, ee0: np.ndarray, t: np.ndarray,
                     beta_colat: float, lam: float, omega: float, phase=(0.5, 0.1, -0.5), cl=0.1, cls=1.0,
                     phi0=0.0) -> np.ndarray:
    """Disk-integrated brightness of the convex model (bright.c without derivatives)."""
    area = basis["darea"] * np.exp(basis["dsph"] @ cg)
    B = _rotation(beta_colat, lam)
    f = np.mod(omega * t + phi0, 2 * np.pi)
    c, s = np.cos(f), np.sin(f)
    R = np.zeros((len(t), 3, 3))
    R[:, 0, 0], R[:, 0, 1], R[:, 1, 0], R[:, 1, 1], R[:, 2, 2] = c, s, -s, c, 1.0
    T = R @ B
    e = np.einsum("pij,pj->pi", T, ee)
    e0 = np.einsum("pij,pj->pi", T, ee0)
    nor = basis["nor"]
    mu = e @ nor.T
    mu0 = e0 @ nor.T
    vis = (mu > 1e-8) & (mu0 > 1e-8)
    dn = np.where(vis, mu + mu0, 1.0)
    sc = np.where(vis, mu * mu0 * (cl + cls / dn), 0.0)
    br = sc @ area
    alpha = np.arccos(np.clip(np.sum(ee * ee0, 1), -1, 1))
    scale = 1 + phase[0] * np.exp(-alpha / phase[1]) + phase[2] * alpha
    return scale * br


def _observable(jd: np.ndarray, incl_deg: float = 0.0) -> np.ndarray:
    """Solar elongation of the asteroid seen from the Earth above 60 deg."""
    sun, earth = orbit_vectors(jd, incl_deg=incl_deg)
    to_ast = -earth
    to_sun = sun - earth
    cosel = np.sum(to_ast * to_sun, 1) / (np.linalg.norm(to_ast, axis=1) * np.linalg.norm(to_sun, axis=1))
    return cosel < 0.5


def _epochs(rng, n: int, span_days: float, incl_deg: float = 0.0) -> np.ndarray:
    out = []
    while len(out) < n:
        jd = JD_EPOCH + rng.uniform(0, span_days, size=4 * n)
        out.extend(jd[_observable(jd, incl_deg)].tolist())
    return np.array(out[:n])


def make_asteroid(seed: int, *, lmax: int, mmax: int, nrows: int, n_dense: int, dense_pts: int, n_sparse: int,
                  noise: float, span_years: float, period_h: float, pole=(120.0, 40.0), cadence_min: float = 5.0,
                  sparse_lc_max: int = 1000, incl_deg: float = 0.0) -> dict:
    """Observations of one synthetic asteroid (no file written)."""
    from . import shape_basis

    rng = np.random.default_rng(seed)
    basis = shape_basis(nrows, lmax, mmax, 1.0, 0.8, 0.6)
    deg = coef_degrees(lmax, mmax)
    # N(0, 0.3/(l+1)) in units of each basis function's rms over the sphere
    # (the Laplace basis is unnormalised, P_l^m grows like (2l-1)!!)
    rms = np.sqrt(np.mean(basis["dsph"] ** 2, axis=0))
    cg = basis["cg_first"] + rng.normal(0.0, 0.3 / (deg + 1)) / rms
    span = span_years * YEAR
    lcs = []
    for _ in range(n_dense):
        t0 = _epochs(rng, 1, span - 1.0, incl_deg)[0]
        lcs.append(t0 + np.arange(dense_pts) * cadence_min / 1440.0)
    if n_sparse:
        sp = np.sort(_epochs(rng, n_sparse, span, incl_deg))
        for k in range(0, n_sparse, sparse_lc_max):
            lcs.append(sp[k:k + sparse_lc_max])
    jd_all = np.concatenate(lcs)
    jd0 = float(int(jd_all.min()))
    lam, beta = np.radians(pole[0]), np.radians(90.0 - pole[1])
    omega = 24 * 2 * np.pi / period_h
    out_lcs = []
    for jd in lcs:
        sun, earth = orbit_vectors(jd, incl_deg=incl_deg)
        ee = earth / np.linalg.norm(earth, axis=1)[:, None]
        ee0 = sun / np.linalg.norm(sun, axis=1)[:, None]
        y = model_brightness(cg, basis, ee, ee0, jd - jd0, beta, lam, omega)
        y = y * (1 + noise * rng.normal(size=len(y)))
        out_lcs.append(dict(jd=jd, y=y, sun=sun, earth=earth))
    return dict(lcs=out_lcs, cg_true=cg, period_h=period_h, pole=pole, lmax=lmax, mmax=mmax, nrows=nrows)


def write_input(path: str, ast: dict, per_start: float, per_end: float, step_coef: float = 0.5,
                stop: float = 50, n_iter_min: int = 10, conw: float = 0.1) -> None:
    """period_search_in with the reference's default header (period_search/period_search_in:1-14);
    conw is the initial convexity weight (0.1 there)."""
    with open(path, "w") as f:
        f.write(f"{per_start!r} {step_coef!r} {per_end!r} 1\tperiod_start period_step period_end fixed/free\n")
        f.write("0.0\t\t\t\tinitial epoch\n")
        f.write("0.0\t\t\t\tinitial rotation angle\n")
        f.write(f"{conw!r}\t\t\t\tconvexity weight\n")
        f.write(f"{ast['lmax']} {ast['mmax']}\t\t\t\tdegree and order of spherical harmonics\n")
        f.write(f"{ast['nrows']}\t\t\t\tnumber of rows\n")
        f.write("0.5\t1\t\t\tscattering parameters\n0.1\t1\n-0.5\t1\n")
        f.write("0.1\t0\t\t\tLS coeff\n")
        f.write(f"{stop!r}\t\t\t\titeration stop condition\n")
        f.write(f"{n_iter_min}\t\t\t\tminimum number of iterations (only if the above value < 1)\n")
        f.write("5\t\t\t\tAlamda multiplicative factor\n")
        f.write("0.1\t\t\t\tAlamda initial value\n")
        f.write(f"{len(ast['lcs'])}\n")
        for lc in ast["lcs"]:
            f.write(f"{len(lc['jd'])} 0 synthetic\n")
            for j in range(len(lc["jd"])):
                s, e = lc["sun"][j], lc["earth"][j]
                f.write(f"{lc['jd'][j]:.6f} {lc['y'][j]:.10e} {s[0]:.10e} {s[1]:.10e} {s[2]:.10e} "
                        f"{e[0]:.10e} {e[1]:.10e} {e[2]:.10e}\n")
        for i in range(len(ast["lcs"])):
            f.write(f"{i + 1} 1\n")


WORKLOADS = {
    "small": dict(seed=0, lmax=6, mmax=6, nrows=6, n_dense=4, dense_pts=60, n_sparse=160, noise=0.02,
                  span_years=8, period_h=7.3, per_start=7.0, per_end=7.6),
    "wu": dict(seed=1, lmax=6, mmax=6, nrows=8, n_dense=6, dense_pts=80, n_sparse=600, noise=0.05,
               span_years=15, period_h=7.3, per_start=2.0, per_end=100.0),
    "big": dict(seed=2, lmax=8, mmax=8, nrows=10, n_dense=20, dense_pts=150, n_sparse=2000, noise=0.02,
                span_years=8, period_h=7.3, per_start=4.0, per_end=12.0),
    # the same asteroid observed over two apparitions 420 days apart, so that
    # the reference's step rule gives 6,720 trial periods over 4-12 h (one
    # apparition, 190 days, would give 3,037)
    "c2": dict(seed=2, lmax=8, mmax=8, nrows=10, n_dense=20, dense_pts=150, n_sparse=2000, noise=0.02,
               span_years=1.15, period_h=7.3, per_start=4.0, per_end=12.0),
}


def write_workload(name: str, path: str, **overrides) -> dict:
    cfg = dict(WORKLOADS[name])
    cfg.update(overrides)
    per_start, per_end = cfg.pop("per_start"), cfg.pop("per_end")
    per_start, per_end = overrides.get("per_start", per_start), overrides.get("per_end", per_end)
    cfg.pop("per_start", None)
    cfg.pop("per_end", None)
    stop = cfg.pop("stop", 50)
    ast = make_asteroid(cfg.pop("seed"), **cfg)
    write_input(path, ast, per_start, per_end, stop=stop)
    return ast


def uneven_asteroid(seed: int) -> dict:
    """Parameters of one asteroid of the uneven multi-asteroid batch: point
    count log-uniform in [200, 8000] (about 30% in dense light curves of
    20-150 points, the rest sparse), lmax = mmax uniform in {4, 6, 8} with
    nrows 6 / 8 / 10, period log-uniform in [2, 50] h, a random pole, 2-5%
    noise, 4-15 years, orbit inclination 0-15 deg."""
    rng = np.random.default_rng(seed)
    n = int(round(np.exp(rng.uniform(np.log(200), np.log(8000)))))
    lmax = int(rng.choice([4, 6, 8]))
    n_dense = int(min(20, max(1, n // 400)))
    dense_pts = int(min(150, max(20, 0.3 * n / n_dense)))
    return dict(seed=seed, lmax=lmax, mmax=lmax, nrows={4: 6, 6: 8, 8: 10}[lmax], n_dense=n_dense,
                dense_pts=dense_pts, n_sparse=max(0, n - n_dense * dense_pts),
                noise=float(rng.uniform(0.02, 0.05)), span_years=float(rng.uniform(4, 15)),
                period_h=float(np.exp(rng.uniform(np.log(2.0), np.log(50.0)))),
                pole=(float(rng.uniform(0, 360)), float(np.degrees(np.arcsin(rng.uniform(-1, 1))))),
                incl_deg=float(rng.uniform(0, 15)))
