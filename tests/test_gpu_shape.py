"""GPU shape setup for many asteroids (ps_shape_basis_many) against the host
shape basis, which is bit-identical to the reference's (test_host.py,
test_oracle_ref.py).

The device follows the host's operation order with FMA contraction off; the
only source of difference is the device libm (sin, cos, acos, atan2, log;
<= 1 ulp from glibc).  The facet tables are smooth functions of those
values, so they agree to a few ulp of their largest entry.  The ellipsoid
coefficients are a least-squares solution whose normal matrix has
condition number up to 3e18 (lmax 10, unnormalised Legendre functions):
ulp-level input differences move the poorly determined coefficients by
much more than an ulp, while the fitted log-curvature (Dsph . cg, what the
coefficients are used for) and the size coefficient c00 stay within about
1e-15 relative (measured on the CPU by perturbing Dsph and the log-curvature
by one ulp: at most 1.5e-15 over these shapes); the test allows 1e-11.  The
end-to-end test checks that fits from the GPU-built tables match fits from
the host-built ones.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (nrows, lmax, mmax): the smallest grid, lmax 0, mmax < lmax, the three
# configuration shapes, the largest facet count with the largest degree
SHAPES = [(1, 0, 0), (2, 2, 1), (3, 3, 0), (6, 4, 4), (8, 6, 6), (10, 8, 8), (11, 10, 3), (11, 10, 10)]
TABLES = ("nor", "darea", "pleg", "fc", "fs", "dsph")


def _specs():
    rng = np.random.default_rng(3)
    specs = []
    for nrows, lmax, mmax in SHAPES:
        for _ in range(2):  # two asteroids per shape: shared tables, own coefficients
            e = float(rng.uniform(0.3, 2.0))
            specs.append((nrows, lmax, mmax, 1.05 * e, 1.00 * e, 0.95 * e))
    order = rng.permutation(len(specs))
    return [specs[i] for i in order]


def test_gpu_shape_tables_match_host():
    import periodsearch_amd as ps

    specs = _specs()
    eng = ps.Engine(0)
    gpu = eng.shape_basis_many(specs)
    eng.close()
    n_bitwise = n_total = 0
    for spec, g in zip(specs, gpu):
        h = ps.shape_basis(*spec)
        for k in TABLES:
            scale = max(1.0, float(np.abs(h[k]).max()))
            err = float(np.abs(g[k] - h[k]).max()) / scale
            assert err <= 1e-13, (spec, k, err)
            n_bitwise += int((g[k] == h[k]).sum())
            n_total += h[k].size
        D = h["dsph"]
        fit_h, fit_g = D @ h["cg_first"], D @ g["cg_first"]
        rel = float(np.abs(fit_g - fit_h).max() / np.abs(fit_h).max())
        assert rel <= 1e-11, (spec, rel)
        # the size coefficient c00 is well determined
        assert abs(g["cg_first"][0] - h["cg_first"][0]) <= 1e-11 * abs(h["cg_first"][0]), spec
    print(f"{len(specs)} asteroids, {len(SHAPES)} shapes: {n_bitwise / n_total:.4f} of table entries bitwise equal")


def test_gpu_shape_many_asteroids_in_one_call_equals_one_by_one():
    """The batch shares tables per shape; each asteroid alone must give the
    same bits (the per-asteroid arithmetic does not depend on the batch)."""
    import periodsearch_amd as ps

    specs = _specs()
    eng = ps.Engine(0)
    together = eng.shape_basis_many(specs)
    for spec, t in zip(specs[:6], together[:6]):
        alone = eng.shape_basis_many([spec])[0]
        for k in TABLES + ("cg_first",):
            assert np.array_equal(alone[k], t[k]), (spec, k)
    eng.close()


def test_gpu_shape_rejects_bad_jobs():
    import periodsearch_amd as ps

    eng = ps.Engine(0)
    for bad in [(0, 2, 2, 1, 1, 1), (4, 2, 3, 1, 1, 1), (4, 11, 2, 1, 1, 1), (12, 2, 2, 1, 1, 1)]:
        with pytest.raises(ps.PsError):
            eng.shape_basis_many([bad])
    assert eng.shape_basis_many([]) == []
    eng.close()


def test_gpu_shape_tables_give_the_same_fits(small_incl_input, tmp_path):
    """Ten fits from GPU-built tables against ten fits from the host-built
    ones, on one asteroid: the same LM trajectories up to the last-bit
    differences of the tables."""
    from periodsearch_amd import Engine, Input, synth

    host = Input(small_incl_input)
    lazy = Input(small_incl_input, defer_shape=True)
    # a second, different asteroid in the same setup call
    other_path = os.path.join(str(tmp_path), "other")
    ast = synth.make_asteroid(**synth.uneven_asteroid(1003))
    synth.write_input(other_path, ast, 5.0, 5.1)
    other = Input(other_path, defer_shape=True)
    eng = Engine(0)
    eng.shape_basis_inputs([lazy, other])
    freq = host.freq_start
    res = []
    for inp in (host, lazy):
        eng.set_problem(inp.problem, keep=inp)
        res.append(eng.fit_batch([freq] * 10, list(range(10)), inp.conw_r0))
    eng.close()
    a, b = res
    rel = np.abs(a["chisq"] - b["chisq"]) / np.abs(a["chisq"])
    print(f"chi2 relative difference: median {np.median(rel):.1e} max {rel.max():.1e}")
    assert np.median(rel) < 1e-9 and rel.max() < 1e-5
    assert np.all(np.abs(a["prd"] - b["prd"]) / a["prd"] < 1e-7)
