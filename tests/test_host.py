"""Host side of the product library: C-ABI exports, input parsing and the
shape basis, checked against the oracle (no GPU needed)."""
import re

import numpy as np
import pytest

import oracle
import periodsearch_amd as ps

ROW_FIELDS = ("period_h", "rms", "chisq", "dark", "lambda", "beta")


def test_every_declared_symbol_is_exported():
    inc = ps.os.path.join(ps.os.path.dirname(ps.HERE), "include")
    hdr = "".join(open(ps.os.path.join(inc, f)).read() for f in ("periodsearch.h", "ps_proposal.h"))
    declared = set(re.findall(r"^\w[\w \*]*?\b(ps_\w+)\(", hdr, re.M))
    assert declared == set(ps.EXPORTED)
    for name in declared:
        assert hasattr(ps.lib, name), name


def test_input_and_basis_match_oracle(small_input):
    inp = ps.Input(small_input)
    o = oracle.Oracle(small_input)
    o.prepare(inp.conw_r0)
    p = inp.problem
    F, nc = p.n_fac, ps.n_coef(p.lmax, p.mmax)
    assert F == o.iget("Numfac") and nc == o.iget("Ncoef")
    assert inp.n_periods == o.n_periods()
    assert inp.escl == o.get("escl") and inp.freq_step == o.get("freq_step")
    arr = lambda ptr, n: np.ctypeslib.as_array(ptr, shape=(n,))
    N = p.n_points
    assert np.array_equal(arr(p.tim, N), o.darr("tim", (20004,))[1:N + 1])
    assert np.array_equal(arr(p.sig, N), o.darr("sig", (20004,))[1:N + 1])
    assert np.array_equal(arr(p.ee, 3 * N).reshape(N, 3), o.darr("ee", (20004, 4))[1:N + 1, 1:4])
    assert np.array_equal(arr(p.ee0, 3 * N).reshape(N, 3), o.darr("ee0", (20004, 4))[1:N + 1, 1:4])
    assert np.array_equal(arr(p.nor, 3 * F).reshape(F, 3), o.darr("Nor", (1001, 4))[1:F + 1, 1:4])
    assert np.array_equal(arr(p.darea, F), o.darr("Darea", (1001,))[1:F + 1])
    assert np.array_equal(arr(p.dsph, F * nc).reshape(F, nc), o.darr("Dsph", (1001, 201))[1:F + 1, 1:nc + 1])
    assert np.array_equal(arr(p.cg_first, nc), o.darr("cg_first", (201,))[1:nc + 1])


def test_shape_basis_invariants():
    b = ps.shape_basis(8, 6, 6, 1.0, 1.0, 1.0)
    nor, da = b["nor"], b["darea"]
    assert nor.shape == (512, 3)
    assert np.allclose(np.linalg.norm(nor, axis=1), 1.0, atol=1e-15)
    # closed polyhedron: sum of area-weighted normals vanishes; total area -> 4 pi from below
    assert np.abs((da[:, None] * nor).sum(0)).max() < 1e-14
    assert 4 * np.pi * 0.98 < da.sum() < 4 * np.pi
    # a sphere has constant log-curvature: only c_00 is non-zero
    cg = b["cg_first"]
    assert abs(cg[0]) < 1e-12 and np.abs(cg[1:]).max() < 1e-12


def test_format_row_matches_reference_printf():
    row = np.zeros(1, dtype=ps.ROW_DTYPE)[0]
    row["period_h"], row["rms"], row["chisq"], row["dark"], row["lambda"], row["beta"] = (
        89.99905543123, 0.1083041, 5.8648571, 0.123, 176.4, 6.49)
    assert ps.format_row(row) == "89.99905543  0.108304  5.864857  0.1  176    6\n"
    row["lambda"] = 2.5
    assert ps.format_row(row).split()[4] == "2"          # printf rounds half to even
    assert ps.format_row(row, round_pole=True).split()[4] == "3"  # round() as the CUDA/OC apps


def test_deferred_input_hands_its_tables_to_the_shape_job(small_input):
    """ps_input_read_deferred leaves the tables zero and ps_input_shape_job
    points at them; filling them through the job's pointers gives the
    eagerly read input's tables (here with the host shape basis)."""
    eager = ps.Input(small_input)
    lazy = ps.Input(small_input, defer_shape=True)
    job = ps.PsShapeJob()
    assert ps.lib.ps_input_shape_job(lazy._h, ps.C.byref(job)) == 0
    p, q = eager.problem, lazy.problem
    F, nc = p.n_fac, ps.n_coef(p.lmax, p.mmax)
    assert (job.nrows, job.lmax, job.mmax) == (6, p.lmax, p.mmax) and 8 * job.nrows ** 2 == F
    assert (job.a, job.b, job.c) == (1.05 * eager.escl, 1.00 * eager.escl, 0.95 * eager.escl)
    arr = lambda ptr, n: np.ctypeslib.as_array(ptr, shape=(n,))
    assert ps.C.cast(job.dsph, ps.C.c_void_p).value == ps.C.cast(q.dsph, ps.C.c_void_p).value
    assert not arr(q.dsph, F * nc).any() and not arr(q.cg_first, nc).any()
    rc = ps.lib.ps_shape_basis(job.nrows, job.lmax, job.mmax, job.a, job.b, job.c, job.nor, job.darea, job.pleg,
                               job.fc, job.fs, job.dsph, job.cg_first)
    assert rc == F
    L, M = p.lmax + 1, p.mmax + 1
    for name, n in (("nor", 3 * F), ("darea", F), ("pleg", F * L * L), ("fc", F * M), ("fs", F * M),
                    ("dsph", F * nc), ("cg_first", nc)):
        assert np.array_equal(arr(getattr(q, name), n), arr(getattr(p, name), n)), name
    # everything else is the same input
    assert lazy.escl == eager.escl and lazy.n_periods == eager.n_periods
    assert np.array_equal(arr(q.tim, p.n_points), arr(p.tim, p.n_points))


def test_read_rows_and_best_row_over_work_units(golden_dir, tmp_path):
    """Post-scan selection: the reference-held sample output parses to its
    134 rows, and the least-chi^2 row over several outputs is the one numpy
    finds; the first of equal minima wins, NaN rows are skipped."""
    import os

    sample = os.path.join(golden_dir, "ps_sample_out")
    rows = ps.read_rows(sample)
    ref = np.loadtxt(sample)
    assert len(rows) == len(ref) == 134
    for k, name in enumerate(ROW_FIELDS):
        assert np.array_equal(rows[name], ref[:, k]), name
    assert ps.best_row(rows) == int(np.argmin(ref[:, 2]))
    # split into three "work units" and select over them
    paths = []
    for k, (a, b) in enumerate(((0, 40), (40, 90), (90, 134))):
        p = str(tmp_path / f"wu{k}")
        with open(p, "w") as f:
            for r in rows[a:b]:
                f.write(ps.format_row(r))
        paths.append(p)
    path, row = ps.best_period(paths)
    i = int(np.argmin(ref[:, 2]))
    assert path == paths[0 if i < 40 else 1 if i < 90 else 2] and row["chisq"] == ref[i, 2]
    # ties: first; NaN: skipped; nothing finite: error
    t = rows[:5].copy()
    t["chisq"] = [3.0, np.nan, 1.0, 1.0, 2.0]
    assert ps.best_row(t) == 2
    t["chisq"] = np.nan
    with pytest.raises(ps.PsError):
        ps.best_row(t)
    # malformed line and capacity
    bad = str(tmp_path / "bad")
    with open(bad, "w") as f:
        f.write("1.0 2.0 3.0 0.1 10 20\n1.0 2.0 x\n")
    with pytest.raises(ps.PsError):
        ps.read_rows(bad)
    with pytest.raises(ps.PsError):
        ps.read_rows(sample, cap=10)
