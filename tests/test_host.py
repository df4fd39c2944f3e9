"""Host side of the product library: C-ABI exports, input parsing and the
shape basis, checked against the oracle (no GPU needed)."""
import re

import numpy as np

import oracle
import periodsearch_amd as ps


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
