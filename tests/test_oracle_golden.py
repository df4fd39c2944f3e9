"""The oracle against the golden vectors of the reference's own unit tests.

Fixtures: tests/golden/arrangers.npz, extracted by extract_arrangers.py from
period_search_pure_cpu_win/UnitTest_pure_cpu_win/arrangers.cpp (nrows=6, 288
facets, Lmax=Mmax=3, the 118-point input of that directory); each test below
follows the matching method of .../InnerFunctionsTests.cpp (cited per test).
The fixtures were printed with %.30f by an MSVC build, so values are compared
to 1e-13 relative / 1e-15 absolute rather than bit for bit; the reference
asserted only ``_x - x < DBL_EPSILON`` (one-sided) for most of them.
"""
import math
import os

import numpy as np
import pytest

import oracle

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "arrangers.npz"))
RT, AT = 1e-13, 1e-15


def g(key):
    return G[key]


@pytest.fixture
def orc():
    return oracle.Oracle()


def _facets(o, nrows=6):
    o.set("Lmax", 3)
    o.set("Mmax", 3)
    o.set("Numfac", 8 * nrows * nrows)
    o.darr("t", (1001,))[:147] = g("get_t/_t")
    o.darr("f", (1001,))[:147] = g("get_f/_f")
    oracle.olib.orc_trifac(o.c, nrows)
    oracle.olib.orc_areanorm(o.c, 146, 288)


def test_trifac(orc):  # InnerFunctionsTests.cpp:165-201
    oracle.olib.orc_trifac(orc.c, 6)
    ifp = orc.iarr("ifp", (1001, 4))
    for k in (1, 2, 3):
        assert np.array_equal(ifp[1:289, k], g(f"get_ifp/_ifp_{k}")[1:289]), k


def test_areanorm(orc):  # InnerFunctionsTests.cpp:203-243
    _facets(orc)
    np.testing.assert_allclose(orc.darr("at", (1001,))[1:289], g("get_at/_at")[1:289], rtol=RT, atol=AT)
    np.testing.assert_allclose(orc.darr("af", (1001,))[1:289], g("get_af/_af")[1:289], rtol=RT, atol=AT)
    np.testing.assert_allclose(orc.darr("Darea", (1001,))[1:289], g("get_darea/__Darea")[:288], rtol=RT, atol=AT)
    np.testing.assert_allclose(orc.darr("Nor", (1001, 4))[1:289, 1], g("get_nor_0/__Nor")[:288], rtol=RT, atol=AT)


def test_sphfunc(orc):  # InnerFunctionsTests.cpp:245-300
    _facets(orc)
    oracle.olib.orc_sphfunc(orc.c, 288)
    P = orc.darr("Pleg", (1001, 11, 11))
    for l in range(4):
        for m in range(4):
            np.testing.assert_allclose(P[1:289, l, m], g(f"get_pleg/_pleg_j{l}_k{m}")[1:289], rtol=RT, atol=1e-14)
    D = orc.darr("Dsph", (1001, 201))
    for k in range(1, 17):
        np.testing.assert_allclose(D[1:147, k], g(f"get_dsph/_dsph_{k}")[1:147], rtol=RT, atol=1e-14)


def test_ellfit(orc):  # InnerFunctionsTests.cpp:356-422
    _facets(orc)
    oracle.olib.orc_sphfunc(orc.c, 288)
    ave, al0 = 0.382008772881355651573898057904, 0.069555451868686835048549710336
    rfit = math.sqrt(2 * ave / (0.5 * math.pi * (1 + math.cos(al0))))
    a0, b0, c0 = 1.05, 1.00, 0.95
    escl = rfit / math.sqrt((a0 * b0 + b0 * c0 + a0 * c0) / 3) * 0.8
    cg = np.zeros(201)
    oracle.olib.orc_ellfit(orc.c, oracle._d(cg), a0 * escl, b0 * escl, c0 * escl, 288, 16)
    np.testing.assert_allclose(cg[1:17], g("get_cg_first/_cg_first_local")[1:17], rtol=1e-12, atol=1e-14)


def _blmat_from(prefix):
    B = np.zeros((4, 4))
    for j in range(4):
        B[:, j] = g(f"{prefix}/_blmat_{j}")
    return B


def _dblm_from(prefix):
    D = np.zeros((3, 4, 4))
    for j in range(4):
        for k in range(4):
            D[:, j, k] = g(f"{prefix}/_dblm_{j}_{k}")
    return D


def test_blmatrix(orc):  # InnerFunctionsTests.cpp:543-667 (pole beta=0, lambda=0)
    oracle.olib.orc_blmatrix(orc.c, math.radians(90.0), 0.0)
    np.testing.assert_allclose(orc.darr("Blmat", (4, 4))[1:, 1:], _blmat_from("get_blmat")[1:, 1:], atol=1e-15)
    np.testing.assert_allclose(orc.darr("Dblm", (3, 4, 4))[1:, 1:, 1:], _dblm_from("get_dblm")[1:, 1:, 1:],
                               atol=1e-15)


def test_phasec(orc):  # InnerFunctionsTests.cpp:669-707
    x1, x2 = g("get_xx1/__xx1"), g("get_xx2/__xx2")
    alpha = math.acos(x1[1] * x2[1] + x1[2] * x2[2] + x1[3] * x2[3])
    cg = g("get_cg24/_cg")
    php = np.array([0.0, cg[20], cg[21], cg[22]])
    dcdp = np.zeros(4)
    oracle.olib.orc_phasec(orc.c, oracle._d(dcdp), alpha, oracle._d(php))
    np.testing.assert_allclose(dcdp[1:4], g("get_dphp/__dphp")[1:4], rtol=RT)
    assert abs(orc.get("Scale") - 1.214621163362573641464337015350) < 1e-15


def test_matrix(orc):  # InnerFunctionsTests.cpp:709-757
    orc.darr("Blmat", (4, 4))[:] = _blmat_from("get_blmat_m")
    orc.darr("Dblm", (3, 4, 4))[:] = _dblm_from("get_dblm_m")
    orc.set("Phi_0", 0.0)
    cg = g("get_cg24/_cg")
    t = g("get_tim/_tim")[1]
    tmat = np.zeros(16)
    dtm = np.zeros(64)
    oracle.olib.orc_matrix(orc.c, cg[19], t, oracle._d(tmat), oracle._d(dtm))
    tm, dt = tmat.reshape(4, 4), dtm.reshape(4, 4, 4)
    for j in range(1, 4):
        np.testing.assert_allclose(tm[1:4, j], g(f"get_tmat/_tmat_{j}")[1:4], rtol=1e-12, atol=1e-15)
        for k in range(1, 4):
            np.testing.assert_allclose(dt[1:4, j, k], g(f"get_dtm/_dtm_{j}_{k}")[1:4], rtol=1e-12, atol=1e-12)


def test_bright(orc):  # InnerFunctionsTests.cpp:759-852: ymod = 0.34207493045125292, dyda to 1e-9
    _facets(orc)
    orc.set("Nphpar", 3)
    Dg = orc.darr("Dg", (1001, 201))
    for k in range(16):
        Dg[1:289, k + 1] = g(f"get_dg/_dg_{k}")[:288]
    orc.darr("Area", (1001,))[1:289] = g("get_area/_area_a")[:288]
    orc.darr("Blmat", (4, 4))[:] = _blmat_from("get_blmat_m")
    orc.darr("Dblm", (3, 4, 4))[:] = _dblm_from("get_dblm_m")
    orc.set("Phi_0", 0.0)
    cg = np.ascontiguousarray(g("get_cg24/_cg"))
    x1 = np.ascontiguousarray(g("get_xx1/__xx1"))
    x2 = np.ascontiguousarray(g("get_xx2/__xx2"))
    t = g("get_tim/_tim")[1]
    dyda = np.zeros(201)
    ymod = oracle.olib.orc_bright(orc.c, oracle._d(x1), oracle._d(x2), t, oracle._d(cg), oracle._d(dyda), 24)
    assert abs(ymod - 0.34207493045125292) < 1e-15
    np.testing.assert_allclose(dyda[1:25], g("get_dyda/__dyda")[:24], rtol=0, atol=1e-9)
