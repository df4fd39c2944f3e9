"""Python bindings of the FP64 CPU oracle (test infrastructure only).

liboracle.so restates the reference CPU app operation for operation
(see ps_oracle.h).  _ref/libpsref.so, when it was built where the reference
sources exist, is the reference's own numeric code behind a small harness
(ref_harness.cpp).  Only tests, the smoke check and the CPU-baseline leg of
bench.py may use this package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
ORACLE_MAIN = os.path.join(HERE, "ps_oracle_main")
REF_LIB = os.path.join(HERE, "_ref", "libpsref.so")

NP = 201  # 1-indexed parameter stride (MAX_N_PAR + 1)
NF = 1001
NL = 11
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class FitResult(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("dev", "chisq", "rchisq", "prd", "la", "be", "dark", "totarea", "alamda")] + [
        ("niter", C.c_int), ("pad", C.c_int)]


def build(force: bool = False) -> None:
    """Compile liboracle.so (always) and _ref/libpsref.so (if the reference exists)."""
    if force or not os.path.exists(ORACLE_LIB) or not os.path.exists(ORACLE_MAIN):
        subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if force or not os.path.exists(REF_LIB):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _load_oracle() -> C.CDLL:
    lib = C.CDLL(ORACLE_LIB)
    vp = C.c_void_p
    sig = {
        "orc_create": (vp, []), "orc_destroy": (None, [vp]),
        "orc_read_input": (C.c_int, [vp, C.c_char_p]),
        "orc_prepare_round": (None, [vp, C.c_double]), "orc_end_round": (None, [vp]),
        "orc_fit": (None, [vp, C.c_double, C.c_int, C.POINTER(FitResult)]),
        "orc_period": (None, [vp, C.c_int, _dp, C.c_int]),
        "orc_precalc": (C.c_double, [vp]),
        "orc_scan_to_file": (C.c_int, [vp, C.c_int, C.c_int, C.c_char_p, C.c_char_p]),
        "orc_n_periods": (C.c_int, [vp]),
        "orc_dptr": (_dp, [vp, C.c_char_p]), "orc_iptr": (_ip, [vp, C.c_char_p]),
        "orc_get": (C.c_double, [vp, C.c_char_p]), "orc_set": (None, [vp, C.c_char_p, C.c_double]),
        "orc_trifac": (None, [vp, C.c_int]), "orc_areanorm": (None, [vp, C.c_int, C.c_int]),
        "orc_sphfunc": (None, [vp, C.c_int]),
        "orc_ellfit": (None, [vp, _dp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]),
        "orc_ludcmp": (None, [_dp, C.c_int, _ip, _dp]), "orc_lubksb": (None, [_dp, C.c_int, _ip, _dp]),
        "orc_curv": (None, [vp, _dp]), "orc_blmatrix": (None, [vp, C.c_double, C.c_double]),
        "orc_matrix": (None, [vp, C.c_double, C.c_double, _dp, _dp]),
        "orc_phasec": (None, [vp, _dp, C.c_double, _dp]),
        "orc_bright": (C.c_double, [vp, _dp, _dp, C.c_double, _dp, _dp, C.c_int]),
        "orc_conv": (C.c_double, [vp, C.c_int, _dp, C.c_int]),
        "orc_mrqcof": (C.c_double, [vp, _dp, _ip, C.c_int, _dp, _dp]),
        "orc_gauss_errc": (C.c_int, [_dp, C.c_int, _dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


olib = _load_oracle() if os.path.exists(ORACLE_LIB) else None


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_ip)


class Oracle:
    """One oracle context (the reference's globals for one problem)."""

    def __init__(self, path: str | None = None):
        if olib is None:
            raise RuntimeError("oracle not built: run `make -C oracle`")
        self.c = olib.orc_create()
        if path is not None:
            rc = olib.orc_read_input(self.c, os.fsencode(path))
            if rc != 0:
                raise RuntimeError(f"oracle cannot read {path}: {rc}")

    def __del__(self):
        if getattr(self, "c", None):
            olib.orc_destroy(self.c)
            self.c = None

    def get(self, name: str) -> float:
        return olib.orc_get(self.c, name.encode())

    def set(self, name: str, v: float) -> None:
        olib.orc_set(self.c, name.encode(), float(v))

    def iget(self, name: str) -> int:
        return int(self.get(name))

    def darr(self, name: str, shape) -> np.ndarray:
        p = olib.orc_dptr(self.c, name.encode())
        n = int(np.prod(shape))
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(shape)

    def iarr(self, name: str, shape) -> np.ndarray:
        p = olib.orc_iptr(self.c, name.encode())
        n = int(np.prod(shape))
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(shape)

    def prepare(self, conw_r: float) -> None:
        olib.orc_prepare_round(self.c, conw_r)

    def end_round(self) -> None:
        olib.orc_end_round(self.c)

    def n_periods(self) -> int:
        return olib.orc_n_periods(self.c)

    def fit(self, freq: float, pole: int) -> FitResult:
        """pole is 1-based as in the reference."""
        r = FitResult()
        olib.orc_fit(self.c, freq, pole, C.byref(r))
        return r

    def period(self, n: int, precalc: bool = False) -> np.ndarray:
        row = np.zeros(6)
        olib.orc_period(self.c, n, _d(row), 1 if precalc else 0)
        return row

    def precalc(self) -> float:
        return olib.orc_precalc(self.c)

    def mrqcof(self, cg1: np.ndarray):
        """cg1: 1-indexed parameter vector (length ma+1)."""
        ma = self.iget("Ncoef") + 8
        ia = self.iarr("ia", (NP,))
        alpha = np.zeros(NP * NP)
        beta = np.zeros(NP)
        cg1 = np.ascontiguousarray(cg1, dtype=np.float64)
        chi = olib.orc_mrqcof(self.c, _d(cg1), _i(np.ascontiguousarray(ia)), ma, _d(alpha), _d(beta))
        mfit = int(ia[1:ma + 1].astype(bool).sum())
        A = alpha.reshape(NP, NP)[1:mfit + 1, 1:mfit + 1].copy()
        return A, beta[1:mfit + 1].copy(), chi

    def initial_cg(self, freq: float, pole: int) -> np.ndarray:
        """Starting parameters of a fit (period_search_BOINC.cpp:819-844), 1-indexed."""
        nc = self.iget("Ncoef")
        ma = nc + 8
        cg = np.zeros(ma + 1)
        cg[1:nc + 1] = self.darr("cg_first", (NP,))[1:nc + 1]
        betas = [0, 0, 0, 0, 0, 60, 60, 60, -60, -60, -60]
        lams = [0, 0, 90, 180, 270, 60, 180, 300, 60, 180, 300]
        cg[nc + 1] = (np.pi / 180) * (90 - betas[pole])
        cg[nc + 2] = (np.pi / 180) * lams[pole]
        cg[nc + 3] = 24 * 2 * np.pi / (1 / freq)
        cg[nc + 4:nc + 7] = self.darr("par", (4,))[1:4]
        cg[nc + 7] = np.log(self.get("cl"))
        cg[nc + 8] = 1.0
        return cg


def load_ref():
    """The reference's own numeric sources behind ref_harness.cpp, or None."""
    if not os.path.exists(REF_LIB):
        return None
    lib = C.CDLL(REF_LIB)
    vp = C.c_void_p

    class RefFit(C.Structure):
        _fields_ = FitResult._fields_

    sig = {
        "ref_set_problem": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, C.c_int, _dp, _dp, _dp, _dp, _dp,
                                      _dp, C.c_double, C.c_double, C.c_double, _ip, C.c_int]),
        "ref_ellfit": (None, [_dp, C.c_double, C.c_double, C.c_double]),
        "ref_set_cg_first": (None, [_dp]),
        "ref_fit": (None, [C.c_double, C.c_double, C.c_double, _dp, _ip, C.c_double, C.c_double, C.c_int,
                           C.c_double, C.POINTER(FitResult)]),
        "ref_curv": (None, [_dp]),
        "ref_bright": (C.c_double, [_dp, _dp, C.c_double, _dp, _dp, C.c_int]),
        "ref_blmatrix": (None, [C.c_double, C.c_double]),
        "ref_mrqcof": (C.c_double, [_dp, C.c_int, _dp, _dp]),
        "ref_gauss_errc": (C.c_int, [_dp, C.c_int, _dp]),
        "ref_dptr": (_dp, [C.c_char_p]),
        "ref_get_scale": (C.c_double, []),
        "ref_ncoef": (C.c_int, []), "ref_numfac": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def ref_load_problem(ref, orc: Oracle) -> None:
    """Copy the oracle context's current round (points incl. conv rows, ia,
    header constants) into the reference harness's globals."""
    nc = orc.iget("Ncoef")
    ma = nc + 8
    lc = orc.iget("Lcurves")
    ndata = orc.iget("ndata")
    lp = np.ascontiguousarray(orc.iarr("Lpoints", (202,)))
    ir = np.ascontiguousarray(orc.iarr("Inrel", (202,)))
    tim = np.ascontiguousarray(orc.darr("tim", (20004,)))
    br = np.ascontiguousarray(orc.darr("brightness", (20004,)))
    sg = np.ascontiguousarray(orc.darr("sig", (20004,)))
    ee = np.ascontiguousarray(orc.darr("ee", (20004 * 4,)))
    ee0 = np.ascontiguousarray(orc.darr("ee0", (20004 * 4,)))
    w = np.ascontiguousarray(orc.darr("Weight", (20004,)))
    ia = np.ascontiguousarray(orc.iarr("ia", (NP,)))
    rc = ref.ref_set_problem(orc.iget("Lmax"), orc.iget("Mmax"), orc.iget("nrows"), lc, _i(lp), _i(ir), ndata,
                             _d(tim), _d(br), _d(sg), _d(ee), _d(ee0), _d(w), orc.get("Phi_0"),
                             orc.get("Alamda_incr"), orc.get("Alamda_start"), _i(ia), ma)
    if rc != 0:
        raise RuntimeError(f"ref_set_problem failed: {rc}")
    cgf = np.zeros(NP)
    ref.ref_ellfit(_d(cgf), orc.get("a_ax"), orc.get("b_ax"), orc.get("c_ax"))


def run_oracle_main(inp: str, out: str, n_from: int | None = None, n_to: int | None = None,
                    conw_r: float | None = None, timeout: float | None = None) -> subprocess.CompletedProcess:
    args = [ORACLE_MAIN, inp, out]
    if n_from is not None:
        args += [str(n_from), str(n_to)]
        if conw_r is not None:
            args += [repr(conw_r)]
    return subprocess.run(args, capture_output=True, text=True, timeout=timeout, check=True)
