"""MI355X period-search engine: Python mirror of the C ABI in include/periodsearch.h
and include/ps_proposal.h.

The compute path is the HIP library ``libpsamd.so`` built in this directory
(``python -m periodsearch_amd.build``).  Importing this package loads it; if it
is missing or cannot be loaded the import fails loudly: there is no CPU
fallback for the fit loop.

The classes follow the roles of the reference's CUDA backend
(period_search_cuda/period_search/start_CUDA.h:2-14):

    Engine(device)            CUDAPrepare's device/context part
    Engine.set_problem(p)     CUDAPrepare + the static uploads of CUDAStart
    Engine.precalc(...)       CUDAPrecalc (one convexity-weight round)
    Engine.scan(...)          CUDAStart (rows as numbers; text via format_row)
    Engine.fit_batch(...)     per-fit results (test seam)
    Engine.eval_mrqcof(...)   one normal-equation build (test seam)

Learned starting points (no counterpart in the reference): Engine.lc_features,
Engine.fit_batch_from and the Mlp network (see periodsearch_amd.proposal).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PSAMD_LIB selects another build of the same library (kernel variants).
LIB_PATH = os.environ.get("PSAMD_LIB") or os.path.join(HERE, "libpsamd.so")
DRIVER_PATH = os.path.join(HERE, "period_search_amd")

PS_PREC_FP64 = 0
PS_PREC_FP32 = 1
PS_PREC_BF16 = 2
PS_PREC_BF16X3 = 3
PS_PREC_BF16X6 = 4
PS_PREC_BF16A = 5

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)


class PsProblem(C.Structure):
    _fields_ = [
        ("n_points", C.c_int), ("n_lc", C.c_int), ("lc_points", _ip), ("lc_inrel", _ip),
        ("tim", _dp), ("brightness", _dp), ("sig", _dp), ("weight", _dp), ("ee", _dp), ("ee0", _dp),
        ("lmax", C.c_int), ("mmax", C.c_int), ("n_fac", C.c_int),
        ("nor", _dp), ("darea", _dp), ("pleg", _dp), ("fc", _dp), ("fs", _dp), ("dsph", _dp), ("cg_first", _dp),
        ("phi0", C.c_double), ("par", C.c_double * 3), ("ia_par", C.c_int * 3), ("cl", C.c_double),
        ("ia_cl", C.c_int), ("ia_prd", C.c_int), ("onlyrel", C.c_int),
        ("alamda_start", C.c_double), ("alamda_incr", C.c_double), ("stop_condition", C.c_double),
        ("n_iter_min", C.c_int), ("n_poles", C.c_int), ("beta_pole", _dp), ("lambda_pole", _dp),
    ]


class PsShapeJob(C.Structure):
    _fields_ = [("nrows", C.c_int), ("lmax", C.c_int), ("mmax", C.c_int),
                ("a", C.c_double), ("b", C.c_double), ("c", C.c_double),
                ("nor", _dp), ("darea", _dp), ("pleg", _dp), ("fc", _dp), ("fs", _dp), ("dsph", _dp),
                ("cg_first", _dp)]


class PsFitOut(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("dev", "chisq", "rchisq", "prd", "la", "be", "dark", "totarea", "alamda")] + [
        ("niter", C.c_int), ("err", C.c_int), ("latency_us", C.c_double)]


class PsRow(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("period_h", "rms", "chisq", "dark", "lambda_", "beta")]


class PsProposalNet(C.Structure):
    _fields_ = [("mlp", C.c_void_p), ("n_in", C.c_int), ("n_out", C.c_int), ("nbins", C.c_int),
                ("max_batch", C.c_int), ("x_mean", _fp), ("x_std", _fp), ("t_mean", _fp), ("t_std", _fp)]


class PsLearnedOpts(C.Structure):
    _fields_ = [("block", C.c_int), ("n_candidates", C.c_int), ("halo", C.c_int), ("mirror", C.c_int),
                ("refine", C.c_int), ("judge", C.c_void_p)]


class PsState(C.Structure):
    _fields_ = [("format", C.c_int), ("nlines", C.c_int), ("new_conw", C.c_int), ("conw_r", C.c_double),
                ("sum_dark", C.c_double), ("test_periods", C.c_int)]


PS_OK, PS_ERR_ARG, PS_ERR_HIP, PS_ERR_LIMIT, PS_ERR_IO, PS_ERR_STATE = 0, -1, -2, -3, -4, -5
PS_STATE_NONE, PS_STATE_BASE, PS_STATE_CUDA, PS_STATE_OC = 0, 1, 2, 3
PS_READ_DEFER_SHAPE, PS_READ_CLAMP_MIN = 1, 2
PROGRESS_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_longlong, C.c_longlong)

FIT_DTYPE = np.dtype([(n, "f8") for n in ("dev", "chisq", "rchisq", "prd", "la", "be", "dark", "totarea", "alamda")]
                     + [("niter", "i4"), ("err", "i4"), ("latency_us", "f8")])
# the fields that a fit's arithmetic fixes (latency_us is a timing)
FIT_RESULT_FIELDS = FIT_DTYPE.names[:-1]
ROW_DTYPE = np.dtype([(n, "f8") for n in ("period_h", "rms", "chisq", "dark", "lambda", "beta")])
assert FIT_DTYPE.itemsize == C.sizeof(PsFitOut)
assert ROW_DTYPE.itemsize == C.sizeof(PsRow)


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"periodsearch_amd: {LIB_PATH} is not built; run `python -m periodsearch_amd.build`")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "ps_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "ps_destroy": (C.c_int, [vp]),
        "ps_last_error": (C.c_char_p, [vp]),
        "ps_set_precision": (C.c_int, [vp, C.c_int]),
        "ps_set_slots": (C.c_int, [vp, C.c_int]),
        "ps_set_groups": (C.c_int, [vp, C.c_int]),
        "ps_set_problem": (C.c_int, [vp, C.POINTER(PsProblem)]),
        "ps_set_problems": (C.c_int, [vp, C.POINTER(PsProblem), C.c_int]),
        "ps_fit_batch_multi": (C.c_int, [vp, _ip, _dp, _ip, C.c_int, _dp, vp]),
        "ps_precalc": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, _dp, _ip]),
        "ps_scan": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, vp]),
        "ps_fit_batch": (C.c_int, [vp, _dp, _ip, C.c_int, C.c_double, vp]),
        "ps_eval_mrqcof": (C.c_int, [vp, _dp, C.c_int, C.c_double, _dp, _dp, _dp, _dp]),
        "ps_dims": (C.c_int, [vp, _ip, _ip, _ip]),
        "ps_last_kernel_ms": (C.c_int, [vp, _dp, C.POINTER(C.c_longlong)]),
        "ps_last_batch": (C.c_int, [vp, _dp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
        "ps_set_profiling": (C.c_int, [vp, C.c_int]),
        "ps_kernel_stats": (C.c_int, [vp, _dp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
        "ps_input_read": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
        "ps_input_read_deferred": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
        "ps_input_read_ex": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(vp)]),
        "ps_state_read": (C.c_int, [C.c_char_p, C.POINTER(PsState)]),
        "ps_state_write": (C.c_int, [C.c_char_p, C.POINTER(PsState)]),
        "ps_precalc_resume": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, _dp, _ip]),
        "ps_set_progress": (C.c_int, [vp, PROGRESS_FN, vp]),
        "ps_aspect_features": (C.c_int, [C.POINTER(PsProblem), _fp]),
        "ps_rows_from_fits": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp]),
        "ps_fit_batch_noalloc": (C.c_int, [vp, vp, vp, C.c_int, C.c_double, vp]),
        "ps_scan_learned": (C.c_int, [vp, C.POINTER(PsProblem), C.POINTER(PsProposalNet), C.POINTER(PsLearnedOpts),
                                      C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, vp, _ip, vp,
                                      C.POINTER(C.c_longlong)]),
        "ps_proposal_open": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.POINTER(PsProposalNet)]),
        "ps_proposal_close": (None, [C.POINTER(PsProposalNet)]),
        "ps_input_shape_job": (C.c_int, [vp, C.POINTER(PsShapeJob)]),
        "ps_shape_basis_many": (C.c_int, [vp, C.c_int, C.POINTER(PsShapeJob)]),
        "ps_input_free": (None, [vp]),
        "ps_input_problem": (C.c_int, [vp, C.POINTER(PsProblem)]),
        "ps_input_scalars": (C.c_int, [vp, _dp, _ip]),
        "ps_format_row": (C.c_int, [C.POINTER(PsRow), C.c_int, C.c_char_p, C.c_int]),
        "ps_best_row": (C.c_int, [vp, C.c_longlong, C.POINTER(C.c_longlong)]),
        "ps_read_rows": (C.c_int, [C.c_char_p, vp, C.c_longlong, C.POINTER(C.c_longlong)]),
        "ps_shape_basis": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                     _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "ps_lc_features": (C.c_int, [vp, _dp, C.c_int, C.c_int, _fp]),
        "ps_fit_batch_from": (C.c_int, [vp, _dp, C.c_int, C.c_double, vp, _dp]),
        "ps_mlp_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
        "ps_mlp_destroy": (C.c_int, [vp]),
        "ps_mlp_last_error": (C.c_char_p, [vp]),
        "ps_mlp_param_count": (C.c_longlong, [vp]),
        "ps_mlp_set_params": (C.c_int, [vp, _fp, C.c_longlong]),
        "ps_mlp_get_params": (C.c_int, [vp, _fp, C.c_longlong]),
        "ps_mlp_forward": (C.c_int, [vp, _fp, C.c_int, _fp]),
        "ps_mlp_gradients": (C.c_int, [vp, _fp, _fp, _fp, C.c_int, _fp, _fp]),
        "ps_mlp_train_step": (C.c_int, [vp, _fp, _fp, _fp, C.c_int, C.c_float, _fp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()

# Symbols declared in include/*.h (checked by the CPU test suite).
EXPORTED = ("ps_create", "ps_destroy", "ps_last_error", "ps_set_precision", "ps_set_slots", "ps_set_problem",
            "ps_set_problems", "ps_fit_batch_multi", "ps_last_batch", "ps_set_groups",
            "ps_precalc", "ps_scan", "ps_fit_batch", "ps_eval_mrqcof", "ps_dims", "ps_last_kernel_ms",
            "ps_set_profiling", "ps_kernel_stats", "ps_input_read", "ps_input_free", "ps_input_problem",
            "ps_input_scalars", "ps_format_row", "ps_shape_basis", "ps_shape_basis_many", "ps_input_read_deferred",
            "ps_input_shape_job", "ps_best_row", "ps_read_rows", "ps_input_read_ex", "ps_state_read",
            "ps_state_write", "ps_precalc_resume", "ps_set_progress", "ps_aspect_features", "ps_scan_learned",
            "ps_rows_from_fits", "ps_fit_batch_noalloc", "ps_proposal_open", "ps_proposal_close",
            "ps_lc_features", "ps_fit_batch_from", "ps_mlp_create", "ps_mlp_destroy", "ps_mlp_last_error",
            "ps_mlp_param_count", "ps_mlp_set_params", "ps_mlp_get_params", "ps_mlp_forward", "ps_mlp_gradients",
            "ps_mlp_train_step")


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def _f(a: np.ndarray):
    return a.ctypes.data_as(_fp)


class PsError(RuntimeError):
    pass


def n_coef(lmax: int, mmax: int) -> int:
    return sum(2 if m else 1 for m in range(mmax + 1) for _ in range(m, lmax + 1))


def shape_basis(nrows: int, lmax: int, mmax: int, a: float, b: float, c: float) -> dict:
    """Gaussian-sphere facet tables and ellipsoid coefficients (host code)."""
    F = 8 * nrows * nrows
    nc = n_coef(lmax, mmax)
    out = dict(nor=np.zeros(F * 3), darea=np.zeros(F), pleg=np.zeros(F * (lmax + 1) ** 2),
               fc=np.zeros(F * (mmax + 1)), fs=np.zeros(F * (mmax + 1)), dsph=np.zeros(F * nc),
               cg_first=np.zeros(nc))
    rc = lib.ps_shape_basis(nrows, lmax, mmax, a, b, c, *(_d(out[k]) for k in
                                                          ("nor", "darea", "pleg", "fc", "fs", "dsph", "cg_first")))
    if rc < 0:
        raise PsError(f"ps_shape_basis failed: {rc}")
    out["nor"] = out["nor"].reshape(F, 3)
    out["dsph"] = out["dsph"].reshape(F, nc)
    out["pleg"] = out["pleg"].reshape(F, lmax + 1, lmax + 1)
    out["fc"] = out["fc"].reshape(F, mmax + 1)
    out["fs"] = out["fs"].reshape(F, mmax + 1)
    return out


def _shape_arrays(nrows: int, lmax: int, mmax: int) -> dict:
    F = 8 * nrows * nrows
    return dict(nor=np.zeros(F * 3), darea=np.zeros(F), pleg=np.zeros(F * (lmax + 1) ** 2),
                fc=np.zeros(F * (mmax + 1)), fs=np.zeros(F * (mmax + 1)), dsph=np.zeros(F * n_coef(lmax, mmax)),
                cg_first=np.zeros(n_coef(lmax, mmax)))


def rows_from_fits(fits: np.ndarray, per: int, n_points: int) -> np.ndarray:
    """Best-of-`per` rows of consecutive fits (ps_rows_from_fits: the
    reference's pole reduction)."""
    fits = np.ascontiguousarray(fits, dtype=FIT_DTYPE)
    n = len(fits) // per
    rows = np.zeros(n, dtype=ROW_DTYPE)
    rc = lib.ps_rows_from_fits(fits.ctypes.data, n, per, n_points, rows.ctypes.data)
    if rc != 0:
        raise PsError(f"ps_rows_from_fits failed: {rc}")
    return rows


def read_rows(path: str, cap: int = 1 << 20) -> np.ndarray:
    """Rows of a period_search_out file (ps_read_rows)."""
    rows = np.zeros(cap, dtype=ROW_DTYPE)
    n = C.c_longlong()
    rc = lib.ps_read_rows(os.fsencode(path), rows.ctypes.data, cap, C.byref(n))
    if rc != 0:
        raise PsError(f"ps_read_rows({path}) failed: {rc} after {n.value} rows")
    return rows[:n.value].copy()


def best_row(rows: np.ndarray) -> int:
    """Index of the least-chi^2 row (ps_best_row): the period selection over
    one or more scans' rows."""
    rows = np.ascontiguousarray(rows, dtype=ROW_DTYPE)
    idx = C.c_longlong()
    rc = lib.ps_best_row(rows.ctypes.data, len(rows), C.byref(idx))
    if rc != 0:
        raise PsError(f"ps_best_row failed: {rc}")
    return idx.value


def best_period(paths) -> tuple:
    """(path, row) of the least-chi^2 row over several period_search_out
    files, e.g. the work units of one asteroid."""
    parts = [read_rows(p) for p in paths]
    owner = np.concatenate([np.full(len(r), k) for k, r in enumerate(parts)])
    rows = np.concatenate(parts)
    i = best_row(rows)
    return paths[int(owner[i])], rows[i]


def state_read(path: str) -> dict:
    """A period_search_state checkpoint in any of the three reference formats
    (ps_state_read); format PS_STATE_NONE when there is none usable."""
    st = PsState()
    rc = lib.ps_state_read(os.fsencode(path), C.byref(st))
    if rc != 0:
        raise PsError(f"ps_state_read failed: {rc}")
    return {k: getattr(st, k) for k, _ in PsState._fields_}


def state_write(path: str, fmt: int, nlines: int, new_conw: int = 0, conw_r: float = 0.0, sum_dark: float = 0.0,
                test_periods: int = 0) -> None:
    st = PsState(fmt, nlines, new_conw, conw_r, sum_dark, test_periods)
    rc = lib.ps_state_write(os.fsencode(path), C.byref(st))
    if rc != 0:
        raise PsError(f"ps_state_write failed: {rc}")


class Input:
    """A parsed period_search_in (period_search/period_search_BOINC.cpp:277-485).

    With ``defer_shape`` the shape basis is left for ``Engine.shape_basis_inputs``
    (the GPU shape setup of many inputs at once); ``clamp_min`` raises times
    and brightnesses below DBL_MIN to DBL_MIN as the optimised CPU app does."""

    def __init__(self, path: str, defer_shape: bool = False, clamp_min: bool = False):
        h = C.c_void_p()
        flags = (PS_READ_DEFER_SHAPE if defer_shape else 0) | (PS_READ_CLAMP_MIN if clamp_min else 0)
        rc = lib.ps_input_read_ex(os.fsencode(path), flags, C.byref(h))
        if rc != 0:
            raise PsError(f"cannot read {path}: {rc}")
        self._h = h
        self.problem = PsProblem()
        lib.ps_input_problem(self._h, C.byref(self.problem))
        s = np.zeros(10)
        nd = C.c_int()
        lib.ps_input_scalars(self._h, _d(s), C.byref(nd))
        (self.per_start, self.per_end, self.per_step_coef, self.jd_min, self.jd_max, self.escl, self.conw,
         self.freq_start, self.freq_end, self.freq_step) = s.tolist()
        self.ndata = nd.value
        self.n_periods = int((self.freq_start - self.freq_end) / self.freq_step) + 1

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.ps_input_free(h)
            self._h = None

    @property
    def conw_r0(self) -> float:
        return self.conw / self.escl / self.escl


class Engine:
    """One engine handle on one GPU."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib.ps_create(device, C.byref(h))
        if rc != 0:
            raise PsError(f"ps_create({device}) failed: {rc} (no HIP device?)")
        self._h = h
        self._keep = None

    def close(self):
        if self._h:
            lib.ps_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc: int, what: str):
        if rc != 0:
            raise PsError(f"{what} failed ({rc}): {lib.ps_last_error(self._h).decode()}")

    def set_precision(self, prec: int):
        self._check(lib.ps_set_precision(self._h, prec), "ps_set_precision")

    def set_slots(self, n: int):
        self._check(lib.ps_set_slots(self._h, n), "ps_set_slots")

    def set_groups(self, n: int):
        self._check(lib.ps_set_groups(self._h, n), "ps_set_groups")

    def set_problem(self, problem: PsProblem, keep=None):
        self._keep = keep  # keep the arrays behind the pointers alive
        self._check(lib.ps_set_problem(self._h, C.byref(problem)), "ps_set_problem")

    def set_problems(self, problems, keep=None):
        """Several asteroids at once (a list of PsProblem)."""
        arr = (PsProblem * len(problems))(*problems)
        self._keep = (keep, arr)
        self._check(lib.ps_set_problems(self._h, arr, len(problems)), "ps_set_problems")

    def fit_batch_multi(self, prob, freq, pole, conw_r) -> np.ndarray:
        """Fits over the problems of set_problems: item i fits problem prob[i]
        at freq[i] from pole[i]; conw_r[k] is problem k's convexity weight."""
        prob = np.ascontiguousarray(prob, dtype=np.int32)
        freq = np.ascontiguousarray(freq, dtype=np.float64)
        pole = np.ascontiguousarray(pole, dtype=np.int32)
        conw = np.ascontiguousarray(conw_r, dtype=np.float64)
        out = np.zeros(len(freq), dtype=FIT_DTYPE)
        self._check(lib.ps_fit_batch_multi(self._h, _i(prob), _d(freq), _i(pole), len(freq), _d(conw),
                                           out.ctypes.data), "ps_fit_batch_multi")
        return out

    def dims(self):
        nc, ma, mf = C.c_int(), C.c_int(), C.c_int()
        self._check(lib.ps_dims(self._h, C.byref(nc), C.byref(ma), C.byref(mf)), "ps_dims")
        return nc.value, ma.value, mf.value

    def fit_batch(self, freq, pole, conw_r: float) -> np.ndarray:
        freq = np.ascontiguousarray(freq, dtype=np.float64)
        pole = np.ascontiguousarray(pole, dtype=np.int32)
        out = np.zeros(len(freq), dtype=FIT_DTYPE)
        self._check(lib.ps_fit_batch(self._h, _d(freq), _i(pole), len(freq), conw_r, out.ctypes.data),
                    "ps_fit_batch")
        return out

    def scan(self, freq_start: float, freq_step: float, n_from: int, n_to: int, conw_r: float) -> np.ndarray:
        rows = np.zeros(n_to - n_from + 1, dtype=ROW_DTYPE)
        self._check(lib.ps_scan(self._h, freq_start, freq_step, n_from, n_to, conw_r, rows.ctypes.data), "ps_scan")
        return rows

    def precalc(self, freq_start: float, freq_end: float, freq_step: float, conw_r: float):
        cw = C.c_double(conw_r)
        nc = C.c_int(0)
        self._check(lib.ps_precalc(self._h, freq_start, freq_end, freq_step, C.byref(cw), C.byref(nc)),
                    "ps_precalc")
        return cw.value, nc.value

    def precalc_resume(self, freq_start: float, freq_end: float, freq_step: float, done_periods: int,
                       sum_dark: float, conw_r: float):
        cw = C.c_double(conw_r)
        nc = C.c_int(0)
        self._check(lib.ps_precalc_resume(self._h, freq_start, freq_end, freq_step, done_periods, sum_dark,
                                          C.byref(cw), C.byref(nc)), "ps_precalc_resume")
        return cw.value, nc.value

    def set_progress(self, fn=None):
        """fn(fits_done, fits_total) between GPU polls of every call; None turns it off."""
        self._progress = PROGRESS_FN(lambda _u, d, t: fn(d, t)) if fn else PROGRESS_FN()
        self._check(lib.ps_set_progress(self._h, self._progress, None), "ps_set_progress")

    def eval_mrqcof(self, cg: np.ndarray, conw_r: float):
        cg = np.ascontiguousarray(np.atleast_2d(cg), dtype=np.float64)
        n = cg.shape[0]
        _, ma, mfit = self.dims()
        assert cg.shape[1] == ma
        alpha = np.zeros((n, mfit, mfit))
        beta = np.zeros((n, mfit))
        chi = np.zeros(n)
        chck = np.zeros((n, 3))
        self._check(lib.ps_eval_mrqcof(self._h, _d(cg), n, conw_r, _d(alpha), _d(beta), _d(chi), _d(chck)),
                    "ps_eval_mrqcof")
        return alpha, beta, chi, chck

    def lc_features(self, freq, nbins: int = 64) -> np.ndarray:
        """Phase-folded light-curve features at trial frequencies (1/h):
        [n][nbins + 2] (bins, mean phase angle, mean |Earth latitude|)."""
        freq = np.ascontiguousarray(freq, dtype=np.float64)
        out = np.zeros((len(freq), nbins + 2), dtype=np.float32)
        self._check(lib.ps_lc_features(self._h, _d(freq), len(freq), nbins, _f(out)), "ps_lc_features")
        return out

    def fit_batch_from(self, cg: np.ndarray, conw_r: float, want_params: bool = False):
        """LM fits from given starting parameters cg[n][ma]; with want_params
        also the final parameters [n][ma]."""
        cg = np.ascontiguousarray(np.atleast_2d(cg), dtype=np.float64)
        _, ma, _ = self.dims()
        assert cg.shape[1] == ma
        out = np.zeros(cg.shape[0], dtype=FIT_DTYPE)
        cgo = np.zeros_like(cg) if want_params else None
        self._check(lib.ps_fit_batch_from(self._h, _d(cg), cg.shape[0], conw_r, out.ctypes.data,
                                          _d(cgo) if want_params else None), "ps_fit_batch_from")
        return (out, cgo) if want_params else out

    def set_profiling(self, on: bool = True):
        self._check(lib.ps_set_profiling(self._h, 1 if on else 0), "ps_set_profiling")

    def kernel_stats(self):
        ms = C.c_double()
        n = C.c_longlong()
        ev = C.c_longlong()
        lib.ps_kernel_stats(self._h, C.byref(ms), C.byref(n), C.byref(ev))
        return dict(model_ms=ms.value, model_launches=n.value, evals=ev.value)

    def last_batch(self):
        """(GPU ms, fits, LM iterations) of the last batch of fits."""
        ms = C.c_double()
        n = C.c_longlong()
        it = C.c_longlong()
        lib.ps_last_batch(self._h, C.byref(ms), C.byref(n), C.byref(it))
        return ms.value, n.value, it.value

    def last_kernel_ms(self):
        ms = C.c_double()
        n = C.c_longlong()
        lib.ps_last_kernel_ms(self._h, C.byref(ms), C.byref(n))
        return ms.value, n.value

    def shape_basis_many(self, specs) -> list:
        """GPU shape setup (ps_shape_basis_many): specs is a list of
        (nrows, lmax, mmax, a, b, c); returns one dict per spec shaped as
        shape_basis's."""
        outs, jobs = [], (PsShapeJob * max(len(specs), 1))()
        for j, (nrows, lmax, mmax, a, b, c) in enumerate(specs):
            o = _shape_arrays(nrows, lmax, mmax)
            outs.append(o)
            jobs[j] = PsShapeJob(nrows, lmax, mmax, a, b, c, *(_d(o[k]) for k in
                                                             ("nor", "darea", "pleg", "fc", "fs", "dsph", "cg_first")))
        self._check(lib.ps_shape_basis_many(self._h, len(specs), jobs), "ps_shape_basis_many")
        for (nrows, lmax, mmax, *_), o in zip(specs, outs):
            F = 8 * nrows * nrows
            o["nor"] = o["nor"].reshape(F, 3)
            o["dsph"] = o["dsph"].reshape(F, n_coef(lmax, mmax))
            o["pleg"] = o["pleg"].reshape(F, lmax + 1, lmax + 1)
            o["fc"] = o["fc"].reshape(F, mmax + 1)
            o["fs"] = o["fs"].reshape(F, mmax + 1)
        return outs

    def shape_basis_inputs(self, inputs) -> None:
        """Fill the shape tables of inputs read with defer_shape=True, all in
        one ps_shape_basis_many call."""
        jobs = (PsShapeJob * max(len(inputs), 1))()
        for j, inp in enumerate(inputs):
            self._check(lib.ps_input_shape_job(inp._h, C.byref(jobs[j])), "ps_input_shape_job")
        self._check(lib.ps_shape_basis_many(self._h, len(inputs), jobs), "ps_shape_basis_many")


class Mlp:
    """The proposal network on the GPU: n_in -> hidden -> hidden -> n_out,
    ReLU, FP32 (include/ps_proposal.h)."""

    def __init__(self, n_in: int, n_hidden: int, n_out: int, max_batch: int = 1024, device: int = 0):
        h = C.c_void_p()
        rc = lib.ps_mlp_create(device, n_in, n_hidden, n_out, max_batch, C.byref(h))
        if rc != 0:
            raise PsError(f"ps_mlp_create failed: {rc}")
        self._h = h
        self.shape = (n_in, n_hidden, n_out)
        self.max_batch = max_batch
        self.n_params = lib.ps_mlp_param_count(h)

    def close(self):
        if self._h:
            lib.ps_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc: int, what: str):
        if rc != 0:
            raise PsError(f"{what} failed ({rc}): {lib.ps_mlp_last_error(self._h).decode()}")

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(lib.ps_mlp_set_params(self._h, _f(p), p.size), "ps_mlp_set_params")

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        self._check(lib.ps_mlp_get_params(self._h, _f(p), p.size), "ps_mlp_get_params")
        return p

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.shape[2]), dtype=np.float32)
        self._check(lib.ps_mlp_forward(self._h, _f(x), x.shape[0], _f(y)), "ps_mlp_forward")
        return y

    def gradients(self, x, t, w):
        x, t, w = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, t, w))
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_float()
        self._check(lib.ps_mlp_gradients(self._h, _f(x), _f(t), _f(w), x.shape[0], _f(g), C.byref(loss)),
                    "ps_mlp_gradients")
        return g, loss.value

    def train_step(self, x, t, w, lr: float) -> float:
        x, t, w = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, t, w))
        loss = C.c_float()
        self._check(lib.ps_mlp_train_step(self._h, _f(x), _f(t), _f(w), x.shape[0], lr, C.byref(loss)),
                    "ps_mlp_train_step")
        return loss.value


def format_row(row, round_pole: bool = False) -> str:
    r = PsRow(*(float(row[k]) for k in ROW_DTYPE.names))
    buf = C.create_string_buffer(256)
    lib.ps_format_row(C.byref(r), 1 if round_pole else 0, buf, 256)
    return buf.value.decode()


def run_scan(inp: Input, eng: Engine, n_from: int = 1, n_to: int | None = None, conw_r: float | None = None,
             batch: int = 512) -> tuple[list[str], float]:
    """Driver loop of period_search_BOINC.cpp:487-931 on the engine: conw
    precalc (unless conw_r is given), then rows n_from..n_to as text."""
    if n_to is None:
        n_to = inp.n_periods
    if conw_r is None:
        conw_r = inp.conw_r0
        new_conw = 0
        while new_conw != 1 and conw_r * inp.escl * inp.escl < 10.0:
            conw_r, new_conw = eng.precalc(inp.freq_start, inp.freq_end, inp.freq_step, conw_r)
    lines = []
    for n0 in range(n_from, n_to + 1, batch):
        n1 = min(n_to, n0 + batch - 1)
        rows = eng.scan(inp.freq_start, inp.freq_step, n0, n1, conw_r)
        for k, r in enumerate(rows):
            r = r.copy()
            if n0 + k == 1:
                r["dark"] = conw_r * inp.escl * inp.escl
            lines.append(format_row(r))
    return lines, conw_r
