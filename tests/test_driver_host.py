"""Host-side driver pieces (no GPU): checkpoint formats, the optional
DBL_MIN clamp of the input reader, the driver's argument checks, the
network-file reader, the NaN-safe pole reduction of the learned scan, and the reference's real
work-unit inputs parsed against the oracle."""
import ctypes
import glob
import os
import subprocess

import numpy as np
import pytest

import oracle
import periodsearch_amd as ps

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_WU = "/root/reference/period_search/input_period_search"


# ---------------------------------------------------------------- checkpoints
@pytest.mark.parametrize("fmt,text", [
    (ps.PS_STATE_BASE, "17"),
    (ps.PS_STATE_CUDA, "17 1 0.00026564558463277498"),
    (ps.PS_STATE_OC, "17 1 0.00026564558463277498 0 0"),
])
def test_state_write_is_the_reference_text(tmp_path, fmt, text):
    p = str(tmp_path / "period_search_state")
    ps.state_write(p, fmt, 17, 1, 0.00026564558463277498)
    assert open(p).read() == text
    assert not os.path.exists(p + ".temp")
    st = ps.state_read(p)
    assert st["format"] == fmt and st["nlines"] == 17
    if fmt >= ps.PS_STATE_CUDA:
        assert st["new_conw"] == 1 and st["conw_r"] == 0.00026564558463277498
    else:
        assert st["new_conw"] == 0 and st["conw_r"] == 0


def test_state_oc_inside_a_precalc_round(tmp_path):
    p = str(tmp_path / "s")
    # what the optimised CPU app writes after 4 of 10 precalc periods
    open(p, "w").write("0 0 0.0010000000000000000208 3.25 4")
    st = ps.state_read(p)
    assert st == dict(format=ps.PS_STATE_OC, nlines=0, new_conw=0, conw_r=0.001, sum_dark=3.25, test_periods=4)


@pytest.mark.parametrize("text,fmt", [
    (None, ps.PS_STATE_NONE), ("", ps.PS_STATE_NONE), ("garbage\n", ps.PS_STATE_NONE),
    ("-4 1 0.1", ps.PS_STATE_NONE), ("5 1 -0.1", ps.PS_STATE_NONE), ("5 1 nan", ps.PS_STATE_NONE),
    ("5 1", ps.PS_STATE_BASE), ("5 1 0.1 2.0", ps.PS_STATE_CUDA), ("5x", ps.PS_STATE_BASE),
])
def test_state_unreadable_or_partial(tmp_path, text, fmt):
    p = str(tmp_path / "s")
    if text is not None:
        open(p, "w").write(text)
    assert ps.state_read(p)["format"] == fmt


# ---------------------------------------------------------------- input
def _write_with_bad_values(src, dst):
    lines = open(src).read().splitlines()
    # first data line of the first light curve: time and brightness set non-positive
    v = lines[16].split()
    v[1] = "-1.0"
    lines[16] = " ".join(v)
    open(dst, "w").write("\n".join(lines) + "\n")


def test_dbl_min_clamp_is_selectable(tmp_path):
    src = os.path.join(GOLDEN, "ps_sample_in")
    dst = str(tmp_path / "in")
    _write_with_bad_values(src, dst)
    arr = lambda ptr, n: np.ctypeslib.as_array(ptr, shape=(n,))
    base = ps.Input(dst)
    clamped = ps.Input(dst, clamp_min=True)
    n = base.problem.n_points
    yb, yc = arr(base.problem.brightness, n), arr(clamped.problem.brightness, n)
    assert yb[0] == -1.0 and yc[0] == np.finfo(float).tiny
    assert np.array_equal(yb[1:], yc[1:])
    # the base reader agrees with the oracle (the base app keeps the value)
    o = oracle.Oracle(dst)
    assert np.array_equal(yb, o.darr("brightness", (20004,))[1:n + 1])


def test_driver_rejects_unknown_options(tmp_path):
    for args in (["--prec", "bf16x7"], ["--state-format", "xml"], ["--bogus"], ["--prec"]):
        r = subprocess.run([ps.DRIVER_PATH, *args], cwd=tmp_path, capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, (args, r.stderr)


# ---------------------------------------------------------------- network file
def _open_net(path):
    net = ps.PsProposalNet()
    rc = ps.lib.ps_proposal_open(os.fsencode(str(path)), 0, 64, ctypes.byref(net))
    if rc == 0:
        ps.lib.ps_proposal_close(ctypes.byref(net))
    return rc


def test_network_file_reader(tmp_path):
    """ps_proposal_open parses the stored (uncompressed) .npz that
    proposal.py writes and refuses other files with PS_ERR_IO before it
    touches a device; a well-formed file reaches the device step (which
    fails here, with no GPU, with another code)."""
    from periodsearch_amd.proposal import DEFAULT_NET

    z = np.load(DEFAULT_NET, allow_pickle=False)
    arrs = {k: z[k] for k in ("params", "x_mean", "x_std", "t_mean", "t_std")}
    ok = tmp_path / "ok.npz"
    np.savez(ok, **arrs)
    assert _open_net(DEFAULT_NET) not in (ps.PS_ERR_IO, ps.PS_ERR_ARG)
    assert _open_net(ok) not in (ps.PS_ERR_IO, ps.PS_ERR_ARG)
    bad = {
        "missing": None,
        "compressed": lambda p: np.savez_compressed(p, **arrs),
        "float64": lambda p: np.savez(p, **{**arrs, "params": arrs["params"].astype(np.float64)}),
        "short_params": lambda p: np.savez(p, **{**arrs, "params": arrs["params"][:-1]}),
        "no_t_std": lambda p: np.savez(p, **{k: v for k, v in arrs.items() if k != "t_std"}),
        "two_dims": lambda p: np.savez(p, **{**arrs, "x_mean": arrs["x_mean"][None, :]}),
        "not_zip": lambda p: open(p, "wb").write(b"\x93NUMPY" + bytes(100)),
    }
    for name, make in bad.items():
        p = tmp_path / f"{name}.npz"
        if make:
            make(p)
        assert _open_net(p) == ps.PS_ERR_IO, name


# ---------------------------------------------------------------- pole reduction
def _fits(devs):
    f = np.zeros(len(devs), dtype=ps.FIT_DTYPE)
    f["dev"] = devs
    f["prd"] = 0.3 + 0.01 * np.arange(len(devs))
    f["la"] = -10.0
    f["be"] = 20.0
    f["dark"] = 1.0
    f["totarea"] = 10.0
    return f


def test_pole_reduction_skips_nan():
    """period_search_BOINC.cpp:913-924 (ps_rows_from_fits, used by the learned
    scan): the first strictly smaller dev wins, starting from 1e40; NaN never
    wins; with none, dev 1e40 and period 0."""
    nan = float("nan")
    rows = ps.rows_from_fits(np.concatenate([_fits([nan, 0.5, 0.2, 0.2]), _fits([nan] * 4),
                                                    _fits([0.3, nan, 0.1, nan])]), 4, 100)
    assert rows["rms"][0] == 0.2 and np.isclose(rows["period_h"][0], 24 * 0.32)
    assert rows["rms"][1] == 1e40 and rows["period_h"][1] == 0.0
    assert rows["rms"][2] == 0.1 and np.isclose(rows["period_h"][2], 24 * 0.32)
    assert rows["lambda"][0] == 350.0 and rows["chisq"][2] == 0.1 ** 2 * 100


# ---------------------------------------------------------------- real work units
def _compare_with_oracle(path):
    inp = ps.Input(path)
    o = oracle.Oracle(path)
    o.prepare(inp.conw_r0)
    p = inp.problem
    N = p.n_points
    arr = lambda ptr, n: np.ctypeslib.as_array(ptr, shape=(n,))
    assert N == o.iget("ndata") - 3
    assert inp.escl == o.get("escl") and inp.freq_step == o.get("freq_step") and inp.n_periods == o.n_periods()
    assert p.lmax == o.iget("Lmax") and p.n_fac == o.iget("Numfac")
    for name, n, key, shape in (("tim", N, "tim", None), ("brightness", N, "brightness", None),
                                ("sig", N, "sig", None), ("weight", N, "Weight", None)):
        assert np.array_equal(arr(getattr(p, name), n), o.darr(key, (20004,))[1:N + 1]), name
    assert np.array_equal(arr(p.ee, 3 * N).reshape(N, 3), o.darr("ee", (20004, 4))[1:N + 1, 1:4])
    assert np.array_equal(arr(p.ee0, 3 * N).reshape(N, 3), o.darr("ee0", (20004, 4))[1:N + 1, 1:4])
    nc = ps.n_coef(p.lmax, p.mmax)
    assert np.array_equal(arr(p.cg_first, nc), o.darr("cg_first", (201,))[1:nc + 1])
    inrel = o.iarr("Inrel", (202,))[1:p.n_lc + 1]
    assert np.array_equal(arr(p.lc_inrel, p.n_lc), inrel)
    return inp


def test_committed_work_units_match_oracle():
    files = sorted(glob.glob(os.path.join(GOLDEN, "wu", "period_search_in_*")))
    assert len(files) == 10
    for f in files:
        _compare_with_oracle(f)


@pytest.mark.skipif(not os.path.isdir(REF_WU), reason="the reference checkout is not on this machine")
def test_all_reference_work_units_match_oracle():
    files = sorted(glob.glob(os.path.join(REF_WU, "period_search_in_*")))
    assert len(files) == 100
    lmax = set()
    for f in files:
        lmax.add(_compare_with_oracle(f).problem.lmax)
    assert lmax == {3}
