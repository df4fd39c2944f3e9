"""Alternative kernel forms of the mixed-precision LM iteration against each
other, at the 5,000-point shape (lmax 8, 800 facets, 87 free parameters):

* the normal equations split over the waves by points (default) against
  the staged kernel (PSAMD_NEQ=staged): one evaluation agrees to the
  summation-order level of its form (FP32 or FP64), and both agree with the
  oracle;
* each fit's points in their own visibility order (default for the mixed
  forms) against input order (PSAMD_ORDER=0): the same model values and
  derivatives, summed in another order;
* the normal equations and the solve in one launch (PSAMD_FUSE=1) against
  two launches: bit for bit the same fits, in FP64 and in a mixed form;
* the mixed forms' solve with diagonal pivots (default) against complete
  pivoting (PSAMD_PIVOT=full), in the four-wave and the one-wave solver: on
  damped normal matrices the two take the same pivots, so the fits are bit
  for bit the same;
* PS_PREC_BF16A against PS_PREC_BF16: beta and chi2 bit for bit, alpha
  within the bound that bf16 rounding of both operands of every product
  allows (2^-8 of sqrt(alpha_jj alpha_kk) per entry, by Cauchy-Schwarz);
* the FP32 and bf16 model kernels: wave-owned points (default; mu and mu0
  on the matrix cores) against waves that share all 64 points
  (PSAMD_MIXED=shared): both agree with the oracle, and with each other to
  the FP32 rounding level (FP32) or within the bf16 form's bar (bf16);
* the FP64 model kernels: the overlapped kernel (default; points in the
  fit's order, zero facet blocks skipped), the same kernel in input order
  (PSAMD_ORDER=0) and the kernel whose facet stage and contraction
  alternate across barriers (PSAMD_MODEL=mfma): one evaluation agrees with
  the oracle and between the kernels to the FP64 summation-order level.

The switches are read when an engine is created, so each engine below is
made with the environment it needs."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big(tmp_path_factory):
    import oracle
    from periodsearch_amd import Input, synth

    path = str(tmp_path_factory.mktemp("forms") / "period_search_in")
    ast = synth.write_workload("big", path, incl_deg=10.0)
    inp = Input(path)
    kt = round((inp.freq_start - 1 / ast["period_h"]) / inp.freq_step) + 1
    orc = oracle.Oracle(path)
    orc.prepare(inp.conw_r0)
    return inp, orc, kt


def _engine(inp, prec, **env):
    from periodsearch_amd import Engine

    old = {k: os.environ.get(k) for k in env}
    os.environ.update({k: str(v) for k, v in env.items()})
    try:
        eng = Engine(0)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    eng.set_precision(prec)
    eng.set_problem(inp.problem, keep=inp)
    return eng


def _starts(inp, orc, kt):
    nc = orc.iget("Ncoef")
    rng = np.random.default_rng(3)
    cgs = []
    for pole in (1, 4, 7, 10):
        cg1 = orc.initial_cg(inp.freq_start - (kt - 1) * inp.freq_step, pole)
        cg1[nc + 1:nc + 7] += rng.normal(0, 1e-3, size=6) * np.maximum(np.abs(cg1[nc + 1:nc + 7]), 0.05)
        cgs.append(cg1)
    return cgs


@pytest.mark.parametrize("form", ["fp64", "fp32", "bf16x6"])
def test_split_and_staged_normal_equations_agree(big, form):
    from periodsearch_amd import PS_PREC_BF16X6, PS_PREC_FP32, PS_PREC_FP64

    inp, orc, kt = big
    prec = {"fp64": PS_PREC_FP64, "fp32": PS_PREC_FP32, "bf16x6": PS_PREC_BF16X6}[form]
    # against the oracle: FP64 to its summation-order level; the mixed forms
    # bounded by the FP32 facet stage (about 3e-7 alpha, 2e-6 beta)
    bar = (1e-12, 1e-12, 1e-12) if form == "fp64" else (1e-5, 1e-4, 1e-6)
    # between kernels: the same Jacobian summed in another order
    sbar = (1e-13, 1e-13, 1e-13) if form == "fp64" else (2e-6, 2e-5, 1e-7)
    cgs = _starts(inp, orc, kt)
    cg = np.array([c[1:] for c in cgs])
    outs = {}
    for name, env in (("split", {}), ("staged", {"PSAMD_NEQ": "staged"}), ("input order", {"PSAMD_ORDER": "0"})):
        eng = _engine(inp, prec, **env)
        outs[name] = eng.eval_mrqcof(cg, inp.conw_r0)[:3]
        eng.close()
    for i, cg1 in enumerate(cgs):
        A, b, c = orc.mrqcof(cg1)
        for name, (alpha, beta, chi) in outs.items():
            ea = np.abs(alpha[i] - A).max() / np.abs(A).max()
            eb = np.abs(beta[i] - b).max() / np.abs(b).max()
            ec = abs(chi[i] - c) / c
            print(f"{form} {name} start {i}: alpha {ea:.2e} beta {eb:.2e} chi2 {ec:.2e} against the oracle")
            assert ea < bar[0] and eb < bar[1] and ec < bar[2], (name, i, ea, eb, ec)
        a0, b0, c0 = (outs["split"][k][i] for k in range(3))
        for name in ("staged", "input order"):
            a1, b1, c1 = (outs[name][k][i] for k in range(3))
            assert np.abs(a1 - a0).max() / np.abs(a0).max() < sbar[0], name
            assert np.abs(b1 - b0).max() / np.abs(b0).max() < sbar[1], name
            assert abs(c1 - c0) / c0 < sbar[2], name


@pytest.mark.parametrize("form", ["fp64", "bf16", "bf16a"])
def test_fused_iteration_kernel_equals_two_launches(big, form):
    from periodsearch_amd import PS_PREC_BF16, PS_PREC_BF16A, PS_PREC_FP64

    inp, orc, kt = big
    prec = {"fp64": PS_PREC_FP64, "bf16": PS_PREC_BF16, "bf16a": PS_PREC_BF16A}[form]
    freq = [inp.freq_start - (kt - 1 + d) * inp.freq_step for d in (0, 1)]
    freqs = [f for f in freq for _ in range(10)]
    poles = list(range(10)) * 2
    res = {}
    for fuse in (0, 1):
        eng = _engine(inp, prec, PSAMD_FUSE=fuse)
        res[fuse] = eng.fit_batch(freqs, poles, inp.conw_r0)
        eng.close()
    from periodsearch_amd import FIT_RESULT_FIELDS

    for k in FIT_RESULT_FIELDS:
        assert np.array_equal(res[0][k], res[1][k]), k


@pytest.mark.parametrize("workload", ["big", "wu"])
def test_diagonal_and_complete_pivots_give_the_same_fits(big, tmp_path, workload):
    # 87 free parameters: the four-wave solver; 55 (the work-unit shape): the one-wave solver
    from periodsearch_amd import FIT_RESULT_FIELDS, PS_PREC_BF16X6, Input, synth

    if workload == "big":
        inp, orc, kt = big
    else:
        path = str(tmp_path / "period_search_in")
        ast = synth.write_workload("wu", path)
        inp = Input(path)
        kt = round((inp.freq_start - 1 / ast["period_h"]) / inp.freq_step) + 1
    freqs = [inp.freq_start - (kt - 1 + d) * inp.freq_step for d in (-2, 0, 3) for _ in range(10)]
    poles = list(range(10)) * 3
    res = {}
    for pivot in ("diag", "full"):
        eng = _engine(inp, PS_PREC_BF16X6, PSAMD_PIVOT=pivot)
        res[pivot] = eng.fit_batch(freqs, poles, inp.conw_r0)
        eng.close()
    for k in FIT_RESULT_FIELDS:
        assert np.array_equal(res["diag"][k], res["full"][k]), k


def test_bf16a_alpha_within_the_bf16_rounding_bound(big):
    # each product w*jn_j*jn_k takes both operands rounded to bf16 (relative
    # error at most 2^-9 each), so it moves by at most (2^-8 + 2^-18) of its
    # size, and sum_p w |jn_j jn_k| <= sqrt(alpha_jj alpha_kk); the FP32 sums
    # of the two forms add their own ~1e-7
    from periodsearch_amd import PS_PREC_BF16, PS_PREC_BF16A

    inp, orc, kt = big
    cgs = _starts(inp, orc, kt)
    cg = np.array([c[1:] for c in cgs])
    outs = {}
    for prec in (PS_PREC_BF16, PS_PREC_BF16A):
        eng = _engine(inp, prec)
        outs[prec] = eng.eval_mrqcof(cg, inp.conw_r0)[:3]
        eng.close()
    (a0, b0, c0), (a1, b1, c1) = outs[PS_PREC_BF16], outs[PS_PREC_BF16A]
    assert np.array_equal(b0, b1) and np.array_equal(c0, c1)
    for i in range(len(cgs)):
        d = np.sqrt(np.abs(np.diag(a0[i])))
        bound = (2.0 ** -8 + 2.0 ** -18) * np.outer(d, d) + 1e-6 * np.abs(a0[i]).max()
        err = np.abs(a1[i] - a0[i])
        print(f"start {i}: alpha moved by {err.max() / np.abs(a0[i]).max():.2e} of its largest entry; "
              f"largest share of the bound {(err / bound).max():.3f}")
        assert np.all(err <= bound), i


@pytest.mark.parametrize("form", ["fp32", "bf16"])
def test_wave_owned_and_shared_point_kernels_agree(big, form):
    from periodsearch_amd import PS_PREC_BF16, PS_PREC_FP32

    inp, orc, kt = big
    prec = {"fp32": PS_PREC_FP32, "bf16": PS_PREC_BF16}[form]
    # bars against the oracle (alpha, beta) and between the kernels (alpha, beta, chi2)
    bar = {"fp32": (1e-5, 1e-4, 2e-6, 2e-5, 1e-7), "bf16": (1e-2, 1.0, 1e-2, 1.0, 1e-6)}[form]
    cgs = _starts(inp, orc, kt)
    cg = np.array([c[1:] for c in cgs])
    outs = {}
    for name, env in (("wave", {}), ("shared", {"PSAMD_MIXED": "shared"})):
        eng = _engine(inp, prec, **env)
        outs[name] = eng.eval_mrqcof(cg, inp.conw_r0)[:3]
        eng.close()
    for i, cg1 in enumerate(cgs):
        A, b, c = orc.mrqcof(cg1)
        for name, (alpha, beta, chi) in outs.items():
            ea = np.abs(alpha[i] - A).max() / np.abs(A).max()
            eb = np.abs(beta[i] - b).max() / np.abs(b).max()
            ec = abs(chi[i] - c) / c
            print(f"{form} {name} start {i}: alpha {ea:.2e} beta {eb:.2e} chi2 {ec:.2e} against the oracle")
            assert ea < bar[0] and eb < bar[1] and ec < 1e-6, (name, i, ea, eb, ec)
        a0, b0, c0 = (outs["wave"][k][i] for k in range(3))
        a1, b1, c1 = (outs["shared"][k][i] for k in range(3))
        da, db, dc = np.abs(a1 - a0).max() / np.abs(a0).max(), np.abs(b1 - b0).max() / np.abs(b0).max(), abs(c1 - c0) / c0
        print(f"{form} start {i}: wave against shared: alpha {da:.2e} beta {db:.2e} chi2 {dc:.2e}")
        assert da < bar[2] and db < bar[3] and dc < bar[4], (i, da, db, dc)


@pytest.mark.parametrize("workload", ["big", "wu"])
def test_fp64_model_kernels_agree(big, tmp_path, workload):
    # 81 coefficients (80 on the matrix cores + 1 vector column) and 49 (48 + 1)
    import oracle
    from periodsearch_amd import PS_PREC_FP64, Input, synth

    if workload == "big":
        inp, orc, kt = big
    else:
        path = str(tmp_path / "period_search_in")
        ast = synth.write_workload("wu", path)
        inp = Input(path)
        kt = round((inp.freq_start - 1 / ast["period_h"]) / inp.freq_step) + 1
        orc = oracle.Oracle(path)
        orc.prepare(inp.conw_r0)
    cgs = _starts(inp, orc, kt)
    cg = np.array([c[1:] for c in cgs])
    outs = {}
    for name, env in (("overlapped", {}), ("input order", {"PSAMD_ORDER": "0"}), ("barrier", {"PSAMD_MODEL": "mfma"})):
        eng = _engine(inp, PS_PREC_FP64, **env)
        outs[name] = eng.eval_mrqcof(cg, inp.conw_r0)[:3]
        eng.close()
    for i, cg1 in enumerate(cgs):
        A, b, c = orc.mrqcof(cg1)
        for name, (alpha, beta, chi) in outs.items():
            ea = np.abs(alpha[i] - A).max() / np.abs(A).max()
            eb = np.abs(beta[i] - b).max() / np.abs(b).max()
            ec = abs(chi[i] - c) / c
            print(f"fp64 {workload} {name} start {i}: alpha {ea:.2e} beta {eb:.2e} chi2 {ec:.2e} against the oracle")
            assert ea < 1e-12 and eb < 1e-12 and ec < 1e-12, (name, i, ea, eb, ec)
        a0, b0, c0 = (outs["overlapped"][k][i] for k in range(3))
        for name in ("input order", "barrier"):
            a1, b1, c1 = (outs[name][k][i] for k in range(3))
            assert np.abs(a1 - a0).max() / np.abs(a0).max() < 1e-13, name
            assert np.abs(b1 - b0).max() / np.abs(b0).max() < 1e-13, name
            assert abs(c1 - c0) / c0 < 1e-13, name
