"""Host-side logic on the CPU: design builder, priors, the C-ABI library's
exports, the batched path's arithmetic (host instantiation) and diagnostics."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

import fixtures as fx
import oracle_lib as O
from ggdmc_amd import _lib, diagnostics, priors, simulate
from ggdmc_amd.model import build_model
from ggdmc_amd.runs import PriorSpec

ROOT = fx.ROOT


# ------------------------------------------------------------ C ABI library
def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "ggdmc_hip.h")).read()
    names = set(re.findall(r"^\s*(?:int|const char|void)\s*\*?\s*(ggdmc_\w+)\s*\(", hdr, re.M))
    assert {"ggdmc_run_subject", "ggdmc_run_hyper", "ggdmc_run", "ggdmc_run_batch",
            "ggdmc_loglike", "ggdmc_last_error"} <= names
    lib = C.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n


def test_without_gpu_calls_fail_loudly():
    if _lib.device_count() > 0:
        pytest.skip("a GPU is present")
    import ggdmc_amd as g
    fix = fx.load("lba_data6")
    with pytest.raises(g.GgdmcError):
        g.loglike(fix.model, fix.sub_data, fix.sub_theta[:, :2].T)


# ------------------------------------------------------------ design builder
@pytest.mark.parametrize("name", fx.FIXTURES)
def test_build_model_reproduces_reference_designs(name):
    fix = fx.load(name)
    sp = json.loads(str(fix.z["design_spec_json"]))
    m = build_model(sp["p_map"], sp["match_map"], sp["factors"], sp["constants"],
                    sp["accumulators"])
    assert m.cell_names == fix.model.cell_names
    assert m.pnames == fix.model.pnames
    assert (m.par_src == fix.model.par_src).all()
    assert (m.par_const == fix.model.par_const).all()
    assert (m.winner == fix.model.winner).all()


# ------------------------------------------------------------------ priors
def test_host_priors_match_oracle():
    fix = fx.load("lba_data6")
    rng = np.random.default_rng(1)
    for pr in (fix.sub_prior, fix.pop_prior, fix.hyper_prior):
        for _ in range(20):
            x = rng.uniform(-0.5, 11, pr.npar)
            a = priors.sumlogprior(pr, x)
            b = O.sumlogprior(fx.oprior(pr), x)
            assert (a == b) or abs(a - b) < 1e-12 * max(1, abs(b))
    phi = fix.z["phi_theta"][:, 0]
    X = fix.z["pop_theta"][:, :, 0]
    assert abs(priors.sumloghlike(fix.pop_prior, phi, X) / fix.z["phi_ll"][0] - 1) < 1e-13


def test_rprior_within_support_and_moments():
    rng = np.random.default_rng(2)
    pr = PriorSpec.tnorm([0.5, 2.0], [1.0, 0.3], 0.0, np.inf)
    x = priors.rprior(pr, 200000, rng)
    assert (x >= 0).all()
    from scipy import stats
    for k in range(2):
        a = (0 - pr.p0[k]) / pr.p1[k]
        m = stats.truncnorm.mean(a, np.inf, loc=pr.p0[k], scale=pr.p1[k])
        assert abs(x[:, k].mean() - m) < 0.01


# -------------------------------------- batched-path arithmetic on the host
_HOST = os.path.join(ROOT, "tests", "native", "liblba_math_host.so")


def _mixed(fix, data, theta):
    L = C.CDLL(_HOST)
    L.host_mixed_terms.restype = C.c_double
    d, i32, i64 = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    L.host_mixed_terms.argtypes = [i32, i32, d, C.c_int32, C.c_int32, d, i64, d, d]
    m = fix.model
    arrs = [np.ascontiguousarray(a, t) for a, t in ((m.winner, np.int32), (m.par_src, np.int32),
                                                   (m.par_const, np.float64))]
    rt = np.ascontiguousarray(data.rt)
    off = np.ascontiguousarray(data.cell_off, np.int64)
    th = np.ascontiguousarray(theta, np.float64)
    return L.host_mixed_terms(arrs[0].ctypes.data_as(i32), arrs[1].ctypes.data_as(i32),
                              arrs[2].ctypes.data_as(d), m.n_cell, m.n_acc,
                              rt.ctypes.data_as(d), off.ctypes.data_as(i64),
                              th.ctypes.data_as(d), None)


@pytest.mark.parametrize("name", fx.FIXTURES)
def test_mixed_arithmetic_on_goldens(name):
    # the batched path's formulation, with libm transcendentals
    fix = fx.load(name)
    e = [abs(_mixed(fix, fix.sub_data, fix.sub_theta[:, k]) / fix.sub_ll[k] - 1)
         for k in range(fix.sub_theta.shape[1])]
    for s in range(0, fix.n_subjects, 3):
        d = fix.pop_data(s)
        for k in range(fix.z["pop_theta"].shape[2]):
            e.append(abs(_mixed(fix, d, fix.z["pop_theta"][s, :, k]) / fix.z["pop_ll"][s, k] - 1))
    e = np.array(e)
    assert e.max() < 1e-7, e.max()


def test_mixed_arithmetic_near_posterior():
    fix = fx.load("lba_data6")
    worst = 0.0
    for s in range(0, fix.n_subjects, 8):
        d = fix.pop_data(s)
        for th in fx.near_thetas(fix, 40, seed=s):
            ref = O.lba_loglike(fix.omodel, d.rt, d.cell_off, th)
            worst = max(worst, abs(_mixed(fix, d, th) / ref - 1))
    assert worst <= 1e-6, worst


def _acc_eval(pars, rt):
    L = C.CDLL(_HOST)
    L.host_acc_eval.argtypes = [C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double)]
    p = (C.c_double * 5)(*pars)
    o = (C.c_double * 6)()
    L.host_acc_eval(p, rt, o)
    return list(o)


def test_ill_conditioned_evaluations_fall_back_to_the_reference_formula():
    # over a broad range of accumulators (negative drifts, narrow start-point
    # ranges, long response times) every evaluation the batched path keeps is
    # within 1e-5 (log) of the reference formula; the rest are flagged and
    # evaluated with the reference formula itself
    rng = np.random.default_rng(5)
    worst_w = worst_l = 0.0
    flagged = 0
    n = 6000
    for _ in range(n):
        A, B, v, s = rng.uniform(0.05, 1.5), rng.uniform(0.05, 2.0), rng.uniform(-3, 12), rng.uniform(0.1, 3)
        t0 = rng.uniform(0.05, 0.5)
        rt = t0 + rng.uniform(0.01, 4.0)
        lw, S, pdf, cdf, ill_w, ill_l = _acc_eval((A, B, v, s, t0), rt)
        flagged += ill_w or ill_l
        if not ill_w:
            worst_w = max(worst_w, abs(lw - np.log(max(pdf, 1e-10))))
        Sn = 1 - max(cdf, 1e-10)
        if not ill_l and Sn > 0 and S > 0:
            worst_l = max(worst_l, abs(np.log(S) - np.log(Sn)))
    assert worst_w < 1e-5 and worst_l < 1e-5, (worst_w, worst_l)
    assert flagged < 0.05 * n


def test_mixed_tail_accuracy():
    # the batched path's normal tail Q(|z|) = exp(-z^2/2) R(|z|) against
    # erfc: the error model takes every computed tail to 1e-9 relative
    # (kTailErr), the Mills-ratio fit to 2.9e-10 on [0, 14]; exp(-z^2/2) to
    # the rounding of its argument (z^2/2 in two roundings) plus 4e-16
    from scipy.special import erfc
    L = C.CDLL(_HOST)
    d = C.POINTER(C.c_double)
    L.host_mixed_tail.argtypes = [d, C.c_int64, d, d]
    z = np.concatenate([np.linspace(-14, 14, 400001), np.geomspace(1e-12, 14, 100000), [0.0, -0.0]])
    e, q = np.empty_like(z), np.empty_like(z)
    L.host_mixed_tail(z.ctypes.data_as(d), len(z), e.ctypes.data_as(d), q.ctypes.data_as(d))
    rq = np.abs(q / (0.5 * erfc(np.abs(z) * np.sqrt(0.5))) - 1)
    re = np.abs(e / np.exp(-0.5 * z * z) - 1)
    assert rq.max() < 4e-10, (rq.max(), z[rq.argmax()])
    assert np.all(re <= 4e-16 + 1.2e-16 * z * z), (re.max(), z[re.argmax()])
    # beyond the range of doubles the tail and exp are 0; NaN propagates through q
    big = np.array([40.0, -1e10, np.inf, np.nan])
    e, q = np.empty_like(big), np.empty_like(big)
    L.host_mixed_tail(big.ctypes.data_as(d), len(big), e.ctypes.data_as(d), q.ctypes.data_as(d))
    assert np.all(e[:3] == 0.0) and np.all(q[:2] < 1e-300) and np.isnan(q[3])


def test_division_from_reciprocal_is_correctly_rounded():
    # the loser survivor's F = (1 - omG) / den divides through the correctly
    # rounded reciprocal with one correction step: bit for bit IEEE division
    # (numpy's) on x in the survivor range and den over [1e-10, 1], with
    # significands of den near all ones and near powers of two
    L = C.CDLL(_HOST)
    d = C.POINTER(C.c_double)
    L.host_div_rn.argtypes = [d, d, C.c_int64, d]
    rng = np.random.default_rng(12)
    n = 2_000_000
    x = np.concatenate([rng.uniform(0.25, 1.5, n), 1.0 - np.ldexp(rng.random(n), -rng.integers(0, 53, n))])
    den = np.concatenate([np.exp(-23.0 * rng.random(n)), 1.0 - np.ldexp(rng.random(n), -rng.integers(0, 60, n))])
    mant = np.concatenate([(1 << 52) - 1 - rng.integers(0, 1024, n // 2), rng.integers(0, 1024, n // 2)])
    expo = (1022 - rng.integers(0, 34, n)).astype(np.uint64)
    edge = ((expo << np.uint64(52)) | mant.astype(np.uint64)).view(np.float64)
    den[:n] = np.where(rng.random(n) < 0.5, den[:n], edge)
    den = np.clip(den, 1e-10, 1.0)
    x = np.concatenate([x, [np.inf, -np.inf, np.nan, 0.0, 1.0]])
    den = np.concatenate([den, [0.5, 0.84, 1.0, 0.3, 1e-10]])
    out = np.empty_like(x)
    L.host_div_rn(x.ctypes.data_as(d), den.ctypes.data_as(d), len(x), out.ctypes.data_as(d))
    ref = x / den
    same = (out.view(np.uint64) == ref.view(np.uint64)) | (np.isnan(out) & np.isnan(ref))
    assert same.all(), (x[~same][:5], den[~same][:5])


def test_tiny_survivors_reproduce_the_reference_rounding():
    # a sharp loser long after its typical finishing time: the reference's
    # survivor 1 - F is quantised by FP64 rounding near F = 1; the batched
    # path reproduces the rounded value bit for bit
    same = 0
    cases = 0
    rng = np.random.default_rng(9)
    for _ in range(400):
        A, B, v, s = rng.uniform(0.2, 1.0), rng.uniform(0.1, 0.6), rng.uniform(6, 12), rng.uniform(0.8, 1.5)
        t0 = rng.uniform(0.1, 0.4)
        rt = t0 + rng.uniform(0.9, 1.6)
        lw, S, pdf, cdf, ill_w, ill_l = _acc_eval((A, B, v, s, t0), rt)
        Sn = 1 - max(cdf, 1e-10)
        if 0 < Sn < 1e-9 and not ill_l:
            cases += 1
            same += S == Sn
    assert cases > 20 and same >= 0.95 * cases, (cases, same)


def test_far_tail_winner_reproduces_the_reference_rounding():
    # a fast response of a winner far from its threshold (z2 = 5 .. 7): the
    # reference's v * (Phi(z1) - Phi(z2)) is the difference of two values
    # rounded near 1 (R's pnorm above 0: 1 - erfc(z / sqrt 2) / 2 in FP64),
    # quantised in steps of v * 2^-53. Where that rounding could move the
    # density by more than 1e-5 (|v| 2^-52 > 1e-5 |numerator|), the batched
    # path reproduces it (within the log's 1e-8) instead of flagging the
    # trial; elsewhere it is within 1e-5
    from scipy.special import erfc

    def ref(A, B, v, s, t):
        b = A + B
        zs, chi = t * s, b - t * v
        z1, z2 = chi / zs, (chi - A) / zs
        Q1, Q2 = 0.5 * erfc(z1 * np.sqrt(0.5)), 0.5 * erfc(z2 * np.sqrt(0.5))
        d1, d2 = np.exp(-0.5 * z1 * z1) / np.sqrt(2 * np.pi), np.exp(-0.5 * z2 * z2) / np.sqrt(2 * np.pi)
        den = max(0.5 * erfc(-v / s * np.sqrt(0.5)), 1e-10)
        num = v * ((1.0 - Q1) - (1.0 - Q2)) + s * (d2 - d1)
        return num / (A * den), 2.0 ** -52 * abs(v) > 1.2e-5 * abs(num)

    rng = np.random.default_rng(21)
    cases = flagged = rounded = 0
    worst = worst_rounded = 0.0
    for _ in range(3000):
        A, B, v, s = rng.uniform(0.2, 1.0), rng.uniform(0.5, 2.0), rng.uniform(3, 12), rng.uniform(0.5, 1.5)
        t0 = rng.uniform(0.1, 0.4)
        t = B / (v + rng.uniform(5.0, 7.0) * s)           # z2 = B / (t s) - v / s
        lw, S, pdf, cdf, ill_w, ill_l = _acc_eval((A, B, v, s, t0), t0 + t)
        p, coarse = ref(A, B, v, s, t)
        if not (2e-10 < p < 1e-6):
            continue
        cases += 1
        if ill_w:
            flagged += 1
            continue
        worst = max(worst, abs(lw - np.log(p)))
        if coarse:
            rounded += 1
            worst_rounded = max(worst_rounded, abs(lw - np.log(p)))
    assert cases > 300, cases
    assert flagged <= 0.01 * cases and rounded > 30, (flagged, rounded, cases)
    assert worst < 1e-5 and worst_rounded < 1e-7, (worst, worst_rounded)


# -------------------------------------------------------------- simulator
def test_simulator_rule_recovers_choice_probabilities():
    # P(response = accumulator 0) = int f_0 S_1 dt from the oracle's density;
    # the generation rule redraws only when every drift is negative, which
    # equals the likelihood's truncated drifts when P(drift < 0) is negligible
    m = build_model(p_map={"A": ["1"], "B": ["1"], "t0": ["1"], "mean_v": ["M"], "sd_v": ["M"],
                           "st0": ["1"]}, match_map={"M": {"s1": "r1", "s2": "r2"}},
                    factors={"S": ["s1", "s2"]}, constants={"st0": 0.0, "sd_v.false": 0.3},
                    accumulators=["r1", "r2"])
    theta = np.array([0.5, 1.0, 1.0, 2.0, 0.3, 0.2])  # A B mean_v.false mean_v.true sd_v.true t0
    assert m.pnames == ["A", "B", "mean_v.false", "mean_v.true", "sd_v.true", "t0"]
    rng = np.random.default_rng(0)
    data = simulate.simulate_subject(m, theta, 40000, rng)
    n = np.diff(data.cell_off)
    # cells: s1.r1 (correct), s1.r2, s2.r1, s2.r2 (correct)
    p_correct = (n[0] + n[3]) / n.sum()
    om = O.Model(m.winner, m.par_src, m.par_const, 6)
    grid = np.linspace(0.2 + 1e-6, 30, 200001)
    off = np.array([0, len(grid), len(grid), len(grid), len(grid)], np.int64)
    f = np.exp(O.lba_trial_terms(om, grid, off, theta))
    p = np.trapezoid(f, grid)
    assert abs(p_correct - p) < 0.006


# ------------------------------------------------------------ diagnostics
def test_psrf_and_mcse_on_known_processes():
    rng = np.random.default_rng(0)
    iid = rng.normal(size=(2, 8, 4000))
    assert np.all(np.abs(diagnostics.psrf(iid) - 1) < 0.01)
    assert abs(diagnostics.mpsrf(iid) - 1) < 0.02
    shifted = iid.copy()
    shifted[:, :4] += 1.0
    assert np.all(diagnostics.psrf(shifted) > 1.1)
    # AR(1) with rho: variance of the mean = sigma^2 / (1 - rho)^2 / n
    rho, n, m = 0.8, 20000, 4
    e = rng.normal(size=(m, n))
    x = np.zeros((m, n))
    for t in range(1, n):
        x[:, t] = rho * x[:, t - 1] + e[:, t]
    se = diagnostics.mcse(x[None, :, :])[0]
    want = 1 / (1 - rho) / np.sqrt(n * m)
    assert abs(se / want - 1) < 0.15


# -------------------------------------------------------------- workloads
def test_replicated_workload_shares_data_with_new_keys():
    # independent runs of one workload (bench.py --runs): same data and start
    # state, Philox keys seed + r, the run order kept job by job
    from ggdmc_amd import workloads

    def ll(model, data, thetas):
        om = O.Model(model.winner, model.par_src, model.par_const, model.npar)
        return np.array([O.lba_loglike(om, data.rt, data.cell_off, t) for t in thetas])

    w = workloads.make("c3", scale=1 / 16, ll_fn=ll)
    r = workloads.replicate(w, 3)
    assert len(r.jobs) == 3 * len(w.jobs)
    assert [j.seed for j in r.jobs] == [w.jobs[0].seed + k for k in range(3)]
    assert all(j.subjects is w.jobs[0].subjects and j.phi_start is w.jobs[0].phi_start for j in r.jobs)
    assert r.n_subjects == 3 * w.n_subjects and r.n_trials == 3 * w.n_trials
    assert workloads.replicate(w, 1) is w


# ------------------------------------------- posterior assembly, initialisers
def _post(run, npar=3, nchain=4, nmc=5):
    """A posterior whose every entry names its (run, parameter, chain, column)."""
    from ggdmc_amd.runs import Posterior
    p, c, m = np.meshgrid(np.arange(npar), np.arange(nchain), np.arange(nmc), indexing="ij")
    th = run * 10000 + p * 1000 + c * 100 + m
    lp = -(run * 10000 + c[0] * 100 + m[0]).astype(float)
    ll = run * 10000 + c[0] * 100 + m[0] + 0.5
    return Posterior(th.astype(float), lp, ll, nmc, 2, nchain, ["a", "b", "c"], 1)


def test_rebuild_posterior_follows_the_reference_layout():
    # R/parallel_helper.R:15-80: run i becomes chain i; its [npar, nchain, nmc]
    # theta is flattened column-major (chain fastest) into nchain*nmc columns;
    # lp / ll matrices are transposed then flattened (column fastest)
    from ggdmc_amd import sampling
    runs = [_post(i) for i in range(3)]
    r = sampling.RebuildPosterior(runs)
    assert r.theta.shape == (3, 3, 20) and r.nchain == 3 and r.nmc == 20 and r.thin == 2
    for i in range(3):
        for p in range(3):
            for c in range(4):
                for m in range(5):
                    assert r.theta[p, i, c + 4 * m] == i * 10000 + p * 1000 + c * 100 + m
                    assert r.summed_log_prior[i, m + 5 * c] == -(i * 10000 + c * 100 + m)
                    assert r.log_likelihoods[i, m + 5 * c] == i * 10000 + c * 100 + m + 0.5
    h = sampling.RebuildHyper([{"phi": x} for x in runs])
    assert np.array_equal(h.theta, r.theta) and np.array_equal(h.log_likelihoods, r.log_likelihoods)
    subs = sampling.RebuildPosteriors([{"phi": None, "subject_theta": [_post(i), _post(i + 5)]}
                                       for i in range(3)])
    assert len(subs) == 2
    assert np.array_equal(subs[0].theta, r.theta)
    assert subs[1].theta[0, 2, 0] == (2 + 5) * 10000
    with pytest.raises(ValueError):
        sampling.RebuildPosteriors([])


def _cpu_loglike(model, data, thetas, precision=None):
    om = O.Model(model.winner, model.par_src, model.par_const, model.npar)
    return np.array([O.lba_loglike(om, data.rt, data.cell_off, t) for t in np.atleast_2d(thetas)])


def test_initialise_phi_keeps_the_subject_prior_density(monkeypatch):
    # R/phi.R:256-332: subject starts come from initialise_theta under p_prior
    # and keep that lp; phi chains are the first h_prior draws with finite
    # log prior and hyper likelihood against chain c's subject starts
    from ggdmc_amd import sampling
    from ggdmc_amd.priors import sumloghlike, sumlogprior
    monkeypatch.setattr(sampling, "loglike", _cpu_loglike)
    fix = fx.load("lba_data6")
    priors = sampling.set_priors(fix.pop_prior, fix.hyper_prior)
    dmis = [(fix.model, fix.pop_data(s)) for s in range(2)]
    (phi, plp, pll), subs = sampling.initialise_phi(4, priors, dmis, seed=11, max_init_attempts=200)
    for th, lp, ll in subs:
        for c in range(4):
            assert lp[c] == sumlogprior(fix.pop_prior, th[:, c])
    for c in range(4):
        assert plp[c] == sumlogprior(fix.hyper_prior, phi[:, c])
        X = np.stack([s[0][:, c] for s in subs])
        assert pll[c] == sumloghlike(fix.pop_prior, phi[:, c], X)
        assert np.isfinite(plp[c]) and np.isfinite(pll[c])


def test_exhausted_initialisation_raises(monkeypatch):
    # R/phi.R:322-325: no valid phi after max_init_attempts draws is an error;
    # StartSampling_hyper initialises through initialise_phi and raises too
    from ggdmc_amd import sampling
    monkeypatch.setattr(sampling, "loglike", _cpu_loglike)
    fix = fx.load("lba_data6")
    P = fix.P
    bad = PriorSpec.uniform(2 * P, -2.0, -1.0)          # every scale negative: NaN likelihood
    priors = sampling.set_priors(fix.pop_prior, bad)
    dmis = [(fix.model, fix.pop_data(0))]
    with pytest.raises(RuntimeError, match="Failed to find valid phi after 7 attempts"):
        sampling.initialise_phi(3, priors, dmis, seed=1, max_init_attempts=7)
    with pytest.raises(RuntimeError, match="Failed to find valid phi"):
        sampling.StartSampling_hyper(np.ones((2, P)), dmis, priors, nchain=3, ncore=1, seed=2,
                                     max_init_attempts=5)


def test_shim_resolves_the_start_column_from_the_samples_object():
    # a fresh samples object (set_up_new_samples + initialise_theta: column 1
    # filled) starts from column 1; a finished posterior from its last column;
    # a partly filled one from its last filled column
    import shim_lib as SH
    L = SH.load()
    th = np.arange(2 * 3, dtype=float).reshape(2, 3)
    for filled, want in ((1, 0), (5, 4), (3, 2)):
        x = SH.r_samples(np.repeat(th[:, :, None], filled, 2), np.zeros((3, filled)),
                         np.zeros((3, filled)), nmc_in=5, filled=filled)[0]
        assert L.shim_start_column(_lib.ptr(x), 2, 3, 5) == want
    x = SH.r_samples(th, np.zeros(3), np.zeros(3), nmc_in=4)[0]
    x[1, 2, 3] = 1.0          # one finite entry does not make a column
    assert L.shim_start_column(_lib.ptr(x), 2, 3, 4) == 0
