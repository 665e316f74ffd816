"""The CPU oracle pinned against the reference's golden vectors and against
what the mathematics fixes (CPU only)."""
import numpy as np
import pytest
from scipy import integrate, stats

import fixtures as fx
import oracle_lib as O


# ---------------------------------------------------------------- Philox
def test_philox_known_answers():
    # Philox4x32-10 known-answer vectors of the Random123 distribution
    assert [hex(x) for x in O.philox([0, 0, 0, 0], [0, 0])] == \
        ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    assert [hex(x) for x in O.philox([0xffffffff] * 4, [0xffffffff] * 2)] == \
        ["0x408f276d", "0x41c83b0e", "0xa20bc7c6", "0x6d5451fd"]
    assert [hex(x) for x in O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                                     [0xa4093822, 0x299f31d0])] == \
        ["0xd16cfe09", "0x94fdcceb", "0x5001e420", "0x24126ea1"]


# -------------------------------------------------------- goldens (likelihood)
@pytest.mark.parametrize("name", fx.FIXTURES)
def test_loglike_goldens(name):
    fix = fx.load(name)
    errs = []
    for k in range(fix.sub_theta.shape[1]):
        v = O.lba_loglike(fix.omodel, fix.sub_data.rt, fix.sub_data.cell_off, fix.sub_theta[:, k])
        errs.append(abs(v / fix.sub_ll[k] - 1))
    for s in range(fix.n_subjects):
        d = fix.pop_data(s)
        for k in range(fix.z["pop_theta"].shape[2]):
            v = O.lba_loglike(fix.omodel, d.rt, d.cell_off, fix.z["pop_theta"][s, :, k])
            errs.append(abs(v / fix.z["pop_ll"][s, k] - 1))
    assert max(errs) < 1e-13, max(errs)


@pytest.mark.parametrize("name", fx.FIXTURES)
def test_prior_goldens(name):
    fix = fx.load(name)
    # subject prior U(0, 10): summed_log_prior of the single-subject start
    for k in range(fix.sub_theta.shape[1]):
        assert abs(O.sumlogprior(fix.oprior, fix.sub_theta[:, k]) - fix.sub_lp[k]) < 1e-12
    # population start: subject lp under the tnorm p_prior, phi lp under h_prior,
    # phi ll = sumloghlike over subjects (src/de.cpp:245-270)
    pp = fx.oprior(fix.pop_prior)
    hp = fx.oprior(fix.hyper_prior)
    th = fix.z["pop_theta"]
    for s in range(fix.n_subjects):
        for k in range(th.shape[2]):
            assert abs(O.sumlogprior(pp, th[s, :, k]) / fix.z["pop_lp"][s, k] - 1) < 1e-13
    for k in range(fix.z["phi_theta"].shape[1]):
        phi = fix.z["phi_theta"][:, k]
        assert abs(O.sumlogprior(hp, phi) - fix.z["phi_lp"][k]) < 1e-12
        v = O.sumloghlike(pp, phi, th[:, :, k])
        assert abs(v / fix.z["phi_ll"][k] - 1) < 1e-13


def test_tnorm_and_uniform_against_scipy():
    rng = np.random.default_rng(0)
    for _ in range(200):
        mu, sd = rng.uniform(-2, 3), rng.uniform(0.1, 3)
        lo, up = 0.0, rng.choice([np.inf, 5.0])
        x = rng.uniform(0, 4)
        pr = O.Prior(np.array([1], np.int32), [mu], [sd], [lo], [up])
        want = stats.truncnorm.logpdf(x, (lo - mu) / sd, (up - mu) / sd, loc=mu, scale=sd)
        assert abs(O.sumlogprior(pr, [x]) - want) < 1e-10 * max(1, abs(want))
    pr = O.Prior(np.array([6], np.int32), [0.0], [10.0], [0.0], [10.0])
    assert O.sumlogprior(pr, [3.0]) == pytest.approx(-np.log(10))
    assert O.sumlogprior(pr, [-0.1]) == -np.inf
    assert O.sumlogprior(pr, [10.1]) == -np.inf


# --------------------------------------------------- likelihood mathematics
def _two_acc(A=0.5, B=1.0, v=(2.0, 1.2), s=(0.8, 1.0), t0=0.2):
    """Two cells (response 0, response 1) of a 2-accumulator model with all
    values as constants; theta is unused."""
    src = np.full((2, 2, 6), -1, np.int32)
    const = np.zeros((2, 2, 6))
    for c in range(2):
        for j in range(2):
            const[c, j] = [A, B, v[j], s[j], 0.0, t0]
    return O.Model(np.array([0, 1], np.int32), src, const, 1)


def test_defective_densities_integrate_to_one():
    # sum over responses of int f_w(t) S_other(t) dt = 1 for positive-truncated drifts
    m = _two_acc()
    t = np.linspace(0.2 + 1e-6, 40.0, 400001)
    tot = 0.0
    for w in (0, 1):
        off = np.array([0, 0, 0], np.int64)
        off[w + 1:] = len(t)
        terms = O.lba_trial_terms(m, t, off, np.zeros(1))
        tot += integrate.trapezoid(np.exp(terms), t)
    assert abs(tot - 1) < 2e-4, tot


def test_cdf_is_integral_of_pdf():
    # winner t <= 0 (t0 of accumulator 1 huge): the term is log(1e-10) + log(1 - F_0(t))
    A, B, v, s, t0 = 0.6, 1.4, 1.7, 0.9, 0.1
    src = np.full((1, 2, 6), -1, np.int32)
    const = np.zeros((1, 2, 6))
    const[0, 0] = [A, B, v, s, 0.0, t0]
    const[0, 1] = [A, B, v, s, 0.0, 1e6]
    m_cdf = O.Model(np.array([1], np.int32), src, const, 1)
    m_pdf = O.Model(np.array([0], np.int32), src, const, 1)   # winner 0: log pdf_0 + log(1-1e-10)
    grid = np.linspace(t0 + 1e-7, 6.0, 200001)
    off = np.array([0, len(grid)], np.int64)
    pdf = np.exp(O.lba_trial_terms(m_pdf, grid, off, np.zeros(1)) - np.log1p(-1e-10))
    F_int = integrate.cumulative_trapezoid(pdf, grid, initial=0.0)
    chk = np.array([0.5, 1.0, 2.0, 4.0])
    terms = O.lba_trial_terms(m_cdf, chk, np.array([0, 4], np.int64), np.zeros(1))
    F = 1 - np.exp(terms - np.log(1e-10))
    assert np.allclose(F, np.interp(chk, grid, F_int), atol=2e-6)


def test_floor_and_dbl_eps_rules():
    fix = fx.load("lba_data6")
    th = fx.truth_vector(fix).copy()
    th[fix.model.pnames.index("t0")] = 50.0
    n = len(fix.sub_data.rt)
    want = n * (np.log(1e-10) + np.log(1 - 1e-10))
    assert abs(O.lba_loglike(fix.omodel, fix.sub_data.rt, fix.sub_data.cell_off, th) - want) < 1e-9
    # a loser whose CDF rounds to 1: the whole trial becomes log(DBL_EPSILON)
    m = _two_acc(A=0.1, B=0.2, v=(0.5, 30.0), s=(1.0, 0.05), t0=0.0)
    terms = O.lba_trial_terms(m, np.array([5.0]), np.array([0, 1, 1], np.int64), np.zeros(1))
    assert terms[0] == pytest.approx(np.log(np.finfo(float).eps))


def test_golden_floored_count_matches_closed_form():
    # the many prior-draw vectors with every trial floored equal the closed form
    fix = fx.load("lba_data6")
    n = len(fix.sub_data.rt)
    closed = n * (np.log(1e-10) + np.log(1 - 1e-10))
    assert np.sum(np.abs(fix.sub_ll - closed) < 1e-6) >= 1


# ------------------------------------------------------------- DE sampler
def _lp(pr, th):
    return np.array([O.sumlogprior(pr, th[:, k]) for k in range(th.shape[1])])


def _flat(n_par):
    """A model with no observations: the likelihood is 0 for any theta."""
    m = O.Model(np.zeros(n_par, np.int32), np.zeros((n_par, 1, 6), np.int32),
                np.zeros((n_par, 1, 6)), n_par)
    off = np.zeros(n_par + 1, np.int64)
    pr = O.Prior(np.full(n_par, 6, np.int32), np.full(n_par, -1e6), np.full(n_par, 1e6),
                 np.full(n_par, -1e6), np.full(n_par, 1e6))
    return m, off, pr


def test_crossover_sequential_partners_and_gamma():
    # flat target, rp = 0: every proposal is accepted and equals
    # theta_i + gamma (theta_m - theta_n) with m != n != i, chains updated in place
    n, P = 3, 1
    m, off, pr = _flat(P)
    th0 = np.array([[0.0, 1.0, 5.0]])
    cfg = O.config(nmc=2, nchain=n, nparameter_de=P, rp=0.0, sweep_groups=n)
    th, lp, ll = O.run_subject(cfg, m, pr, np.zeros(0), off, th0, _lp(pr, th0), np.zeros(n),
                               O.MODEL_NORMAL)
    g = 2.38 / np.sqrt(2 * P)
    x = th0[0].copy()
    out = th[0, :, 1]
    c0 = {x[0] + g * (x[1] - x[2]), x[0] - g * (x[1] - x[2])}
    assert min(abs(out[0] - v) for v in c0) < 1e-12
    x[0] = out[0]
    c1 = {x[1] + g * (x[0] - x[2]), x[1] - g * (x[0] - x[2])}
    assert min(abs(out[1] - v) for v in c1) < 1e-12
    x[1] = out[1]
    c2 = {x[2] + g * (x[0] - x[1]), x[2] - g * (x[0] - x[1])}
    assert min(abs(out[2] - v) for v in c2) < 1e-12


def test_two_groups_use_partners_outside_the_group():
    n, P = 4, 1
    m, off, pr = _flat(P)
    th0 = np.array([[0.0, 1.0, 3.0, 7.0]])
    cfg = O.config(nmc=2, nchain=n, nparameter_de=P, rp=0.0, sweep_groups=2)
    th, _, _ = O.run_subject(cfg, m, pr, np.zeros(0), off, th0, _lp(pr, th0), np.zeros(n),
                             O.MODEL_NORMAL)
    g = 2.38 / np.sqrt(2)
    x = th0[0]
    out = th[0, :, 1]
    for i in (0, 1):   # group {0, 1} draws from {2, 3}
        assert min(abs(out[i] - (x[i] + s * g * (x[2] - x[3]))) for s in (1, -1)) < 1e-12
    for i in (2, 3):   # group {2, 3} draws from the updated {0, 1}
        assert min(abs(out[i] - (x[i] + s * g * (out[0] - out[1]))) for s in (1, -1)) < 1e-12


def test_migration_ring_copies_the_lowest_member():
    # flat target, rp = 0, migration every iteration: the sorted subset's ring
    # copies S_k into S_{k+1}, so every member ends equal to the lowest one
    n, P = 9, 1
    m, off, pr = _flat(P)
    th0 = np.arange(n, dtype=float)[None, :] * 10
    cfg = O.config(nmc=2, nchain=n, nparameter_de=P, rp=0.0, sub_migration_prob=1.0,
                   sweep_groups=n)
    for seed in range(20):
        cfg.seed = seed
        th, _, _ = O.run_subject(cfg, m, pr, np.zeros(0), off, th0, _lp(pr, th0), np.zeros(n),
                                 O.MODEL_NORMAL)
        out = th[0, :, 1]
        changed = np.nonzero(out != th0[0])[0]
        assert len(changed) >= 1
        src = out[changed[0]]
        assert np.all(out[changed] == src)
        lowest = int(np.nonzero(th0[0] == src)[0][0])
        assert lowest < changed.min() and out[lowest] == src


def test_nan_posterior_is_rejected():
    n, P = 5, 1
    m, off, _ = _flat(P)
    pr = O.Prior(np.array([1], np.int32), [0.0], [-1.0], [-1e6], [1e6])   # scale < 0 -> NaN
    th0 = np.arange(n, dtype=float)[None, :]
    cfg = O.config(nmc=5, nchain=n, nparameter_de=P, sweep_groups=n)
    th, _, _ = O.run_subject(cfg, m, pr, np.zeros(0), off, th0, np.zeros(n), np.zeros(n),
                             O.MODEL_NORMAL)
    assert np.all(th[0, :, -1] == th0[0])


@pytest.mark.parametrize("groups", [None, 2])
def test_normal_mean_posterior(groups):
    # x_ck ~ N(mu_c, 1), flat prior: posterior mu_c ~ N(mean_c, 1/n_c)
    rng = np.random.default_rng(4)
    ncell = 2
    x = [rng.normal(1.0, 1, 40), rng.normal(-0.5, 1, 90)]
    obs = np.concatenate(x)
    off = np.array([0, 40, 130], np.int64)
    m = O.Model(np.zeros(ncell, np.int32), np.zeros((ncell, 1, 6), np.int32),
                np.zeros((ncell, 1, 6)), ncell)
    pr = O.Prior(np.full(ncell, 6, np.int32), [-10, -10], [10, 10], [-10, -10], [10, 10])
    n = 12
    th0 = rng.normal(0, 1, (ncell, n))
    lp0 = np.array([O.sumlogprior(pr, th0[:, k]) for k in range(n)])
    ll0 = np.array([np.sum([-0.5 * np.sum((x[c] - th0[c, k]) ** 2) - 0.5 * len(x[c]) * np.log(2 * np.pi)
                            for c in range(ncell)]) for k in range(n)])
    cfg = O.config(nmc=4000, nchain=n, nparameter_de=ncell, sub_migration_prob=0.05,
                   sweep_groups=groups)
    th, _, _ = O.run_subject(cfg, m, pr, obs, off, th0, lp0, ll0, O.MODEL_NORMAL)
    draws = th[:, :, 500:].reshape(ncell, -1)
    for c in range(ncell):
        mean, sd = x[c].mean(), 1 / np.sqrt(len(x[c]))
        # 500 effective draws is conservative for 12 chains x 3500 iterations
        assert abs(draws[c].mean() - mean) < 4 * sd / np.sqrt(500)
        assert abs(draws[c].std() / sd - 1) < 0.1
