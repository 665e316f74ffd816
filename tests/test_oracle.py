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


# ------------------------------------------- hierarchical / hyper-only samplers
# The oracle's run_hier / run_hyper (oracle/oracle.cpp) against a brute-force
# pure-Python replay written from src/de.cpp (tests/replay.py), which
# re-derives every proposal and accept from the Philox words and the previous
# state; and against invariants of the reference's wiring that a plausible
# mistake (no refresh, a refreshed subject prior, the destination's context in
# migration, subjects before the population) breaks.
import replay as R  # noqa: E402


def test_replay_philox_known_answers():
    assert R.philox((0, 0, 0, 0), (0, 0)) == (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)
    assert R.philox((0xffffffff,) * 4, (0xffffffff,) * 2) == \
        (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)


def _tiny_hier(nchain=5, S=2, per_cell=2, seed=21):
    """lba_data6 design, S subjects of per_cell trials per cell, a start near
    the generating values (phi [2P, n], subjects [S, P, n] with lp under phi
    chain c and ll from the oracle)."""
    fix = fx.load("lba_data6")
    P = fix.P
    rts, offs = [], []
    for s in range(S):
        d = fix.pop_data(s)
        co = d.cell_off
        take = [np.asarray(d.rt[co[c]:co[c + 1]][:per_cell]) for c in range(len(co) - 1)]
        rts.append(np.concatenate(take))
        offs.append(np.concatenate([[0], np.cumsum([len(t) for t in take])]).astype(np.int64))
    rng = np.random.default_rng(seed)
    tv = fx.truth_vector(fix)
    sc = np.array([fx.BV_SCALE[p] for p in fix.model.pnames])
    phi = np.concatenate([tv[:, None] * np.exp(rng.normal(0, 0.1, (P, nchain))),
                          sc[:, None] * np.exp(rng.normal(0, 0.1, (P, nchain)))])
    pp, hp = fix.pop_prior, fix.hyper_prior
    ppr = (pp.dist, pp.p0, pp.p1, pp.lower, pp.upper)
    hpr = (hp.dist, hp.p0, hp.p1, hp.lower, hp.upper)
    sub = np.empty((S, P, nchain))
    slp = np.empty((S, nchain))
    sll = np.empty((S, nchain))
    for s in range(S):
        sub[s] = tv[:, None] * np.exp(rng.normal(0, 0.1, (P, nchain)))
        for c in range(nchain):
            slp[s, c] = R.sumlogprior(R.phi_prior(phi[:, c], ppr), sub[s, :, c])
            sll[s, c] = O.lba_loglike(fix.omodel, rts[s], offs[s], sub[s, :, c])
    plp = np.array([R.sumlogprior(hpr, phi[:, c]) for c in range(nchain)])
    pll = np.array([R.sumloghlike(ppr, phi[:, c], [sub[s, :, c] for s in range(S)])
                    for c in range(nchain)])
    return dict(fix=fix, P=P, S=S, n=nchain, rts=rts, offs=offs, phi=phi, plp=plp, pll=pll,
                sub=sub, slp=slp, sll=sll, ppr=ppr, hpr=hpr)


def _oracle_hier(t, cfg):
    rt = np.concatenate(t["rts"])
    so = np.concatenate([[0], np.cumsum([len(r) for r in t["rts"]])]).astype(np.int64)
    co = np.stack(t["offs"])
    fix = t["fix"]
    return O.run_hier(cfg, fix.omodel, fx.oprior(fix.pop_prior), fx.oprior(fix.hyper_prior), rt,
                      so, co, t["phi"], t["plp"], t["pll"], t["sub"], t["slp"], t["sll"])


def _replay_hier(t, cfg, nmc):
    fix = t["fix"]
    rcfg = dict(seed=cfg.seed, iter_offset=cfg.iter_offset, rp=cfg.rp,
                gamma=cfg.gamma_precursor / np.sqrt(2.0 * cfg.nparameter_de),
                pop_mig=cfg.pop_migration_prob, sub_mig=cfg.sub_migration_prob,
                groups=cfg.sweep_groups, hblocked=bool(cfg.is_hblocked),
                pblocked=bool(cfg.is_pblocked))
    phi = R.Level(t["phi"], t["plp"], t["pll"])
    subs = [R.Level(t["sub"][s], t["slp"][s], t["sll"][s]) for s in range(t["S"])]

    def sll(s, x):
        return O.lba_loglike(fix.omodel, t["rts"][s], t["offs"][s], x)

    states = [(phi.copy(), [x.copy() for x in subs])]
    for it in range(1, nmc):
        R.hier_iteration(rcfg, phi, subs, t["ppr"], t["hpr"], sll, it)
        states.append((phi.copy(), [x.copy() for x in subs]))
    return states


def _same_as_replay(out, states):
    pth, plp, pll, sth, slp, sll = out
    for k, (phi, subs) in enumerate(states):
        np.testing.assert_allclose(pth[:, :, k], phi.th, rtol=1e-13, atol=0, err_msg="phi %d" % k)
        np.testing.assert_allclose(plp[:, k], phi.lp, rtol=1e-13, err_msg="phi lp %d" % k)
        np.testing.assert_allclose(pll[:, k], phi.ll, rtol=1e-12, err_msg="phi ll %d" % k)
        for s, x in enumerate(subs):
            np.testing.assert_allclose(sth[s, :, :, k], x.th, rtol=1e-13, atol=0,
                                       err_msg="subject %d, %d" % (s, k))
            np.testing.assert_allclose(slp[s, :, k], x.lp, rtol=1e-12, err_msg="lp %d, %d" % (s, k))
            np.testing.assert_allclose(sll[s, :, k], x.ll, rtol=1e-13, err_msg="ll %d, %d" % (s, k))


def _moves(out):
    """Fractions of (chain, iteration) steps in which the population and the
    subject states changed."""
    pth, _, _, sth, _, _ = out
    pm = np.mean(np.any(np.diff(pth, axis=2) != 0, axis=0))
    sm = np.mean(np.any(np.diff(sth, axis=3) != 0, axis=1))
    return pm, sm


@pytest.mark.parametrize("groups,mig", [(5, 0.25), (2, 0.25), (5, 1.0)])
def test_oracle_hier_follows_the_replay(groups, mig):
    t = _tiny_hier()
    nmc = 30
    cfg = O.config(nmc=nmc, nchain=t["n"], nparameter_de=2 * t["P"], sweep_groups=groups,
                   pop_migration_prob=mig, sub_migration_prob=mig, seed=77, iter_offset=3)
    out = _oracle_hier(t, cfg)
    _same_as_replay(out, _replay_hier(t, cfg, nmc))
    pm, sm = _moves(out)
    assert pm > 0.05 and sm > 0.05, (pm, sm)     # the trajectories exercise accepts


def test_oracle_hier_blocked_follows_the_replay():
    t = _tiny_hier(S=2, per_cell=1)
    nmc = 6
    cfg = O.config(nmc=nmc, nchain=t["n"], nparameter_de=2 * t["P"], sweep_groups=2,
                   pop_migration_prob=0.1, sub_migration_prob=0.1, is_hblocked=True,
                   is_pblocked=True, seed=5)
    out = _oracle_hier(t, cfg)
    _same_as_replay(out, _replay_hier(t, cfg, nmc))
    pm, sm = _moves(out)
    assert pm > 0.05 and sm > 0.05, (pm, sm)


@pytest.mark.parametrize("blocked,mig", [(False, 0.3), (True, 0.3)])
def test_oracle_hyper_follows_the_replay(blocked, mig):
    t = _tiny_hier(nchain=6, S=8)
    data = np.stack([t["sub"][s][:, 0] for s in range(t["S"])])     # [S, P] subject vectors
    fix = t["fix"]
    nmc = 25 if not blocked else 6
    cfg = O.config(nmc=nmc, nchain=t["n"], nparameter_de=2 * t["P"], sweep_groups=t["n"],
                   sub_migration_prob=mig, is_pblocked=blocked, seed=13)
    pll = np.array([R.sumloghlike(t["ppr"], t["phi"][:, c], list(data)) for c in range(t["n"])])
    th, lp, ll = O.run_hyper(cfg, fx.oprior(fix.pop_prior), fx.oprior(fix.hyper_prior), data,
                             t["phi"], t["plp"], pll)
    rcfg = dict(seed=cfg.seed, iter_offset=0, rp=cfg.rp, mig=mig, groups=t["n"],
                gamma=cfg.gamma_precursor / np.sqrt(2.0 * cfg.nparameter_de), blocked=blocked)
    states = R.run_single(rcfg, R.Level(t["phi"], t["plp"], pll),
                          lambda x, i: R.sumlogprior(t["hpr"], x),
                          lambda x, i: R.sumloghlike(t["ppr"], x, list(data)), nmc)
    for k, L in enumerate(states):
        np.testing.assert_allclose(th[:, :, k], L.th, rtol=1e-13, atol=0)
        np.testing.assert_allclose(lp[:, k], L.lp, rtol=1e-13)
        np.testing.assert_allclose(ll[:, k], L.ll, rtol=1e-12)
    assert np.mean(np.any(np.diff(th, axis=2) != 0, axis=0)) > 0.05


def test_oracle_subject_run_follows_the_replay():
    fix = fx.load("lba_data6")
    start = fx.near_start(fix, 7, seed=3)
    nmc = 25
    cfg = O.config(nmc=nmc, nchain=7, nparameter_de=fix.P, sweep_groups=7,
                   sub_migration_prob=0.3, seed=99)
    th, lp, ll = O.run_subject(cfg, fix.omodel, fix.oprior, fix.sub_data.rt,
                               fix.sub_data.cell_off, *start)
    sp = fix.sub_prior
    rcfg = dict(seed=99, iter_offset=0, rp=cfg.rp, mig=0.3, groups=7, blocked=False,
                gamma=2.38 / np.sqrt(2.0 * fix.P))
    states = R.run_single(rcfg, R.Level(*start),
                          lambda x, i: R.sumlogprior((sp.dist, sp.p0, sp.p1, sp.lower, sp.upper), x),
                          lambda x, i: O.lba_loglike(fix.omodel, fix.sub_data.rt,
                                                     fix.sub_data.cell_off, x), nmc)
    for k, L in enumerate(states):
        np.testing.assert_allclose(th[:, :, k], L.th, rtol=1e-13, atol=0)
        np.testing.assert_allclose(ll[:, k], L.ll, rtol=1e-13)
    assert np.mean(np.any(np.diff(th, axis=2) != 0, axis=0)) > 0.05


def test_oracle_hier_reference_wiring_invariants():
    """Wiring of run_hchains, column by column of one oracle run (crossover
    only; thin 1): column k holds the state after iteration k."""
    t = _tiny_hier(nchain=6, S=3)
    nmc = 40
    cfg = O.config(nmc=nmc, nchain=t["n"], nparameter_de=2 * t["P"], sweep_groups=t["n"],
                   seed=8)
    pth, plp, pll, sth, slp, sll = _oracle_hier(t, cfg)
    S, n = t["S"], t["n"]
    stale = 0
    for k in range(1, nmc):
        for c in range(n):
            # the population step refreshes chain c's likelihood from the
            # subjects' columns of the previous iteration before it uses it
            # (src/de.cpp:397-398): whatever it decided, the stored ll is the
            # hyper likelihood of its state against those columns
            prev_cols = [sth[s, :, c, k - 1] for s in range(S)]
            assert abs(pll[c, k] - R.sumloghlike(t["ppr"], pth[:, c, k], prev_cols)) <= 1e-11
            for s in range(S):
                x = sth[s, :, c, k]
                under_new = R.sumlogprior(R.phi_prior(pth[:, c, k], t["ppr"]), x)
                if np.any(x != sth[s, :, c, k - 1]):
                    # an accepted subject proposal takes its prior under this
                    # iteration's population state: the population step ran
                    # first (src/de.cpp:283-374, :599-600)
                    assert abs(slp[s, c, k] - under_new) <= 1e-11
                else:
                    # a rejected one keeps the cached lp, not refreshed when
                    # the population state moved (src/de.cpp:577, :606)
                    assert slp[s, c, k] == slp[s, c, k - 1]
                    if np.any(pth[:, c, k] != pth[:, c, k - 1]) and \
                            abs(slp[s, c, k] - under_new) > 1e-9 * abs(under_new):
                        stale += 1
    assert stale > 10


def test_oracle_hier_migration_takes_the_source_context():
    """Migration every iteration at both levels (src/de.cpp:467-565,
    :615-665). Population ring over the sorted subset S: step k refreshes the
    cached likelihoods of S_k and S_k+1 from their own subject columns, then
    proposes a jittered copy of S_k for S_k+1 whose likelihood is taken against
    S_k's subject columns. So after the ring: chains outside S keep their
    cached likelihood; S_1.. (each refreshed again as the source of the next
    step) hold the hyper likelihood of their state against their own columns;
    S_0, written by the last step, holds, if that step accepted, the hyper
    likelihood against the last member's columns. Subject ring: an accepted
    copy's prior is under the source chain's population state."""
    t = _tiny_hier(nchain=6, S=2)
    nmc = 30
    cfg = O.config(nmc=nmc, nchain=t["n"], nparameter_de=2 * t["P"], sweep_groups=t["n"],
                   pop_migration_prob=1.0, sub_migration_prob=1.0, rp=0.05, seed=4)
    pth, plp, pll, sth, slp, sll = _oracle_hier(t, cfg)
    S, n = t["S"], t["n"]
    seen_pop = seen_sub = 0

    def hll(c, k, cols_chain):
        return R.sumloghlike(t["ppr"], pth[:, c, k], [sth[s, :, cols_chain, k - 1] for s in range(S)])

    for k in range(1, nmc):
        ring = R.subset(R.Stream(cfg.seed, k, R.UNIT_PHI, 0), n)
        for c in range(n):
            if c not in ring:
                assert np.array_equal(pth[:, c, k], pth[:, c, k - 1]) and pll[c, k] == pll[c, k - 1]
        for j, c in enumerate(ring):
            if j > 0:
                assert abs(pll[c, k] - hll(c, k, c)) <= 1e-11
            elif np.any(pth[:, c, k] != pth[:, c, k - 1]):
                last = ring[-1]
                want, own = hll(c, k, last), hll(c, k, c)
                assert abs(pll[c, k] - want) <= 1e-11
                if abs(want - own) > 1e-9:
                    seen_pop += 1
            else:
                assert abs(pll[c, k] - hll(c, k, c)) <= 1e-11
        for s in range(S):
            ring = R.subset(R.Stream(cfg.seed, k, s, 0), n)
            for j, dst in enumerate(ring):
                src = ring[j - 1]
                x = sth[s, :, dst, k]
                if np.any(x != sth[s, :, dst, k - 1]):
                    want = R.sumlogprior(R.phi_prior(pth[:, src, k], t["ppr"]), x)
                    own = R.sumlogprior(R.phi_prior(pth[:, dst, k], t["ppr"]), x)
                    assert abs(slp[s, dst, k] - want) <= 1e-11
                    if abs(want - own) > 1e-9:
                        seen_sub += 1
    assert seen_pop > 3 and seen_sub > 3, (seen_pop, seen_sub)
