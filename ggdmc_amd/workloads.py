"""The benchmark / test workloads (synthetic data, fixed seeds).

  c1  single subject, 2 accumulators, 6 free parameters, 768 trials, 18 chains
  c2  single subject, B x v design (13 parameters), 768 trials, 39 chains
  c3  hierarchical B x v: 32 subjects x 768 trials, 26 population parameters, 39 chains
  c4  256 subjects with ragged trial counts U{96..2048} (hierarchical) plus
      64 independent single-subject B x v fits of 768 trials, 39 chains
  c5  hierarchical 4-accumulator LBA (B by stimulus, v / sd_v / t0 by match):
      1024 subjects x 2048 trials, 11 parameters, 33 chains

Designs, priors and generating values follow the reference's data generators
(tests/testthat/Group1/data/1_lba_model1.r, 6_lba_Bv_model.r): uniform(0, 10)
single-subject priors, tnorm(lower 0) population distributions, uniform(0, 10)
hyper priors.
"""
from dataclasses import dataclass, replace

import numpy as np

from . import _lib
from .model import build_model
from .runs import JobSpec, PriorSpec
from .priors import sumloghlike, sumlogprior
from .simulate import rtnorm, simulate_subject

BV_SPEC = dict(
    p_map={"A": ["1"], "B": ["S", "COLOR"], "t0": ["1"], "mean_v": ["NOISE", "M"],
           "sd_v": ["M"], "st0": ["1"]},
    match_map={"M": {"left": "z_key", "right": "x_key"}},
    factors={"S": ["left", "right"], "COLOR": ["red", "blue"],
             "NOISE": ["high", "moderate", "low"]},
    constants={"st0": 0.0, "sd_v.false": 1.0},
    accumulators=["z_key", "x_key"])
BV_MEAN = {"A": 0.25, "B.left.red": 2.33, "B.right.red": 2.13, "B.left.blue": 2.55,
           "B.right.blue": 2.35, "t0": 0.05, "mean_v.high.true": 2.1,
           "mean_v.moderate.true": 3.1, "mean_v.low.true": 4.1, "mean_v.high.false": 1.82,
           "mean_v.moderate.false": 1.70, "mean_v.low.false": 1.5, "sd_v.true": 0.2}
BV_SCALE = {"A": 0.2, "B.left.red": .23, "B.right.red": .23, "B.left.blue": .255,
            "B.right.blue": .235, "t0": 0.05, "mean_v.high.true": .25,
            "mean_v.moderate.true": .35, "mean_v.low.true": .45, "mean_v.high.false": .182,
            "mean_v.moderate.false": .170, "mean_v.low.false": .15, "sd_v.true": 0.2}

C1_SPEC = dict(
    p_map={"A": ["1"], "B": ["1"], "t0": ["1"], "mean_v": ["M"], "sd_v": ["M"], "st0": ["1"]},
    match_map={"M": {"s1": "r1", "s2": "r2"}},
    factors={"S": ["s1", "s2"]},
    constants={"st0": 0.0, "sd_v.false": 1.0},
    accumulators=["r1", "r2"])
C1_TRUE = {"A": 0.25, "B": 2.33, "t0": 0.05, "mean_v.true": 3.1, "mean_v.false": 1.7,
           "sd_v.true": 0.2}

C5_SPEC = dict(
    p_map={"A": ["1"], "B": ["S"], "t0": ["M"], "mean_v": ["M"], "sd_v": ["M"], "st0": ["1"]},
    match_map={"M": {"s1": "r1", "s2": "r2", "s3": "r3", "s4": "r4"}},
    factors={"S": ["s1", "s2", "s3", "s4"]},
    constants={"st0": 0.0},
    accumulators=["r1", "r2", "r3", "r4"])
C5_MEAN = {"A": 0.3, "B.s1": 1.4, "B.s2": 1.5, "B.s3": 1.6, "B.s4": 1.7, "mean_v.true": 3.0,
           "mean_v.false": 1.0, "sd_v.true": 0.4, "sd_v.false": 1.0, "t0.true": 0.15,
           "t0.false": 0.2}
C5_SCALE = {"A": 0.05, "B.s1": 0.15, "B.s2": 0.15, "B.s3": 0.15, "B.s4": 0.15,
            "mean_v.true": 0.3, "mean_v.false": 0.2, "sd_v.true": 0.05, "sd_v.false": 0.1,
            "t0.true": 0.02, "t0.false": 0.02}


@dataclass
class Workload:
    name: str
    description: str
    model: object
    jobs: list
    nchain: int
    iterations: int
    sub_migration_prob: float = 0.0
    pop_migration_prob: float = 0.0

    @property
    def n_subjects(self):
        return sum(1 if j.kind == _lib.KIND_SUBJECT else len(j.subjects or []) for j in self.jobs)

    @property
    def n_trials(self):
        t = 0
        for j in self.jobs:
            if j.kind == _lib.KIND_SUBJECT:
                t += len(j.subject.rt)
            elif j.subjects:
                t += sum(len(s.rt) for s in j.subjects)
        return t

    def config(self, nmc, precision=_lib.MIXED_BATCHED, **kw):
        from .runs import make_config
        return make_config(nmc=nmc, nchain=self.nchain, precision=precision,
                           sub_migration_prob=self.sub_migration_prob,
                           pop_migration_prob=self.pop_migration_prob, **kw)


def _vec(names, d):
    return np.array([d[p] for p in names])


def _gpu_ll(model, data, thetas):
    """Start-state log-likelihoods [n] of thetas [n, P] (GPU, FP64 arithmetic)."""
    from .runs import loglike
    return loglike(model, data, thetas, _lib.FP64_PLAIN)


def _subject_start(model, prior, data, truth, nchain, rng, ll_fn, scale=0.05):
    """Start near the generating values; lp from the prior, ll from ll_fn."""
    th = truth[:, None] * np.exp(rng.normal(0, scale, (len(truth), nchain)))
    lp = np.array([sumlogprior(prior, th[:, c]) for c in range(nchain)])
    ll = np.asarray(ll_fn(model, data, th.T))
    return th, lp, ll


def _hier_job(model, pop_mean, pop_scale, n_trials_list, nchain, seed, data_seed, ll_fn):
    rng = np.random.default_rng(data_seed)
    pn = model.pnames
    P = len(pn)
    mean = _vec(pn, pop_mean)
    scale = _vec(pn, pop_scale)
    n_cond = model.n_cell // model.n_acc
    subjects, thetas = [], []
    for ntr in n_trials_list:
        th = rtnorm(mean, scale, 0.0, rng)
        thetas.append(th)
        subjects.append(simulate_subject(model, th, max(1, int(round(ntr / n_cond))), rng))
    p_prior = PriorSpec.tnorm(np.zeros(P), 10.0, 0.0, np.inf, pn)
    hp_names = ["loc_" + p for p in pn] + ["sca_" + p for p in pn]
    h_prior = PriorSpec.uniform(2 * P, 0.0, 10.0, hp_names)
    # start: phi near (mean, scale); subject chains near their generating values
    phi = np.concatenate([mean[:, None] * np.exp(rng.normal(0, 0.05, (P, nchain))),
                          scale[:, None] * np.exp(rng.normal(0, 0.05, (P, nchain)))])
    starts = []
    for th, d in zip(thetas, subjects):
        x = th[:, None] * np.exp(rng.normal(0, 0.05, (P, nchain)))
        lp = np.empty(nchain)
        ll = np.empty(nchain)
        for c in range(nchain):
            pc = PriorSpec.tnorm(phi[:P, c], phi[P:, c], 0.0, np.inf)
            lp[c] = sumlogprior(pc, x[:, c])
        ll = np.asarray(ll_fn(model, d, x.T))
        starts.append((x, lp, ll))
    plp = np.array([sumlogprior(h_prior, phi[:, c]) for c in range(nchain)])
    pll = np.array([sumloghlike(p_prior, phi[:, c], np.stack([s[0][:, c] for s in starts]))
                    for c in range(nchain)])
    job = JobSpec(_lib.KIND_HIER, seed, p_prior, h_prior=h_prior, phi_start=(phi, plp, pll),
                  subjects=subjects, subject_starts=starts, nparameter_de=2 * P)
    return job, thetas


def _single_job(model, truth_d, n_trials, nchain, seed, data_seed, ll_fn):
    rng = np.random.default_rng(data_seed)
    pn = model.pnames
    truth = _vec(pn, truth_d)
    n_cond = model.n_cell // model.n_acc
    data = simulate_subject(model, truth, n_trials // n_cond, rng)
    prior = PriorSpec.uniform(len(pn), 0.0, 10.0, pn)
    start = _subject_start(model, prior, data, truth, nchain, rng, ll_fn)
    return JobSpec(_lib.KIND_SUBJECT, seed, prior, start=start, subject=data,
                   nparameter_de=len(pn)), truth


def make(name, scale=1.0, ll_fn=None):
    """Builds workload c1..c5. scale < 1 shrinks subject / trial counts (tests).
    ll_fn(model, data, thetas [n, P]) -> ll [n] evaluates start states (default:
    the library's FP64 likelihood on the GPU)."""
    name = name.lower()
    ll_fn = ll_fn or _gpu_ll
    if name == "c1":
        m = build_model(**C1_SPEC)
        job, _ = _single_job(m, C1_TRUE, 768, 18, 9032, 9033, ll_fn)
        return Workload("c1", "single subject 2-acc LBA, 6 params, 768 trials, 18 chains",
                        m, [job], 18, 500)
    if name == "c2":
        m = build_model(**BV_SPEC)
        job, _ = _single_job(m, BV_MEAN, 768, 39, 9032, 9034, ll_fn)
        return Workload("c2", "single subject B x v LBA, 13 params, 768 trials, 39 chains",
                        m, [job], 39, 2000, sub_migration_prob=0.06)
    if name == "c3":
        m = build_model(**BV_SPEC)
        S = max(2, int(round(32 * scale)))
        job, _ = _hier_job(m, BV_MEAN, BV_SCALE, [768] * S, 39, 9032, 9035, ll_fn)
        return Workload("c3", "hierarchical B x v LBA, %d subjects x 768 trials, 13 params "
                        "(26 population), 39 chains" % S, m, [job], 39, 3000)
    if name == "c4":
        m = build_model(**BV_SPEC)
        rng = np.random.default_rng(9036)
        S = max(2, int(round(256 * scale)))
        F = max(1, int(round(64 * scale)))
        ntr = rng.integers(96, 2049, S)
        job, _ = _hier_job(m, BV_MEAN, BV_SCALE, list(ntr), 39, 9032, 9036, ll_fn)
        jobs = [job] + [_single_job(m, BV_MEAN, 768, 39, k + 1, 10000 + k, ll_fn)[0]
                        for k in range(F)]
        return Workload("c4", "%d ragged subjects U{96..2048} (hierarchical) + %d single-subject "
                        "B x v fits, 39 chains" % (S, F), m, jobs, 39, 1000)
    if name == "c5":
        m = build_model(**C5_SPEC)
        S = max(2, int(round(1024 * scale)))
        T = max(64, int(round(2048 * scale)))
        job, _ = _hier_job(m, C5_MEAN, C5_SCALE, [T] * S, 33, 9032, 9037, ll_fn)
        return Workload("c5", "hierarchical 4-acc LBA, %d subjects x %d trials, 11 params, "
                        "33 chains" % (S, T), m, [job], 33, 5000)
    raise ValueError("unknown workload " + name)


def replicate(w, runs):
    """The workload `runs` times over in one batch: every job repeated with
    the same data and start state under Philox keys seed + r (r = 0..runs-1),
    the way independent runs of one model are launched side by side."""
    if runs <= 1:
        return w
    jobs = [replace(j, seed=j.seed + r) for r in range(runs) for j in w.jobs]
    return Workload("%sx%d" % (w.name, runs), "%d independent runs of %s" % (runs, w.description),
                    w.model, jobs, w.nchain, w.iterations, w.sub_migration_prob,
                    w.pop_migration_prob)
