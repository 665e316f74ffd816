"""The R-facing sampling API of ggdmc, on the HIP engine.

Mirrors R/sampling.R (StartSampling :247, RestartSampling :308,
StartSampling_subject :446, RestartSampling_subject :507, StartSampling_hyper
:615, RestartSampling_hyper :681) and the initialisers of R/phi.R
(initialise_theta :141, initialise_phi :256), with the same argument names,
defaults and meaning. Differences from the R functions:
  * the `ncore` independent runs (parallel_lapply's forked processes,
    R/sampling.R:13-121) run as one batch on the GPU;
  * draws come from Philox streams keyed by each run's seed; a restart
    continues the stream after the iterations already run (iter_offset)
    instead of replaying it;
  * `precision` / `sweep_groups` select the engine path (default: batched
    mixed precision with two chain groups; FP64_PLAIN with
    sweep_groups = nchain is the reference's arithmetic and order).
A "dmi" here is (LBAModel, SubjectData); "priors" is a dict with p_prior,
h_prior (or None) and nparameter, like the S4 prior object.
"""
import numpy as np

from . import _lib
from .priors import logdprior, rprior, sumloghlike, sumlogprior
from .runs import JobSpec, PriorSpec, Posterior, Session, format_trace, loglike, make_config


def set_priors(p_prior, h_prior=None):
    """ggdmcPrior::set_priors: nparameter is 2P with a population prior."""
    n = p_prior.npar if h_prior is None else h_prior.npar
    return {"p_prior": p_prior, "h_prior": h_prior, "nparameter": n}


def _get_nchain(priors, nchain):
    return int(nchain) if nchain is not None else 3 * int(priors["nparameter"])  # :1-10


def _run_seeds(seed, ncore):
    """set_configs (R/model-class.R:1490-1542): one seed per independent run."""
    ss = np.random.SeedSequence(seed if seed is not None else None)
    return [int(s.generate_state(2, np.uint32).view(np.uint64)[0]) for s in ss.spawn(ncore)]


# ------------------------------------------------------------ initialisers
def initialise_theta(nchain, priors, dmi, seed=None, max_init_attempts=1000, batch=16):
    """R/phi.R:141-204: per chain, the first prior draw with finite summed log
    prior and log-likelihood. Candidates are drawn `batch` per chain at a time
    and their likelihoods evaluated on the GPU in one call.
    Returns the start state (theta [P, nchain], lp, ll)."""
    model, data = dmi
    p_prior = priors["p_prior"]
    rng = np.random.default_rng(seed)
    P = p_prior.npar
    th = np.full((P, nchain), np.nan)
    lp = np.full(nchain, -np.inf)
    ll = np.full(nchain, -np.inf)
    todo = list(range(nchain))
    attempts = 0
    while todo and attempts < max_init_attempts:
        k = min(batch, max_init_attempts - attempts)
        cand = rprior(p_prior, len(todo) * k, rng)
        lps = np.sum(logdprior(p_prior, cand), axis=1)
        lls = loglike(model, data, cand, _lib.FP64_PLAIN)
        left = []
        for i, c in enumerate(todo):
            rows = range(i * k, (i + 1) * k)
            ok = [r for r in rows if np.isfinite(lps[r]) and np.isfinite(lls[r])]
            if ok:
                r = ok[0]
                th[:, c], lp[c], ll[c] = cand[r], lps[r], lls[r]
            else:
                left.append(c)
        todo = left
        attempts += k
    if todo:
        raise RuntimeError("Chain %d: Failed to find valid theta after %d attempts"
                           % (todo[0] + 1, max_init_attempts))
    return th, lp, ll


def initialise_phi(nchain, priors, dmis, seed=None, max_init_attempts=1000, rng=None):
    """R/phi.R:256-332: subjects from initialise_theta under p_prior (seed +
    subject index, 1-based), then per chain the first h_prior draw with finite
    log prior and sumloghlike against the subjects' start values of that chain;
    an error when max_init_attempts draws of a chain all fail. The subjects'
    start lp stays the p_prior density initialise_theta gave it (the reference
    does not re-evaluate it under the phi chain it will be paired with).
    Returns (phi start, [subject starts])."""
    if priors["h_prior"] is None:
        raise ValueError("hyper prior is NULL")
    seed = 0 if seed is None else int(seed)
    subs = [initialise_theta(nchain, priors, d, seed + s + 1, max_init_attempts)
            for s, d in enumerate(dmis)]
    rng = rng if rng is not None else np.random.default_rng(seed)
    hp, pp = priors["h_prior"], priors["p_prior"]
    phi = np.full((hp.npar, nchain), np.nan)
    plp = np.full(nchain, -np.inf)
    pll = np.full(nchain, -np.inf)
    for c in range(nchain):
        X = np.stack([s[0][:, c] for s in subs])
        for _ in range(max_init_attempts):
            cand = rprior(hp, 1, rng)[0]
            a = sumlogprior(hp, cand)
            b = sumloghlike(pp, cand, X)
            if np.isfinite(a) and np.isfinite(b):
                phi[:, c], plp[c], pll[c] = cand, a, b
                break
        else:
            raise RuntimeError("Chain %d: Failed to find valid phi after %d attempts"
                               % (c + 1, max_init_attempts))
    return (phi, plp, pll), subs


# ------------------------------------------------------------ fits
class Fit(list):
    """The list of ncore run results with the attributes R attaches
    (configs, dmis, priors; R/sampling.R:299-303)."""
    configs = None
    dmis = None
    priors = None
    kind = None
    iterations_done = 0


def _config(nmc, nchain, thin, kw):
    return make_config(nmc=nmc, nchain=nchain, thin=thin,
                       pop_migration_prob=kw.get("pop_migration_prob", 0.0),
                       sub_migration_prob=kw.get("sub_migration_prob", 0.0),
                       gamma_precursor=kw.get("gamma_precursor", 2.38), rp=kw.get("rp", 0.001),
                       is_hblocked=kw.get("is_hblocked", False),
                       is_pblocked=kw.get("is_pblocked", False),
                       precision=kw.get("precision", _lib.MIXED_BATCHED),
                       sweep_groups=kw.get("sweep_groups", 0),
                       iter_offset=kw.get("iter_offset", 0),
                       pop_debug=kw.get("pop_debug", False), sub_debug=kw.get("sub_debug", False))


def _execute(kind, model, jobs, cfg):
    """Runs the jobs to completion; with a debug flag set, prints their
    decisions as the reference does (src/de.cpp:85-106)."""
    s = Session(cfg, model, jobs)
    try:
        s.advance(s.total_iters)
        if cfg.pop_debug or cfg.sub_debug:
            rec, dropped = s.trace()
            if len(rec):
                print(format_trace(rec))
            if dropped:
                print("(%d further decisions not recorded)" % dropped)
        return s.download()
    finally:
        s.close()


def StartSampling_subject(dmi, priors, samples_list=None, nmc=500, nchain=None, thin=1,
                          max_init_attempts=1000, sub_migration_prob=0.0, gamma_precursor=2.38,
                          rp=0.001, is_pblocked=False, sub_debug=False, ncore=3, seed=None,
                          precision=_lib.MIXED_BATCHED, sweep_groups=0):
    """R/sampling.R:446-502: ncore independent single-subject runs."""
    model, data = dmi
    nchain = _get_nchain(priors, nchain)
    seeds = _run_seeds(seed, ncore)
    if samples_list is None:
        samples_list = [initialise_theta(nchain, priors, dmi, s, max_init_attempts) for s in seeds]
    kw = dict(sub_migration_prob=sub_migration_prob, gamma_precursor=gamma_precursor, rp=rp,
              is_pblocked=is_pblocked, precision=precision, sweep_groups=sweep_groups,
              sub_debug=sub_debug)
    cfg = _config(nmc, nchain, thin, kw)
    jobs = [JobSpec(_lib.KIND_SUBJECT, sd, priors["p_prior"], start=st, subject=data,
                    nparameter_de=priors["nparameter"]) for sd, st in zip(seeds, samples_list)]
    res = _execute(_lib.KIND_SUBJECT, model, jobs, cfg)
    out = Fit(r["subjects"][0] for r in res)
    out.configs = dict(kw, nmc=nmc, nchain=nchain, thin=thin, seeds=seeds)
    out.dmis, out.priors, out.kind = dmi, priors, _lib.KIND_SUBJECT
    out.iterations_done = nmc * thin - 1
    return out


def RestartSampling_subject(samples_list, nmc=500, thin=1, **kw):
    """R/sampling.R:507-610: continue every run from its last stored state."""
    return _restart(samples_list, nmc, thin, kw)


def StartSampling(dmis, priors, samples_list=None, nmc=500, nchain=None, thin=1,
                  max_init_attempts=1000, pop_migration_prob=0.0, sub_migration_prob=0.0,
                  gamma_precursor=2.38, rp=0.001, is_hblocked=False, is_pblocked=False,
                  pop_debug=False, sub_debug=False, ncore=3, seed=None,
                  precision=_lib.MIXED_BATCHED, sweep_groups=0):
    """R/sampling.R:247-304: ncore independent hierarchical runs; each result
    is {'phi': Posterior, 'subject_theta': [Posterior]}."""
    model = dmis[0][0]
    nchain = _get_nchain(priors, nchain)
    seeds = _run_seeds(seed, ncore)
    if samples_list is None:
        samples_list = [initialise_phi(nchain, priors, dmis, s % 1000000, max_init_attempts)
                        for s in seeds]
    kw = dict(pop_migration_prob=pop_migration_prob, sub_migration_prob=sub_migration_prob,
              gamma_precursor=gamma_precursor, rp=rp, is_hblocked=is_hblocked,
              is_pblocked=is_pblocked, precision=precision, sweep_groups=sweep_groups,
              pop_debug=pop_debug, sub_debug=sub_debug)
    cfg = _config(nmc, nchain, thin, kw)
    jobs = [JobSpec(_lib.KIND_HIER, sd, priors["p_prior"], h_prior=priors["h_prior"],
                    phi_start=st[0], subjects=[d[1] for d in dmis], subject_starts=st[1],
                    nparameter_de=priors["nparameter"]) for sd, st in zip(seeds, samples_list)]
    res = _execute(_lib.KIND_HIER, model, jobs, cfg)
    out = Fit({"phi": r["phi"], "subject_theta": r["subjects"]} for r in res)
    out.configs = dict(kw, nmc=nmc, nchain=nchain, thin=thin, seeds=seeds)
    out.dmis, out.priors, out.kind = dmis, priors, _lib.KIND_HIER
    out.iterations_done = nmc * thin - 1
    return out


def RestartSampling(samples_list, nmc=500, thin=1, **kw):
    """R/sampling.R:308-441."""
    return _restart(samples_list, nmc, thin, kw)


def StartSampling_hyper(hyper_dmi, dmis, priors, samples_list=None, nmc=500, nchain=None, thin=1,
                        max_init_attempts=1000, sub_migration_prob=0.0, gamma_precursor=2.38,
                        rp=0.001, is_pblocked=False, sub_debug=False, ncore=3, seed=None,
                        precision=_lib.MIXED_BATCHED, sweep_groups=0):
    """R/sampling.R:615-676: population parameters of fixed subject vectors
    hyper_dmi [S, P] (the hyper_dmi data). As in the reference, the start of
    each run is initialise_phi's phi on the subjects `dmis`
    (R/sampling.R:651-656), lp / ll as initialise_phi computed them, so an
    exhausted initialisation raises."""
    nchain = _get_nchain(priors, nchain)
    seeds = _run_seeds(seed, ncore)
    hp, pp = priors["h_prior"], priors["p_prior"]
    if hp is None:
        raise ValueError("h_prior has not found")
    if samples_list is None:
        samples_list = [initialise_phi(nchain, priors, dmis, sd % 1000000, max_init_attempts)[0]
                        for sd in seeds]
    kw = dict(sub_migration_prob=sub_migration_prob, gamma_precursor=gamma_precursor, rp=rp,
              is_pblocked=is_pblocked, precision=precision, sweep_groups=sweep_groups,
              sub_debug=sub_debug)
    cfg = _config(nmc, nchain, thin, kw)
    hd = np.asarray(hyper_dmi, float)
    # the start's cached likelihood is initialise_phi's (against the subjects'
    # start values), handed to run_hyper unchanged as the reference does
    jobs = [JobSpec(_lib.KIND_HYPER, sd, pp, h_prior=hp, phi_start=st, hyper_data=hd,
                    nparameter_de=priors["nparameter"]) for sd, st in zip(seeds, samples_list)]
    res = _execute(_lib.KIND_HYPER, None, jobs, cfg)
    out = Fit(r["phi"] for r in res)
    out.configs = dict(kw, nmc=nmc, nchain=nchain, thin=thin, seeds=seeds)
    out.dmis, out.priors, out.kind = hd, priors, _lib.KIND_HYPER
    out.iterations_done = nmc * thin - 1
    return out


def RestartSampling_hyper(samples_list, nmc=500, thin=1, **kw):
    """R/sampling.R:681-811."""
    return _restart(samples_list, nmc, thin, kw)


def _restart(prev, nmc, thin, kw):
    """Continue each run from its last stored column; unchanged flags keep the
    previous configuration (R/sampling.R:369-421)."""
    if not isinstance(prev, Fit) or prev.configs is None:
        raise ValueError("Must provide a previously fitted samples")
    cfgd = dict(prev.configs)
    cfgd.update({k: v for k, v in kw.items() if k in cfgd or k in (
        "pop_migration_prob", "sub_migration_prob", "gamma_precursor", "rp", "is_hblocked",
        "is_pblocked", "precision", "sweep_groups", "pop_debug", "sub_debug")})
    seeds = kw.get("seeds", cfgd["seeds"])
    cfgd["iter_offset"] = prev.iterations_done
    cfg = _config(nmc, cfgd["nchain"], thin, cfgd)
    priors = prev.priors
    if prev.kind == _lib.KIND_SUBJECT:
        model, data = prev.dmis
        jobs = [JobSpec(_lib.KIND_SUBJECT, sd, priors["p_prior"], start=p.last_state(),
                        subject=data, nparameter_de=priors["nparameter"])
                for sd, p in zip(seeds, prev)]
        res = [r["subjects"][0] for r in _execute(prev.kind, model, jobs, cfg)]
    elif prev.kind == _lib.KIND_HIER:
        model = prev.dmis[0][0]
        jobs = [JobSpec(_lib.KIND_HIER, sd, priors["p_prior"], h_prior=priors["h_prior"],
                        phi_start=p["phi"].last_state(), subjects=[d[1] for d in prev.dmis],
                        subject_starts=[s.last_state() for s in p["subject_theta"]],
                        nparameter_de=priors["nparameter"]) for sd, p in zip(seeds, prev)]
        res = [{"phi": r["phi"], "subject_theta": r["subjects"]}
               for r in _execute(prev.kind, model, jobs, cfg)]
    else:
        jobs = [JobSpec(_lib.KIND_HYPER, sd, priors["p_prior"], h_prior=priors["h_prior"],
                        phi_start=p.last_state(), hyper_data=np.asarray(prev.dmis, float),
                        nparameter_de=priors["nparameter"]) for sd, p in zip(seeds, prev)]
        res = [r["phi"] for r in _execute(prev.kind, None, jobs, cfg)]
    out = Fit(res)
    out.configs = dict(cfgd, nmc=nmc, thin=thin, seeds=seeds)
    out.dmis, out.priors, out.kind = prev.dmis, priors, prev.kind
    out.iterations_done = prev.iterations_done + nmc * thin - 1
    return out


def _param_array(posts):
    """.GetParamArray (R/parallel_helper.R:15-53): run i's theta [npar, nchain,
    nmc] flattened column-major to [npar, nchain * nmc] (chain fastest) becomes
    column i of [npar, nruns, nchain * nmc]."""
    p0 = posts[0]
    npar, nchain, nmc = p0.theta.shape
    out = np.empty((npar, len(posts), nchain * nmc))
    for i, p in enumerate(posts):
        out[:, i, :] = p.theta.reshape(npar, nchain * nmc, order="F")
    return out


def _like_matrix(posts, is_prior):
    """.get_likelihood_matrix (R/parallel_helper.R:1-12): each run's
    [nchain, nmc] matrix transposed and flattened column-major (iteration
    fastest), one row per run: [nruns, nmc * nchain]. The order differs from
    the parameter array's (chain fastest); kept as the reference has it."""
    rows = [(p.summed_log_prior if is_prior else p.log_likelihoods).T.reshape(-1, order="F")
            for p in posts]
    return np.stack(rows)


def _rebuild(posts):
    p0 = posts[0]
    return Posterior(_param_array(posts), _like_matrix(posts, True), _like_matrix(posts, False),
                     p0.nmc * p0.nchain, p0.thin, len(posts), p0.pnames, p0.start)


def RebuildPosterior(fits):
    """R/parallel_helper.R:59-80: the ncore single-level runs as one posterior
    whose chains are the runs and whose nmc is nmc x nchain of one run."""
    return _rebuild([f for f in fits])


def RebuildHyper(fits):
    """R/parallel_helper.R:84-106: the population-level posteriors of ncore
    hierarchical runs (each fit holds 'phi'), rebuilt as RebuildPosterior."""
    return _rebuild([f["phi"] for f in fits])


def RebuildPosteriors(fits):
    """R/parallel_helper.R:162-208: one rebuilt posterior per subject of ncore
    hierarchical runs (each fit holds 'subject_theta')."""
    if not isinstance(fits, (list, tuple)) or len(fits) == 0:
        raise ValueError("Input must be a non-empty list of model fits")
    nsub = len(fits[0]["subject_theta"])
    return [_rebuild([f["subject_theta"][i] for f in fits]) for i in range(nsub)]
