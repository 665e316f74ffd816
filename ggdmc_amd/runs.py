"""Run-level API over the C ABI: the counterparts of the reference's Rcpp
exports run_subject / run_hyper / run (src/de2R.cpp:8-171) plus batches of
independent runs and device-resident sessions.

Arrays follow R's layouts: a start state is theta [P, nchain] (+ lp, ll
[nchain]); a posterior holds theta [P, nchain, nmc], summed_log_prior and
log_likelihoods [nchain, nmc] (posterior S4 slots, R/model-class.R:238-254).
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import Keep, check, ptr


@dataclass
class LBAModel:
    """Flattened LBA design (see include/ggdmc_hip.h ggdmc_model)."""
    winner: np.ndarray          # [n_cell] int32
    par_src: np.ndarray         # [n_cell, n_acc, 6] int32
    par_const: np.ndarray       # [n_cell, n_acc, 6] float64
    pnames: list = field(default_factory=list)
    cell_names: list = field(default_factory=list)

    @property
    def n_cell(self):
        return int(self.par_src.shape[0])

    @property
    def n_acc(self):
        return int(self.par_src.shape[1])

    @property
    def npar(self):
        return len(self.pnames) if self.pnames else int(self.par_src.max()) + 1

    def c_struct(self, keep):
        w = keep.arr(self.winner, np.int32)
        s = keep.arr(self.par_src, np.int32)
        c = keep.arr(self.par_const, np.float64)
        m = _lib.Model(self.n_acc, self.n_cell, self.npar, ptr(w, _lib._i32), ptr(s, _lib._i32),
                       ptr(c))
        keep.items.append(m)
        return m


@dataclass
class PriorSpec:
    """One prior per parameter (ggdmcPrior p_prior / h_prior entries)."""
    dist: np.ndarray
    p0: np.ndarray
    p1: np.ndarray
    lower: np.ndarray
    upper: np.ndarray
    pnames: list = field(default_factory=list)

    @classmethod
    def uniform(cls, npar, lo=0.0, hi=10.0, pnames=None):
        return cls(np.full(npar, _lib.DIST_UNIF, np.int32), np.full(npar, lo), np.full(npar, hi),
                   np.full(npar, lo), np.full(npar, hi), list(pnames or []))

    @classmethod
    def tnorm(cls, p0, p1, lower=0.0, upper=np.inf, pnames=None):
        p0 = np.asarray(p0, float)
        n = len(p0)
        return cls(np.full(n, _lib.DIST_TNORM, np.int32), p0, np.broadcast_to(p1, n).astype(float),
                   np.broadcast_to(lower, n).astype(float),
                   np.broadcast_to(upper, n).astype(float), list(pnames or []))

    @property
    def npar(self):
        return len(self.dist)

    def c_struct(self, keep):
        arrs = [keep.arr(self.dist, np.int32)] + [keep.arr(a, np.float64) for a in
                                                  (self.p0, self.p1, self.lower, self.upper)]
        p = _lib.Prior(self.npar, ptr(arrs[0], _lib._i32), *[ptr(a) for a in arrs[1:]])
        keep.items.append(p)
        return p


@dataclass
class SubjectData:
    """dmi@data of one subject: RTs sorted by cell, cell_off [n_cell + 1]."""
    rt: np.ndarray
    cell_off: np.ndarray

    @classmethod
    def from_cells(cls, cells):
        rt = np.concatenate([np.asarray(c, float) for c in cells]) if cells else np.zeros(0)
        off = np.concatenate([[0], np.cumsum([len(c) for c in cells])]).astype(np.int64)
        return cls(rt, off)

    def c_struct(self, keep):
        rt = keep.arr(self.rt, np.float64)
        off = keep.arr(self.cell_off, np.int64)
        s = _lib.Subject(len(rt), ptr(rt), ptr(off, _lib._i64))
        keep.items.append(s)
        return s


@dataclass
class Posterior:
    """posterior S4 slots (R/model-class.R:238-254)."""
    theta: np.ndarray             # [P, nchain, nmc]
    summed_log_prior: np.ndarray  # [nchain, nmc]
    log_likelihoods: np.ndarray   # [nchain, nmc]
    nmc: int
    thin: int
    nchain: int
    pnames: list = field(default_factory=list)
    start: int = 1

    @property
    def npar(self):
        return self.theta.shape[0]

    def last_state(self):
        """The state a restart continues from (last stored column)."""
        return (self.theta[:, :, -1].copy(), self.summed_log_prior[:, -1].copy(),
                self.log_likelihoods[:, -1].copy())


def _state(keep, theta, lp, ll):
    th = keep.arr(np.asarray(theta, float).T, np.float64)   # [nchain][P] == R [P x nchain]
    a = keep.arr(lp, np.float64)
    b = keep.arr(ll, np.float64)
    st = _lib.State(ptr(th), ptr(a), ptr(b))
    keep.items.append(st)
    return st


class _Out:
    def __init__(self, keep, P, nchain, nmc):
        self.th = np.empty((nmc, nchain, P))
        self.lp = np.empty((nmc, nchain))
        self.ll = np.empty((nmc, nchain))
        keep.items.extend([self.th, self.lp, self.ll])
        self.c = _lib.Samples(ptr(self.th), ptr(self.lp), ptr(self.ll))

    def posterior(self, cfg, pnames=None):
        return Posterior(self.th.transpose(2, 1, 0).copy(), self.lp.T.copy(), self.ll.T.copy(),
                         cfg.nmc, cfg.thin, cfg.nchain, list(pnames or []))


def make_config(nmc, nchain, thin=1, pop_migration_prob=0.0, sub_migration_prob=0.0,
                gamma_precursor=2.38, rp=0.001, is_hblocked=False, is_pblocked=False,
                precision=_lib.MIXED_BATCHED, sweep_groups=0, trials_per_tile=0,
                chains_per_item=0, launch_mode=0, iter_offset=0, pop_debug=False,
                sub_debug=False, waves_per_simd=0):
    """ggdmc_config. trials_per_tile / chains_per_item / launch_mode 0 let the
    library autotune the work split; launch_mode 1 = a subject launch per chain
    group, 2 = one for every group, 3 = one for every group and 32 iterations,
    4 = as 3 with an item holding its chain range of every group.
    waves_per_simd (0 = autotuned, 3 or 4) picks the build of the subject kernel
    for launch modes 3 and 4. None of them changes a result bit. pop_debug /
    sub_debug record the decisions of the population level / of the subjects
    (Session.trace)."""
    return _lib.Config(int(nmc), int(thin), int(nchain), float(pop_migration_prob),
                       float(sub_migration_prob), float(gamma_precursor), float(rp),
                       int(bool(is_hblocked)), int(bool(is_pblocked)), int(precision),
                       int(sweep_groups), int(trials_per_tile), int(chains_per_item),
                       int(launch_mode), int(iter_offset), int(bool(pop_debug)),
                       int(bool(sub_debug)), int(waves_per_simd))


def format_trace(records):
    """The decision records as src/de.cpp prints them with debug on
    (update_theta :85-106, the population steps :441-463 / :533-558): one
    line per decision, '[C i]: lp, ll, (new vs old posterior), [ratio]-> accepted'
    or 'rejected' ('rejected (mh nan)' for a NaN ratio), migration steps as
    '[C src] vs [C dst]'; a header per run, iteration and level."""
    lines = []
    last = None
    for r in records:
        key = (int(r["job"]), int(r["iteration"]), int(r["subject"]))
        if key != last:
            lvl = "population" if r["subject"] < 0 else "subject %d" % r["subject"]
            lines.append("run %d, iteration %d, %s:" % (key[0], key[1], lvl))
            last = key
        who = ("  [C%d]: " % r["target"]) if r["kind"] == 0 else \
            ("  [C%d] vs [C%d]: " % (r["chain"], r["target"]))
        if r["accepted"] == 1:
            post = r["tmp_lp"] + r["tmp_ll"]
            ratio = np.exp(post - r["cur_post"]) if np.isfinite(post - r["cur_post"]) else np.inf
            lines.append("%s%.6g, %.6g, (%.6g vs %.6g), \t[%.6g]-> accepted"
                         % (who, r["tmp_lp"], r["tmp_ll"], post, r["cur_post"], ratio))
        else:
            lines.append(who + ("rejected (mh nan)" if r["accepted"] < 0 else "rejected"))
    return "\n".join(lines)


@dataclass
class JobSpec:
    """One independent run of a batch (ggdmc_job)."""
    kind: int
    seed: int
    p_prior: PriorSpec
    start: tuple = None                 # SUBJECT: (theta [P,n], lp, ll)
    subject: SubjectData = None
    h_prior: PriorSpec = None
    phi_start: tuple = None             # HYPER / HIER: (theta [2P,n], lp, ll)
    subjects: list = None               # HIER: [SubjectData]
    subject_starts: list = None         # HIER: [(theta, lp, ll)]
    hyper_data: np.ndarray = None       # HYPER: [S, P]
    nparameter_de: int = 0


class Session:
    """A batch of runs resident on the GPU (ggdmc_session_*)."""

    def __init__(self, cfg, model, jobs):
        L = _lib.lib()
        self.cfg = cfg
        self.keep = Keep()
        self.jobs = jobs
        k = self.keep
        P = model.npar if model is not None else jobs[0].p_prior.npar
        self.P = P
        cm = model.c_struct(k) if model is not None else None
        cjobs = (_lib.Job * len(jobs))()
        self.outs = []
        for i, j in enumerate(jobs):
            cj = cjobs[i]
            cj.kind = j.kind
            cj.seed = int(j.seed) & 0xFFFFFFFFFFFFFFFF
            cj.nparameter_de = int(j.nparameter_de)
            cj.p_prior = C.pointer(j.p_prior.c_struct(k))
            o = {"phi": None, "subjects": []}
            if j.kind == _lib.KIND_SUBJECT:
                subs = (_lib.Subject * 1)(j.subject.c_struct(k))
                sts = (_lib.State * 1)(_state(k, *j.start))
                out = _Out(k, P, cfg.nchain, cfg.nmc)
                outs = (_lib.Samples * 1)(out.c)
                k.items.extend([subs, sts, outs])
                cj.nsubject = 1
                cj.subjects = subs
                cj.subject_starts = sts
                cj.subject_outs = outs
                o["subjects"].append(out)
            else:
                cj.h_prior = C.pointer(j.h_prior.c_struct(k))
                cj.phi_start = C.pointer(_state(k, *j.phi_start))
                pout = _Out(k, 2 * P, cfg.nchain, cfg.nmc)
                cj.phi_out = C.pointer(pout.c)
                o["phi"] = pout
                if j.kind == _lib.KIND_HIER:
                    S = len(j.subjects)
                    subs = (_lib.Subject * S)(*[s.c_struct(k) for s in j.subjects])
                    sts = (_lib.State * S)(*[_state(k, *st) for st in j.subject_starts])
                    outs_l = [_Out(k, P, cfg.nchain, cfg.nmc) for _ in range(S)]
                    outs = (_lib.Samples * S)(*[x.c for x in outs_l])
                    k.items.extend([subs, sts, outs])
                    cj.nsubject = S
                    cj.subjects = subs
                    cj.subject_starts = sts
                    cj.subject_outs = outs
                    o["subjects"] = outs_l
                else:
                    hd = k.arr(j.hyper_data, np.float64)
                    cj.nsubject = hd.shape[0]
                    cj.hyper_data = ptr(hd)
            self.outs.append(o)
        k.items.extend([cm, cjobs])
        h = C.c_void_p()
        check(L.ggdmc_session_create(C.byref(cfg), C.byref(cm) if cm is not None else None,
                                     cjobs, len(jobs), C.byref(h)))
        self.h = h
        self.total_iters = cfg.nmc * cfg.thin - 1

    def advance(self, n_iter):
        check(_lib.lib().ggdmc_session_advance(self.h, int(n_iter)))
        ms = C.c_float()
        check(_lib.lib().ggdmc_session_elapsed_ms(self.h, C.byref(ms)))
        return ms.value

    def profile(self, n_iter):
        ms = C.c_double()
        n = C.c_int32()
        check(_lib.lib().ggdmc_session_profile(self.h, int(n_iter), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def trial_evals(self):
        v = C.c_double()
        check(_lib.lib().ggdmc_session_trial_evals(self.h, C.byref(v)))
        return v.value

    def trace(self):
        """Decision records of the last advance (pop_debug / sub_debug), in the
        reference's print order, and how many were dropped beyond the cap."""
        p = C.POINTER(_lib.TraceRecord)()
        n = C.c_int64()
        dr = C.c_int64()
        check(_lib.lib().ggdmc_session_trace(self.h, C.byref(p), C.byref(n), C.byref(dr)))
        return _lib.trace_array(p, n.value), dr.value

    def decomposition(self):
        """(chains per item or -1 for one item per subject, trials per item or 0
        for whole subjects, items of the largest launch, launch mode 1 to 4 as
        make_config's launch_mode)."""
        k = C.c_int32()
        t = C.c_int32()
        m = C.c_int32()
        lm = C.c_int32()
        check(_lib.lib().ggdmc_session_decomposition(self.h, C.byref(k), C.byref(t), C.byref(m),
                                                     C.byref(lm)))
        return k.value, t.value, m.value, lm.value

    def waves_per_simd(self):
        """Waves per SIMD the subject kernel of this session was built for (3 or 4)."""
        w = C.c_int32()
        check(_lib.lib().ggdmc_session_waves_per_simd(self.h, C.byref(w)))
        return w.value

    def download(self):
        """Returns per job {'phi': Posterior|None, 'subjects': [Posterior]}."""
        check(_lib.lib().ggdmc_session_download(self.h))
        res = []
        for j, o in zip(self.jobs, self.outs):
            res.append({"phi": o["phi"].posterior(self.cfg, j.h_prior.pnames if j.h_prior else None)
                        if o["phi"] is not None else None,
                        "subjects": [x.posterior(self.cfg, j.p_prior.pnames) for x in o["subjects"]]})
        return res

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().ggdmc_session_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def run_batch(cfg, model, jobs):
    """All jobs to completion in one session (ggdmc_run_batch)."""
    s = Session(cfg, model, jobs)
    try:
        s.advance(s.total_iters)
        return s.download()
    finally:
        s.close()


def run_subject(cfg, model, p_prior, data, start, seed):
    """run_subject (src/de2R.cpp:8-23)."""
    return run_batch(cfg, model, [JobSpec(_lib.KIND_SUBJECT, seed, p_prior, start=start,
                                          subject=data)])[0]["subjects"][0]


def run_hyper(cfg, p_prior, h_prior, hyper_data, start, seed):
    """run_hyper (src/de2R.cpp:30-47)."""
    return run_batch(cfg, None, [JobSpec(_lib.KIND_HYPER, seed, p_prior, h_prior=h_prior,
                                         phi_start=start, hyper_data=hyper_data)])[0]["phi"]


def run(cfg, model, p_prior, h_prior, subjects, phi_start, subject_starts, seed):
    """run (src/de2R.cpp:123-171): returns {'phi': Posterior, 'subject_theta': [Posterior]}."""
    r = run_batch(cfg, model, [JobSpec(_lib.KIND_HIER, seed, p_prior, h_prior=h_prior,
                                       phi_start=phi_start, subjects=subjects,
                                       subject_starts=subject_starts)])[0]
    return {"phi": r["phi"], "subject_theta": r["subjects"]}


def loglike(model, data, thetas, precision=_lib.MIXED_BATCHED, noalloc=False):
    """sumloglike of parameter vectors thetas [n, P] for one subject
    (ggdmc_loglike; noalloc: ggdmc_loglike_noalloc, staged in the library's
    fixed device array instead of a per-call allocation)."""
    k = Keep()
    cm = model.c_struct(k)
    cs = data.c_struct(k)
    th = k.arr(np.atleast_2d(thetas), np.float64)
    out = np.empty(th.shape[0])
    fn = _lib.lib().ggdmc_loglike_noalloc if noalloc else _lib.lib().ggdmc_loglike
    check(fn(C.byref(cm), C.byref(cs), ptr(th), th.shape[0], int(precision), ptr(out)))
    return out
