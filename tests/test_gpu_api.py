"""The R-facing API mirror (R/sampling.R, R/phi.R) on the GPU engine: shapes,
initialisation, restarts, and that restarting continues the chains."""
import numpy as np
import pytest

import fixtures as fx
import oracle_lib as O
from ggdmc_amd import PriorSpec, sampling, workloads

pytestmark = pytest.mark.gpu


def test_initialise_theta_matches_oracle(gpu):
    fix = fx.load("lba_data6")
    priors = sampling.set_priors(fix.sub_prior)
    th, lp, ll = sampling.initialise_theta(12, priors, (fix.model, fix.sub_data), seed=3)
    assert np.isfinite(lp).all() and np.isfinite(ll).all()
    for c in range(12):
        ref = O.lba_loglike(fix.omodel, fix.sub_data.rt, fix.sub_data.cell_off, th[:, c])
        assert abs(ll[c] / ref - 1) < 1e-12
        assert abs(lp[c] - O.sumlogprior(fix.oprior, th[:, c])) < 1e-12


def test_start_and_restart_subject(gpu):
    w = workloads.make("c1")
    job = w.jobs[0]
    priors = sampling.set_priors(job.p_prior)
    dmi = (w.model, job.subject)
    fits = sampling.StartSampling_subject(dmi, priors, nmc=40, thin=2, ncore=3, seed=9032,
                                          sub_migration_prob=0.05)
    assert len(fits) == 3
    for f in fits:
        assert f.theta.shape == (6, 18, 40)
        assert np.isfinite(f.log_likelihoods).all()
    more = sampling.RestartSampling_subject(fits, nmc=20, thin=1)
    assert more.iterations_done == 79 + 19
    for a, b in zip(fits, more):
        # the restart begins at the previous run's last stored state
        assert np.array_equal(b.theta[:, :, 0], a.theta[:, :, -1])
        assert np.array_equal(b.log_likelihoods[:, 0], a.log_likelihoods[:, -1])
    # the chains stay finite, and at least one chain moved after the restart
    assert any(np.any(b.theta[:, :, -1] != b.theta[:, :, 0]) for b in more)


def test_start_hierarchical_and_hyper(gpu):
    w = workloads.make("c3", scale=0.125)      # 4 subjects
    job = w.jobs[0]
    priors = sampling.set_priors(job.p_prior, job.h_prior)
    dmis = [(w.model, d) for d in job.subjects]
    fits = sampling.StartSampling(dmis, priors, nmc=10, nchain=20, ncore=2, seed=11,
                                  pop_migration_prob=0.05, sub_migration_prob=0.05)
    assert len(fits) == 2
    assert fits[0]["phi"].theta.shape == (26, 20, 10)
    assert len(fits[0]["subject_theta"]) == 4
    assert np.isfinite(fits[0]["phi"].log_likelihoods).all()
    again = sampling.RestartSampling(fits, nmc=5)
    assert np.array_equal(again[1]["phi"].theta[:, :, 0], fits[1]["phi"].theta[:, :, -1])
    data = np.stack([s[0][:, 0] for s in job.subject_starts])
    hyper = sampling.StartSampling_hyper(data, dmis, priors, nmc=15, nchain=20, ncore=2, seed=5)
    assert hyper[0].theta.shape == (26, 20, 15)
    assert np.isfinite(hyper[0].log_likelihoods).all()


def test_errors_come_back_as_exceptions(gpu):
    w = workloads.make("c1")
    job = w.jobs[0]
    bad = gpu.make_config(nmc=5, nchain=2)               # fewer than three chains
    with pytest.raises(gpu.GgdmcError, match="three or more chains"):
        gpu.run_subject(bad, w.model, job.p_prior, job.subject, job.start, seed=1)
    wrong = PriorSpec.uniform(5)                          # wrong parameter count
    with pytest.raises(gpu.GgdmcError):
        gpu.run_subject(gpu.make_config(nmc=5, nchain=18), w.model, wrong, job.subject,
                        job.start, seed=1)


def test_session_reports_its_decomposition_and_kernel_build(gpu):
    # the decomposition and kernel build a configuration asks for are the ones
    # the session runs; the 4-wave build exists for several iterations per
    # launch only, and a value other than 0, 3 or 4 is refused
    w = workloads.make("c3", scale=1 / 8, ll_fn=lambda m, d, th: np.zeros(len(th)))
    for k, lm, wps, expect in ((2, 3, 4, 4), (1, 4, 4, 4), (2, 4, 3, 3), (2, 1, 4, 3)):
        cfg = w.config(nmc=4, chains_per_item=k, launch_mode=lm, waves_per_simd=wps)
        s = gpu.Session(cfg, w.model, w.jobs)
        try:
            kci, tpt, nitems, mode = s.decomposition()
            assert (kci, tpt, mode) == (k, 0, lm)
            assert s.waves_per_simd() == expect
        finally:
            s.close()
    with pytest.raises(gpu.GgdmcError, match="waves_per_simd"):
        gpu.Session(w.config(nmc=4, waves_per_simd=2), w.model, w.jobs)


def _trace_run(gpu, cfg, model, jobs):
    s = gpu.Session(cfg, model, jobs)
    try:
        s.advance(s.total_iters)
        rec, dropped = s.trace()
        return rec, dropped, s.download()
    finally:
        s.close()


def _check_level(rec, th, lp, ll, nchain, nmc, population=False):
    # every decision of every iteration, in the reference's order: a
    # crossover sweep has one record per chain in chain order; a migration
    # ring one per step, each target written once. An accepted decision is a
    # changed state carrying the proposal's lp / ll; the comparison is always
    # with the target's posterior before the iteration
    assert np.all(np.diff(rec["iteration"].astype(int)) >= 0)
    for it in range(1, nmc):
        r = rec[rec["iteration"] == it]
        assert len(r) > 0
        if r["kind"][0] == 0:
            assert list(r["step"]) == list(range(nchain)) and np.all(r["kind"] == 0)
        else:
            assert list(r["step"]) == list(range(len(r))) and np.all(r["kind"] == 1)
            assert len(set(r["target"])) == len(r)
        for x in r:
            t = x["target"]
            changed = bool(np.any(th[:, t, it] != th[:, t, it - 1]))
            assert changed == (x["accepted"] == 1)
            if population:
                # the population level compares with its refreshed likelihood
                # (src/de.cpp:397, :494-500), which later ring steps refresh again
                if changed:
                    assert x["tmp_lp"] == lp[t, it]
                continue
            assert x["cur_post"] == lp[t, it - 1] + ll[t, it - 1]
            if changed:
                assert x["tmp_lp"] == lp[t, it] and x["tmp_ll"] == ll[t, it]


def test_debug_trace_follows_the_trajectory(gpu):
    fix = fx.load("lba_data6")
    n, nmc = 12, 8
    start = fx.near_start(fix, n, seed=7)
    job = gpu.JobSpec(gpu.KIND_SUBJECT, 9032, fix.sub_prior, start=start, subject=fix.sub_data)
    for prec, groups in ((gpu.FP64_PLAIN, n), (gpu.FP64_BATCHED, 2), (gpu.MIXED_BATCHED, 2)):
        cfg = gpu.make_config(nmc=nmc, nchain=n, precision=prec, sweep_groups=groups,
                              sub_migration_prob=0.3, sub_debug=True)
        rec, dropped, res = _trace_run(gpu, cfg, fix.model, [job])
        post = res[0]["subjects"][0]
        assert dropped == 0 and np.all(rec["subject"] == 0)
        _check_level(rec, post.theta, post.summed_log_prior, post.log_likelihoods, n, nmc)
        assert "accepted" in gpu.runs.format_trace(rec)
    # no debug flag, no records
    rec, _, _ = _trace_run(gpu, gpu.make_config(nmc=3, nchain=n), fix.model, [job])
    assert len(rec) == 0


def test_debug_trace_of_a_hierarchical_run(gpu):
    fix = fx.load("lba_data6")
    n, nmc, S = 10, 6, 4
    phi, subs = fx.hier_start(fix, S, n, seed=3)
    job = gpu.JobSpec(gpu.KIND_HIER, 77, fix.pop_prior, h_prior=fix.hyper_prior, phi_start=phi,
                      subjects=[fix.pop_data(s) for s in range(S)], subject_starts=subs,
                      nparameter_de=2 * fix.P)
    for prec, groups in ((gpu.FP64_PLAIN, n), (gpu.MIXED_BATCHED, 2)):
        cfg = gpu.make_config(nmc=nmc, nchain=n, precision=prec, sweep_groups=groups,
                              pop_migration_prob=0.3, sub_migration_prob=0.3, pop_debug=True,
                              sub_debug=True)
        rec, dropped, res = _trace_run(gpu, cfg, fix.model, [job])
        assert dropped == 0
        # the population level and the first three subjects only (src/de.cpp:327)
        assert set(rec["subject"]) == {-1, 0, 1, 2}
        p = res[0]["phi"]
        _check_level(rec[rec["subject"] == -1], p.theta, p.summed_log_prior, p.log_likelihoods, n, nmc,
                     population=True)
        for s in range(3):
            q = res[0]["subjects"][s]
            r = rec[rec["subject"] == s]
            for it in range(1, nmc):
                for x in r[r["iteration"] == it]:
                    t = x["target"]
                    assert bool(np.any(q.theta[:, t, it] != q.theta[:, t, it - 1])) == (x["accepted"] == 1)
