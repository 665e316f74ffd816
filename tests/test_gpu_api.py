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
