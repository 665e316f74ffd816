"""Statistical parity of the batched mixed-precision sampler (two chain
groups) with the oracle's FP64 sampler in the reference's sequential order,
batching invariance, and full-size property checks."""
import numpy as np
import pytest

import fixtures as fx
import oracle_lib as O
from ggdmc_amd import diagnostics, workloads

pytestmark = pytest.mark.gpu


def _oracle_ll(model, data, thetas):
    om = O.Model(model.winner, model.par_src, model.par_const, model.npar)
    return np.array([O.lba_loglike(om, data.rt, data.cell_off, t) for t in thetas])


def _moments_agree(a, b, burn):
    """posterior mean within 5 MCSE and SD within 15% per parameter."""
    sa, sb = diagnostics.summary(a, burn), diagnostics.summary(b, burn)
    z = np.abs(sa["mean"] - sb["mean"]) / np.sqrt(sa["ts_se"] ** 2 + sb["ts_se"] ** 2)
    r = sa["sd"] / sb["sd"]
    return z, r


def test_single_subject_posterior_matches_oracle(gpu):
    w = workloads.make("c2", ll_fn=_oracle_ll)
    job = w.jobs[0]
    nmc = 1500
    cfg = w.config(nmc=nmc, precision=gpu.MIXED_BATCHED, sweep_groups=2)
    runs = gpu.run_batch(cfg, w.model, [gpu.JobSpec(job.kind, 100 + k, job.p_prior, start=job.start,
                                                    subject=job.subject, nparameter_de=13)
                                        for k in range(3)])
    g_theta = np.concatenate([r["subjects"][0].theta for r in runs], axis=1)
    ocfg = O.config(nmc=nmc, nchain=w.nchain, nparameter_de=13, sub_migration_prob=0.06,
                    sweep_groups=w.nchain, seed=7)
    om = O.Model(w.model.winner, w.model.par_src, w.model.par_const, 13)
    o_theta, _, _ = O.run_subject(ocfg, om, fx.oprior(job.p_prior), job.subject.rt,
                                  job.subject.cell_off, *job.start)
    burn = 500
    rhat = diagnostics.psrf(g_theta, burn)
    z, r = _moments_agree(g_theta, o_theta, burn)
    print("R-hat max %.3f; mean z max %.2f; sd ratio %.3f..%.3f" % (rhat.max(), z.max(), r.min(), r.max()))
    assert rhat.max() < 1.1
    assert z.max() < 5.0
    assert np.all(np.abs(r - 1) < 0.15)


def test_hierarchical_posterior_matches_oracle(gpu):
    w = workloads.make("c3", scale=0.25, ll_fn=_oracle_ll)     # 8 subjects x 768 trials
    job = w.jobs[0]
    nmc = 800
    cfg = w.config(nmc=nmc, precision=gpu.MIXED_BATCHED, sweep_groups=2)
    r = gpu.run_batch(cfg, w.model, [job])[0]
    ocfg = O.config(nmc=nmc, nchain=w.nchain, nparameter_de=26, sweep_groups=w.nchain, seed=5)
    om = O.Model(w.model.winner, w.model.par_src, w.model.par_const, 13)
    datas = job.subjects
    rt = np.concatenate([d.rt for d in datas])
    so = np.concatenate([[0], np.cumsum([len(d.rt) for d in datas])])
    co = np.stack([d.cell_off for d in datas])
    o = O.run_hier(ocfg, om, fx.oprior(job.p_prior), fx.oprior(job.h_prior), rt, so, co,
                   *job.phi_start, np.stack([s[0] for s in job.subject_starts]),
                   np.stack([s[1] for s in job.subject_starts]),
                   np.stack([s[2] for s in job.subject_starts]))
    burn = 300
    # population locations and one subject's parameters
    zl, rl = _moments_agree(r["phi"].theta[:13], o[0][:13], burn)
    zs, rs = _moments_agree(r["subject_theta"][0].theta, o[3][0], burn)
    print("phi loc z max %.2f, sd ratio %.2f..%.2f; subject z max %.2f"
          % (zl.max(), rl.min(), rl.max(), zs.max()))
    assert diagnostics.psrf(r["phi"].theta[:13], burn).max() < 1.1
    assert zl.max() < 5.0 and zs.max() < 5.0
    assert np.all(np.abs(rl - 1) < 0.25) and np.all(np.abs(rs - 1) < 0.25)


def test_batched_job_equals_job_alone(gpu):
    # a run's trajectory does not depend on the other runs of its batch
    w = workloads.make("c4", scale=1 / 16, ll_fn=_oracle_ll)    # 16 ragged subjects + 4 fits
    cfg = w.config(nmc=6, precision=gpu.MIXED_BATCHED, sweep_groups=2, trials_per_tile=256,
                   sub_migration_prob=0.2, pop_migration_prob=0.2)
    together = gpu.run_batch(cfg, w.model, w.jobs)
    for k in (0, len(w.jobs) - 1):
        alone = gpu.run_batch(cfg, w.model, [w.jobs[k]])[0]
        for a, b in zip(alone["subjects"], together[k]["subjects"]):
            assert np.array_equal(a.theta, b.theta)
            assert np.array_equal(a.log_likelihoods, b.log_likelihoods)
        if alone["phi"] is not None:
            assert np.array_equal(alone["phi"].theta, together[k]["phi"].theta)


def test_scale_workload_full_size(gpu):
    # 1024 subjects x 2048 trials, 4 accumulators: every stored state's
    # log-likelihood within the per-proposal tolerance of FP64
    w = workloads.make("c5")
    cfg = w.config(nmc=4, thin=2, precision=gpu.MIXED_BATCHED)
    s = gpu.Session(cfg, w.model, w.jobs)
    s.advance(s.total_iters)
    res = s.download()[0]
    s.close()
    worst = 0.0
    moved = 0
    job = w.jobs[0]
    for i in range(0, 1024, 97):
        post = res["subjects"][i]
        ref = gpu.loglike(w.model, job.subjects[i], post.theta[:, :, -1].T, gpu.FP64_PLAIN)
        worst = max(worst, float(np.max(np.abs(post.log_likelihoods[:, -1] / ref - 1))))
        moved += int(np.sum(post.theta[:, :, -1] != job.subject_starts[i][0]))
    print("c5 stored-LL max rel %.2e, moved %d" % (worst, moved))
    assert worst <= 1e-6
    assert moved > 0
    assert np.all(np.isfinite(res["phi"].theta))
