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


def _last(post):
    return post.theta[:, :, -1], post.summed_log_prior[:, -1], post.log_likelihoods[:, -1]


def test_single_subject_posterior_matches_oracle(gpu):
    # The reference's workflow (tests/testthat/Group5/*.r): a burn-in with
    # migration, then sampling without it. Migration is not a reversible move,
    # so its stationary bias depends on the sweep order; the migration-free
    # sampling phase of the two-group batched sampler (mixed precision) and of
    # the oracle in the reference's sequential order share the posterior.
    w = workloads.make("c2", ll_fn=_oracle_ll)
    job = w.jobs[0]
    n1, n2 = 1000, 3000
    cfg1 = gpu.make_config(nmc=n1, nchain=w.nchain, sub_migration_prob=0.06,
                           precision=gpu.MIXED_BATCHED, sweep_groups=2)
    cfg2 = gpu.make_config(nmc=n2, nchain=w.nchain, precision=gpu.MIXED_BATCHED, sweep_groups=2,
                           iter_offset=n1)
    specs = [gpu.JobSpec(job.kind, 100 + k, job.p_prior, start=job.start, subject=job.subject,
                         nparameter_de=13) for k in range(3)]
    burn = gpu.run_batch(cfg1, w.model, specs)
    specs2 = [gpu.JobSpec(job.kind, 100 + k, job.p_prior, start=_last(b["subjects"][0]),
                          subject=job.subject, nparameter_de=13) for k, b in enumerate(burn)]
    runs = gpu.run_batch(cfg2, w.model, specs2)
    g_theta = np.concatenate([r["subjects"][0].theta for r in runs], axis=1)
    om = O.Model(w.model.winner, w.model.par_src, w.model.par_const, 13)
    o1 = O.run_subject(O.config(nmc=n1, nchain=w.nchain, nparameter_de=13, sub_migration_prob=0.06,
                                sweep_groups=w.nchain, seed=7),
                       om, fx.oprior(job.p_prior), job.subject.rt, job.subject.cell_off, *job.start)
    o_theta, _, _ = O.run_subject(O.config(nmc=n2, nchain=w.nchain, nparameter_de=13,
                                           sweep_groups=w.nchain, seed=7, iter_offset=n1),
                                  om, fx.oprior(job.p_prior), job.subject.rt, job.subject.cell_off,
                                  o1[0][:, :, -1], o1[1][:, -1], o1[2][:, -1])
    rhat = diagnostics.psrf(g_theta, 0)
    z, r = _moments_agree(g_theta, o_theta, 0)
    print("R-hat max %.3f; mean z max %.2f; sd ratio %.3f..%.3f" % (rhat.max(), z.max(), r.min(), r.max()))
    assert rhat.max() < 1.1
    assert z.max() < 5.0
    assert np.all(np.abs(r - 1) < 0.15)


def test_hierarchical_posterior_matches_reference_order(gpu):
    # 32 subjects x 768 trials: the two-group mixed-precision sampler against
    # the FP64 batched sampler in the reference's sequential order (whose
    # trajectory follows the oracle draw for draw, test_gpu_sampler.py), three
    # independent runs each, no migration
    w = workloads.make("c3")
    job = w.jobs[0]
    # the first half of each run discarded; the batched runs are long enough
    # for the slowly mixing population level to reach R-hat < 1.1
    nmc, burn = 10000, 5000
    out = {}
    for name, prec, G, thin in (("mixed", gpu.MIXED_BATCHED, 2, 16),
                                ("fp64_seq", gpu.FP64_BATCHED, w.nchain, 4)):
        cfg = gpu.make_config(nmc=nmc, thin=thin, nchain=w.nchain, precision=prec, sweep_groups=G)
        specs = [gpu.JobSpec(job.kind, 500 + 10 * k + G, job.p_prior, h_prior=job.h_prior,
                             subjects=job.subjects, phi_start=job.phi_start,
                             subject_starts=job.subject_starts, nparameter_de=26)
                 for k in range(3)]
        rs = gpu.run_batch(cfg, w.model, specs)
        out[name] = (np.concatenate([r["phi"].theta for r in rs], axis=1),
                     np.concatenate([r["subjects"][0].theta for r in rs], axis=1))
    for name in out:
        print(name, "phi R-hat", np.round(diagnostics.psrf(out[name][0], burn), 3))
    zl, rl = _moments_agree(out["mixed"][0][:13], out["fp64_seq"][0][:13], burn)
    zs, rs_ = _moments_agree(out["mixed"][1], out["fp64_seq"][1], burn)
    print("phi loc z max %.2f, sd ratio %.2f..%.2f; subject z max %.2f, sd ratio %.2f..%.2f"
          % (zl.max(), rl.min(), rl.max(), zs.max(), rs_.min(), rs_.max()))
    # the population level mixes slowly in either order (locations are tied to
    # the subjects' mean): 160000 iterations of the batched sampler reach
    # R-hat < 1.1 on every population and subject parameter
    r_mixed = diagnostics.psrf(out["mixed"][0], burn).max()
    r_mixed_sub = diagnostics.psrf(out["mixed"][1], burn).max()
    print("mixed R-hat max: population %.3f, subject 0 %.3f" % (r_mixed, r_mixed_sub))
    assert r_mixed < 1.1 and r_mixed_sub < 1.1, (r_mixed, r_mixed_sub)
    assert zl.max() < 5.0 and zs.max() < 5.0
    assert np.all(np.abs(rl - 1) < 0.25) and np.all(np.abs(rs_ - 1) < 0.25)


@pytest.mark.parametrize("launch_mode", [1, 2, 3, 4])
def test_batched_job_equals_job_alone(gpu, launch_mode):
    # a run's trajectory does not depend on the other runs of its batch
    w = workloads.make("c4", scale=1 / 16, ll_fn=_oracle_ll)    # 16 ragged subjects + 4 fits
    # (items spanning every group, launch mode 4, take whole subjects)
    cfg = gpu.make_config(nmc=6, nchain=w.nchain, precision=gpu.MIXED_BATCHED, sweep_groups=2,
                          trials_per_tile=256 if launch_mode < 4 else 0, sub_migration_prob=0.2,
                          pop_migration_prob=0.2, launch_mode=launch_mode)
    together = gpu.run_batch(cfg, w.model, w.jobs)
    for k in (0, len(w.jobs) - 1):
        alone = gpu.run_batch(cfg, w.model, [w.jobs[k]])[0]
        for a, b in zip(alone["subjects"], together[k]["subjects"]):
            assert np.array_equal(a.theta, b.theta)
            assert np.array_equal(a.log_likelihoods, b.log_likelihoods)
        if alone["phi"] is not None:
            assert np.array_equal(alone["phi"].theta, together[k]["phi"].theta)


def test_mixed_batch_schedules_agree(gpu):
    # hierarchical run + independent fits in one batch: a launch per chain group
    # and one launch for every group, with whole or tiled subjects, give the
    # same bits (the waits between groups only order the work)
    w = workloads.make("c4", scale=1 / 16, ll_fn=_oracle_ll)
    ref = None
    for k, tpt, lm, wps in ((0, 0, 1, 3), (2, 0, 2, 3), (0, 256, 2, 3), (4, 128, 2, 3), (2, 0, 3, 3),
                            (0, 256, 3, 3), (1, 0, 4, 3), (4, 0, 4, 3), (4, 0, 3, 4), (2, 0, 4, 4)):
        cfg = gpu.make_config(nmc=6, nchain=w.nchain, precision=gpu.MIXED_BATCHED, sweep_groups=2,
                              chains_per_item=k, trials_per_tile=tpt, launch_mode=lm,
                              waves_per_simd=wps,
                              sub_migration_prob=0.2, pop_migration_prob=0.2)
        res = gpu.run_batch(cfg, w.model, w.jobs)
        if ref is None:
            ref = res
            continue
        for a, b in zip(res, ref):
            for x, y in zip(a["subjects"], b["subjects"]):
                assert np.array_equal(x.theta, y.theta), (k, tpt, lm)
                assert np.array_equal(x.log_likelihoods, y.log_likelihoods), (k, tpt, lm)
                assert np.array_equal(x.summed_log_prior, y.summed_log_prior), (k, tpt, lm)
            if a["phi"] is not None:
                assert np.array_equal(a["phi"].theta, b["phi"].theta), (k, tpt, lm)
                assert np.array_equal(a["phi"].log_likelihoods, b["phi"].log_likelihoods), (k, tpt, lm)


@pytest.mark.parametrize("launch_mode", [1, 2, 3, 4])
def test_batch_of_hierarchical_runs_equals_each_alone(gpu, launch_mode):
    # three hierarchical runs (same data, different Philox keys) in one batch:
    # each gives the bits it gives alone, phi and subjects, with migration
    w = workloads.replicate(workloads.make("c3", scale=1 / 8, ll_fn=_oracle_ll), 3)
    cfg = gpu.make_config(nmc=6, nchain=w.nchain, precision=gpu.MIXED_BATCHED, sweep_groups=2,
                          sub_migration_prob=0.2, pop_migration_prob=0.2,
                          launch_mode=launch_mode)
    together = gpu.run_batch(cfg, w.model, w.jobs)
    for k in range(len(w.jobs)):
        alone = gpu.run_batch(cfg, w.model, [w.jobs[k]])[0]
        assert np.array_equal(alone["phi"].theta, together[k]["phi"].theta)
        assert np.array_equal(alone["phi"].log_likelihoods, together[k]["phi"].log_likelihoods)
        for a, b in zip(alone["subjects"], together[k]["subjects"]):
            assert np.array_equal(a.theta, b.theta)
            assert np.array_equal(a.log_likelihoods, b.log_likelihoods)
    # different keys give different trajectories
    assert not np.array_equal(together[0]["phi"].theta, together[1]["phi"].theta)


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


def test_batching_workload_full_size(gpu):
    # the batching workload at its configured size: 256 subjects with ragged
    # trial counts U{96..2048} in one hierarchical run plus 64 independent
    # single-subject fits, 39 chains, one session (the reference runs them as
    # separate processes, R/sampling.R:13-121). Every fit and a spread of
    # subjects (the shortest and longest included): stored log-likelihoods
    # within the per-proposal tolerance of an FP64 recomputation, every state
    # finite, chains moving
    w = workloads.make("c4")
    hier = w.jobs[0]
    ntr = np.array([len(d.rt) for d in hier.subjects])
    assert len(hier.subjects) == 256 and len(w.jobs) == 65
    assert ntr.min() < 150 and ntr.max() > 1950
    cfg = w.config(nmc=9, thin=4, precision=gpu.MIXED_BATCHED)     # 35 iterations
    s = gpu.Session(cfg, w.model, w.jobs)
    s.advance(s.total_iters)
    res = s.download()
    print("c4 decomposition", s.decomposition())
    s.close()
    check = sorted(set(list(range(0, 256, 17)) + [int(ntr.argmin()), int(ntr.argmax())]))
    pairs = [(hier.subjects[i], res[0]["subjects"][i], hier.subject_starts[i][0]) for i in check]
    pairs += [(j.subject, r["subjects"][0], j.start[0]) for j, r in zip(w.jobs[1:], res[1:])]
    worst = 0.0
    moved = 0
    for data, post, th0 in pairs:
        assert np.all(np.isfinite(post.theta)) and np.all(np.isfinite(post.log_likelihoods))
        ref = gpu.loglike(w.model, data, post.theta[:, :, -1].T, gpu.FP64_PLAIN)
        worst = max(worst, float(np.max(np.abs(post.log_likelihoods[:, -1] / ref - 1))))
        moved += int(np.any(post.theta[:, :, -1] != th0))
    print("c4: %d states checked, stored-LL max rel %.2e, %d of %d runs moved"
          % (39 * len(pairs), worst, moved, len(pairs)))
    assert worst <= 1e-6
    assert moved == len(pairs)
    assert np.all(np.isfinite(res[0]["phi"].theta))


def test_ragged_batch_follows_oracle(gpu):
    # FP64 arithmetic, two chain groups, trial tiles of 256: a hierarchical run
    # whose subjects have 96 and 2048 trials (and two in between) batched with
    # two single-subject fits follows the oracle draw for draw, every run
    w = workloads.make("c4", scale=1 / 32, ll_fn=_oracle_ll)     # for its model and two fits
    from ggdmc_amd.workloads import BV_MEAN, BV_SCALE, _hier_job
    hier, _ = _hier_job(w.model, BV_MEAN, BV_SCALE, [96, 2048, 700, 1500], 12, 9032, 77, _oracle_ll)
    fits = [gpu.JobSpec(j.kind, j.seed, j.p_prior, start=(j.start[0][:, :12], j.start[1][:12],
                                                             j.start[2][:12]),
                        subject=j.subject, nparameter_de=13) for j in w.jobs[1:3]]
    jobs = [hier] + fits
    cfg = gpu.make_config(nmc=5, nchain=12, precision=gpu.FP64_BATCHED, sweep_groups=2,
                          trials_per_tile=256, sub_migration_prob=0.2, pop_migration_prob=0.2)
    res = gpu.run_batch(cfg, w.model, jobs)
    om = O.Model(w.model.winner, w.model.par_src, w.model.par_const, 13)
    ocfg = lambda npde, seed: O.config(nmc=5, nchain=12, nparameter_de=npde, sweep_groups=2,  # noqa: E731
                                       sub_migration_prob=0.2, pop_migration_prob=0.2, seed=seed)
    rt = np.concatenate([d.rt for d in hier.subjects])
    so = np.concatenate([[0], np.cumsum([len(d.rt) for d in hier.subjects])])
    co = np.stack([d.cell_off for d in hier.subjects])
    o = O.run_hier(ocfg(26, hier.seed), om, fx.oprior(hier.p_prior), fx.oprior(hier.h_prior), rt, so,
                   co, *hier.phi_start, np.stack([s[0] for s in hier.subject_starts]),
                   np.stack([s[1] for s in hier.subject_starts]),
                   np.stack([s[2] for s in hier.subject_starts]))
    tol = dict(rtol=1e-9, atol=1e-12)
    assert np.allclose(res[0]["phi"].theta, o[0], **tol)
    assert np.allclose(res[0]["phi"].log_likelihoods, o[2], **tol)
    for s in range(4):
        assert np.allclose(res[0]["subjects"][s].theta, o[3][s], **tol)
        assert np.allclose(res[0]["subjects"][s].log_likelihoods, o[5][s], **tol)
    for j, r in zip(fits, res[1:]):
        ot, olp, oll = O.run_subject(ocfg(13, j.seed), om, fx.oprior(j.p_prior), j.subject.rt,
                                     j.subject.cell_off, *j.start)
        assert np.allclose(r["subjects"][0].theta, ot, **tol)
        assert np.allclose(r["subjects"][0].log_likelihoods, oll, **tol)
    assert np.any(res[0]["subjects"][1].theta[:, :, -1] != hier.subject_starts[1][0])
