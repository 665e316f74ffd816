"""Sampler parity: with FP64 arithmetic the HIP sampler follows the oracle's
trajectory draw for draw (same Philox streams, same chain-group order), for
the plain path (the reference's sequential order) and the batched kernels;
the mixed-precision batched path keeps stored log-likelihoods within the
per-proposal tolerance of an FP64 recomputation."""
import numpy as np
import pytest

import fixtures as fx
import oracle_lib as O

pytestmark = pytest.mark.gpu

RT = 1e-9   # trajectory tolerance (FP64 vs FP64: summation order only)


def _cfg(g, **kw):
    base = dict(nmc=6, nchain=12, thin=1, sub_migration_prob=0.25)
    base.update(kw)
    return g.make_config(**base)


def _ocfg(fix, cfg, nparameter_de=None, seed=9032):
    return O.config(nmc=cfg.nmc, nchain=cfg.nchain, thin=cfg.thin,
                    nparameter_de=nparameter_de or fix.P,
                    pop_migration_prob=cfg.pop_migration_prob,
                    sub_migration_prob=cfg.sub_migration_prob,
                    gamma_precursor=cfg.gamma_precursor, rp=cfg.rp,
                    is_hblocked=bool(cfg.is_hblocked), is_pblocked=bool(cfg.is_pblocked),
                    sweep_groups=cfg.sweep_groups or cfg.nchain, seed=seed,
                    iter_offset=cfg.iter_offset)


def _subject_case(g, precision, groups, tpt=0, blocked=False, mig=0.25, thin=1, nmc=6):
    fix = fx.load("lba_data6")
    n = 12
    start = fx.near_start(fix, n, seed=7)
    cfg = _cfg(g, nchain=n, precision=precision, sweep_groups=groups, trials_per_tile=tpt,
               is_pblocked=blocked, sub_migration_prob=mig, thin=thin, nmc=nmc)
    post = g.run_subject(cfg, fix.model, fix.sub_prior, fix.sub_data, start, seed=9032)
    ocfg = _ocfg(fix, cfg)
    ocfg.sweep_groups = groups or n
    oth, olp, oll = O.run_subject(ocfg, fix.omodel, fix.oprior, fix.sub_data.rt,
                                  fix.sub_data.cell_off, *start)
    return post, oth, olp, oll, start


def _same(post, oth, olp, oll):
    assert np.allclose(post.theta, oth, rtol=RT, atol=1e-12)
    assert np.allclose(post.summed_log_prior, olp, rtol=RT, atol=1e-12)
    assert np.allclose(post.log_likelihoods, oll, rtol=RT, atol=1e-12)


def test_plain_fp64_follows_oracle(gpu):
    post, oth, olp, oll, start = _subject_case(gpu, gpu.FP64_PLAIN, 12)
    _same(post, oth, olp, oll)
    assert not np.allclose(oth[:, :, -1], start[0])     # the chains moved


def test_batched_fp64_two_groups_follows_oracle(gpu):
    _same(*_subject_case(gpu, gpu.FP64_BATCHED, 2)[:4])


def test_batched_fp64_sequential_groups_follows_oracle(gpu):
    _same(*_subject_case(gpu, gpu.FP64_BATCHED, 12)[:4])


def test_batched_multitile_follows_oracle(gpu):
    # 768 trials in tiles of 128: one launch per group, last-arriving tile accepts
    _same(*_subject_case(gpu, gpu.FP64_BATCHED, 2, tpt=128)[:4])


def test_blocked_follows_oracle(gpu):
    _same(*_subject_case(gpu, gpu.FP64_PLAIN, 12, blocked=True, nmc=3)[:4])
    _same(*_subject_case(gpu, gpu.FP64_BATCHED, 3, blocked=True, nmc=3)[:4])


def test_thinning_follows_oracle(gpu):
    _same(*_subject_case(gpu, gpu.FP64_BATCHED, 2, thin=3, nmc=4)[:4])


def _hier_case(g, precision, groups, S=6, n=10, nmc=5, tpt=0, blocked=False):
    fix = fx.load("lba_data6")
    (phi, plp, pll), subs = fx.hier_start(fix, S, n, seed=11)
    cfg = g.make_config(nmc=nmc, nchain=n, pop_migration_prob=0.3, sub_migration_prob=0.2,
                        precision=precision, sweep_groups=groups, trials_per_tile=tpt,
                        is_hblocked=blocked, is_pblocked=blocked)
    datas = [fix.pop_data(s) for s in range(S)]
    r = g.run(cfg, fix.model, fix.pop_prior, fix.hyper_prior, datas, (phi, plp, pll), subs,
              seed=4242)
    ocfg = _ocfg(fix, cfg, nparameter_de=2 * fix.P, seed=4242)
    ocfg.sweep_groups = groups or n
    rt = np.concatenate([d.rt for d in datas])
    so = np.concatenate([[0], np.cumsum([len(d.rt) for d in datas])])
    co = np.stack([d.cell_off for d in datas])
    o = O.run_hier(ocfg, fix.omodel, fx.oprior(fix.pop_prior), fx.oprior(fix.hyper_prior), rt, so,
                   co, phi, plp, pll, np.stack([s[0] for s in subs]),
                   np.stack([s[1] for s in subs]), np.stack([s[2] for s in subs]))
    return r, o


def _same_hier(r, o):
    pth, plp, pll, sth, slp, sll = o
    assert np.allclose(r["phi"].theta, pth, rtol=RT, atol=1e-12)
    assert np.allclose(r["phi"].summed_log_prior, plp, rtol=RT, atol=1e-12)
    assert np.allclose(r["phi"].log_likelihoods, pll, rtol=1e-8, atol=1e-9)
    for s, post in enumerate(r["subject_theta"]):
        assert np.allclose(post.theta, sth[s], rtol=RT, atol=1e-12)
        assert np.allclose(post.summed_log_prior, slp[s], rtol=RT, atol=1e-12)
        assert np.allclose(post.log_likelihoods, sll[s], rtol=RT, atol=1e-12)


def test_hier_plain_follows_oracle(gpu):
    _same_hier(*_hier_case(gpu, gpu.FP64_PLAIN, 10))


def test_hier_batched_follows_oracle(gpu):
    _same_hier(*_hier_case(gpu, gpu.FP64_BATCHED, 2))
    _same_hier(*_hier_case(gpu, gpu.FP64_BATCHED, 2, tpt=256))


def test_hier_blocked_follows_oracle(gpu):
    _same_hier(*_hier_case(gpu, gpu.FP64_BATCHED, 2, nmc=2, blocked=True))


def test_hyper_follows_oracle(gpu):
    fix = fx.load("lba_data6")
    S, n, P = 32, 10, fix.P
    rng = np.random.default_rng(2)
    data = fx.truth_vector(fix)[None, :] * np.exp(rng.normal(0, 0.1, (S, P)))
    (phi, _, _), _ = fx.hier_start(fix, 0, n, seed=3)
    hp = fx.oprior(fix.hyper_prior)
    pp = fx.oprior(fix.pop_prior)
    plp = np.array([O.sumlogprior(hp, phi[:, c]) for c in range(n)])
    pll = np.array([O.sumloghlike(pp, phi[:, c], data) for c in range(n)])
    for prec, G in ((gpu.FP64_PLAIN, n), (gpu.FP64_BATCHED, 2)):
        cfg = gpu.make_config(nmc=8, nchain=n, sub_migration_prob=0.2, precision=prec,
                              sweep_groups=G)
        post = gpu.run_hyper(cfg, fix.pop_prior, fix.hyper_prior, data, (phi, plp, pll), seed=77)
        ocfg = _ocfg(fix, cfg, nparameter_de=2 * P, seed=77)
        ocfg.sweep_groups = G
        oth, olp, oll = O.run_hyper(ocfg, pp, hp, data, phi, plp, pll)
        assert np.allclose(post.theta, oth, rtol=RT, atol=1e-12)
        assert np.allclose(post.log_likelihoods, oll, rtol=1e-8, atol=1e-9)


def test_mixed_hier_stored_ll_within_tolerance(gpu):
    fix = fx.load("lba_data6")
    S, n = 8, 12
    (phi, plp, pll), subs = fx.hier_start(fix, S, n, seed=5)
    cfg = gpu.make_config(nmc=20, nchain=n, precision=gpu.MIXED_BATCHED, sweep_groups=2)
    r = gpu.run(cfg, fix.model, fix.pop_prior, fix.hyper_prior,
                [fix.pop_data(s) for s in range(S)], (phi, plp, pll), subs, seed=99)
    worst = 0.0
    moved = 0
    for s, post in enumerate(r["subject_theta"]):
        d = fix.pop_data(s)
        for k in (5, 19):
            for c in range(n):
                ref = O.lba_loglike(fix.omodel, d.rt, d.cell_off, post.theta[:, c, k])
                worst = max(worst, abs(post.log_likelihoods[c, k] / ref - 1))
        moved += int(np.sum(post.theta[:, :, -1] != subs[s][0]))
    print("mixed hierarchical: stored-LL max rel %.2e" % worst)
    assert worst <= 1e-6
    assert moved > 0


@pytest.mark.parametrize("precision", ["MIXED_BATCHED", "FP64_BATCHED"])
def test_decomposition_never_changes_results(gpu, precision):
    # work items (chain ranges x trial tiles, or one item per subject over every
    # group) only change the schedule: log-likelihoods are canonical chunk sums
    fix = fx.load("lba_data6")
    S, n = 5, 12
    (phi, plp, pll), subs = fx.hier_start(fix, S, n, seed=21)
    datas = [fix.pop_data(s) for s in range(S)]
    prec = getattr(gpu, precision)
    ref = None
    # (chains per item, trials per item, launch mode: 1 = a launch per chain
    # group, 2 = every group in one launch, later groups waiting for earlier,
    # 3 = as 2 with several iterations per launch, 4 = as 3 with an item
    # holding its chain range of every group)
    # (the fourth entry: the subject kernel built for that many waves per SIMD)
    for k, tpt, lm, w in ((-1, 0, 1, 3), (0, 0, 1, 3), (1, 0, 1, 3), (4, 128, 1, 3), (2, 256, 1, 3),
                          (0, 384, 1, 3), (0, 0, 2, 3), (1, 0, 2, 3), (4, 128, 2, 3), (2, 256, 2, 3),
                          (0, 0, 3, 3), (1, 0, 3, 3), (2, 0, 3, 3), (4, 128, 3, 3), (0, 0, 4, 3),
                          (1, 0, 4, 3), (2, 0, 4, 3), (5, 0, 4, 3), (1, 0, 3, 4), (2, 0, 4, 4),
                          (4, 128, 3, 4)):
        cfg = gpu.make_config(nmc=6, nchain=n, pop_migration_prob=0.2, sub_migration_prob=0.3,
                              precision=prec, sweep_groups=2, chains_per_item=k,
                              trials_per_tile=tpt, launch_mode=lm, waves_per_simd=w)
        r = gpu.run(cfg, fix.model, fix.pop_prior, fix.hyper_prior, datas, (phi, plp, pll), subs,
                    seed=31)
        if ref is None:
            ref = r
            continue
        assert np.array_equal(r["phi"].theta, ref["phi"].theta), (k, tpt, lm, w)
        assert np.array_equal(r["phi"].log_likelihoods, ref["phi"].log_likelihoods), (k, tpt, lm, w)
        for a, b in zip(r["subject_theta"], ref["subject_theta"]):
            assert np.array_equal(a.theta, b.theta), (k, tpt, lm, w)
            assert np.array_equal(a.log_likelihoods, b.log_likelihoods), (k, tpt, lm, w)
            assert np.array_equal(a.summed_log_prior, b.summed_log_prior), (k, tpt, lm, w)


def test_loglike_api_equals_sampler_values(gpu):
    # ggdmc_loglike uses the sampler's chunk sums: a stored state's LL is
    # reproduced bit for bit
    fix = fx.load("lba_data6")
    n = 12
    start = fx.near_start(fix, n, seed=3)
    cfg = _cfg(gpu, nchain=n, precision=gpu.MIXED_BATCHED, sweep_groups=2, nmc=8)
    post = gpu.run_subject(cfg, fix.model, fix.sub_prior, fix.sub_data, start, seed=5)
    th = post.theta[:, :, -1].T
    ll = gpu.loglike(fix.model, fix.sub_data, th, gpu.MIXED_BATCHED)
    moved = post.log_likelihoods[:, -1] != start[2]
    assert moved.any()
    assert np.array_equal(ll[moved], post.log_likelihoods[moved, -1])


def test_three_chains_follow_oracle(gpu):
    # the fewest chains the reference accepts (de.cpp:4-28): each chain's two
    # partners are the other two; batched with three groups and plain
    fix = fx.load("lba_data6")
    n = 3
    start = fx.near_start(fix, n, seed=13)
    for prec in (gpu.FP64_BATCHED, gpu.FP64_PLAIN):
        cfg = _cfg(gpu, nchain=n, precision=prec, sweep_groups=3, nmc=8, sub_migration_prob=0.3)
        post = gpu.run_subject(cfg, fix.model, fix.sub_prior, fix.sub_data, start, seed=9032)
        ocfg = _ocfg(fix, cfg)
        oth, olp, oll = O.run_subject(ocfg, fix.omodel, fix.oprior, fix.sub_data.rt,
                                      fix.sub_data.cell_off, *start)
        _same(post, oth, olp, oll)
    with pytest.raises(gpu.GgdmcError):
        gpu.run_subject(_cfg(gpu, nchain=n, precision=gpu.FP64_BATCHED, sweep_groups=2),
                        fix.model, fix.sub_prior, fix.sub_data, start, seed=1)


def test_long_subject_follows_oracle(gpu):
    # 100,608 trials (each cell of the fixture repeated 131 times): 786 chunks
    # of 128 trials, more than the 682 chunk sums per chain an item holds in
    # LDS with 6 chains per item, so the sums run in batches
    fix = fx.load("lba_data6")
    d0 = fix.sub_data
    cells = [np.tile(d0.rt[d0.cell_off[c]:d0.cell_off[c + 1]], 131)
             for c in range(len(d0.cell_off) - 1)]
    data = fx.SubjectData.from_cells(cells)
    assert len(data.rt) == 100608
    n = 12
    start = fx.near_start(fix, n, seed=17, data=data)
    ocfg = None
    for k in (6, 0):
        cfg = _cfg(gpu, nchain=n, precision=gpu.FP64_BATCHED, sweep_groups=2, nmc=3,
                   chains_per_item=k, launch_mode=1)
        post = gpu.run_subject(cfg, fix.model, fix.sub_prior, data, start, seed=9032)
        if ocfg is None:
            ocfg = _ocfg(fix, cfg)
            ocfg.sweep_groups = 2
            ref = O.run_subject(ocfg, fix.omodel, fix.oprior, data.rt, data.cell_off, *start)
        _same(post, *ref)


def test_hier_with_an_empty_subject_follows_oracle(gpu):
    # a subject with no trials contributes log-likelihood 0: its chains sample
    # the population distribution; the other subjects are unaffected
    fix = fx.load("lba_data6")
    S, n = 3, 10
    (phi, plp, pll), subs = fx.hier_start(fix, S, n, seed=19)
    empty = fx.SubjectData(np.zeros(0), np.zeros(len(fix.sub_data.cell_off), np.int64))
    datas = [fix.pop_data(0), empty, fix.pop_data(2)]
    subs[1] = (subs[1][0], subs[1][1], np.zeros(n))
    for prec, G in ((gpu.FP64_BATCHED, 2), (gpu.FP64_PLAIN, n)):
        cfg = gpu.make_config(nmc=6, nchain=n, pop_migration_prob=0.3, sub_migration_prob=0.2,
                              precision=prec, sweep_groups=G)
        r = gpu.run(cfg, fix.model, fix.pop_prior, fix.hyper_prior, datas, (phi, plp, pll), subs,
                    seed=555)
        ocfg = _ocfg(fix, cfg, nparameter_de=2 * fix.P, seed=555)
        ocfg.sweep_groups = G
        rt = np.concatenate([d.rt for d in datas])
        so = np.concatenate([[0], np.cumsum([len(d.rt) for d in datas])])
        co = np.stack([d.cell_off for d in datas])
        o = O.run_hier(ocfg, fix.omodel, fx.oprior(fix.pop_prior), fx.oprior(fix.hyper_prior), rt,
                       so, co, phi, plp, pll, np.stack([s[0] for s in subs]),
                       np.stack([s[1] for s in subs]), np.stack([s[2] for s in subs]))
        _same_hier(r, o)
        assert np.all(r["subject_theta"][1].log_likelihoods == 0.0)
        assert not np.array_equal(r["subject_theta"][1].theta[:, :, -1], subs[1][0])


@pytest.mark.parametrize("launch_mode", [2, 3, 4])
def test_advance_in_pieces_equals_one_advance(gpu, launch_mode):
    # a session advanced in uneven pieces (a restart's pattern; with several
    # iterations per launch the pieces end mid-launch) gives the bits of one
    # advance over the same iterations, stored columns included
    fix = fx.load("lba_data6")
    S, n = 4, 10
    (phi, plp, pll), subs = fx.hier_start(fix, S, n, seed=23)
    datas = [fix.pop_data(s) for s in range(S)]
    cfg = gpu.make_config(nmc=29, thin=2, nchain=n, pop_migration_prob=0.1, sub_migration_prob=0.1,
                          precision=gpu.MIXED_BATCHED, sweep_groups=2, chains_per_item=2,
                          launch_mode=launch_mode)
    job = gpu.JobSpec(gpu.KIND_HIER, 77, fix.pop_prior, h_prior=fix.hyper_prior, subjects=datas,
                      phi_start=(phi, plp, pll), subject_starts=subs, nparameter_de=2 * fix.P)
    out = []
    for pieces in ((57,), (5, 7, 13, 31, 1)):
        s = gpu.Session(cfg, fix.model, [job])
        for p in pieces:
            s.advance(p)
        out.append(s.download()[0])
        s.close()
    a, b = out
    assert np.array_equal(a["phi"].theta, b["phi"].theta)
    assert np.array_equal(a["phi"].log_likelihoods, b["phi"].log_likelihoods)
    for x, y in zip(a["subjects"], b["subjects"]):
        assert np.array_equal(x.theta, y.theta)
        assert np.array_equal(x.log_likelihoods, y.log_likelihoods)
        assert np.array_equal(x.summed_log_prior, y.summed_log_prior)
    assert not np.array_equal(a["subjects"][0].theta[:, :, -1], subs[0][0])


def test_three_chain_groups_in_every_launch_mode(gpu):
    # three chain groups: with every group in one launch a group waits for all
    # of the previous group's items, also inside an item that holds its chain
    # range of every group; the schedule changes no bit
    fix = fx.load("lba_data6")
    S, n = 4, 10
    (phi, plp, pll), subs = fx.hier_start(fix, S, n, seed=37)
    datas = [fix.pop_data(s) for s in range(S)]
    ref = None
    for k, lm in ((0, 1), (1, 2), (1, 3), (1, 4), (2, 4), (4, 4)):
        cfg = gpu.make_config(nmc=8, nchain=n, pop_migration_prob=0.2, sub_migration_prob=0.2,
                              precision=gpu.FP64_BATCHED, sweep_groups=3, chains_per_item=k,
                              launch_mode=lm)
        r = gpu.run(cfg, fix.model, fix.pop_prior, fix.hyper_prior, datas, (phi, plp, pll), subs,
                    seed=43)
        if ref is None:
            ref = r
            continue
        assert np.array_equal(r["phi"].theta, ref["phi"].theta), (k, lm)
        for a, b in zip(r["subject_theta"], ref["subject_theta"]):
            assert np.array_equal(a.theta, b.theta), (k, lm)
            assert np.array_equal(a.log_likelihoods, b.log_likelihoods), (k, lm)
            assert np.array_equal(a.summed_log_prior, b.summed_log_prior), (k, lm)
    assert not np.array_equal(ref["subject_theta"][0].theta[:, :, -1], subs[0][0])


def test_hblocked_population_in_every_launch_mode(gpu):
    # blocked population sweeps: every iteration takes the whole-run population
    # step (one workgroup of the launch); with several iterations per launch it
    # waits for the whole run's previous iteration. Follows the oracle, and
    # every launch mode gives the same bits.
    fix = fx.load("lba_data6")
    S, n = 4, 10
    (phi, plp, pll), subs = fx.hier_start(fix, S, n, seed=29)
    datas = [fix.pop_data(s) for s in range(S)]
    ref = None
    for lm in (1, 2, 3, 4):
        cfg = gpu.make_config(nmc=6, nchain=n, pop_migration_prob=0.2, sub_migration_prob=0.2,
                              precision=gpu.FP64_BATCHED, sweep_groups=2, is_hblocked=True,
                              chains_per_item=2, launch_mode=lm)
        r = gpu.run(cfg, fix.model, fix.pop_prior, fix.hyper_prior, datas, (phi, plp, pll), subs,
                    seed=91)
        if ref is None:
            ocfg = _ocfg(fix, cfg, nparameter_de=2 * fix.P, seed=91)
            ocfg.sweep_groups = 2
            rt = np.concatenate([d.rt for d in datas])
            so = np.concatenate([[0], np.cumsum([len(d.rt) for d in datas])])
            co = np.stack([d.cell_off for d in datas])
            o = O.run_hier(ocfg, fix.omodel, fx.oprior(fix.pop_prior), fx.oprior(fix.hyper_prior),
                           rt, so, co, phi, plp, pll, np.stack([s[0] for s in subs]),
                           np.stack([s[1] for s in subs]), np.stack([s[2] for s in subs]))
            _same_hier(r, o)
            ref = r
            continue
        assert np.array_equal(r["phi"].theta, ref["phi"].theta), lm
        assert np.array_equal(r["phi"].log_likelihoods, ref["phi"].log_likelihoods), lm
        for a, b in zip(r["subject_theta"], ref["subject_theta"]):
            assert np.array_equal(a.theta, b.theta), lm
            assert np.array_equal(a.log_likelihoods, b.log_likelihoods), lm


def test_hyper_and_hierarchical_batch_in_every_launch_mode(gpu):
    # a hyper-only run (fixed data rows, whole-run population step every
    # iteration) in one batch with a hierarchical run: every launch mode gives
    # the same bits
    fix = fx.load("lba_data6")
    S, n, P = 4, 10, fix.P
    (phi, plp, pll), subs = fx.hier_start(fix, S, n, seed=31)
    datas = [fix.pop_data(s) for s in range(S)]
    rng = np.random.default_rng(5)
    hdata = fx.truth_vector(fix)[None, :] * np.exp(rng.normal(0, 0.1, (24, P)))
    hp, pp = fx.oprior(fix.hyper_prior), fx.oprior(fix.pop_prior)
    hlp = np.array([O.sumlogprior(hp, phi[:, c]) for c in range(n)])
    hll = np.array([O.sumloghlike(pp, phi[:, c], hdata) for c in range(n)])
    jobs = [gpu.JobSpec(gpu.KIND_HIER, 41, fix.pop_prior, h_prior=fix.hyper_prior, subjects=datas,
                        phi_start=(phi, plp, pll), subject_starts=subs, nparameter_de=2 * P),
            gpu.JobSpec(gpu.KIND_HYPER, 43, fix.pop_prior, h_prior=fix.hyper_prior,
                        phi_start=(phi, hlp, hll), hyper_data=hdata, nparameter_de=2 * P)]
    ref = None
    for lm in (1, 2, 3, 4):
        cfg = gpu.make_config(nmc=6, nchain=n, pop_migration_prob=0.2, sub_migration_prob=0.2,
                              precision=gpu.MIXED_BATCHED, sweep_groups=2, chains_per_item=2,
                              launch_mode=lm)
        res = gpu.run_batch(cfg, fix.model, jobs)
        if ref is None:
            ref = res
            continue
        for a, b in zip(res, ref):
            assert np.array_equal(a["phi"].theta, b["phi"].theta), lm
            assert np.array_equal(a["phi"].log_likelihoods, b["phi"].log_likelihoods), lm
            for x, y in zip(a["subjects"], b["subjects"]):
                assert np.array_equal(x.theta, y.theta), lm
                assert np.array_equal(x.log_likelihoods, y.log_likelihoods), lm
    assert not np.array_equal(ref[1]["phi"].theta[:, :, -1], phi)
