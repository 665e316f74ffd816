"""The Rcpp shim's ABI path (tests/native/shim_abi.c: R column-major buffers,
the start column resolved from the samples object, R matrices transposed,
one job through ggdmc_run_batch or the direct entry points) against the
oracle's trajectories, with FP64 arithmetic in the reference's order."""
import ctypes as C

import numpy as np
import pytest

import fixtures as fx
import oracle_lib as O
import shim_lib as SH
from ggdmc_amd import _lib

pytestmark = pytest.mark.gpu
RT = 1e-9


def _cfg(gpu, n, nmc, **kw):
    return gpu.make_config(nmc=nmc, nchain=n, precision=gpu.FP64_BATCHED, sweep_groups=n, **kw)


def _ocfg(cfg, npde, seed):
    return O.config(nmc=cfg.nmc, nchain=cfg.nchain, nparameter_de=npde,
                    pop_migration_prob=cfg.pop_migration_prob,
                    sub_migration_prob=cfg.sub_migration_prob, sweep_groups=cfg.nchain, seed=seed)


def _hier_inputs(fix, S, n, seed):
    phi, subs = fx.hier_start(fix, S, n, seed=seed)
    return phi, subs


def _run_hier(gpu, L, fix, cfg, S, phi_s, sub_s, nmc_in, seed, direct):
    k = _lib.Keep()
    P = fix.P
    cm = fix.model.c_struct(k)
    pp = fix.pop_prior.c_struct(k)
    hp = fix.hyper_prior.c_struct(k)
    subj = (_lib.Subject * S)(*[fix.pop_data(s).c_struct(k) for s in range(S)])
    th_p = (_lib._d * S)(*[_lib.ptr(x[0]) for x in sub_s])
    lp_p = (_lib._d * S)(*[_lib.ptr(x[1]) for x in sub_s])
    ll_p = (_lib._d * S)(*[_lib.ptr(x[2]) for x in sub_s])
    phi_o, phi_c = SH.r_out(2 * P, cfg.nchain, cfg.nmc, k.items)
    outs = [SH.r_out(P, cfg.nchain, cfg.nmc, k.items) for _ in range(S)]
    sub_c = (_lib.Samples * S)(*[o[1] for o in outs])
    rc = L.shim_run(C.byref(cfg), C.byref(cm), C.byref(pp), C.byref(hp), _lib.KIND_HIER, S, subj,
                    None, P, _lib.ptr(phi_s[0]), _lib.ptr(phi_s[1]), _lib.ptr(phi_s[2]), th_p, lp_p,
                    ll_p, nmc_in, seed, 2 * P, direct, C.byref(phi_c), sub_c)
    _lib.check(rc)
    return phi_o, [o[0] for o in outs]


def _oracle_hier(fix, cfg, S, phi, subs, seed):
    rt = np.concatenate([fix.pop_data(s).rt for s in range(S)])
    so = np.concatenate([[0], np.cumsum([len(fix.pop_data(s).rt) for s in range(S)])])
    co = np.stack([fix.pop_data(s).cell_off for s in range(S)])
    return O.run_hier(_ocfg(cfg, 2 * fix.P, seed), fix.omodel, fx.oprior(fix.pop_prior),
                      fx.oprior(fix.hyper_prior), rt, so, co, *phi,
                      np.stack([s[0] for s in subs]), np.stack([s[1] for s in subs]),
                      np.stack([s[2] for s in subs]))


def _same(got, want):
    for g, w in zip(got, want):
        assert np.allclose(g, w, rtol=RT, atol=1e-12)


def test_shim_hierarchical_fresh_then_restart(gpu):
    L = SH.load()
    fix = fx.load("lba_data6")
    S, n, nmc, seed = 3, 8, 5, 4242
    phi, subs = _hier_inputs(fix, S, n, seed=5)
    cfg = _cfg(gpu, n, nmc, pop_migration_prob=0.2, sub_migration_prob=0.2)
    # a fresh samples object (column 1 filled, 7 columns) through ggdmc_run_batch
    phi_s = SH.r_samples(phi[0], phi[1], phi[2], nmc_in=7)
    sub_s = [SH.r_samples(*x, nmc_in=7) for x in subs]
    got_phi, got_sub = _run_hier(gpu, L, fix, cfg, S, phi_s, sub_s, 7, seed, direct=0)
    o = _oracle_hier(fix, cfg, S, phi, subs, seed)
    _same(got_phi, o[:3])
    for s in range(S):
        _same(got_sub[s], (o[3][s], o[4][s], o[5][s]))
    assert np.any(got_phi[0][:, :, -1] != phi[0])
    # restart from the finished posteriors: their last column, through ggdmc_run
    got2_phi, got2_sub = _run_hier(gpu, L, fix, cfg, S, got_phi, got_sub, nmc, seed + 1, direct=1)
    last = lambda x: (x[0][:, :, -1], x[1][:, -1], x[2][:, -1])  # noqa: E731
    o2 = _oracle_hier(fix, cfg, S, last(got_phi), [last(x) for x in got_sub], seed + 1)
    _same(got2_phi, o2[:3])
    for s in range(S):
        _same(got2_sub[s], (o2[3][s], o2[4][s], o2[5][s]))


@pytest.mark.parametrize("direct", [0, 1])
def test_shim_hyper_from_an_r_matrix(gpu, direct):
    L = SH.load()
    fix = fx.load("lba_data6")
    n, nmc, seed, S = 8, 6, 31, 12
    phi, subs = _hier_inputs(fix, S, n, seed=6)
    data = np.stack([x[0][:, 0] for x in subs])            # [S x P]
    rmat = np.asfortranarray(data)                          # as R holds hyper_dmi@data
    opp = fx.oprior(fix.pop_prior)
    pll = np.array([O.sumloghlike(opp, phi[0][:, c], data) for c in range(n)])
    cfg = _cfg(gpu, n, nmc, sub_migration_prob=0.2)
    k = _lib.Keep()
    pp, hp = fix.pop_prior.c_struct(k), fix.hyper_prior.c_struct(k)
    phi_s = SH.r_samples(phi[0], phi[1], pll, nmc_in=3)
    phi_o, phi_c = SH.r_out(2 * fix.P, n, nmc, k.items)
    _lib.check(L.shim_run(C.byref(cfg), None, C.byref(pp), C.byref(hp), _lib.KIND_HYPER, S, None,
                          _lib.ptr(rmat), fix.P, _lib.ptr(phi_s[0]), _lib.ptr(phi_s[1]),
                          _lib.ptr(phi_s[2]), None, None, None, 3, seed, 2 * fix.P, direct,
                          C.byref(phi_c), None))
    o = O.run_hyper(_ocfg(cfg, 2 * fix.P, seed), opp, fx.oprior(fix.hyper_prior), data,
                    phi[0], phi[1], pll)
    _same(phi_o, o)


def test_shim_subject_run(gpu):
    L = SH.load()
    fix = fx.load("lba_data6")
    n, nmc, seed = 9, 6, 17
    start = fx.near_start(fix, n, seed=2)
    cfg = _cfg(gpu, n, nmc, sub_migration_prob=0.3)
    k = _lib.Keep()
    cm, pr = fix.model.c_struct(k), fix.sub_prior.c_struct(k)
    subj = (_lib.Subject * 1)(fix.sub_data.c_struct(k))
    s = SH.r_samples(*start, nmc_in=4)
    outs = SH.r_out(fix.P, n, nmc, k.items)
    sub_c = (_lib.Samples * 1)(outs[1])
    one = lambda a: (_lib._d * 1)(_lib.ptr(a))  # noqa: E731
    _lib.check(L.shim_run(C.byref(cfg), C.byref(cm), C.byref(pr), None, _lib.KIND_SUBJECT, 1, subj,
                          None, fix.P, None, None, None, one(s[0]), one(s[1]), one(s[2]), 4, seed,
                          fix.P, 1, None, sub_c))
    o = O.run_subject(_ocfg(cfg, fix.P, seed), fix.omodel, fix.oprior, fix.sub_data.rt,
                      fix.sub_data.cell_off, *start)
    _same(outs[0], o)
