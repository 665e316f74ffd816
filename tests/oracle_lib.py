"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg only, as the checker or the timed CPU baseline.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)
_u32 = C.POINTER(C.c_uint32)

MODEL_LBA = 0
MODEL_NORMAL = 1


class OracleModel(C.Structure):
    _fields_ = [("n_acc", C.c_int32), ("n_cell", C.c_int32), ("npar", C.c_int32),
                ("winner", _i32), ("par_src", _i32), ("par_const", _d)]


class OraclePrior(C.Structure):
    _fields_ = [("npar", C.c_int32), ("dist", _i32), ("p0", _d), ("p1", _d),
                ("lower", _d), ("upper", _d)]


class OracleConfig(C.Structure):
    _fields_ = [("nmc", C.c_int32), ("thin", C.c_int32), ("nchain", C.c_int32),
                ("nparameter_de", C.c_int32),
                ("pop_migration_prob", C.c_double), ("sub_migration_prob", C.c_double),
                ("gamma_precursor", C.c_double), ("rp", C.c_double),
                ("is_hblocked", C.c_int32), ("is_pblocked", C.c_int32),
                ("sweep_groups", C.c_int32), ("seed", C.c_uint64),
                ("iter_offset", C.c_uint64)]


def _ptr(a, t):
    return a.ctypes.data_as(t)


def load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    lib.oracle_lba_loglike.restype = C.c_double
    lib.oracle_lba_loglike.argtypes = [C.POINTER(OracleModel), _d, _i64, _d]
    lib.oracle_lba_trial_terms.restype = None
    lib.oracle_lba_trial_terms.argtypes = [C.POINTER(OracleModel), _d, _i64, _d, _d]
    lib.oracle_sumlogprior.restype = C.c_double
    lib.oracle_sumlogprior.argtypes = [C.POINTER(OraclePrior), _d]
    lib.oracle_sumloghlike.restype = C.c_double
    lib.oracle_sumloghlike.argtypes = [C.POINTER(OraclePrior), _d, C.c_int32, _d, C.c_int64]
    lib.oracle_philox4x32_10.restype = None
    lib.oracle_philox4x32_10.argtypes = [_u32, _u32, _u32]
    lib.oracle_run_subject.restype = C.c_int
    lib.oracle_run_subject.argtypes = [C.POINTER(OracleConfig), C.c_int32,
                                       C.POINTER(OracleModel), C.POINTER(OraclePrior),
                                       _d, _i64, _d, _d, _d, _d, _d, _d]
    lib.oracle_run_hyper.restype = C.c_int
    lib.oracle_run_hyper.argtypes = [C.POINTER(OracleConfig), C.POINTER(OraclePrior),
                                     C.POINTER(OraclePrior), C.c_int32, _d, _d, _d, _d,
                                     _d, _d, _d]
    lib.oracle_run_hier.restype = C.c_int
    lib.oracle_run_hier.argtypes = [C.POINTER(OracleConfig), C.POINTER(OracleModel),
                                    C.POINTER(OraclePrior), C.POINTER(OraclePrior), C.c_int32,
                                    _d, _i64, _i64, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d,
                                    _d, _d]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


class Model:
    """Keeps the design arrays alive behind an OracleModel struct."""

    def __init__(self, winner, par_src, par_const, npar):
        self.winner = np.ascontiguousarray(winner, dtype=np.int32)
        self.par_src = np.ascontiguousarray(par_src, dtype=np.int32)
        self.par_const = np.ascontiguousarray(par_const, dtype=np.float64)
        n_cell, n_acc = self.par_src.shape[:2]
        self.s = OracleModel(n_acc, n_cell, int(npar), _ptr(self.winner, _i32),
                             _ptr(self.par_src, _i32), _ptr(self.par_const, _d))


class Prior:
    def __init__(self, dist, p0, p1, lower, upper):
        self.dist = np.ascontiguousarray(dist, dtype=np.int32)
        self.p0 = np.ascontiguousarray(p0, dtype=np.float64)
        self.p1 = np.ascontiguousarray(p1, dtype=np.float64)
        self.lower = np.ascontiguousarray(lower, dtype=np.float64)
        self.upper = np.ascontiguousarray(upper, dtype=np.float64)
        self.s = OraclePrior(len(self.dist), _ptr(self.dist, _i32), _ptr(self.p0, _d),
                             _ptr(self.p1, _d), _ptr(self.lower, _d), _ptr(self.upper, _d))


def config(nmc, nchain, nparameter_de, thin=1, pop_migration_prob=0.0,
           sub_migration_prob=0.0, gamma_precursor=2.38, rp=0.001, is_hblocked=False,
           is_pblocked=False, sweep_groups=None, seed=9032, iter_offset=0):
    return OracleConfig(nmc, thin, nchain, nparameter_de, pop_migration_prob,
                        sub_migration_prob, gamma_precursor, rp, int(is_hblocked),
                        int(is_pblocked), nchain if sweep_groups is None else sweep_groups,
                        seed, iter_offset)


def lba_loglike(model, rt, cell_off, theta):
    rt = np.ascontiguousarray(rt, dtype=np.float64)
    cell_off = np.ascontiguousarray(cell_off, dtype=np.int64)
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    return lib().oracle_lba_loglike(C.byref(model.s), _ptr(rt, _d), _ptr(cell_off, _i64),
                                    _ptr(theta, _d))


def lba_trial_terms(model, rt, cell_off, theta):
    rt = np.ascontiguousarray(rt, dtype=np.float64)
    cell_off = np.ascontiguousarray(cell_off, dtype=np.int64)
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    out = np.empty(int(cell_off[-1]))
    lib().oracle_lba_trial_terms(C.byref(model.s), _ptr(rt, _d), _ptr(cell_off, _i64),
                                 _ptr(theta, _d), _ptr(out, _d))
    return out


def sumlogprior(prior, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().oracle_sumlogprior(C.byref(prior.s), _ptr(x, _d))


def sumloghlike(prior, phi, subj):
    """subj: [S, P] subject parameter vectors."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    subj = np.ascontiguousarray(subj, dtype=np.float64)
    return lib().oracle_sumloghlike(C.byref(prior.s), _ptr(phi, _d), subj.shape[0],
                                    _ptr(subj, _d), subj.shape[1])


def philox(ctr, key):
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().oracle_philox4x32_10(_ptr(c, _u32), _ptr(k, _u32), _ptr(o, _u32))
    return o


def run_subject(cfg, model, prior, rt, cell_off, theta0, lp0, ll0, model_kind=MODEL_LBA):
    """Returns theta [P, nchain, nmc], lp [nchain, nmc], ll [nchain, nmc]."""
    P, n = theta0.shape
    nmc = cfg.nmc
    th = np.empty((nmc, n, P))
    lp = np.empty((nmc, n))
    ll = np.empty((nmc, n))
    rt = np.ascontiguousarray(rt, dtype=np.float64)
    cell_off = np.ascontiguousarray(cell_off, dtype=np.int64)
    t0 = np.ascontiguousarray(theta0.T, dtype=np.float64)
    lp0 = np.ascontiguousarray(lp0, dtype=np.float64)
    ll0 = np.ascontiguousarray(ll0, dtype=np.float64)
    rc = lib().oracle_run_subject(C.byref(cfg), model_kind, C.byref(model.s), C.byref(prior.s),
                                  _ptr(rt, _d), _ptr(cell_off, _i64), _ptr(t0, _d),
                                  _ptr(lp0, _d), _ptr(ll0, _d), _ptr(th, _d), _ptr(lp, _d),
                                  _ptr(ll, _d))
    if rc != 0:
        raise ValueError("oracle_run_subject failed: %d" % rc)
    return th.transpose(2, 1, 0), lp.T, ll.T


def run_hyper(cfg, p_prior, h_prior, data, theta0, lp0, ll0):
    P2, n = theta0.shape
    nmc = cfg.nmc
    th = np.empty((nmc, n, P2))
    lp = np.empty((nmc, n))
    ll = np.empty((nmc, n))
    data = np.ascontiguousarray(data, dtype=np.float64)
    t0 = np.ascontiguousarray(theta0.T, dtype=np.float64)
    lp0 = np.ascontiguousarray(lp0, dtype=np.float64)
    ll0 = np.ascontiguousarray(ll0, dtype=np.float64)
    rc = lib().oracle_run_hyper(C.byref(cfg), C.byref(p_prior.s), C.byref(h_prior.s),
                                data.shape[0], _ptr(data, _d), _ptr(t0, _d), _ptr(lp0, _d),
                                _ptr(ll0, _d), _ptr(th, _d), _ptr(lp, _d), _ptr(ll, _d))
    if rc != 0:
        raise ValueError("oracle_run_hyper failed: %d" % rc)
    return th.transpose(2, 1, 0), lp.T, ll.T


def run_hier(cfg, model, p_prior, h_prior, rt, subject_off, cell_off, phi0, phi_lp0, phi_ll0,
             sub0, sub_lp0, sub_ll0):
    """sub0: [S, P, nchain]; returns (phi_theta [2P,n,nmc], phi_lp, phi_ll,
    sub_theta [S,P,n,nmc], sub_lp [S,n,nmc], sub_ll [S,n,nmc])."""
    S, P, n = sub0.shape
    nmc = cfg.nmc
    pth = np.empty((nmc, n, 2 * P))
    plp = np.empty((nmc, n))
    pll = np.empty((nmc, n))
    sth = np.empty((S, nmc, n, P))
    slp = np.empty((S, nmc, n))
    sll = np.empty((S, nmc, n))
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in
            (rt, phi0.T, phi_lp0, phi_ll0, sub0.transpose(0, 2, 1), sub_lp0, sub_ll0)]
    so = np.ascontiguousarray(subject_off, dtype=np.int64)
    co = np.ascontiguousarray(cell_off, dtype=np.int64)
    rc = lib().oracle_run_hier(C.byref(cfg), C.byref(model.s), C.byref(p_prior.s),
                               C.byref(h_prior.s), S, _ptr(args[0], _d), _ptr(so, _i64),
                               _ptr(co, _i64), _ptr(args[1], _d), _ptr(args[2], _d),
                               _ptr(args[3], _d), _ptr(args[4], _d), _ptr(args[5], _d),
                               _ptr(args[6], _d), _ptr(pth, _d), _ptr(plp, _d), _ptr(pll, _d),
                               _ptr(sth, _d), _ptr(slp, _d), _ptr(sll, _d))
    if rc != 0:
        raise ValueError("oracle_run_hier failed: %d" % rc)
    return (pth.transpose(2, 1, 0), plp.T, pll.T, sth.transpose(0, 3, 2, 1),
            slp.transpose(0, 2, 1), sll.transpose(0, 2, 1))
