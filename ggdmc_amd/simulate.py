"""Synthetic LBA data for the benchmark workloads.

Generation rule (the LBA simulation rule, as in lbaModel::simulate used by
tests/testthat/Group1/data/*_lba_*.r): per trial and accumulator j, start
a_j ~ U(0, A_j), threshold b_j = A_j + B_j, drift v_j ~ N(mean_v_j, sd_v_j)
with all drifts redrawn when none is positive; RT = t0 + min over positive
drifts of (b_j - a_j) / v_j; the response is the argmin.
"""
import numpy as np

from .runs import SubjectData

FAMILIES = ("A", "B", "mean_v", "sd_v", "st0", "t0")


def conditions(model):
    """Stimulus conditions of a design: {condition: {accumulator: cell}}; a
    cell name is '<factor levels>.<response>'."""
    acc_names = None
    out = {}
    for c, name in enumerate(model.cell_names):
        cond, resp = name.rsplit(".", 1)
        out.setdefault(cond, {})[int(model.winner[c])] = c
    return out, acc_names


def acc_params(model, theta, cell):
    """[n_acc, 6] parameter values of a cell."""
    src = model.par_src[cell]
    const = model.par_const[cell]
    return np.where(src >= 0, np.asarray(theta)[np.maximum(src, 0)], const)


def simulate_condition(par, n, rng):
    """n trials of one condition: (rt [n], response [n])."""
    A, B, v, s, _, t0 = (par[:, k] for k in range(6))
    n_acc = par.shape[0]
    rts = np.empty(n)
    resp = np.empty(n, dtype=np.int64)
    todo = np.arange(n)
    while len(todo):
        drift = rng.normal(v[None, :], s[None, :], (len(todo), n_acc))
        ok = (drift > 0).any(axis=1)
        idx = todo[ok]
        d = drift[ok]
        start = rng.uniform(0, 1, d.shape) * A[None, :]
        with np.errstate(divide="ignore", invalid="ignore"):
            tt = np.where(d > 0, (A + B)[None, :] - start, np.inf) / np.where(d > 0, d, 1.0)
        r = np.argmin(tt, axis=1)
        rts[idx] = t0[r] + tt[np.arange(len(idx)), r]
        resp[idx] = r
        todo = todo[~ok]
    return rts, resp


def simulate_subject(model, theta, n_per_condition, rng):
    """SubjectData with n_per_condition trials in every stimulus condition."""
    conds, _ = conditions(model)
    cells = [[] for _ in range(model.n_cell)]
    for cond in sorted(conds):
        accs = conds[cond]
        par = acc_params(model, theta, next(iter(accs.values())))
        rts, resp = simulate_condition(par, n_per_condition, rng)
        for t, r in zip(rts, resp):
            if int(r) in accs:
                cells[accs[int(r)]].append(t)
    return SubjectData.from_cells([np.array(c) for c in cells])


def rtnorm(mean, sd, lower, rng, size=None):
    """Normal(mean, sd) truncated below at `lower` (rejection)."""
    mean = np.asarray(mean, float)
    sd = np.asarray(sd, float)
    out = rng.normal(mean, sd, size)
    bad = out < lower
    while bad.any():
        out[bad] = rng.normal(np.broadcast_to(mean, out.shape)[bad], np.broadcast_to(sd, out.shape)[bad])
        bad = out < lower
    return out
