"""Prior densities for run set-up (start states, initialisation).

Host-side counterparts of ggdmcPrior's dprior / rprior as the reference's R
code uses them (R/phi.R:170-173 initialise_theta, :299-304 initialise_phi,
:37-48 .sumloghlike). dist_id 6 = uniform(p0, p1); dist_id 1 = normal(p0, p1)
truncated to [lower, upper]; densities on the log scale.
"""
import numpy as np
from scipy.special import log_ndtr, ndtr

from . import _lib


def logdprior(prior, x):
    """Per-parameter log densities of x [P] (or [n, P])."""
    x = np.asarray(x, float)
    out = np.empty(np.broadcast(x, prior.p0).shape)
    uni = prior.dist == _lib.DIST_UNIF
    tn = prior.dist == _lib.DIST_TNORM
    with np.errstate(all="ignore"):
        u = np.where((x >= prior.p0) & (x <= prior.p1), -np.log(prior.p1 - prior.p0), -np.inf)
        z = (x - prior.p0) / prior.p1
        a = (prior.lower - prior.p0) / prior.p1
        b = (prior.upper - prior.p0) / prior.p1
        Z = np.where(a > 0, ndtr(-a) - ndtr(-b), ndtr(b) - ndtr(a))
        t = -0.5 * z * z - 0.5 * np.log(2 * np.pi) - np.log(prior.p1) - np.log(Z)
        t = np.where((x < prior.lower) | (x > prior.upper), -np.inf, t)
        t = np.where(prior.p1 > 0, t, np.nan)
    out[...] = np.where(uni, u, np.where(tn, t, np.nan))
    return np.where(np.isnan(x), np.nan, out)


def sumlogprior(prior, x):
    return float(np.sum(logdprior(prior, x)))


def sumloghlike(p_prior, phi, subject_thetas):
    """sum over subjects of log tnorm(theta_s; loc = phi[:P], scale = phi[P:])
    (src/de.cpp:245-270; R/phi.R:37-48). subject_thetas [S, P]."""
    P = p_prior.npar
    from .runs import PriorSpec
    pr = PriorSpec.tnorm(phi[:P], phi[P:], p_prior.lower, p_prior.upper)
    return float(np.sum(logdprior(pr, np.asarray(subject_thetas))))


def rprior(prior, n, rng):
    """n draws [n, P] (rprior): uniform, or tnorm by inversion."""
    P = prior.npar
    out = np.empty((n, P))
    for k in range(P):
        if prior.dist[k] == _lib.DIST_UNIF:
            out[:, k] = rng.uniform(prior.p0[k], prior.p1[k], n)
        else:
            a = ndtr((prior.lower[k] - prior.p0[k]) / prior.p1[k])
            b = ndtr((prior.upper[k] - prior.p0[k]) / prior.p1[k])
            u = rng.uniform(a, b, n)
            from scipy.special import ndtri
            out[:, k] = prior.p0[k] + prior.p1[k] * ndtri(u)
    return out


__all__ = ["logdprior", "sumlogprior", "sumloghlike", "rprior", "log_ndtr"]
