"""Convergence and Monte-Carlo error diagnostics of posterior samples.

Counterparts of the reference's gelman (R/model-class.R:1593-1736: PSRF per
parameter and the multivariate PSRF of Brooks & Gelman 1998, from coda's
gelman.diag) and of the summary's time-series standard error
(spectrum0.ar, :257-275), written from the published formulas.
"""
import numpy as np


def _seqs(theta, start=0):
    """theta [P, nchain, nmc] -> [m sequences, n draws, P] (chains as sequences)."""
    x = np.asarray(theta)[:, :, start:]
    return np.transpose(x, (1, 2, 0))


def psrf(theta, start=0):
    """Potential scale reduction factor per parameter (Gelman & Rubin 1992),
    upper-bound free point estimate sqrt(((n-1)/n W + B/n) / W)."""
    x = _seqs(theta, start)
    m, n, _ = x.shape
    means = x.mean(axis=1)
    W = x.var(axis=1, ddof=1).mean(axis=0)
    B = n * means.var(axis=0, ddof=1)
    V = (n - 1) / n * W + B / n
    return np.sqrt(V / W)


def mpsrf(theta, start=0):
    """Multivariate PSRF (Brooks & Gelman 1998): sqrt((n-1)/n + (m+1)/m lambda_max),
    lambda_max the largest eigenvalue of W^-1 B/n."""
    x = _seqs(theta, start)
    m, n, P = x.shape
    means = x.mean(axis=1)
    W = np.zeros((P, P))
    for j in range(m):
        d = x[j] - means[j]
        W += d.T @ d / (n - 1)
    W /= m
    Bn = np.cov(means.T, ddof=1).reshape(P, P)
    lam = np.max(np.real(np.linalg.eigvals(np.linalg.solve(W, Bn))))
    return float(np.sqrt((n - 1) / n + (m + 1) / m * lam))


def spectrum0_ar(y, max_order=None):
    """Spectral density at frequency 0 of a series from a fitted AR model
    (Yule-Walker, order by AIC), as coda::spectrum0.ar; returns (spec, order)."""
    y = np.asarray(y, float)
    y = y - y.mean()
    n = len(y)
    max_order = max_order or min(n - 1, int(10 * np.log10(n)))
    acov = np.array([np.dot(y[:n - k], y[k:]) / n for k in range(max_order + 1)])
    if acov[0] <= 0:
        return 0.0, 0
    best = (np.log(acov[0]) * n + 0, 0, acov[0], np.zeros(0))
    phi = np.zeros(0)
    v = acov[0]
    for p in range(1, max_order + 1):
        # Levinson-Durbin
        k = (acov[p] - np.dot(phi, acov[p - 1:0:-1])) / v
        phi = np.concatenate([phi - k * phi[::-1], [k]])
        v = v * (1 - k * k)
        if v <= 0:
            break
        aic = n * np.log(v) + 2 * p
        if aic < best[0]:
            best = (aic, p, v, phi.copy())
    _, order, var, coef = best
    return float(var / (1 - coef.sum()) ** 2), order


def mcse(theta, start=0):
    """Time-series standard error of the posterior mean per parameter: the
    per-chain AR spectral variance pooled over chains."""
    x = _seqs(theta, start)
    m, n, P = x.shape
    out = np.empty(P)
    for p in range(P):
        s = np.mean([spectrum0_ar(x[j, :, p])[0] for j in range(m)])
        out[p] = np.sqrt(s / (n * m))
    return out


def summary(theta, start=0):
    """mean, sd, naive SE and time-series SE per parameter (summary,
    R/model-class.R:370-424)."""
    x = _seqs(theta, start).reshape(-1, np.asarray(theta).shape[0])
    return {"mean": x.mean(axis=0), "sd": x.std(axis=0, ddof=1),
            "naive_se": x.std(axis=0, ddof=1) / np.sqrt(x.shape[0]),
            "ts_se": mcse(theta, start)}
