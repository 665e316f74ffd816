"""Brute-force replay of the DE-MCMC sampler in pure Python (test
infrastructure), written from the reference's src/de.cpp and independent of
oracle/oracle.cpp: it re-derives every proposal and accept decision of an
iteration from the Philox words and the previous state. Used to pin the
oracle's samplers (tests/test_oracle.py), whose only shared pieces are the
trial log-likelihood (pinned by the reference's golden vectors) and the
Philox stream layout both sides define for this engine:

  key     = (seed & 0xffffffff, seed >> 32)
  counter = (iteration + iter_offset, unit, sweep << 16 | slot, block)
  unit    = subject index (0 for single-level runs), 0xffffffff for the
            population level of a hierarchical run
  slot    = chain i (crossover: words 0,1 partners, 2,3 accept U, 4+j jitter
            of the j-th perturbed parameter); 0x8000 | k (migration ring step
            k: words 0,1 U, 4+j jitter); 0x7fff (migration subset: words 0,1
            size U, 2+k Fisher-Yates draw k); 0x7ffe (level decision, words 0,1)

Reference semantics restated here (src/de.cpp):
  run_chains :201-242   one decision per iteration; migration, or crossover
                        (blocked: one crossover per parameter)
  run_hchains :272-383  population step first (blocked: a decision per
                        parameter), then every subject (blocked: a decision per
                        parameter), then the stores
  crossover :111-155 / :385-465 / :567-613, in a chain-group sweep: chains of
                        group g take two distinct partners outside g (G = nchain
                        is the reference's sequential in-place order); the
                        population level first refreshes the chain's cached
                        likelihood (:397-398); the subject level takes its prior
                        from population chain i (:599-600) and compares with its
                        cached lp (:577), which is not refreshed
  migration :157-199 / :467-565 / :615-665: sorted random subset, a jittered
                        copy of S_k proposed for S_{k+1} and judged against
                        S_{k+1}'s cached posterior; the population level
                        refreshes both likelihoods first (:494-500) and takes the
                        proposal's likelihood from S_k's subject columns
                        (:519-520); the subject level takes its prior from
                        population chain S_k (:646-649)
  update_theta :81-108  ratio = exp(post' - post), NaN rejects, else U < ratio
"""
import math

import numpy as np

MASK = 0xFFFFFFFF
UNIT_PHI = 0xFFFFFFFF
SLOT_DECISION = 0x7FFE
SLOT_SUBSET = 0x7FFF
SLOT_MIGRATE = 0x8000


def philox(ctr, key):
    """Philox4x32-10 (Salmon et al., SC'11), Python integers."""
    c0, c1, c2, c3 = ctr
    k0, k1 = key
    for r in range(10):
        if r:
            k0 = (k0 + 0x9E3779B9) & MASK
            k1 = (k1 + 0xBB67AE85) & MASK
        p0 = 0xD2511F53 * c0
        p1 = 0xCD9E8D57 * c2
        c0, c1, c2, c3 = ((p1 >> 32) ^ c1 ^ k0) & MASK, p1 & MASK, ((p0 >> 32) ^ c3 ^ k1) & MASK, \
            p0 & MASK
    return c0, c1, c2, c3


class Stream:
    def __init__(self, seed, it, unit, sweep=0):
        self.key = (seed & MASK, (seed >> 32) & MASK)
        self.it, self.unit, self.sweep = it, unit, sweep
        self._cache = {}

    def at(self, sweep):
        return Stream((self.key[1] << 32) | self.key[0], self.it, self.unit, sweep)

    def word(self, slot, w):
        b = w >> 2
        k = (slot, b)
        if k not in self._cache:
            self._cache[k] = philox((self.it & MASK, self.unit, ((self.sweep << 16) | slot) & MASK, b),
                                    self.key)
        return self._cache[k][w & 3]


def u53(a, b):
    return ((a >> 5) * 67108864.0 + (b >> 6)) / 9007199254740992.0


def jitter(w, rp):
    return rp * (2.0 * ((w + 0.5) * 2.0 ** -32) - 1.0)


def below(w, n):
    return (w * n) >> 32


def accept(tpost, cur, U):
    d = tpost - cur
    if math.isnan(d):
        return False
    try:
        ratio = math.exp(d)
    except OverflowError:
        ratio = math.inf
    return U < ratio


# ---------------------------------------------------------------- densities
def _pnorm(z):
    """Phi as R's pnorm rounds it (upper tail complemented above 0)."""
    if z == math.inf:
        return 1.0
    if z == -math.inf:
        return 0.0
    r = 0.7071067811865476   # 1/sqrt(2) rounded to FP64
    return 1.0 - 0.5 * math.erfc(z * r) if z > 0 else 0.5 * math.erfc(-z * r)


def log_prior1(dist, p0, p1, lo, up, x):
    """ggdmcPrior log densities: 6 uniform(p0, p1); 1 normal(p0, p1) on [lo, up]."""
    if math.isnan(x):
        return x
    if dist == 6:
        return -math.inf if (x < p0 or x > p1) else -math.log(p1 - p0)
    if dist == 1:
        if not p1 > 0:
            return math.nan
        if x < lo or x > up:
            return -math.inf
        a, b = (lo - p0) / p1, (up - p0) / p1
        Z = (_pnorm(-a) - _pnorm(-b)) if a > 0 else (_pnorm(b) - _pnorm(a))
        z = (x - p0) / p1
        return -0.5 * z * z - 0.918938533204672741780329736406 - math.log(p1) - math.log(Z)
    return math.nan


def sumlogprior(prior, x):
    """prior: (dist, p0, p1, lower, upper) arrays."""
    dist, p0, p1, lo, up = prior
    s = 0.0
    for k in range(len(dist)):
        s += log_prior1(int(dist[k]), p0[k], p1[k], lo[k], up[k], float(x[k]))
    return s


def sumloghlike(p_prior, phi, subj_cols):
    """src/de.cpp:245-270: sum over subjects of log tnorm(theta_s; head(phi),
    tail(phi)) with p_prior's bounds."""
    dist, _, _, lo, up = p_prior
    P = len(dist)
    out = 0.0
    for x in subj_cols:
        v = 0.0
        for k in range(P):
            v += log_prior1(int(dist[k]), phi[k], phi[P + k], lo[k], up[k], float(x[k]))
        out += v
    return out


def phi_prior(phi, p_prior):
    """The subject prior under population state phi (src/de.cpp:599-600)."""
    dist, _, _, lo, up = p_prior
    P = len(dist)
    return (dist, phi[:P], phi[P:], lo, up)


# --------------------------------------------------------------- one level
class Level:
    """State [np x n] (columns are chains) with cached lp / ll."""

    def __init__(self, th, lp, ll):
        self.th = np.array(th, float)
        self.lp = np.array(lp, float)
        self.ll = np.array(ll, float)

    def copy(self):
        return Level(self.th, self.lp, self.ll)


def subset(st, n):
    """get_subchains (src/de.cpp:62-78) on the subset slot."""
    U = u53(st.word(SLOT_SUBSET, 0), st.word(SLOT_SUBSET, 1))
    ns = min(max(int(math.ceil(n * U)), 2), n)
    a = list(range(n))
    for k in range(ns):
        j = k + below(st.word(SLOT_SUBSET, 2 + k), n - k)
        a[k], a[j] = a[j], a[k]
    return sorted(a[:ns])


def crossover(L, st, gamma, rp, groups, para, logprior, loglike, refresh=None):
    """logprior(x, i), loglike(x, i): proposal terms made for chain i;
    refresh(x, i): the population level's recomputed cached likelihood."""
    n = L.th.shape[1]
    np_ = L.th.shape[0]
    for g in range(groups):
        lo, hi = (g * n) // groups, ((g + 1) * n) // groups
        pool = n - (hi - lo)
        for i in range(lo, hi):
            if refresh is not None:
                L.ll[i] = refresh(L.th[:, i], i)
            cur = L.ll[i] + L.lp[i]
            r0 = below(st.word(i, 0), pool)
            r1 = below(st.word(i, 1), pool - 1)
            if r1 >= r0:
                r1 += 1
            m0 = r0 if r0 < lo else r0 + (hi - lo)
            m1 = r1 if r1 < lo else r1 + (hi - lo)
            U = u53(st.word(i, 2), st.word(i, 3))
            tmp = L.th[:, i].copy()
            if para >= 0:
                tmp[para] = tmp[para] + (jitter(st.word(i, 4), rp) +
                                         gamma * (L.th[para, m0] - L.th[para, m1]))
            else:
                for j in range(np_):
                    tmp[j] = tmp[j] + (jitter(st.word(i, 4 + j), rp) +
                                       gamma * (L.th[j, m0] - L.th[j, m1]))
            tlp = logprior(tmp, i)
            tll = loglike(tmp, i)
            if accept(tlp + tll, cur, U):
                L.th[:, i] = tmp
                L.lp[i] = tlp
                L.ll[i] = tll


def migration(L, st, rp, para, logprior, loglike, refresh=None):
    n = L.th.shape[1]
    np_ = L.th.shape[0]
    S = subset(st, n)
    ns = len(S)
    for k in range(ns):
        src, dst = S[k], S[(k + 1) % ns]
        if refresh is not None:
            L.ll[src] = refresh(L.th[:, src], src)
            L.ll[dst] = refresh(L.th[:, dst], dst)
        slot = SLOT_MIGRATE | k
        tmp = L.th[:, src].copy()
        if para >= 0:
            tmp[para] = tmp[para] + jitter(st.word(slot, 4), rp)
        else:
            for j in range(np_):
                tmp[j] = tmp[j] + jitter(st.word(slot, 4 + j), rp)
        tlp = logprior(tmp, src)
        tll = loglike(tmp, src)
        cur = L.lp[dst] + L.ll[dst]
        U = u53(st.word(slot, 0), st.word(slot, 1))
        if accept(tlp + tll, cur, U):
            L.th[:, dst] = tmp
            L.lp[dst] = tlp
            L.ll[dst] = tll


def decide(st):
    return u53(st.word(SLOT_DECISION, 0), st.word(SLOT_DECISION, 1))


def run_chains_step(L, st, gamma, rp, mig, groups, blocked, logprior, loglike):
    """One iteration of run_chains (src/de.cpp:201-242)."""
    if decide(st.at(0)) < mig:
        migration(L, st.at(0), rp, -1, logprior, loglike)
    elif blocked:
        for p in range(L.th.shape[0]):
            crossover(L, st.at(p), gamma, rp, groups, p, logprior, loglike)
    else:
        crossover(L, st.at(0), gamma, rp, groups, -1, logprior, loglike)


def hlevel_step(L, st, gamma, rp, mig, groups, blocked, logprior, loglike, refresh=None):
    """One level of run_hchains (src/de.cpp:283-372): a decision per
    parameter when blocked, else one."""
    if blocked:
        for p in range(L.th.shape[0]):
            s = st.at(p)
            if decide(s) < mig:
                migration(L, s, rp, p, logprior, loglike, refresh)
            else:
                crossover(L, s, gamma, rp, groups, p, logprior, loglike, refresh)
    else:
        s = st.at(0)
        if decide(s) < mig:
            migration(L, s, rp, -1, logprior, loglike, refresh)
        else:
            crossover(L, s, gamma, rp, groups, -1, logprior, loglike, refresh)


# ------------------------------------------------------------------- runs
def run_single(cfg, L0, logprior, loglike, nstore):
    """run_subject / run_hyper: states after iterations 1..nstore-1 (thin 1).
    cfg: dict(seed, gamma, rp, mig, groups, blocked, iter_offset)."""
    L = L0.copy()
    out = [L.copy()]
    for it in range(1, nstore):
        st = Stream(cfg["seed"], it + cfg["iter_offset"], 0)
        run_chains_step(L, st, cfg["gamma"], cfg["rp"], cfg["mig"], cfg["groups"],
                        cfg["blocked"], logprior, loglike)
        out.append(L.copy())
    return out


def hier_iteration(cfg, phi, subs, p_prior, h_prior, subject_ll, it):
    """One iteration of run_hchains, in place. subject_ll(s, x): the
    likelihood of subject s's parameters x."""
    S = len(subs)

    def cols(i):
        return [subs[s].th[:, i] for s in range(S)]

    st = Stream(cfg["seed"], it + cfg["iter_offset"], UNIT_PHI)
    hlevel_step(phi, st, cfg["gamma"], cfg["rp"], cfg["pop_mig"], cfg["groups"],
                cfg["hblocked"],
                logprior=lambda x, i: sumlogprior(h_prior, x),
                loglike=lambda x, i: sumloghlike(p_prior, x, cols(i)),
                refresh=lambda x, i: sumloghlike(p_prior, x, cols(i)))
    for s in range(S):
        ss = Stream(cfg["seed"], it + cfg["iter_offset"], s)
        hlevel_step(subs[s], ss, cfg["gamma"], cfg["rp"], cfg["sub_mig"], cfg["groups"],
                    cfg["pblocked"],
                    logprior=lambda x, i: sumlogprior(phi_prior(phi.th[:, i], p_prior), x),
                    loglike=lambda x, i, s=s: subject_ll(s, x))
