/*
 * CPU FP64 oracle for the DE-MCMC hot path of ggdmc (test infrastructure; see
 * oracle.h for who may load it and which reference lines each part follows).
 *
 * Everything here is deliberately plain: scalar double arithmetic, one trial
 * at a time, chains visited in order. It is the checker the HIP library is
 * compared with, and the single-threaded CPU baseline that bench.py times.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace {

const double kDblEps = 2.220446049250313e-16;    /* R's .Machine$double.eps */
const double kFloor = 1e-10;                     /* LBA pdf / CDF floor     */
const double kInvSqrt2Pi = 0.398942280401432677939946059934;

/* ---------------------------------------------------------------- Philox */
/* Philox4x32-10: J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11. Multipliers and Weyl
 * key increments are the published ones. */
inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
}

void philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k[0] += 0x9E3779B9u;
            k[1] += 0xBB67AE85u;
        }
        philox_round(c, k);
    }
    std::memcpy(out, c, sizeof(c));
}

/* Stream layout shared with the HIP sampler: counter = (iteration, unit,
 * sweep << 16 | slot, block); key = seed. */
const uint32_t kUnitPhi = 0xFFFFFFFFu;
const uint32_t kSlotDecision = 0x7FFEu;
const uint32_t kSlotSubset = 0x7FFFu;
const uint32_t kSlotMigrate = 0x8000u;

struct Stream {
    uint32_t key[2];
    uint32_t iter, unit, sweep;
    void block(uint32_t slot, uint32_t b, uint32_t out[4]) const {
        const uint32_t ctr[4] = {iter, unit, (sweep << 16) | slot, b};
        philox(ctr, key, out);
    }
    /* word w of the slot's sequence of blocks */
    uint32_t word(uint32_t slot, uint32_t w) const {
        uint32_t o[4];
        block(slot, w >> 2, o);
        return o[w & 3];
    }
};

inline double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) *
           (1.0 / 9007199254740992.0);
}
inline double u32_open(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386962890625e-10; }
inline uint32_t below(uint32_t w, uint32_t n) { return (uint32_t)(((uint64_t)w * n) >> 32); }

/* ------------------------------------------------------- normal helpers */
/* Phi as R's pnorm (nmath/pnorm.c) rounds it: the lower tail directly, above
 * 0 as 1 - the accurate upper tail. */
inline double pnorm_(double z) {
    return z > 0.0 ? 1.0 - 0.5 * std::erfc(z * M_SQRT1_2) : 0.5 * std::erfc(-z * M_SQRT1_2);
}
inline double dnorm_(double z) { return kInvSqrt2Pi * std::exp(-0.5 * z * z); }

/* ------------------------------------------------------ LBA likelihood */
struct AccPar {
    double A, b, v, s, t0, den;
};

void resolve(const oracle_model *m, const double *theta, int c, int j, AccPar *o, bool *st0_nonzero) {
    double fam[6];
    for (int f = 0; f < 6; ++f) {
        const int idx = (c * m->n_acc + j) * 6 + f;
        const int src = m->par_src[idx];
        fam[f] = src >= 0 ? theta[src] : m->par_const[idx];
    }
    o->A = fam[0];
    o->b = fam[0] + fam[1];
    o->v = fam[2];
    o->s = fam[3];
    o->t0 = fam[5];
    o->den = std::max(pnorm_(o->v / o->s), kFloor);
    if (fam[4] != 0.0) *st0_nonzero = true;
}

/* Defective density pieces of one accumulator (truncated-normal drift,
 * start ~ U(0, A), threshold b), naive FP64 as the fixtures were computed. */
inline void lba_pdf_cdf(const AccPar &p, double t, double *pdf, double *cdf) {
    if (!(t > 0.0)) {
        *pdf = 0.0;
        *cdf = 0.0;
        return;
    }
    const double zs = t * p.s;
    const double chi = p.b - t * p.v;
    const double chimA = chi - p.A;
    const double z1 = chi / zs;
    const double z2 = chimA / zs;
    const double P1 = pnorm_(z1), P2 = pnorm_(z2);
    const double d1 = dnorm_(z1), d2 = dnorm_(z2);
    *pdf = (p.v * (P1 - P2) + p.s * (d2 - d1)) / (p.A * p.den);
    *cdf = (1.0 + (zs * (d2 - d1) + chimA * P2 - chi * P1) / p.A) / p.den;
}

void lba_terms(const oracle_model *m, const double *rt, const int64_t *cell_off,
               const double *theta, double *out, double *sum) {
    std::vector<AccPar> par(m->n_acc);
    bool st0_bad = false;
    double acc = 0.0;
    for (int c = 0; c < m->n_cell; ++c) {
        for (int j = 0; j < m->n_acc; ++j) resolve(m, theta, c, j, &par[j], &st0_bad);
        const int w = m->winner[c];
        for (int64_t k = cell_off[c]; k < cell_off[c + 1]; ++k) {
            double pdf, cdf;
            lba_pdf_cdf(par[w], rt[k] - par[w].t0, &pdf, &cdf);
            double dens = std::max(pdf, kFloor);
            if (std::isnan(pdf)) dens = pdf;
            for (int j = 0; j < m->n_acc; ++j) {
                if (j == w) continue;
                lba_pdf_cdf(par[j], rt[k] - par[j].t0, &pdf, &cdf);
                const double F = std::isnan(cdf) ? cdf : std::max(cdf, kFloor);
                dens *= (1.0 - F);
            }
            /* R/phi.R:3-13: a trial density <= 0 is replaced by DBL_EPSILON;
             * NaN passes through (na.rm = TRUE drops it from the test). */
            const double term = std::log(dens <= 0.0 ? kDblEps : dens);
            if (out) out[k] = term;
            acc += term;
        }
    }
    if (st0_bad) acc = std::numeric_limits<double>::quiet_NaN();
    if (sum) *sum = acc;
}

/* ---------------------------------------------------------------- priors */
double log_prior1(int dist, double p0, double p1, double lo, double up, double x) {
    const double ninf = -std::numeric_limits<double>::infinity();
    if (std::isnan(x)) return x;
    if (dist == 6) { /* uniform(p0, p1) */
        if (x < p0 || x > p1) return ninf;
        return -std::log(p1 - p0);
    }
    if (dist == 1) { /* normal(p0, p1) truncated to [lo, up] */
        if (!(p1 > 0.0)) return std::numeric_limits<double>::quiet_NaN();
        if (x < lo || x > up) return ninf;
        const double z = (x - p0) / p1;
        /* normaliser Phi(b) - Phi(a); written as Q(a) - Q(b) when a > 0 so
         * that a lower bound far in the upper tail does not cancel to 0 */
        const double a = (lo - p0) / p1, b = (up - p0) / p1;
        const double Z = a > 0.0 ? pnorm_(-a) - pnorm_(-b) : pnorm_(b) - pnorm_(a);
        return -0.5 * z * z - 0.918938533204672741780329736406 - std::log(p1) - std::log(Z);
    }
    return std::numeric_limits<double>::quiet_NaN();
}

double sumlogprior(const oracle_prior *p, const double *x) {
    double s = 0.0;
    for (int k = 0; k < p->npar; ++k)
        s += log_prior1(p->dist[k], p->p0[k], p->p1[k], p->lower[k], p->upper[k], x[k]);
    return s;
}

/* sum over subjects of the log density of their parameters under phi
 * (src/de.cpp:245-270; p0 := head(phi), p1 := tail(phi)). */
double sumloghlike(const oracle_prior *pp, const double *phi, int S, const double *subj,
                   int64_t stride) {
    const int P = pp->npar;
    double out = 0.0;
    for (int s = 0; s < S; ++s) {
        double v = 0.0;
        for (int k = 0; k < P; ++k)
            v += log_prior1(pp->dist[k], phi[k], phi[P + k], pp->lower[k], pp->upper[k],
                            subj[s * stride + k]);
        out += v;
    }
    return out;
}

/* ------------------------------------------------------------ sampler */
/* One sampled level: state [np x n] plus cached lp / ll per chain. */
struct Level {
    int np, n;
    std::vector<double> th, lp, ll;
    double *col(int i) { return &th[(size_t)i * np]; }
};

struct LevelOps {
    /* log prior / log likelihood of a proposal made for chain `chain` */
    virtual double logprior(const double *x, int chain) = 0;
    virtual double loglike(const double *x, int chain) = 0;
    /* phi level: recompute the cached likelihood of chain i before it is used */
    virtual bool refreshes() const { return false; }
    virtual double refresh(const double *x, int chain) { (void)x; (void)chain; return 0.0; }
    virtual ~LevelOps() {}
};

void store(const Level &L, int k, double *th_out, double *lp_out, double *ll_out) {
    std::memcpy(th_out + (size_t)k * L.np * L.n, L.th.data(), sizeof(double) * L.np * L.n);
    std::memcpy(lp_out + (size_t)k * L.n, L.lp.data(), sizeof(double) * L.n);
    std::memcpy(ll_out + (size_t)k * L.n, L.ll.data(), sizeof(double) * L.n);
}

/* Metropolis accept (src/de.cpp:81-108): ratio = exp(post' - post); NaN
 * rejects; otherwise accept iff U < ratio. */
inline bool accept(double tmp_post, double cur_post, double U) {
    const double ratio = std::exp(tmp_post - cur_post);
    if (std::isnan(ratio)) return false;
    return U < ratio;
}

/* DE crossover sweep (src/de.cpp:111-155, :385-465, :567-613) over G chain
 * groups: chains of group g propose from two distinct partners outside g.
 * G = n gives the reference's in-place sequential order. para < 0: joint. */
void crossover(Level &L, LevelOps &ops, const Stream &st, double gamma, double rp,
               int groups, int para) {
    std::vector<double> tmp(L.np);
    const int n = L.n;
    for (int g = 0; g < groups; ++g) {
        const int lo = (g * n) / groups, hi = ((g + 1) * n) / groups;
        const int pool = n - (hi - lo);
        for (int i = lo; i < hi; ++i) {
            if (ops.refreshes()) L.ll[i] = ops.refresh(L.col(i), i);
            const double cur = L.ll[i] + L.lp[i];
            uint32_t b0[4];
            st.block((uint32_t)i, 0, b0);
            uint32_t r0 = below(b0[0], (uint32_t)pool);
            uint32_t r1 = below(b0[1], (uint32_t)(pool - 1));
            if (r1 >= r0) ++r1;
            const int m0 = (int)r0 < lo ? (int)r0 : (int)r0 + (hi - lo);
            const int m1 = (int)r1 < lo ? (int)r1 : (int)r1 + (hi - lo);
            const double U = u53(b0[2], b0[3]);
            const double *th0 = L.col(m0), *th1 = L.col(m1);
            std::memcpy(tmp.data(), L.col(i), sizeof(double) * L.np);
            if (para >= 0) {
                const double jit = rp * (2.0 * u32_open(st.word((uint32_t)i, 4)) - 1.0);
                tmp[para] += jit + gamma * (th0[para] - th1[para]);
            } else {
                for (int j = 0; j < L.np; ++j) {
                    const double jit = rp * (2.0 * u32_open(st.word((uint32_t)i, 4 + j)) - 1.0);
                    tmp[j] += jit + gamma * (th0[j] - th1[j]);
                }
            }
            const double tlp = ops.logprior(tmp.data(), i);
            const double tll = ops.loglike(tmp.data(), i);
            if (accept(tlp + tll, cur, U)) {
                std::memcpy(L.col(i), tmp.data(), sizeof(double) * L.np);
                L.lp[i] = tlp;
                L.ll[i] = tll;
            }
        }
    }
}

/* Migration subset (src/de.cpp:62-78): size ceil(n*U) clamped to [2, n],
 * a uniform random subset of that size, sorted ascending. */
std::vector<int> subchains(const Stream &st, int n) {
    const double U = u53(st.word(kSlotSubset, 0), st.word(kSlotSubset, 1));
    int ns = (int)std::ceil(n * U);
    ns = std::min(std::max(ns, 2), n);
    std::vector<int> a(n);
    for (int k = 0; k < n; ++k) a[k] = k;
    for (int k = 0; k < ns; ++k) {
        const int j = k + (int)below(st.word(kSlotSubset, 2 + k), (uint32_t)(n - k));
        std::swap(a[k], a[j]);
    }
    std::vector<int> out(a.begin(), a.begin() + ns);
    std::sort(out.begin(), out.end());
    return out;
}

/* Migration ring (src/de.cpp:157-199, :467-565, :615-665): a jittered copy of
 * chain S_k is proposed for slot S_{k+1}, judged against S_{k+1}'s cached
 * posterior; prior / likelihood use the source chain's context. */
void migration(Level &L, LevelOps &ops, const Stream &st, double rp, int para) {
    const std::vector<int> S = subchains(st, L.n);
    const int ns = (int)S.size();
    std::vector<double> tmp(L.np);
    for (int k = 0; k < ns; ++k) {
        const int src = S[k];
        const int dst = S[(k + 1) % ns];
        if (ops.refreshes()) {
            L.ll[src] = ops.refresh(L.col(src), src);
            L.ll[dst] = ops.refresh(L.col(dst), dst);
        }
        const uint32_t slot = kSlotMigrate | (uint32_t)k;
        std::memcpy(tmp.data(), L.col(src), sizeof(double) * L.np);
        if (para >= 0) {
            tmp[para] += rp * (2.0 * u32_open(st.word(slot, 4)) - 1.0);
        } else {
            for (int j = 0; j < L.np; ++j)
                tmp[j] += rp * (2.0 * u32_open(st.word(slot, 4 + j)) - 1.0);
        }
        const double tlp = ops.logprior(tmp.data(), src);
        const double tll = ops.loglike(tmp.data(), src);
        const double cur = L.lp[dst] + L.ll[dst];
        const double U = u53(st.word(slot, 0), st.word(slot, 1));
        if (accept(tlp + tll, cur, U)) {
            std::memcpy(L.col(dst), tmp.data(), sizeof(double) * L.np);
            L.lp[dst] = tlp;
            L.ll[dst] = tll;
        }
    }
}

inline double decision_u(const Stream &st) {
    return u53(st.word(kSlotDecision, 0), st.word(kSlotDecision, 1));
}

/* One level step: joint (one decision) or blocked (one decision per
 * parameter), as run_chains / run_hchains do. single_level_blocking mirrors
 * run_chains, whose blocked branch has no per-parameter migration. */
void level_step(Level &L, LevelOps &ops, Stream st, double gamma, double rp, double mig_prob,
                int groups, bool blocked, bool single_level) {
    if (!blocked) {
        st.sweep = 0;
        if (decision_u(st) < mig_prob)
            migration(L, ops, st, rp, -1);
        else
            crossover(L, ops, st, gamma, rp, groups, -1);
        return;
    }
    if (single_level) { /* src/de.cpp:208-236 */
        st.sweep = 0;
        if (decision_u(st) < mig_prob) {
            migration(L, ops, st, rp, -1);
            return;
        }
        for (int p = 0; p < L.np; ++p) {
            st.sweep = (uint32_t)p;
            crossover(L, ops, st, gamma, rp, groups, p);
        }
        return;
    }
    for (int p = 0; p < L.np; ++p) { /* src/de.cpp:289-306, :339-358 */
        st.sweep = (uint32_t)p;
        if (decision_u(st) < mig_prob)
            migration(L, ops, st, rp, p);
        else
            crossover(L, ops, st, gamma, rp, groups, p);
    }
}

bool valid_cfg(const oracle_config *c, int np) {
    if (c->nchain <= 2 || c->nmc < 1 || c->thin < 1 || np < 1) return false;
    if (c->sweep_groups < 2 || c->sweep_groups > c->nchain) return false;
    /* the largest group must leave two partners */
    const int gmax = (c->nchain + c->sweep_groups - 1) / c->sweep_groups;
    return c->nchain - gmax >= 2;
}

Level make_level(int np, int n, const double *th, const double *lp, const double *ll) {
    Level L;
    L.np = np;
    L.n = n;
    L.th.assign(th, th + (size_t)np * n);
    L.lp.assign(lp, lp + n);
    L.ll.assign(ll, ll + n);
    return L;
}

Stream make_stream(const oracle_config *c, uint32_t unit) {
    Stream s;
    s.key[0] = (uint32_t)c->seed;
    s.key[1] = (uint32_t)(c->seed >> 32);
    s.iter = 0;
    s.unit = unit;
    s.sweep = 0;
    return s;
}

struct LbaOps : LevelOps {
    const oracle_model *m;
    const oracle_prior *prior;
    const double *rt;
    const int64_t *cell_off;
    double logprior(const double *x, int) override { return sumlogprior(prior, x); }
    double loglike(const double *x, int) override {
        double s;
        lba_terms(m, rt, cell_off, x, nullptr, &s);
        return s;
    }
};

struct NormalOps : LevelOps {
    const oracle_prior *prior;
    const double *x;
    const int64_t *cell_off;
    int ncell;
    double logprior(const double *th, int) override { return sumlogprior(prior, th); }
    double loglike(const double *th, int) override {
        double s = 0.0;
        for (int c = 0; c < ncell; ++c)
            for (int64_t k = cell_off[c]; k < cell_off[c + 1]; ++k) {
                const double d = x[k] - th[c];
                s += -0.5 * d * d - 0.918938533204672741780329736406;
            }
        return s;
    }
};

struct HyperOps : LevelOps { /* run_hyper: data rows are fixed subject vectors */
    const oracle_prior *p_prior, *h_prior;
    const double *data;
    int S;
    double logprior(const double *x, int) override { return sumlogprior(h_prior, x); }
    double loglike(const double *x, int) override {
        return sumloghlike(p_prior, x, S, data, p_prior->npar);
    }
};

} // namespace

/* ================================================================ C ABI */
extern "C" {

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox(ctr, key, out);
}

double oracle_lba_loglike(const oracle_model *m, const double *rt, const int64_t *cell_off,
                          const double *theta) {
    double s;
    lba_terms(m, rt, cell_off, theta, nullptr, &s);
    return s;
}

void oracle_lba_trial_terms(const oracle_model *m, const double *rt, const int64_t *cell_off,
                            const double *theta, double *out) {
    lba_terms(m, rt, cell_off, theta, out, nullptr);
}

double oracle_sumlogprior(const oracle_prior *p, const double *x) { return sumlogprior(p, x); }

double oracle_sumloghlike(const oracle_prior *p, const double *phi, int32_t S, const double *subj,
                          int64_t stride) {
    return sumloghlike(p, phi, S, subj, stride);
}

static int run_single(const oracle_config *cfg, LevelOps &ops, int np, const double *init_theta,
                      const double *init_lp, const double *init_ll, double *theta_out,
                      double *lp_out, double *ll_out) {
    if (!valid_cfg(cfg, np)) return -1;
    Level L = make_level(np, cfg->nchain, init_theta, init_lp, init_ll);
    const double gamma = cfg->gamma_precursor / std::sqrt(2.0 * cfg->nparameter_de);
    store(L, 0, theta_out, lp_out, ll_out);
    Stream st = make_stream(cfg, 0u);
    const long nsample = (long)cfg->nmc * cfg->thin;
    for (long it = 1; it < nsample; ++it) {
        st.iter = (uint32_t)(it + cfg->iter_offset);
        level_step(L, ops, st, gamma, cfg->rp, cfg->sub_migration_prob, cfg->sweep_groups,
                   cfg->is_pblocked != 0, true);
        if (it % cfg->thin == 0) store(L, (int)(it / cfg->thin), theta_out, lp_out, ll_out);
    }
    return 0;
}

int oracle_run_subject(const oracle_config *cfg, int32_t model_kind, const oracle_model *m,
                       const oracle_prior *prior, const double *rt, const int64_t *cell_off,
                       const double *init_theta, const double *init_lp, const double *init_ll,
                       double *theta_out, double *lp_out, double *ll_out) {
    if (model_kind == ORACLE_MODEL_LBA) {
        LbaOps ops;
        ops.m = m;
        ops.prior = prior;
        ops.rt = rt;
        ops.cell_off = cell_off;
        return run_single(cfg, ops, m->npar, init_theta, init_lp, init_ll, theta_out, lp_out,
                          ll_out);
    }
    if (model_kind == ORACLE_MODEL_NORMAL) {
        NormalOps ops;
        ops.prior = prior;
        ops.x = rt;
        ops.cell_off = cell_off;
        ops.ncell = m->n_cell;
        return run_single(cfg, ops, m->npar, init_theta, init_lp, init_ll, theta_out, lp_out,
                          ll_out);
    }
    return -2;
}

int oracle_run_hyper(const oracle_config *cfg, const oracle_prior *p_prior,
                     const oracle_prior *h_prior, int32_t nsubject, const double *data,
                     const double *init_theta, const double *init_lp, const double *init_ll,
                     double *theta_out, double *lp_out, double *ll_out) {
    HyperOps ops;
    ops.p_prior = p_prior;
    ops.h_prior = h_prior;
    ops.data = data;
    ops.S = nsubject;
    return run_single(cfg, ops, h_prior->npar, init_theta, init_lp, init_ll, theta_out, lp_out,
                      ll_out);
}

namespace {
struct PhiOps : LevelOps { /* phi level of run_hchains */
    const oracle_prior *p_prior, *h_prior;
    std::vector<Level> *subj;
    int S;
    double hll(const double *phi, int chain) {
        const int P = p_prior->npar;
        double out = 0.0;
        for (int s = 0; s < S; ++s) {
            const double *x = (*subj)[s].col(chain);
            double v = 0.0;
            for (int k = 0; k < P; ++k)
                v += log_prior1(p_prior->dist[k], phi[k], phi[P + k], p_prior->lower[k],
                                p_prior->upper[k], x[k]);
            out += v;
        }
        return out;
    }
    double logprior(const double *x, int) override { return sumlogprior(h_prior, x); }
    double loglike(const double *x, int chain) override { return hll(x, chain); }
    bool refreshes() const override { return true; }
    double refresh(const double *x, int chain) override { return hll(x, chain); }
};

struct SubjOps : LevelOps { /* subject | phi level of run_hchains */
    const oracle_model *m;
    const oracle_prior *p_prior;
    Level *phi;
    const double *rt;
    const int64_t *cell_off;
    double logprior(const double *x, int chain) override {
        const int P = p_prior->npar;
        const double *ph = phi->col(chain);
        double s = 0.0;
        for (int k = 0; k < P; ++k)
            s += log_prior1(p_prior->dist[k], ph[k], ph[P + k], p_prior->lower[k],
                            p_prior->upper[k], x[k]);
        return s;
    }
    double loglike(const double *x, int) override {
        double s;
        lba_terms(m, rt, cell_off, x, nullptr, &s);
        return s;
    }
};
} // namespace

int oracle_run_hier(const oracle_config *cfg, const oracle_model *m, const oracle_prior *p_prior,
                    const oracle_prior *h_prior, int32_t S, const double *rt,
                    const int64_t *subject_off, const int64_t *cell_off,
                    const double *phi_init_theta, const double *phi_init_lp,
                    const double *phi_init_ll, const double *subj_init_theta,
                    const double *subj_init_lp, const double *subj_init_ll,
                    double *phi_theta_out, double *phi_lp_out, double *phi_ll_out,
                    double *subj_theta_out, double *subj_lp_out, double *subj_ll_out) {
    const int P = m->npar;
    const int n = cfg->nchain;
    if (!valid_cfg(cfg, P) || h_prior->npar != 2 * P || p_prior->npar != P || S < 1) return -1;
    Level phi = make_level(2 * P, n, phi_init_theta, phi_init_lp, phi_init_ll);
    std::vector<Level> subj(S);
    for (int s = 0; s < S; ++s)
        subj[s] = make_level(P, n, subj_init_theta + (size_t)s * P * n, subj_init_lp + (size_t)s * n,
                             subj_init_ll + (size_t)s * n);
    const double gamma = cfg->gamma_precursor / std::sqrt(2.0 * cfg->nparameter_de);
    const long nsample = (long)cfg->nmc * cfg->thin;
    const size_t sub_th = (size_t)P * n * cfg->nmc, sub_ch = (size_t)n * cfg->nmc;

    PhiOps pops;
    pops.p_prior = p_prior;
    pops.h_prior = h_prior;
    pops.subj = &subj;
    pops.S = S;
    std::vector<SubjOps> sops(S);
    for (int s = 0; s < S; ++s) {
        sops[s].m = m;
        sops[s].p_prior = p_prior;
        sops[s].phi = &phi;
        sops[s].rt = rt + subject_off[s];
        sops[s].cell_off = cell_off + (size_t)s * (m->n_cell + 1);
    }
    store(phi, 0, phi_theta_out, phi_lp_out, phi_ll_out);
    for (int s = 0; s < S; ++s)
        store(subj[s], 0, subj_theta_out + s * sub_th, subj_lp_out + s * sub_ch,
              subj_ll_out + s * sub_ch);
    for (long it = 1; it < nsample; ++it) {
        Stream st = make_stream(cfg, kUnitPhi);
        st.iter = (uint32_t)(it + cfg->iter_offset);
        level_step(phi, pops, st, gamma, cfg->rp, cfg->pop_migration_prob, cfg->sweep_groups,
                   cfg->is_hblocked != 0, false);
        for (int s = 0; s < S; ++s) {
            Stream ss = st;
            ss.unit = (uint32_t)s;
            level_step(subj[s], sops[s], ss, gamma, cfg->rp, cfg->sub_migration_prob,
                       cfg->sweep_groups, cfg->is_pblocked != 0, false);
        }
        if (it % cfg->thin == 0) {
            const int k = (int)(it / cfg->thin);
            for (int s = 0; s < S; ++s)
                store(subj[s], k, subj_theta_out + s * sub_th, subj_lp_out + s * sub_ch,
                      subj_ll_out + s * sub_ch);
            store(phi, k, phi_theta_out, phi_lp_out, phi_ll_out);
        }
    }
    return 0;
}

} // extern "C"
