/*
 * HIP kernels of the DE-MCMC sampler (gfx950, wave64).
 *
 *  subject_sweep_kernel  one workgroup per work item (subject, chain range of a
 *                        chain group, trial tile): DE crossover proposals
 *                        (src/de.cpp:567-613 / :111-155), LBA log-likelihood as
 *                        canonical chunk sums, Metropolis accept (:81-108) by the
 *                        item itself or by the last-arriving tile; migration ring
 *                        (src/de.cpp:615-665 / :157-199) and the deferred
 *                        thinned store of the previous iteration by the
 *                        subject's first item.
 *  phi_kernel            per-(chain, parameter) sufficient statistics of the
 *                        subject states, then the population-level step
 *                        (src/de.cpp:385-465, :467-565) with sumloghlike
 *                        (:245-270) from them.
 *  plain_*_kernel        the reference arithmetic and sequential order: one
 *                        launch per chain step (the FP64 comparator path).
 *  loglike_kernel        sumloglike of given parameter vectors.
 */
#include <hip/hip_runtime.h>

#include "engine.h"
#include "lba_math.h"
#include "philox.h"
#include "../../include/ggdmc_hip.h"

namespace ggdmc {

/* 2^(j/64) for exp2_neg, staged into LDS by every kernel that evaluates the
 * batched-path likelihood (load_exp2_table before its first barrier). */
__shared__ double g_exp2tab[64];

__device__ __forceinline__ void load_exp2_table(const Dev &d) {
    for (int j = threadIdx.x; j < 64; j += blockDim.x) g_exp2tab[j] = d.exp2tab[j];
}

struct DevOps {
    static __device__ __forceinline__ double exp2_frac(int j) { return g_exp2tab[j]; }
    static __device__ __forceinline__ float exp2(float x) { return __builtin_amdgcn_exp2f(x); }
    static __device__ __forceinline__ float log2(float x) { return __builtin_amdgcn_logf(x); }
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ double erfc(double x) { return ::erfc(x); }
    /* R's pnorm: the lower tail accurately, above 0 as 1 - upper tail */
    static __device__ __forceinline__ double pnorm(double z) {
        return z > 0.0 ? 1.0 - 0.5 * ::erfc(z * kSqrt1_2) : 0.5 * ::erfc(-z * kSqrt1_2);
    }
    static __device__ __forceinline__ double dnorm(double z) {
        return kInvSqrt2Pi * ::exp(-0.5 * z * z);
    }
};

enum { MODE_MIXED = 0, MODE_NAIVE = 1 };

constexpr int kAccKDoubles = sizeof(AccK) / sizeof(double);

/* --------------------------------------------------------------- likelihood */
template <int MODE>
__device__ __forceinline__ double trial_term(const AccK *kc, int n_acc, int w, double rt) {
    return MODE == MODE_MIXED ? mixed_trial_term<DevOps>(kc, n_acc, w, rt)
                              : naive_trial_term<DevOps>(kc, n_acc, w, rt);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

/* Sum over the 8 lanes of an aligned lane group (fixed xor tree). */
__device__ __forceinline__ double sum8(double v) {
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) v += __shfl_xor(v, off, 8);
    return v;
}

/* Terms t[0..m) summed by 8 lanes (lane l takes terms l, l + 8, ... in order,
 * then the xor tree); every lane of the group gets the sum. */
__device__ __forceinline__ double group8_sum(const double *t, int m, int l) {
    double v = 0.0;
    for (int q = l; q < m; q += 8) v += t[q];
    return sum8(v);
}

/* Deterministic block sum (fixed tree); all threads get the result. */
__device__ double block_sum(double v, double *red) {
    const int tid = threadIdx.x;
    v = wave_sum(v);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    __syncthreads();
    return s;
}

/* Canonical chunk sums: C[c * c_stride + jj] = sum of the trial terms of chunk
 * j_lo + jj of the subject whose trials are [t0, t_end), for the kb chains
 * whose constants are acck[c][n_cond][n_acc]. A wave evaluates one
 * (chain, chunk) unit at a time: lane l adds the terms of trials
 * t0 + j*kChunk + l and + 64 + l (zero past the end), then the fixed xor
 * butterfly. Chains with active[c] == 0 (prior density zero) get 0. */
/* Trials of a work item: rt[i - origin], code[i - origin] for trial i (the
 * item's trials staged in LDS, or the global arrays with origin 0). */
struct TrialView {
    const double *rt;
    const uint16_t *code;
    int64_t origin;
};

template <int MODE>
__device__ void chunk_sums(const Dev &d, const TrialView &tv, int64_t t0, int64_t t_end, int j_lo, int nj,
                           int kb, const int *active, const double *acck, double *C, int c_stride) {
    static_assert(kChunk == 128, "two trials per lane");
    const int lane = threadIdx.x & 63;
    /* the wave index is uniform: keep the unit loop scalar */
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nw = blockDim.x >> 6;
    const int per_chain = d.n_cond * d.n_acc;
    for (int u = wave; u < kb * nj; u += nw) {
        const int c = u / nj, jj = u - c * nj;
        double v = 0.0;
        if (active[c]) {
            const AccK *kc = reinterpret_cast<const AccK *>(acck) + (size_t)c * per_chain;
            const int64_t i1 = t0 + (int64_t)(j_lo + jj) * kChunk + lane, i2 = i1 + 64;
            double a = 0.0, b = 0.0;
            const bool h1 = i1 < t_end, h2 = i2 < t_end;
            /* lanes past the end evaluate the chunk's first trial and drop it,
             * so that both evaluations are straight-line code */
            const int64_t r1 = (h1 ? i1 : t0 + (int64_t)(j_lo + jj) * kChunk) - tv.origin;
            const int64_t r2 = (h2 ? i2 : t0 + (int64_t)(j_lo + jj) * kChunk) - tv.origin;
            const int code1 = tv.code[r1], code2 = tv.code[r2];
            const double rt1 = tv.rt[r1], rt2 = tv.rt[r2];
            const AccK *k1 = kc + (code1 >> 3) * d.n_acc, *k2 = kc + (code2 >> 3) * d.n_acc;
            if (MODE == MODE_MIXED) {
                bool ill1 = false, ill2 = false;
                const double ma = mixed_trial_term_flagged<DevOps>(k1, d.n_acc, code1 & 7, rt1, ill1);
                const double mb = mixed_trial_term_flagged<DevOps>(k2, d.n_acc, code2 & 7, rt2, ill2);
                a = h1 ? ma : 0.0;
                b = h2 ? mb : 0.0;
                /* rare: ill-conditioned trials take the reference formula */
                if (h1 && ill1) a = naive_trial_term<DevOps>(k1, d.n_acc, code1 & 7, rt1);
                if (h2 && ill2) b = naive_trial_term<DevOps>(k2, d.n_acc, code2 & 7, rt2);
            } else {
                if (h1) a = naive_trial_term<DevOps>(k1, d.n_acc, code1 & 7, rt1);
                if (h2) b = naive_trial_term<DevOps>(k2, d.n_acc, code2 & 7, rt2);
            }
            v = a + b;
        }
        v = wave_sum(v);
        if (lane == 0) C[(size_t)c * c_stride + jj] = v;
    }
}

/* ------------------------------------------------------------------ priors */
/* Phi at +-inf without evaluating erfc (the same values: 1 and 0). */
__device__ __forceinline__ double pnorm_bound(double z) {
    return z == __builtin_inf() ? 1.0 : (z == -__builtin_inf() ? 0.0 : DevOps::pnorm(z));
}

/* Phi(b) - Phi(a) for a <= b; as Q(a) - Q(b) when a > 0 (no cancellation). */
__device__ __forceinline__ double tnorm_mass(double a, double b) {
    return a > 0.0 ? pnorm_bound(-a) - pnorm_bound(-b) : pnorm_bound(b) - pnorm_bound(a);
}

__device__ double log_prior1(int dist, double p0, double p1, double lo, double up, double x) {
    if (x != x) return x;
    if (dist == GGDMC_DIST_UNIF) {
        if (x < p0 || x > p1) return -__builtin_inf();
        return -::log(p1 - p0);
    }
    if (dist == GGDMC_DIST_TNORM) {
        if (!(p1 > 0.0)) return __builtin_nan("");
        if (x < lo || x > up) return -__builtin_inf();
        const double z = (x - p0) / p1;
        const double Z = tnorm_mass((lo - p0) / p1, (up - p0) / p1);
        return -0.5 * z * z - 0.918938533204672741780329736406 - ::log(p1) - ::log(Z);
    }
    return __builtin_nan("");
}

/* Parameter p's term of the subject-level log prior for chain `chain`: the
 * run's fixed prior, or for hierarchical runs tnorm(loc = phi head,
 * scale = phi tail) (src/de.cpp:599-600). */
__device__ __forceinline__ double subject_prior_term(const Dev &d, const JobDev &job, int chain, int p,
                                                     double x) {
    if (job.kind == GGDMC_KIND_HIER) {
        const double *ph = d.phi + ((size_t)job.phi_index * d.nchain + chain) * (2 * d.P);
        return log_prior1(GGDMC_DIST_TNORM, ph[p], ph[d.P + p], d.pr_lo[job.prior_off + p],
                          d.pr_up[job.prior_off + p], x);
    }
    const int k = job.prior_off + p;
    return log_prior1(d.pr_dist[k], d.pr_p0[k], d.pr_p1[k], d.pr_lo[k], d.pr_up[k], x);
}

/* The same term from cached logarithms (log scale, log normaliser): for
 * hierarchical runs from the population step's cache of the chain's phi,
 * for fixed priors from the session's; same operations, same values. */
__device__ __forceinline__ double subject_prior_term_cached(const Dev &d, const JobDev &job, int chain,
                                                            int p, double x) {
    if (x != x) return x;
    if (job.kind == GGDMC_KIND_HIER) {
        const size_t r = ((size_t)job.phi_index * d.nchain + chain) * d.P + p;
        const double *ph = d.phi + ((size_t)job.phi_index * d.nchain + chain) * (2 * d.P);
        const double mu = ph[p], sg = ph[d.P + p];
        if (!(sg > 0.0)) return __builtin_nan("");
        if (x < d.pr_lo[job.prior_off + p] || x > d.pr_up[job.prior_off + p]) return -__builtin_inf();
        const double z = (x - mu) / sg;
        return -0.5 * z * z - 0.918938533204672741780329736406 - d.pcache[2 * r] - d.pcache[2 * r + 1];
    }
    const int k = job.prior_off + p;
    const int dist = d.pr_dist[k];
    const double p0 = d.pr_p0[k], p1 = d.pr_p1[k];
    if (dist == GGDMC_DIST_UNIF) {
        if (x < p0 || x > p1) return -__builtin_inf();
        return -d.scache[2 * k];
    }
    if (dist == GGDMC_DIST_TNORM) {
        if (!(p1 > 0.0)) return __builtin_nan("");
        if (x < d.pr_lo[k] || x > d.pr_up[k]) return -__builtin_inf();
        const double z = (x - p0) / p1;
        return -0.5 * z * z - 0.918938533204672741780329736406 - d.scache[2 * k] - d.scache[2 * k + 1];
    }
    return __builtin_nan("");
}

/* sumlogprior (src/de.cpp:144, :186) of x: terms added in parameter order. */
__device__ double subject_logprior(const Dev &d, const JobDev &job, int chain, const double *x) {
    double s = 0.0;
    for (int p = 0; p < d.P; ++p) s += subject_prior_term(d, job, chain, p, x[p]);
    return s;
}

__device__ __forceinline__ bool mh_accept(double tpost, double cur, double U) {
    const double ratio = ::exp(tpost - cur);
    if (ratio != ratio) return false;
    return U < ratio;
}

/* -------------------------------------------------------- LDS carve-up */
struct SweepLds {
    double *th, *lp, *ll;   /* [n*P], [n], [n]: the subject's chain states */
    double *prop, *tm;      /* [kmax*P]: proposals, their prior terms */
    double *jit;            /* [kmax*P]: jitter of each proposal parameter */
    double *pri;            /* [kmax*P*4]: prior location, scale, log scale, log mass */
    double *bnd;            /* [P*2]: prior bounds */
    double *trt;            /* [tstage]: the item's response times */
    uint16_t *tcd;          /* [tstage]: the item's trial codes */
    double *tlp, *U, *tll;  /* [kmax] */
    double *acck;           /* [chain_batch * per_chain] accumulator constants */
    double *csum;           /* [kmax * cmax] chunk sums */
    int *m0, *m1, *active, *acc; /* [kmax] */
    int *sub;               /* [n] migration subset */
    int *pdist;             /* [P] prior distribution of each parameter */
    int *flag;              /* [4] */
};

__host__ __device__ inline size_t sweep_lds_bytes(int n, int P, int kmax, int per_chain,
                                                  int chain_batch, int cmax, int tstage) {
    const size_t dbl = (size_t)n * P + 2 * (size_t)n + 7 * (size_t)kmax * P + 2 * (size_t)P +
                       3 * (size_t)kmax + (size_t)chain_batch * per_chain * kAccKDoubles +
                       (size_t)kmax * cmax;
    const size_t ints = 4 * (size_t)kmax + n + P + 4;
    return (dbl + (size_t)tstage) * sizeof(double) + ints * sizeof(int) +
           (((size_t)tstage * sizeof(uint16_t) + 7) & ~(size_t)7);
}

__device__ SweepLds carve(char *base, int n, int P, int kmax, int per_chain, int chain_batch,
                          int cmax, int tstage) {
    SweepLds L;
    double *p = reinterpret_cast<double *>(base);
    L.th = p; p += (size_t)n * P;
    L.lp = p; p += n;
    L.ll = p; p += n;
    L.prop = p; p += (size_t)kmax * P;
    L.tm = p; p += (size_t)kmax * P;
    L.jit = p; p += (size_t)kmax * P;
    L.pri = p; p += (size_t)kmax * P * 4;
    L.bnd = p; p += (size_t)P * 2;
    L.tlp = p; p += kmax;
    L.U = p; p += kmax;
    L.tll = p; p += kmax;
    L.acck = p; p += (size_t)chain_batch * per_chain * kAccKDoubles;
    L.csum = p; p += (size_t)kmax * cmax;
    L.trt = p; p += tstage;
    L.tcd = reinterpret_cast<uint16_t *>(p);
    p += ((size_t)tstage * sizeof(uint16_t) + 7) / 8;
    int *q = reinterpret_cast<int *>(p);
    L.m0 = q; q += kmax;
    L.m1 = q; q += kmax;
    L.active = q; q += kmax;
    L.acc = q; q += kmax;
    L.sub = q; q += n;
    L.pdist = q; q += P;
    L.flag = q;
    return L;
}

/* dst[i] = src[i], i < count, by the whole workgroup with four loads in flight
 * per thread (global -> LDS staging). */
template <class T>
__device__ __forceinline__ void stage_copy(T *dst, const T *src, int count) {
    const int bs = blockDim.x;
    for (int i0 = threadIdx.x; i0 < count; i0 += 4 * bs) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * bs;
            v[u] = i < count ? src[i] : T();
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * bs;
            if (i < count) dst[i] = v[u];
        }
    }
}

/* Fills the accumulator constants of proposals [cb, cb+kb). */
__device__ void build_acck(const Dev &d, const double *prop, int P, int cb, int kb, double *acck) {
    const int per_chain = d.n_cond * d.n_acc;
    AccK *K = reinterpret_cast<AccK *>(acck);
    for (int e = threadIdx.x; e < kb * per_chain; e += blockDim.x) {
        const int c = e / per_chain, r = e - c * per_chain;
        make_acck<DevOps>(prop + (size_t)(cb + c) * P, d.par_src + r * 6, d.par_const + r * 6, K[e]);
    }
}

/* Prior terms of proposals L.prop[0..k) for chains i0 + c, one thread per
 * (chain, parameter), summed per chain in parameter order into L.tlp. */
__device__ void proposal_priors(const Dev &d, const JobDev &job, const SweepLds &L, int i0, int k) {
    const int P = d.P;
    for (int e = threadIdx.x; e < k * P; e += blockDim.x) {
        const int c = e / P, p = e - c * P;
        L.tm[e] = subject_prior_term_cached(d, job, i0 + c, p, L.prop[e]);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < k; c += blockDim.x) {
        double s = 0.0;
        for (int p = 0; p < P; ++p) s += L.tm[c * P + p];
        L.tlp[c] = s;
        L.active[c] = s > -__builtin_inf();
    }
    __syncthreads();
}

/* Crossover proposals of chains i0 .. i0+k-1 of group [lo, hi) (src/de.cpp:583-597,
 * :119-135) in three stages separated by barriers:
 *  1. draws and constants that do not depend on the states: partners, accept U
 *     and jitter per (chain, Philox block); the prior's location, scale and
 *     cached logarithms per (chain, parameter) and its bounds per parameter;
 *  2. per (chain, parameter): the proposal value and its prior term;
 *  3. per chain: the prior terms summed by 8 lanes; per (chain, condition,
 *     accumulator): the likelihood constants when the chains fit one batch.
 * para >= 0: blocked update of that parameter only. */
__device__ void proposal_stage1(const Dev &d, const JobDev &job, const SweepLds &L, const Stream &st,
                                int lo, int hi, int i0, int k, int para) {
    const int P = d.P, n = d.nchain;
    const int nb = para >= 0 ? 1 : (P + 3) >> 2;
    for (int e = threadIdx.x; e < k * nb; e += blockDim.x) {
        const int c = e / nb, b = e - c * nb;
        const int i = i0 + c;
        if (b == 0) {
            int m0, m1;
            double U;
            partners(st, i, lo, hi, n, m0, m1, U);
            L.m0[c] = m0;
            L.m1[c] = m1;
            L.U[c] = U;
        }
        const U4 r = st.block((uint32_t)i, 1u + (uint32_t)b);
        if (para >= 0) {
            L.jit[c * P + para] = jitter_word(r.x[0], d.rp);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int p = 4 * b + j;
                if (p < P) L.jit[c * P + p] = jitter_word(r.x[j], d.rp);
            }
        }
    }
    const bool hier = job.kind == GGDMC_KIND_HIER;
    for (int e = threadIdx.x; e < k * P; e += blockDim.x) {
        const int c = e / P, p = e - c * P;
        double *q = L.pri + (size_t)e * 4;
        if (hier) {
            const size_t r = ((size_t)job.phi_index * n + i0 + c) * P + p;
            const double *ph = d.phi + ((size_t)job.phi_index * n + i0 + c) * (2 * P);
            q[0] = ph[p];
            q[1] = ph[P + p];
            q[2] = d.pcache[2 * r];
            q[3] = d.pcache[2 * r + 1];
        } else {
            const int kk = job.prior_off + p;
            q[0] = d.pr_p0[kk];
            q[1] = d.pr_p1[kk];
            q[2] = d.scache[2 * kk];
            q[3] = d.scache[2 * kk + 1];
        }
    }
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        L.bnd[2 * p] = d.pr_lo[job.prior_off + p];
        L.bnd[2 * p + 1] = d.pr_up[job.prior_off + p];
        L.pdist[p] = hier ? GGDMC_DIST_TNORM : d.pr_dist[job.prior_off + p];
    }
}

/* Prior term of x from the staged constants q = (location or lower end,
 * scale or upper end, log scale or log width, log mass): the operations of
 * log_prior1 with the logarithms cached. */
__device__ __forceinline__ double staged_prior_term(int dist, const double *q, double lo, double up,
                                                    double x) {
    if (x != x) return x;
    if (dist == GGDMC_DIST_UNIF) {
        if (x < q[0] || x > q[1]) return -__builtin_inf();
        return -q[2];
    }
    if (dist == GGDMC_DIST_TNORM) {
        if (!(q[1] > 0.0)) return __builtin_nan("");
        if (x < lo || x > up) return -__builtin_inf();
        const double z = (x - q[0]) / q[1];
        return -0.5 * z * z - 0.918938533204672741780329736406 - q[2] - q[3];
    }
    return __builtin_nan("");
}

__device__ void proposal_stage2(const Dev &d, const JobDev &job, const SweepLds &L, int i0, int k,
                                int para) {
    const int P = d.P;
    for (int e = threadIdx.x; e < k * P; e += blockDim.x) {
        const int c = e / P, p = e - c * P;
        const double xi = L.th[(size_t)(i0 + c) * P + p];
        double x = xi;
        if (para < 0 || p == para)
            x = xi + (L.jit[e] + job.gamma * (L.th[(size_t)L.m0[c] * P + p] - L.th[(size_t)L.m1[c] * P + p]));
        L.prop[e] = x;
        L.tm[e] = staged_prior_term(L.pdist[p], L.pri + (size_t)e * 4, L.bnd[2 * p], L.bnd[2 * p + 1], x);
    }
}

/* Returns whether the likelihood constants of all k chains were built. */
__device__ bool proposal_stage3(const Dev &d, const SweepLds &L, int k, int chain_batch) {
    const int P = d.P;
    for (int e = threadIdx.x; e < ((k + 7) & ~7) * 8; e += blockDim.x) {
        const int c = e >> 3, l = e & 7;
        const double v = group8_sum(L.tm + (size_t)(c < k ? c : 0) * P, P, l);
        if (c < k && l == 0) {
            L.tlp[c] = v;
            L.active[c] = v > -__builtin_inf();
        }
    }
    const bool all = k <= chain_batch;
    if (all) build_acck(d, L.prop, P, 0, k, L.acck);
    return all;
}

/* Log-likelihood of proposals L.prop[0..k) over chunks [j_lo, j_lo + nj) of
 * the subject's trials [t0, t_end). to_global: each chunk sum goes to
 * d.csum[(cbase + j) * n + i0 + c]; otherwise their sequential sum goes to
 * L.tll[c]. */
template <int MODE>
__device__ void proposals_loglike(const Dev &d, const SweepLds &L, const TrialView &tv, int k,
                                  int chain_batch, int cmax, int64_t t0, int64_t t_end, int j_lo,
                                  int nj, bool to_global, int64_t cbase, int i0, bool prebuilt = false) {
    const int P = d.P, n = d.nchain;
    for (int c = threadIdx.x; c < k; c += blockDim.x) L.tll[c] = 0.0;
    for (int cb = 0; cb < k; cb += chain_batch) {
        const int kb = min(chain_batch, k - cb);
        if (!(prebuilt && cb == 0)) {
            __syncthreads();
            build_acck(d, L.prop, P, cb, kb, L.acck);
        }
        __syncthreads();
        for (int jb = 0; jb < nj; jb += cmax) {
            const int nbat = min(cmax, nj - jb);
            chunk_sums<MODE>(d, tv, t0, t_end, j_lo + jb, nbat, kb, L.active + cb, L.acck, L.csum, cmax);
            __syncthreads();
            if (to_global) {
                for (int e = threadIdx.x; e < kb * nbat; e += blockDim.x) {
                    const int c = e / nbat, jj = e - c * nbat;
                    d.csum[(size_t)(cbase + j_lo + jb + jj) * n + i0 + cb + c] = L.csum[(size_t)c * cmax + jj];
                }
            } else {
                for (int c = threadIdx.x; c < kb; c += blockDim.x) {
                    double s = L.tll[cb + c];
                    for (int jj = 0; jj < nbat; ++jj) s += L.csum[(size_t)c * cmax + jj];
                    L.tll[cb + c] = s;
                }
            }
            __syncthreads();
        }
    }
}

/* Metropolis accept of proposals c < k for chains i0 + c (src/de.cpp:81-108):
 * state updated in LDS and in global memory. */
__device__ void accept_proposals(const Dev &d, const SweepLds &L, int x, int i0, int k) {
    const int n = d.nchain, P = d.P;
    for (int c = threadIdx.x; c < k; c += blockDim.x) {
        const int i = i0 + c;
        const double cur = L.lp[i] + L.ll[i];
        const bool a = mh_accept(L.tlp[c] + L.tll[c], cur, L.U[c]);
        L.acc[c] = a;
        if (a) {
            L.lp[i] = L.tlp[c];
            L.ll[i] = L.tll[c];
            d.lp[(size_t)x * n + i] = L.tlp[c];
            d.ll[(size_t)x * n + i] = L.tll[c];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < k * P; e += blockDim.x) {
        const int c = e / P;
        if (!L.acc[c]) continue;
        const size_t r = (size_t)(i0 + c) * P + (e - c * P);
        L.th[r] = L.prop[e];
        d.theta[(size_t)x * n * P + r] = L.prop[e];
    }
    __syncthreads();
}

/* Migration ring of subject x over all its trials (src/de.cpp:615-665,
 * :157-199), in one workgroup; state updated in LDS and global memory. */
template <int MODE>
__device__ void subject_migration(const Dev &d, const SweepLds &L, const JobDev &job, int x,
                                  const Stream &st, int para, int cmax, const TrialView &tv) {
    const int n = d.nchain, P = d.P;
    if (threadIdx.x == 0) L.flag[0] = subchains(st, n, L.sub);
    __syncthreads();
    const int ns = L.flag[0];
    const int64_t t0 = d.subj_off[x], t1 = d.subj_off[x + 1];
    const int nj = (int)((t1 - t0 + kChunk - 1) / kChunk);
    for (int k = 0; k < ns; ++k) {
        const int src = L.sub[k], dst = L.sub[(k + 1) % ns];
        const uint32_t slot = kSlotMigrate | (uint32_t)k;
        const int nb = para >= 0 ? 1 : (P + 3) >> 2;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            const U4 r = st.block(slot, 1u + (uint32_t)b);
            const double *xs = L.th + (size_t)src * P;
            if (para >= 0) {
                for (int p = 0; p < P; ++p) L.prop[p] = xs[p];
                L.prop[para] += jitter_word(r.x[0], d.rp);
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int p = 4 * b + q;
                    if (p < P) L.prop[p] = xs[p] + jitter_word(r.x[q], d.rp);
                }
            }
        }
        if (threadIdx.x == 0) {
            const U4 b0 = st.block(slot, 0);
            L.U[0] = u53(b0.x[0], b0.x[1]);
        }
        __syncthreads();
        proposal_priors(d, job, L, src, 1);
        proposals_loglike<MODE>(d, L, tv, 1, 1, cmax, t0, t1, 0, nj, false, 0, 0);
        if (threadIdx.x == 0) {
            if (L.active[0]) atomicAdd(d.evals + x, (unsigned long long)(t1 - t0));
            const double cur = L.lp[dst] + L.ll[dst];
            const bool a = mh_accept(L.tlp[0] + L.tll[0], cur, L.U[0]);
            L.flag[1] = a;
            if (a) {
                L.lp[dst] = L.tlp[0];
                L.ll[dst] = L.tll[0];
                d.lp[(size_t)x * n + dst] = L.tlp[0];
                d.ll[(size_t)x * n + dst] = L.tll[0];
            }
        }
        __syncthreads();
        if (L.flag[1]) {
            for (int p = threadIdx.x; p < P; p += blockDim.x) {
                L.th[(size_t)dst * P + p] = L.prop[p];
                d.theta[((size_t)x * n + dst) * P + p] = L.prop[p];
            }
        }
        __syncthreads();
    }
}

/* Thinned store of subject x's state (src: LDS copy or global state). */
__device__ void store_subject(const Dev &d, const double *th, const double *lp, const double *ll,
                              int x, uint32_t iter) {
    if (iter % (uint32_t)d.thin != 0) return;
    const int k = (int)(iter / (uint32_t)d.thin);
    if (k >= d.nmc) return;
    const int n = d.nchain, P = d.P;
    double *ot = d.out_theta + ((size_t)x * d.nmc + k) * n * P;
    for (int e = threadIdx.x; e < n * P; e += blockDim.x) ot[e] = th[e];
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        d.out_lp[((size_t)x * d.nmc + k) * n + c] = lp[c];
        d.out_ll[((size_t)x * d.nmc + k) * n + c] = ll[c];
    }
}

/* ------------------------------------------------------------ sweep kernel */
template <int MODE>
__global__ __launch_bounds__(kBlock) void subject_sweep_kernel(Dev d, SweepArgs a) {
    extern __shared__ __attribute__((aligned(16))) char lds_raw[];
    const int4 item = a.items[blockIdx.x];
    const int x = item.x;
    const int c0 = item.y >> 16, c1 = item.y & 0xFFFF;
    const int t = item.z >> 16, ntile = item.z & 0xFFFF;
    const int n = d.nchain, P = d.P, G = d.G;
    const SweepLds L = carve(lds_raw, n, P, a.kmax, d.n_cond * d.n_acc, a.chain_batch, a.cmax, a.tstage);
    load_exp2_table(d);
    /* loads first (they depend on the item only), then the draws */
    stage_copy(L.th, d.theta + (size_t)x * n * P, n * P);
    stage_copy(L.lp, d.lp + (size_t)x * n, n);
    stage_copy(L.ll, d.ll + (size_t)x * n, n);
    /* the item's trials; staged in LDS when they fit */
    const int64_t s0 = d.subj_off[x], s1 = d.subj_off[x + 1];
    const int64_t t_lo = ntile > 1 ? s0 + (int64_t)t * d.tpt : s0;
    const int64_t t_hi = ntile > 1 ? min(s1, t_lo + (int64_t)d.tpt) : s1;
    TrialView tv{d.rt, d.tcode, 0};
    if (a.tstage > 0) {
        stage_copy(L.trt, d.rt + t_lo, (int)(t_hi - t_lo));
        stage_copy(L.tcd, d.tcode + t_lo, (int)(t_hi - t_lo));
        tv = TrialView{L.trt, L.tcd, t_lo};
    }
    const JobDev job = d.subj_jobs[x];
    const uint32_t iter = *d.iter_base + a.iter_off;
    const bool single = job.kind != GGDMC_KIND_HIER;
    const bool blocked = d.blocked_sub != 0;
    const uint64_t seed = ((uint64_t)job.k1 << 32) | job.k0;
    const uint32_t unit = (uint32_t)d.subj_unit[x];

    /* level decision (src/de.cpp:208-236 single level, :333-372 subjects) */
    const int dsweep = (blocked && single) ? 0 : a.sweep;
    const Stream sd = make_stream(seed, iter + d.philox_offset, unit, (uint32_t)dsweep);
    const bool migrate = decision_u(sd) < d.sub_mig;
    const int para = blocked ? a.sweep : -1;
    const Stream st = make_stream(seed, iter + d.philox_offset, unit, (uint32_t)a.sweep);
    auto group_k = [&](int g, int &lo, int &hi, int &i0) {
        lo = group_lo(g, n, G);
        hi = group_lo(g + 1, n, G);
        i0 = lo + c0;
        return min(c1 == kAllChains ? hi - lo : c1, hi - lo) - c0;
    };
    if (!migrate) {
        /* stage 1 of the first group overlaps the state loads */
        int lo, hi, i0;
        const int k = group_k(a.g_begin, lo, hi, i0);
        if (k > 0) proposal_stage1(d, job, L, st, lo, hi, i0, k, para);
    }
    __syncthreads();
    /* deferred store: the first launch of iteration i stores iteration i - 1 */
    if (a.store_prev && c0 == 0 && t == 0 && iter > 1) store_subject(d, L.th, L.lp, L.ll, x, iter - 1);

    if (migrate) {
        const bool act = blocked && single ? a.sweep == 0 : true;
        if (c0 == 0 && t == 0 && act && a.g_begin == 0) {
            const int para = (blocked && !single) ? a.sweep : -1;
            subject_migration<MODE>(d, L, job, x, sd, para, a.cmax,
                                    ntile == 1 && a.tstage > 0 ? tv : TrialView{d.rt, d.tcode, 0});
        }
    } else {
        const int j_lo = (int)((t_lo - s0) / kChunk);
        const int nj = (int)((t_hi - t_lo + kChunk - 1) / kChunk);
        for (int g = a.g_begin; g < a.g_end; ++g) {
            int lo, hi, i0;
            const int k = group_k(g, lo, hi, i0);
            if (k <= 0) continue;
            if (g != a.g_begin) {
                proposal_stage1(d, job, L, st, lo, hi, i0, k, para);
                __syncthreads();
            }
            proposal_stage2(d, job, L, i0, k, para);
            __syncthreads();
            const bool built = proposal_stage3(d, L, k, a.chain_batch);
            __syncthreads();
            proposals_loglike<MODE>(d, L, tv, k, a.chain_batch, a.cmax, s0, s1, j_lo, nj, ntile > 1,
                                    d.chunk_off[x], i0, built);
            if (threadIdx.x == 0) {
                int na = 0;
                for (int c = 0; c < k; ++c) na += L.active[c];
                atomicAdd(d.evals + x, (unsigned long long)na * (unsigned long long)(t_hi - t_lo));
            }
            if (ntile > 1) {
                /* the last-arriving tile of (subject, chain range) accepts */
                __threadfence();
                __syncthreads();
                uint32_t *ctr = d.tile_ctr + (blockIdx.x - t);
                if (threadIdx.x == 0) {
                    const uint32_t old = atomicAdd(ctr, 1u);
                    L.flag[3] = old == (uint32_t)(ntile - 1);
                }
                __syncthreads();
                if (!L.flag[3]) continue;
                __threadfence();
                if (threadIdx.x == 0) *ctr = 0u;
                const int64_t cb = d.chunk_off[x], ncb = d.chunk_off[x + 1] - cb;
                for (int c = threadIdx.x; c < k; c += blockDim.x) {
                    double s = 0.0;
                    for (int64_t j = 0; j < ncb; ++j) s += d.csum[(size_t)(cb + j) * n + i0 + c];
                    L.tll[c] = s;
                }
                __syncthreads();
            }
            accept_proposals(d, L, x, i0, k);
        }
    }
}

/* ------------------------------------------------------------- phi level */
struct PhiLds {
    double *ph, *plp, *pll;  /* [n*2P], [n], [n]: population states */
    double *prop;            /* [ngmax*2P] proposals of a chain group */
    double *tm;              /* [(ngmax+1)*4P] prior and likelihood terms */
    double *pc;              /* [n*P*2] log scale, log normaliser of every chain's phi */
    double *pp;              /* [(ngmax+1)*P*2] the same for the proposals */
    double *U;               /* [ngmax] */
    int *m0, *m1, *acc;      /* [ngmax] */
    int *sub;                /* [n] */
    int *flag;               /* [4] */
};

__host__ __device__ inline size_t phi_lds_bytes(int n, int P, int G) {
    const size_t ng = (size_t)(n + G - 1) / G;
    const size_t dbl = (size_t)n * 2 * P + 2 * (size_t)n + ng * 2 * P + (ng + 1) * 4 * P +
                       (size_t)n * P * 2 + (ng + 1) * P * 2 + ng;
    return dbl * sizeof(double) + (3 * ng + n + 4) * sizeof(int);
}

__device__ PhiLds phi_carve(char *base, int n, int P, int G) {
    const int ng = (n + G - 1) / G;
    PhiLds L;
    double *p = reinterpret_cast<double *>(base);
    L.ph = p; p += (size_t)n * 2 * P;
    L.plp = p; p += n;
    L.pll = p; p += n;
    L.prop = p; p += (size_t)ng * 2 * P;
    L.tm = p; p += (size_t)(ng + 1) * 4 * P;
    L.pc = p; p += (size_t)n * P * 2;
    L.pp = p; p += (size_t)(ng + 1) * P * 2;
    L.U = p; p += ng;
    int *q = reinterpret_cast<int *>(p);
    L.m0 = q; q += ng;
    L.m1 = q; q += ng;
    L.acc = q; q += ng;
    L.sub = q; q += n;
    L.flag = q;
    return L;
}

/* One population parameter's share of sumloghlike from its sufficient
 * statistics (mean, M2, min, max of the S subject values):
 *   sum_s log tnorm(x_s; mu, sg, lo, up)
 *     = -(M2 + S (mean - mu)^2) / (2 sg^2) - S (log sqrt(2 pi) + log sg + log Z)
 * NaN for sg <= 0 or NaN inputs, -inf when a subject value is outside
 * [lo, up]; summing the P shares in parameter order gives sumloghlike
 * (src/de.cpp:245-270). */
__device__ double hll_term(const Dev &d, const JobDev &job, int p, double mu, double sg,
                           const double *lc, const double *st, int S) {
    const double mean = st[0], M2 = st[1], mn = st[2], mx = st[3];
    if (!(sg > 0.0) || mean != mean || mu != mu) return __builtin_nan("");
    const double lo = d.pr_lo[job.prior_off + p], up = d.pr_up[job.prior_off + p];
    if (mn < lo || mx > up) return -__builtin_inf();
    const double dm = mean - mu;
    return -(M2 + (double)S * dm * dm) / (2.0 * sg * sg) -
           (double)S * (0.918938533204672741780329736406 + lc[0] + lc[1]);
}

/* lc = (log sg, log Z) of parameter p's truncated normal with location mu
 * and scale sg on the subject prior's bounds (Z its mass). */
__device__ __forceinline__ void phi_log_consts(const Dev &d, const JobDev &job, int p, double mu,
                                               double sg, double *lc) {
    const double lo = d.pr_lo[job.prior_off + p], up = d.pr_up[job.prior_off + p];
    lc[0] = ::log(sg);
    lc[1] = ::log(tnorm_mass((lo - mu) / sg, (up - mu) / sg));
}

__device__ __forceinline__ double hyper_prior_term(const Dev &d, const JobDev &job, int q, double x) {
    const int k = job.hprior_off + q;
    return log_prior1(d.pr_dist[k], d.pr_p0[k], d.pr_p1[k], d.pr_lo[k], d.pr_up[k], x);
}

/* The same from the session's cached logarithms of the fixed prior. */
__device__ __forceinline__ double hyper_prior_term_cached(const Dev &d, const JobDev &job, int q, double x) {
    if (x != x) return x;
    const int k = job.hprior_off + q;
    const double p0 = d.pr_p0[k], p1 = d.pr_p1[k];
    if (d.pr_dist[k] == GGDMC_DIST_UNIF) {
        if (x < p0 || x > p1) return -__builtin_inf();
        return -d.scache[2 * k];
    }
    if (d.pr_dist[k] == GGDMC_DIST_TNORM) {
        if (!(p1 > 0.0)) return __builtin_nan("");
        if (x < d.pr_lo[k] || x > d.pr_up[k]) return -__builtin_inf();
        const double z = (x - p0) / p1;
        return -0.5 * z * z - 0.918938533204672741780329736406 - d.scache[2 * k] - d.scache[2 * k + 1];
    }
    return __builtin_nan("");
}



/* sumlogprior of a population state x (src/de.cpp:426, :518): 2P terms in order. */
__device__ double hyper_logprior(const Dev &d, const JobDev &job, const double *x) {
    double s = 0.0;
    for (int q = 0; q < 2 * d.P; ++q) s += hyper_prior_term(d, job, q, x[q]);
    return s;
}

/* Sufficient statistics (mean, M2, min, max) of the run's S subject values
 * of every (chain, parameter) pair q = chain * P + parameter, in a fixed
 * order: subjects in blocks of kStatBlock (in subject order); a block's values
 * are held in registers, summed in order, and its M2 taken about the block
 * mean; blocks combine in order as
 *   sum = sum_j sum_j,  mean = sum / S,  M2 = sum_j (M2_j + n_j (mean_j - mean)^2)
 * (one block: the two-pass M2). Step 1: block statistics of items
 * e = q * nblk + j in [e0, e1), one thread each, into bst[e][4]. */
__device__ void block_stats(const Dev &d, const JobDev &job, int e0, int e1, double *bst) {
    const int P = d.P, n = d.nchain, S = job.nsubject;
    const int nblk = (S + kStatBlock - 1) / kStatBlock;
    const size_t stride = (size_t)n * P;
    for (int e = e0 + (int)threadIdx.x; e < e1; e += blockDim.x) {
        const int q = e / nblk, j = e - q * nblk;
        const int s0 = j * kStatBlock, nb = min(kStatBlock, S - s0);
        const double *src = d.theta + (size_t)(job.first_subject + s0) * stride + q;
        double v[kStatBlock];
#pragma unroll
        for (int i = 0; i < kStatBlock; ++i) v[i] = i < nb ? src[i * stride] : 0.0;
        double sum = 0.0, mn = __builtin_inf(), mx = -__builtin_inf();
        bool nan = false;
#pragma unroll
        for (int i = 0; i < kStatBlock; ++i) {
            if (i < nb) {
                sum += v[i];
                mn = fmin(mn, v[i]);
                mx = fmax(mx, v[i]);
                nan |= v[i] != v[i];
            }
        }
        const double mean = sum / (double)nb;
        double m2 = 0.0;
#pragma unroll
        for (int i = 0; i < kStatBlock; ++i) {
            if (i < nb) {
                const double dx = v[i] - mean;
                m2 += dx * dx;
            }
        }
        double *o = bst + (size_t)e * 4;
        o[0] = nan ? __builtin_nan("") : sum;
        o[1] = m2;
        o[2] = mn;
        o[3] = mx;
    }
}

/* Step 2: combine the blocks of pairs [q0, q1) in order into out[q][4] =
 * (mean, M2, min, max). */
__device__ void combine_stats(const JobDev &job, int q0, int q1, const double *bst, double *out) {
    const int S = job.nsubject;
    const int nblk = (S + kStatBlock - 1) / kStatBlock;
    for (int q = q0 + (int)threadIdx.x; q < q1; q += blockDim.x) {
        const double *b = bst + (size_t)q * nblk * 4;
        double sum = 0.0, mn = __builtin_inf(), mx = -__builtin_inf();
        for (int j = 0; j < nblk; ++j) {
            sum += b[j * 4];
            mn = fmin(mn, b[j * 4 + 2]);
            mx = fmax(mx, b[j * 4 + 3]);
        }
        const double mean = sum / (double)S;
        double m2 = 0.0;
        for (int j = 0; j < nblk; ++j) {
            const int nb = min(kStatBlock, S - j * kStatBlock);
            const double dm = b[j * 4] / (double)nb - mean;
            m2 += b[j * 4 + 1] + (double)nb * dm * dm;
        }
        double *o = out + (size_t)q * 4;
        o[0] = mean;
        o[1] = m2;
        o[2] = mn;
        o[3] = mx;
    }
}

/* Migration ring of population state f (src/de.cpp:467-565): per ring step the
 * proposal, its prior and likelihood terms, and the refreshed likelihoods of
 * source and target (hierarchical runs, src/de.cpp:494-500) in parallel, the
 * sums and the accept by one thread. */
__device__ void phi_migration(const Dev &d, const JobDev &job, const PhiLds &L, const Stream &sd,
                              int para, const double *stats, bool hier) {
    const int n = d.nchain, P = d.P, P2 = 2 * P, S = job.nsubject;
    if (threadIdx.x == 0) L.flag[0] = subchains(sd, n, L.sub);
    __syncthreads();
    const int ns = L.flag[0];
    for (int k = 0; k < ns; ++k) {
        const int src = L.sub[k], dst = L.sub[(k + 1) % ns];
        const uint32_t slot = kSlotMigrate | (uint32_t)k;
        const double *xs = L.ph + (size_t)src * P2;
        const int nb = para >= 0 ? 1 : (P2 + 3) >> 2;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            const U4 r = sd.block(slot, 1u + (uint32_t)b);
            if (para >= 0) {
                for (int q = 0; q < P2; ++q) L.prop[q] = xs[q];
                L.prop[para] += jitter_word(r.x[0], d.rp);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int q = 4 * b + j;
                    if (q < P2) L.prop[q] = xs[q] + jitter_word(r.x[j], d.rp);
                }
            }
        }
        __syncthreads();
        /* tm: [0, 2P) prior of the proposal, [2P, 3P) its likelihood shares,
         * [3P, 4P) and [4P, 5P) the refreshed shares of src and dst */
        const int nterm = hier ? 5 * P : 3 * P;
        const double *st_src = stats + (size_t)(hier ? src : 0) * P * 4;
        for (int e = threadIdx.x; e < nterm; e += blockDim.x) {
            double v;
            if (e < P2) {
                v = hyper_prior_term_cached(d, job, e, L.prop[e]);
            } else if (e < 3 * P) {
                const int p = e - P2;
                double *lc = L.pp + (size_t)p * 2;
                phi_log_consts(d, job, p, L.prop[p], L.prop[P + p], lc);
                v = hll_term(d, job, p, L.prop[p], L.prop[P + p], lc, st_src + p * 4, S);
            } else {
                const int which = e < 4 * P ? src : dst;
                const int p = e - (e < 4 * P ? 3 * P : 4 * P);
                const double *ph = L.ph + (size_t)which * P2;
                v = hll_term(d, job, p, ph[p], ph[P + p], L.pc + ((size_t)which * P + p) * 2,
                             stats + ((size_t)which * P + p) * 4, S);
            }
            L.tm[e] = v;
        }
        __syncthreads();
        if (threadIdx.x < 8) {
            /* sums by 8 lanes; the refreshed shares, then the accept, by lane 0 */
            const int l = threadIdx.x;
            const double tlp = group8_sum(L.tm, P2, l);
            const double tll = group8_sum(L.tm + P2, P, l);
            const double ra = hier ? group8_sum(L.tm + 3 * P, P, l) : 0.0;
            const double rb = hier ? group8_sum(L.tm + 4 * P, P, l) : 0.0;
            if (l == 0) {
                if (hier) {
                    L.pll[src] = ra;
                    L.pll[dst] = rb;
                }
                const double cur = L.plp[dst] + L.pll[dst];
                const U4 b0 = sd.block(slot, 0);
                const bool a = mh_accept(tlp + tll, cur, u53(b0.x[0], b0.x[1]));
                L.flag[1] = a;
                if (a) {
                    L.plp[dst] = tlp;
                    L.pll[dst] = tll;
                }
            }
        }
        __syncthreads();
        if (L.flag[1]) {
            for (int q = threadIdx.x; q < P2; q += blockDim.x) L.ph[(size_t)dst * P2 + q] = L.prop[q];
            for (int q = threadIdx.x; q < 2 * P; q += blockDim.x) L.pc[(size_t)dst * P * 2 + q] = L.pp[q];
        }
        __syncthreads();
    }
}

/* Crossover of chain group [lo, hi) of a population state (src/de.cpp:385-465):
 * partners per chain, jitter per (chain, Philox block), prior terms per
 * (chain, parameter of 2P), likelihood shares per (chain, parameter) for the
 * proposal and the refreshed current state (src/de.cpp:397), sums per chain. */
__device__ void phi_crossover(const Dev &d, const JobDev &job, const PhiLds &L, const Stream &st,
                              int lo, int hi, int para, const double *stats, bool hier) {
    const int n = d.nchain, P = d.P, P2 = 2 * P, S = job.nsubject, ng = hi - lo;
    /* partners, accept U and jitter: one thread per (chain, Philox block) */
    const int nb = para >= 0 ? 1 : (P2 + 3) >> 2;
    for (int e = threadIdx.x; e < ng * nb; e += blockDim.x) {
        const int c = e / nb, b = e - c * nb;
        const int i = lo + c;
        int m0, m1;
        double U;
        partners(st, i, lo, hi, n, m0, m1, U);
        if (b == 0) L.U[c] = U;
        const double *xi = L.ph + (size_t)i * P2;
        const double *x0 = L.ph + (size_t)m0 * P2, *x1 = L.ph + (size_t)m1 * P2;
        double *x = L.prop + (size_t)c * P2;
        const U4 r = st.block((uint32_t)i, 1u + (uint32_t)b);
        if (para >= 0) {
            for (int q = 0; q < P2; ++q) x[q] = xi[q];
            x[para] += jitter_word(r.x[0], d.rp) + job.gamma * (x0[para] - x1[para]);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int q = 4 * b + j;
                if (q < P2) x[q] = xi[q] + (jitter_word(r.x[j], d.rp) + job.gamma * (x0[q] - x1[q]));
            }
        }
    }
    __syncthreads();
    /* per chain c, tm[c*4P + ...]: [0, 2P) prior, [2P, 3P) proposal shares,
     * [3P, 4P) refreshed current shares (src/de.cpp:397); one loop per kind so
     * that a wave runs one kind of term */
    for (int e = threadIdx.x; e < ng * P2; e += blockDim.x) {
        const int c = e / P2, q = e - c * P2;
        L.tm[(size_t)c * 4 * P + q] = hyper_prior_term_cached(d, job, q, L.prop[e]);
    }
    for (int e = threadIdx.x; e < ng * P; e += blockDim.x) {
        const int c = e / P, p = e - c * P;
        const int i = lo + c;
        const double *x = L.prop + (size_t)c * P2;
        double *lc = L.pp + (size_t)e * 2;
        phi_log_consts(d, job, p, x[p], x[P + p], lc);
        L.tm[(size_t)c * 4 * P + P2 + p] =
            hll_term(d, job, p, x[p], x[P + p], lc, stats + ((size_t)(hier ? i : 0) * P + p) * 4, S);
    }
    if (hier) {
        for (int e = threadIdx.x; e < ng * P; e += blockDim.x) {
            const int c = e / P, p = e - c * P;
            const int i = lo + c;
            const double *cur = L.ph + (size_t)i * P2;
            L.tm[(size_t)c * 4 * P + 3 * P + p] = hll_term(d, job, p, cur[p], cur[P + p],
                                                           L.pc + ((size_t)i * P + p) * 2,
                                                           stats + ((size_t)i * P + p) * 4, S);
        }
    }
    __syncthreads();
    /* sums by 8 lanes per chain, accept by the group's first lane */
    for (int e = threadIdx.x; e < ((ng + 7) & ~7) * 8; e += blockDim.x) {
        const int c = e >> 3, l = e & 7;
        const bool on = c < ng;
        const double *tc = L.tm + (size_t)(on ? c : 0) * 4 * P;
        const double tlp = group8_sum(tc, P2, l);
        const double tll = group8_sum(tc + P2, P, l);
        const double ref = hier ? group8_sum(tc + 3 * P, P, l) : 0.0;
        if (on && l == 0) {
            const int i = lo + c;
            if (hier) L.pll[i] = ref;
            const double cur = L.pll[i] + L.plp[i];
            const bool a = mh_accept(tlp + tll, cur, L.U[c]);
            L.acc[c] = a;
            if (a) {
                L.plp[i] = tlp;
                L.pll[i] = tll;
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ng * P2; e += blockDim.x) {
        const int c = e / P2;
        if (L.acc[c]) {
            L.ph[(size_t)lo * P2 + e] = L.prop[e];
            L.pc[(size_t)lo * P * 2 + e] = L.pp[e];   /* P2 == 2P entries per chain */
        }
    }
    __syncthreads();
}

/* Population step of state f for iteration iter in one workgroup, from the
 * per-chain statistics in d.stats (src/de.cpp:272-383 decisions). */
__device__ void phi_update(const Dev &d, int f, const JobDev &job, uint32_t iter, const PhiLds &L) {
    const int n = d.nchain, P = d.P, P2 = 2 * P, G = d.G;
    const bool hier = job.kind == GGDMC_KIND_HIER;
    const double *stats = d.stats + (size_t)f * n * P * 4;
    for (int e = threadIdx.x; e < n * P2; e += blockDim.x) L.ph[e] = d.phi[(size_t)f * n * P2 + e];
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        L.plp[c] = d.phi_lp[(size_t)f * n + c];
        L.pll[c] = d.phi_ll[(size_t)f * n + c];
    }
    __syncthreads();
    /* log scale / log mass of every chain's subject prior: kept from the last
     * population step, computed afresh after a start or restore */
    const bool have = d.pcache_valid[f] != 0;
    for (int e = threadIdx.x; e < n * P; e += blockDim.x) {
        const int c = e / P, p = e - c * P;
        if (have) {
            L.pc[2 * e] = d.pcache[((size_t)f * n * P + e) * 2];
            L.pc[2 * e + 1] = d.pcache[((size_t)f * n * P + e) * 2 + 1];
        } else {
            phi_log_consts(d, job, p, L.ph[(size_t)c * P2 + p], L.ph[(size_t)c * P2 + P + p],
                           L.pc + (size_t)e * 2);
        }
    }
    __syncthreads();
    const uint32_t unit = hier ? kUnitPhi : 0u;
    const double mig = hier ? d.pop_mig : d.sub_mig;
    const bool blocked = hier ? d.blocked_pop != 0 : d.blocked_sub != 0;
    const uint64_t seed = ((uint64_t)job.k1 << 32) | job.k0;
    const int nsweep = blocked ? P2 : 1;
    for (int sw = 0; sw < nsweep; ++sw) {
        /* run_chains (single level) decides once per iteration; run_hchains
         * per parameter */
        const int dsweep = (blocked && !hier) ? 0 : sw;
        const Stream sd = make_stream(seed, iter + d.philox_offset, unit, (uint32_t)dsweep);
        if (decision_u(sd) < mig) {
            phi_migration(d, job, L, sd, (blocked && hier) ? sw : -1, stats, hier);
            if (!hier && blocked) break;
            continue;
        }
        const Stream st = make_stream(seed, iter + d.philox_offset, unit, (uint32_t)sw);
        for (int g = 0; g < G; ++g)
            phi_crossover(d, job, L, st, group_lo(g, n, G), group_lo(g + 1, n, G), blocked ? sw : -1,
                          stats, hier);
    }
    for (int e = threadIdx.x; e < n * P2; e += blockDim.x) {
        d.phi[(size_t)f * n * P2 + e] = L.ph[e];
        d.pcache[(size_t)f * n * P2 + e] = L.pc[e];   /* n * P * 2 == n * P2 */
    }
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        d.phi_lp[(size_t)f * n + c] = L.plp[c];
        d.phi_ll[(size_t)f * n + c] = L.pll[c];
    }
    if (threadIdx.x == 0) d.pcache_valid[f] = 1;
    if (iter % (uint32_t)d.thin == 0) {
        const int k = (int)(iter / (uint32_t)d.thin);
        if (k < d.nmc) {
            double *ot = d.phi_out_theta + ((size_t)f * d.nmc + k) * n * P2;
            for (int e = threadIdx.x; e < n * P2; e += blockDim.x) ot[e] = L.ph[e];
            for (int c = threadIdx.x; c < n; c += blockDim.x) {
                d.phi_out_lp[((size_t)f * d.nmc + k) * n + c] = L.plp[c];
                d.phi_out_ll[((size_t)f * d.nmc + k) * n + c] = L.pll[c];
            }
        }
    }
}

/* grid = n_phi * nwg: workgroup (f, b) computes the block statistics of items
 * [b * blockDim, (b + 1) * blockDim) of a hierarchical run; the last
 * workgroup of each run combines them and performs the population step.
 * Hyper-only runs have fixed statistics: their workgroup b == 0 steps
 * directly. */
__global__ __launch_bounds__(kPhiBlock) void phi_kernel(Dev d, uint32_t iter_off, int nwg) {
    extern __shared__ __attribute__((aligned(16))) char lds_raw[];
    const int n = d.nchain, P = d.P;
    const int f = blockIdx.x / nwg, b = blockIdx.x - f * nwg;
    const JobDev job = d.phi_jobs[f];
    const uint32_t iter = *d.iter_base + iter_off;
    const PhiLds L = phi_carve(lds_raw, n, P, d.G);
    if (job.kind != GGDMC_KIND_HIER) {
        if (b != 0) return;
        phi_update(d, f, job, iter, L);
        return;
    }
    const int nitem = n * P * ((job.nsubject + kStatBlock - 1) / kStatBlock);
    double *bst = d.bstats + (size_t)f * d.bstats_stride;
    if (b * (int)blockDim.x < nitem)
        block_stats(d, job, b * (int)blockDim.x, min(nitem, (b + 1) * (int)blockDim.x), bst);
    if (nwg > 1) {
        /* the last workgroup of the run continues */
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t old = atomicAdd(d.stats_ctr + f, 1u);
            L.flag[0] = old == (uint32_t)(nwg - 1);
        }
        __syncthreads();
        if (!L.flag[0]) return;
        __threadfence();
        if (threadIdx.x == 0) d.stats_ctr[f] = 0u;
    } else {
        __syncthreads();
    }
    combine_stats(job, 0, n * P, bst, d.stats + (size_t)f * n * P * 4);
    __syncthreads();
    phi_update(d, f, job, iter, L);
}

/* ------------------------------------------------------- plain FP64 path */
/* Step j of the sequential sweep of every subject: crossover of chain j, or
 * migration ring step j (src/de.cpp:567-665, :111-199), reference arithmetic.
 * One workgroup per subject over all its trials. */
__global__ __launch_bounds__(kBlock) void plain_subject_kernel(Dev d, uint32_t iter, int sweep, int j) {
    __shared__ double prop[64];
    __shared__ double red[kWaves];
    __shared__ AccK acck[256];
    __shared__ int sub[kMaxChains];
    __shared__ int info[4];
    __shared__ double misc[2];
    const int s = blockIdx.x;
    const JobDev job = d.jobs[d.subj_job[s]];
    const int n = d.nchain, P = d.P;
    const bool single = job.kind != GGDMC_KIND_HIER;
    const bool blocked = d.blocked_sub != 0;
    const uint64_t seed = ((uint64_t)job.k1 << 32) | job.k0;
    if (!blocked && sweep > 0) return;
    const int dsweep = (blocked && single) ? 0 : sweep;
    const Stream sd = make_stream(seed, iter + d.philox_offset, (uint32_t)d.subj_unit[s], (uint32_t)dsweep);
    const bool migrate = decision_u(sd) < d.sub_mig;
    double *th = d.theta + (size_t)s * n * P;
    if (threadIdx.x == 0) {
        info[0] = 0; /* active */
        if (migrate) {
            const bool act = blocked && single ? sweep == 0 : true;
            const int ns = act ? subchains(sd, n, sub) : 0;
            if (j < ns) {
                const int src = sub[j], dst = sub[(j + 1) % ns];
                const int para = (blocked && !single) ? sweep : -1;
                const uint32_t slot = kSlotMigrate | (uint32_t)j;
                for (int p = 0; p < P; ++p) prop[p] = th[src * P + p];
                if (para >= 0) prop[para] += jitter(sd, slot, 0, d.rp);
                else
                    for (int p = 0; p < P; ++p) prop[p] += jitter(sd, slot, p, d.rp);
                const U4 b = sd.block(slot, 0);
                misc[0] = u53(b.x[0], b.x[1]);
                misc[1] = subject_logprior(d, job, src, prop);
                info[0] = 1;
                info[1] = dst;
            }
        } else {
            const int para = blocked ? sweep : -1;
            const Stream st = make_stream(seed, iter + d.philox_offset, (uint32_t)d.subj_unit[s], (uint32_t)sweep);
            int m0, m1;
            double U;
            partners(st, j, j, j + 1, n, m0, m1, U);
            for (int p = 0; p < P; ++p) prop[p] = th[j * P + p];
            if (para >= 0) {
                prop[para] += jitter(st, (uint32_t)j, 0, d.rp) + job.gamma * (th[m0 * P + para] - th[m1 * P + para]);
            } else {
                for (int p = 0; p < P; ++p)
                    prop[p] += jitter(st, (uint32_t)j, p, d.rp) + job.gamma * (th[m0 * P + p] - th[m1 * P + p]);
            }
            misc[0] = U;
            misc[1] = subject_logprior(d, job, j, prop);
            info[0] = 1;
            info[1] = j;
        }
    }
    __syncthreads();
    if (!info[0]) return;
    for (int e = threadIdx.x; e < d.n_cond * d.n_acc; e += blockDim.x)
        make_acck<DevOps>(prop, d.par_src + e * 6, d.par_const + e * 6, acck[e]);
    __syncthreads();
    const int64_t t_lo = d.subj_off[s], t_hi = d.subj_off[s + 1];
    double acc = 0.0;
    for (int64_t i = t_lo + threadIdx.x; i < t_hi; i += blockDim.x) {
        const int code = d.tcode[i];
        acc += trial_term<MODE_NAIVE>(acck + (code >> 3) * d.n_acc, d.n_acc, code & 7, d.rt[i]);
    }
    const double tll = block_sum(acc, red);
    if (threadIdx.x == 0) {
        const int dst = info[1];
        atomicAdd(d.evals + s, (unsigned long long)(t_hi - t_lo));
        const double cur = d.lp[(size_t)s * n + dst] + d.ll[(size_t)s * n + dst];
        if (mh_accept(misc[1] + tll, cur, misc[0])) {
            for (int p = 0; p < P; ++p) th[dst * P + p] = prop[p];
            d.lp[(size_t)s * n + dst] = misc[1];
            d.ll[(size_t)s * n + dst] = tll;
        }
    }
}

/* Naive sumloghlike over subject states of chain c (src/de.cpp:245-270). */
__device__ double plain_hyper_ll(const Dev &d, const JobDev &job, const double *phi, int c,
                                 const double *hyper_rows, double *red) {
    const int P = d.P, n = d.nchain, S = job.nsubject;
    double acc = 0.0;
    for (int e = threadIdx.x; e < S * P; e += blockDim.x) {
        const int s = e / P, p = e - s * P;
        const double x = hyper_rows ? hyper_rows[(size_t)s * P + p]
                                    : d.theta[((size_t)(job.first_subject + s) * n + c) * P + p];
        acc += log_prior1(GGDMC_DIST_TNORM, phi[p], phi[P + p], d.pr_lo[job.prior_off + p],
                          d.pr_up[job.prior_off + p], x);
    }
    return block_sum(acc, red);
}

/* Step j of the population sweep (src/de.cpp:385-565), reference order;
 * one workgroup per population state. hyper_rows: [S x P] data of hyper-only
 * runs (null for hierarchical runs). */
__global__ __launch_bounds__(kBlock) void plain_phi_kernel(Dev d, uint32_t iter, int sweep, int j,
                                                          const double *hyper_rows,
                                                          const int64_t *hyper_off) {
    __shared__ double prop[128];
    __shared__ double red[kWaves];
    __shared__ int sub[kMaxChains];
    __shared__ int info[4];
    __shared__ double misc[4];
    const int f = blockIdx.x;
    const JobDev job = d.jobs[d.phi_job[f]];
    const int n = d.nchain, P = d.P, P2 = 2 * P;
    const bool hier = job.kind == GGDMC_KIND_HIER;
    const bool blocked = hier ? d.blocked_pop != 0 : d.blocked_sub != 0;
    const double mig = hier ? d.pop_mig : d.sub_mig;
    const uint64_t seed = ((uint64_t)job.k1 << 32) | job.k0;
    const uint32_t unit = hier ? kUnitPhi : 0u;
    if (!blocked && sweep > 0) return;
    const int dsweep = (blocked && !hier) ? 0 : sweep;
    const Stream sd = make_stream(seed, iter + d.philox_offset, unit, (uint32_t)dsweep);
    const bool migrate = decision_u(sd) < mig;
    double *ph = d.phi + (size_t)f * n * P2;
    double *plp = d.phi_lp + (size_t)f * n, *pll = d.phi_ll + (size_t)f * n;
    const double *rows = hier ? nullptr : hyper_rows + hyper_off[f];
    if (threadIdx.x == 0) {
        info[0] = 0;
        if (migrate) {
            const bool act = (blocked && !hier) ? sweep == 0 : true;
            const int ns = act ? subchains(sd, n, sub) : 0;
            if (j < ns) {
                info[0] = 2;
                info[1] = sub[j];
                info[2] = sub[(j + 1) % ns];
            }
        } else {
            info[0] = 1;
            info[1] = j;
            info[2] = j;
        }
    }
    __syncthreads();
    if (!info[0]) return;
    const int src = info[1], dst = info[2];
    if (hier) {
        /* refresh the cached likelihoods (src/de.cpp:397-398, :494-500) */
        const double a = plain_hyper_ll(d, job, ph + (size_t)src * P2, src, nullptr, red);
        if (threadIdx.x == 0) pll[src] = a;
        if (info[0] == 2) {
            const double b = plain_hyper_ll(d, job, ph + (size_t)dst * P2, dst, nullptr, red);
            if (threadIdx.x == 0) pll[dst] = b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int para = blocked && (hier || info[0] == 1) ? sweep : -1;
        if (info[0] == 2) {
            const uint32_t slot = kSlotMigrate | (uint32_t)j;
            for (int p = 0; p < P2; ++p) prop[p] = ph[(size_t)src * P2 + p];
            const int mp = (blocked && hier) ? sweep : -1;
            if (mp >= 0) prop[mp] += jitter(sd, slot, 0, d.rp);
            else
                for (int p = 0; p < P2; ++p) prop[p] += jitter(sd, slot, p, d.rp);
            const U4 b = sd.block(slot, 0);
            misc[0] = u53(b.x[0], b.x[1]);
        } else {
            const Stream st = make_stream(seed, iter + d.philox_offset, unit, (uint32_t)sweep);
            int m0, m1;
            double U;
            partners(st, j, j, j + 1, n, m0, m1, U);
            for (int p = 0; p < P2; ++p) prop[p] = ph[(size_t)j * P2 + p];
            if (para >= 0) {
                prop[para] += jitter(st, (uint32_t)j, 0, d.rp) +
                              job.gamma * (ph[(size_t)m0 * P2 + para] - ph[(size_t)m1 * P2 + para]);
            } else {
                for (int p = 0; p < P2; ++p)
                    prop[p] += jitter(st, (uint32_t)j, p, d.rp) +
                               job.gamma * (ph[(size_t)m0 * P2 + p] - ph[(size_t)m1 * P2 + p]);
            }
            misc[0] = U;
        }
        misc[1] = hyper_logprior(d, job, prop);
    }
    __syncthreads();
    const double tll = plain_hyper_ll(d, job, prop, src, rows, red);
    if (threadIdx.x == 0) {
        const double cur = plp[dst] + pll[dst];
        if (mh_accept(misc[1] + tll, cur, misc[0])) {
            for (int p = 0; p < P2; ++p) ph[(size_t)dst * P2 + p] = prop[p];
            plp[dst] = misc[1];
            pll[dst] = tll;
        }
    }
}

/* Thinned store of every subject and (unless subjects_only) population
 * state: the plain path after each iteration, the batched path for the last
 * iteration of an advance (whose subject store is otherwise deferred). */
__global__ void store_kernel(Dev d, uint32_t iter, int subjects_only) {
    if (iter % (uint32_t)d.thin != 0) return;
    const int k = (int)(iter / (uint32_t)d.thin);
    if (k >= d.nmc) return;
    const int n = d.nchain, P = d.P;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t e = tid; e < (size_t)d.S * n * P; e += stride) {
        const size_t s = e / ((size_t)n * P), r = e - s * n * P;
        d.out_theta[(s * d.nmc + k) * n * P + r] = d.theta[e];
    }
    for (size_t e = tid; e < (size_t)d.S * n; e += stride) {
        const size_t s = e / n, r = e - s * n;
        d.out_lp[(s * d.nmc + k) * n + r] = d.lp[e];
        d.out_ll[(s * d.nmc + k) * n + r] = d.ll[e];
    }
    if (subjects_only) return;
    const int P2 = 2 * P;
    for (size_t e = tid; e < (size_t)d.n_phi * n * P2; e += stride) {
        const size_t f = e / ((size_t)n * P2), r = e - f * n * P2;
        d.phi_out_theta[(f * d.nmc + k) * n * P2 + r] = d.phi[e];
    }
    for (size_t e = tid; e < (size_t)d.n_phi * n; e += stride) {
        const size_t f = e / n, r = e - f * n;
        d.phi_out_lp[(f * d.nmc + k) * n + r] = d.phi_lp[e];
        d.phi_out_ll[(f * d.nmc + k) * n + r] = d.phi_ll[e];
    }
}

/* Fixed statistics of hyper-only data rows [S x P] for population state f
 * (the data of run_hyper do not change, so they are reduced once): one
 * thread per parameter, rows in order, two passes. */
__global__ void hyper_data_stats_kernel(Dev d, const double *rows, int S, int f) {
    const int P = d.P;
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        double sum = 0.0, mn = __builtin_inf(), mx = -__builtin_inf();
        bool nan = false;
        for (int s = 0; s < S; ++s) {
            const double x = rows[(size_t)s * P + p];
            sum += x;
            mn = fmin(mn, x);
            mx = fmax(mx, x);
            nan |= x != x;
        }
        const double mean = sum / (double)S;
        double m2 = 0.0;
        for (int s = 0; s < S; ++s) {
            const double dx = rows[(size_t)s * P + p] - mean;
            m2 += dx * dx;
        }
        double *o = d.stats + (size_t)f * d.nchain * P * 4 + p * 4;
        o[0] = nan ? __builtin_nan("") : mean;
        o[1] = m2;
        o[2] = mn;
        o[3] = mx;
    }
}

/* ----------------------------------------------------------- sumloglike */
/* out[q] = log-likelihood of thetas[q] over subject s's trials, with the
 * sampler's canonical chunk sums (bit-identical to the sampler's values). */
template <int MODE>
__global__ __launch_bounds__(kBlock) void loglike_kernel(Dev d, const double *thetas, int s, int cmax,
                                                        double *out) {
    extern __shared__ __attribute__((aligned(16))) char lds_raw[];
    const int P = d.P;
    const SweepLds L = carve(lds_raw, 1, P, 1, d.n_cond * d.n_acc, 1, cmax, 0);
    const int q = blockIdx.x;
    load_exp2_table(d);
    for (int p = threadIdx.x; p < P; p += blockDim.x) L.prop[p] = thetas[(size_t)q * P + p];
    if (threadIdx.x == 0) L.active[0] = 1;
    __syncthreads();
    const int64_t t0 = d.subj_off[s], t1 = d.subj_off[s + 1];
    const int nj = (int)((t1 - t0 + kChunk - 1) / kChunk);
    proposals_loglike<MODE>(d, L, TrialView{d.rt, d.tcode, 0}, 1, 1, cmax, t0, t1, 0, nj, false, 0, 0);
    if (threadIdx.x == 0) out[q] = L.tll[0];
}

} // namespace ggdmc

/* ------------------------------------------------------ launch wrappers */
using namespace ggdmc;

extern "C" {

size_t ggdmc_internal_sweep_lds(int n, int P, int kmax, int per_chain, int chain_batch, int cmax,
                                int tstage) {
    return sweep_lds_bytes(n, P, kmax, per_chain, chain_batch, cmax, tstage);
}
size_t ggdmc_internal_phi_lds(int n, int P, int G) { return phi_lds_bytes(n, P, G); }

hipError_t ggdmc_internal_launch_sweep(const Dev *d, const SweepArgs *a, int mode, size_t lds,
                                       hipStream_t stream) {
    const dim3 grid(a->n_items), block(kBlock);
    if (mode == GGDMC_MIXED_BATCHED)
        subject_sweep_kernel<MODE_MIXED><<<grid, block, lds, stream>>>(*d, *a);
    else
        subject_sweep_kernel<MODE_NAIVE><<<grid, block, lds, stream>>>(*d, *a);
    return hipGetLastError();
}

hipError_t ggdmc_internal_launch_phi(const Dev *d, uint32_t iter_off, size_t lds, hipStream_t stream) {
    /* items of the largest run: pairs x subject blocks */
    const int nwg = (int)((d->bstats_stride / 4 + kPhiBlock - 1) / kPhiBlock);
    phi_kernel<<<dim3(d->n_phi * std::max(nwg, 1)), dim3(kPhiBlock), lds, stream>>>(*d, iter_off,
                                                                                   std::max(nwg, 1));
    return hipGetLastError();
}

hipError_t ggdmc_internal_launch_plain_subject(const Dev *d, uint32_t iter, int sweep, int j,
                                               hipStream_t stream) {
    plain_subject_kernel<<<dim3(d->S), dim3(kBlock), 0, stream>>>(*d, iter, sweep, j);
    return hipGetLastError();
}

hipError_t ggdmc_internal_launch_plain_phi(const Dev *d, uint32_t iter, int sweep, int j,
                                           const double *rows, const int64_t *rows_off,
                                           hipStream_t stream) {
    plain_phi_kernel<<<dim3(d->n_phi), dim3(kBlock), 0, stream>>>(*d, iter, sweep, j, rows, rows_off);
    return hipGetLastError();
}

hipError_t ggdmc_internal_launch_store(const Dev *d, uint32_t iter, int subjects_only,
                                      hipStream_t stream) {
    store_kernel<<<dim3(256), dim3(kBlock), 0, stream>>>(*d, iter, subjects_only);
    return hipGetLastError();
}

hipError_t ggdmc_internal_launch_hyper_stats(const Dev *d, const double *rows, int S, int f,
                                             hipStream_t stream) {
    hyper_data_stats_kernel<<<dim3(1), dim3(kBlock), 0, stream>>>(*d, rows, S, f);
    return hipGetLastError();
}

hipError_t ggdmc_internal_launch_loglike(const Dev *d, const double *thetas, int nq, int s,
                                         int mode, int cmax, size_t lds, hipStream_t stream,
                                         double *out) {
    if (mode == GGDMC_MIXED_BATCHED)
        loglike_kernel<MODE_MIXED><<<dim3(nq), dim3(kBlock), lds, stream>>>(*d, thetas, s, cmax, out);
    else
        loglike_kernel<MODE_NAIVE><<<dim3(nq), dim3(kBlock), lds, stream>>>(*d, thetas, s, cmax, out);
    return hipGetLastError();
}

} // extern "C"
