/*
 * HIP kernels of the DE-MCMC sampler (gfx950, wave64).
 *
 *  subject_sweep_kernel  one workgroup per work item (subject, chain range of a
 *                        chain group, trial tile): DE crossover proposals
 *                        (src/de.cpp:567-613 / :111-155), LBA log-likelihood as
 *                        canonical chunk sums, Metropolis accept (:81-108) by the
 *                        item itself or by the last-arriving tile; migration ring
 *                        (src/de.cpp:615-665 / :157-199) and the deferred
 *                        thinned store of the previous iteration, every item
 *                        storing the chains it updates. Leading workgroups run
 *                        the population step of the launch's chains; with
 *                        every chain group in one launch later groups wait for
 *                        earlier ones, and with several iterations per launch
 *                        (MULTI) each iteration's segment waits for the
 *                        previous one, subject by subject.
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
#include "fp64_math.h"
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
    /* the library's own FP64 functions (fp64_math.h), no math-library code */
    static __device__ __forceinline__ double qtail(double a) { return qtail_acc(a); }
    /* R's pnorm: the lower tail accurately, above 0 as 1 - upper tail */
    static __device__ __forceinline__ double pnorm(double z) { return pnorm_acc(z); }
    static __device__ __forceinline__ double dnorm(double z) { return kInvSqrt2Pi * exp_acc(-0.5 * z * z); }
    static __device__ __forceinline__ double log(double x) { return log_acc(x); }
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

/* The terms a lane adds to its chunk sum: trials i1 and i1 + 64 of the
 * chunk (zero past t_end). MODE_MIXED: ill1 / ill2 report trials whose mixed
 * evaluation is ill-conditioned, and fix: those trials are evaluated with the
 * reference formula instead (the fix-up pass). */
template <int MODE, bool FIX>
__device__ __forceinline__ void lane_terms(const Dev &d, const TrialView &tv, const AccK *kc, int64_t base, int64_t t_end,
                                           double &a, double &b, bool &ill1, bool &ill2) {
    const int lane = threadIdx.x & 63;
    const int64_t i1 = base + lane, i2 = i1 + 64;
    const bool h1 = i1 < t_end, h2 = i2 < t_end;
    /* lanes past the end evaluate the chunk's first trial and drop it, so
     * that both evaluations are straight-line code */
    const int64_t r1 = (h1 ? i1 : base) - tv.origin;
    const int64_t r2 = (h2 ? i2 : base) - tv.origin;
    const int code1 = tv.code[r1], code2 = tv.code[r2];
    const double rt1 = tv.rt[r1], rt2 = tv.rt[r2];
    const AccK *k1 = kc + (code1 >> 3) * d.n_acc, *k2 = kc + (code2 >> 3) * d.n_acc;
    a = b = 0.0;
    ill1 = ill2 = false;
    if (MODE == MODE_MIXED) {
        double ma, mb;
        mixed_trial_term_flagged2<DevOps>(k1, code1 & 7, rt1, k2, code2 & 7, rt2, d.n_acc, ma, mb, ill1, ill2);
        ill1 = h1 && ill1;
        ill2 = h2 && ill2;
        a = h1 ? ma : 0.0;
        b = h2 ? mb : 0.0;
        if constexpr (FIX) {
            if (ill1) a = naive_trial_term<DevOps>(k1, d.n_acc, code1 & 7, rt1);
            if (ill2) b = naive_trial_term<DevOps>(k2, d.n_acc, code2 & 7, rt2);
        }
    } else {
        if (h1) a = naive_trial_term<DevOps>(k1, d.n_acc, code1 & 7, rt1);
        if (h2) b = naive_trial_term<DevOps>(k2, d.n_acc, code2 & 7, rt2);
    }
}

/* Canonical chunk sums (see engine.h): C[c * c_stride + jj] for chains c < kb
 * and chunks j_lo + jj < j_lo + nj of the trials [t0, t_end). A wave takes one
 * (chain, chunk) unit at a time; chains with active[c] == 0 get 0.
 * MODE_MIXED: a unit holding an ill-conditioned trial is marked in pend (LDS
 * bits, one per unit) and summed by a second pass after the others, which
 * evaluates its flagged trials with the reference formula: the same per-lane
 * terms, hence the same bits, as evaluating them in place, without the
 * reference formula's code (and registers) inside the main loop. */
template <int MODE>
__device__ __forceinline__ void chunk_sums(const Dev &d, const TrialView &tv, int64_t t0, int64_t t_end, int j_lo, int nj,
                                           int kb, const int *active, const double *acck, double *C, int c_stride,
                                           uint32_t *pend) {
    static_assert(kChunk == 128, "two trials per lane");
    const int lane = threadIdx.x & 63;
    /* the wave index is uniform: keep the unit loop scalar */
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nw = blockDim.x >> 6;
    const int per_chain = d.n_cond * d.n_acc;
    const int nu = kb * nj;
    if (MODE == MODE_MIXED) {
        for (int w = threadIdx.x; w < (nu + 31) / 32; w += blockDim.x) pend[w] = 0u;
        __syncthreads();
    }
    for (int u = wave; u < nu; u += nw) {
        const int c = u / nj, jj = u - c * nj;
        double v = 0.0;
        if (active[c]) {
            const AccK *kc = reinterpret_cast<const AccK *>(acck) + (size_t)c * per_chain;
            double a, b;
            bool ill1, ill2;
            lane_terms<MODE, false>(d, tv, kc, t0 + (int64_t)(j_lo + jj) * kChunk, t_end, a, b, ill1, ill2);
            if (MODE == MODE_MIXED && __builtin_expect(__builtin_amdgcn_ballot_w64(ill1 || ill2) != 0, 0)) {
                if (lane == 0) atomicOr(pend + (u >> 5), 1u << (u & 31));
                continue;
            }
            v = a + b;
        }
        v = wave_sum(v);
        if (lane == 0) C[(size_t)c * c_stride + jj] = v;
    }
    if (MODE == MODE_MIXED) {
        __syncthreads();
        for (int u = wave; u < nu; u += nw) {
            if (__builtin_expect(((pend[u >> 5] >> (u & 31)) & 1u) == 0, 1)) continue;
            const int c = u / nj, jj = u - c * nj;
            const AccK *kc = reinterpret_cast<const AccK *>(acck) + (size_t)c * per_chain;
            double a, b;
            bool ill1, ill2;
            lane_terms<MODE, true>(d, tv, kc, t0 + (int64_t)(j_lo + jj) * kChunk, t_end, a, b, ill1, ill2);
            const double v = wave_sum(a + b);
            if (lane == 0) C[(size_t)c * c_stride + jj] = v;
        }
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
        return -log_acc(p1 - p0);
    }
    if (dist == GGDMC_DIST_TNORM) {
        if (!(p1 > 0.0)) return __builtin_nan("");
        if (x < lo || x > up) return -__builtin_inf();
        const double z = (x - p0) / p1;
        const double Z = tnorm_mass((lo - p0) / p1, (up - p0) / p1);
        return -0.5 * z * z - 0.918938533204672741780329736406 - log_acc(p1) - log_acc(Z);
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
    const double ratio = exp_acc(tpost - cur);
    if (ratio != ratio) return false;
    return U < ratio;
}

/* ------------------------------------------------------------ decision trace
 * With a debug flag set, every Metropolis decision of a traced level is
 * appended to d.trace (at most d.trace_cap records); the host orders them. */
__device__ __forceinline__ bool subject_traced(const Dev &d, int x) {
    return d.trace != nullptr && d.trace_sub &&
           (d.subj_jobs[x].kind != GGDMC_KIND_HIER || d.subj_unit[x] <= 2);
}

__device__ __forceinline__ bool phi_traced(const Dev &d, const JobDev &job) {
    return d.trace != nullptr && (job.kind == GGDMC_KIND_HIER ? d.trace_pop : d.trace_sub);
}

__device__ __forceinline__ void trace_put(const Dev &d, int job, int subject, const Stream &st, int kind,
                                       int step, int chain, int target, double tlp, double tll,
                                       double cur, bool acc) {
    const uint32_t k = atomicAdd(d.trace_n, 1u);
    if (k >= d.trace_cap) return;
    const double diff = tlp + tll - cur;
    ggdmc_trace_record r;
    r.job = job;
    r.subject = subject;
    r.iteration = st.iter - d.philox_offset;
    r.sweep = (int32_t)st.sweep;
    r.kind = kind;
    r.step = step;
    r.chain = chain;
    r.target = target;
    r.tmp_lp = tlp;
    r.tmp_ll = tll;
    r.cur_post = cur;
    r.accepted = acc ? 1 : (diff != diff ? -1 : 0);
    r.reserved = 0;
    d.trace[k] = r;
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
    uint32_t *pend;         /* [(kmax*cmax + 31)/32] units of a chunk-sum batch to fix up */
    double *pcst;           /* [per_chain*6] design constants (par_const) */
    int32_t *psrc;          /* [per_chain*6] design sources (par_src) */
};

__host__ __device__ inline size_t sweep_lds_bytes(int n, int P, int kmax, int per_chain,
                                                  int chain_batch, int cmax, int tstage) {
    const size_t dbl = (size_t)n * P + 2 * (size_t)n + 7 * (size_t)kmax * P + 2 * (size_t)P +
                       3 * (size_t)kmax + (size_t)chain_batch * per_chain * kAccKDoubles +
                       (size_t)kmax * cmax;
    const size_t ints = 4 * (size_t)kmax + n + P + 4 + (size_t)per_chain * 6 + ((size_t)kmax * cmax + 31) / 32;
    return (dbl + (size_t)tstage + (size_t)per_chain * 6) * sizeof(double) + ints * sizeof(int) +
           (((size_t)tstage * sizeof(uint16_t) + 7) & ~(size_t)7);
}

__device__ __forceinline__ SweepLds carve(char *base, int n, int P, int kmax, int per_chain, int chain_batch,
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
    L.pcst = p; p += (size_t)per_chain * 6;
    L.tcd = reinterpret_cast<uint16_t *>(p);
    p += ((size_t)tstage * sizeof(uint16_t) + 7) / 8;
    int *q = reinterpret_cast<int *>(p);
    L.m0 = q; q += kmax;
    L.m1 = q; q += kmax;
    L.active = q; q += kmax;
    L.acc = q; q += kmax;
    L.sub = q; q += n;
    L.pdist = q; q += P;
    L.flag = q; q += 4;
    L.psrc = q; q += per_chain * 6;
    L.pend = reinterpret_cast<uint32_t *>(q);
    return L;
}

/* dst[i] = src[i], start <= i < count, by the whole workgroup with four
 * loads in flight per thread (global -> LDS staging). Loads use clamped
 * indices, so no branch separates them. */
/* *p, or with AG a device-scope load: data written earlier in the same launch
 * by a workgroup on another XCD (several iterations per launch). */
template <bool AG, class T>
__device__ __forceinline__ T ld(const T *p) {
    if constexpr (AG) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
}

template <class T, bool AG = false>
__device__ __forceinline__ void stage_copy(T *dst, const T *src, int count, int start = 0) {
    const int bs = blockDim.x;
    for (int i0 = start + (int)threadIdx.x; i0 < count; i0 += 4 * bs) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = ld<AG>(src + min(i0 + u * bs, count - 1));
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * bs;
            if (i < count) dst[i] = v[u];
        }
    }
}

/* One array of a multi-array staging step: load() issues this thread's
 * first U elements (clamped indices, no branches), store() writes them and
 * copies any elements beyond U * blockDim. Calling every piece's load()
 * before any store() keeps all of a thread's loads in flight together, so
 * the step costs one memory latency however many arrays it stages. */
template <class T, int U, bool AG = false>
struct Piece {
    T *dst;
    const T *src;
    int count;
    T v[U];
    __device__ __forceinline__ void load() {
        if (count <= 0) return;
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = ld<AG>(src + min((int)threadIdx.x + u * (int)blockDim.x, count - 1));
    }
    __device__ __forceinline__ void store() {
        if (count <= 0) return;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = (int)threadIdx.x + u * (int)blockDim.x;
            if (i < count) dst[i] = v[u];
        }
        if (count > U * (int)blockDim.x) stage_copy<T, AG>(dst, src, count, U * (int)blockDim.x);
    }
};

/* Fills the accumulator constants of proposals [cb, cb+kb). */
__device__ __forceinline__ void build_acck(const Dev &d, const SweepLds &L, const double *prop, int P, int cb, int kb,
                                           double *acck) {
    const int per_chain = d.n_cond * d.n_acc;
    AccK *K = reinterpret_cast<AccK *>(acck);
    for (int e = threadIdx.x; e < kb * per_chain; e += blockDim.x) {
        const int c = e / per_chain, r = e - c * per_chain;
        make_acck<DevOps>(prop + (size_t)(cb + c) * P, L.psrc + r * 6, L.pcst + r * 6, K[e]);
    }
}

/* Prior terms of proposals L.prop[0..k) for chains i0 + c, one thread per
 * (chain, parameter), summed per chain in parameter order into L.tlp. */
__device__ __forceinline__ void proposal_priors(const Dev &d, const JobDev &job, const SweepLds &L, int i0, int k) {
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
__device__ __forceinline__ void proposal_stage1(const Dev &d, const JobDev &job, const SweepLds &L, const Stream &st,
                                int lo, int hi, int i0, int k, int para, bool late) {
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
    /* late: the population step is still running, the priors come after the
     * likelihoods (late_prior) */
    for (int e = threadIdx.x; e < (late ? 0 : k * P); e += blockDim.x) {
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

__device__ __forceinline__ void proposal_stage2(const Dev &d, const JobDev &job, const SweepLds &L, int i0, int k,
                                int para, bool late) {
    const int P = d.P;
    for (int e = threadIdx.x; e < k * P; e += blockDim.x) {
        const int c = e / P, p = e - c * P;
        const double xi = L.th[(size_t)(i0 + c) * P + p];
        double x = xi;
        if (para < 0 || p == para)
            x = xi + (L.jit[e] + job.gamma * (L.th[(size_t)L.m0[c] * P + p] - L.th[(size_t)L.m1[c] * P + p]));
        L.prop[e] = x;
        /* late: only the support (the prior's other factors come later and
         * can only turn a finite term into NaN, which rejects as well) */
        L.tm[e] = late ? (x != x ? x : (x < L.bnd[2 * p] || x > L.bnd[2 * p + 1] ? -__builtin_inf() : 0.0))
                       : staged_prior_term(L.pdist[p], L.pri + (size_t)e * 4, L.bnd[2 * p], L.bnd[2 * p + 1], x);
    }
}

/* Returns whether the likelihood constants of all k chains were built. */
__device__ __forceinline__ bool proposal_stage3(const Dev &d, const SweepLds &L, int k, int chain_batch) {
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
    if (all) build_acck(d, L, L.prop, P, 0, k, L.acck);
    return all;
}

/* Log-likelihood of proposals L.prop[0..k) over chunks [j_lo, j_lo + nj) of
 * the subject's trials [t0, t_end). to_global: each chunk sum goes to
 * d.csum[(cbase + j) * n + i0 + c]; otherwise their sequential sum goes to
 * L.tll[c]. */
template <int MODE>
__device__ __forceinline__ void proposals_loglike(const Dev &d, const SweepLds &L, const TrialView &tv, int k,
                                  int chain_batch, int cmax, int64_t t0, int64_t t_end, int j_lo,
                                  int nj, bool to_global, int64_t cbase, int i0, bool prebuilt = false) {
    const int P = d.P, n = d.nchain;
    for (int c = threadIdx.x; c < k; c += blockDim.x) L.tll[c] = 0.0;
    for (int cb = 0; cb < k; cb += chain_batch) {
        const int kb = min(chain_batch, k - cb);
        if (!(prebuilt && cb == 0)) {
            __syncthreads();
            build_acck(d, L, L.prop, P, cb, kb, L.acck);
        }
        __syncthreads();
        for (int jb = 0; jb < nj; jb += cmax) {
            const int nbat = min(cmax, nj - jb);
            chunk_sums<MODE>(d, tv, t0, t_end, j_lo + jb, nbat, kb, L.active + cb, L.acck, L.csum, cmax, L.pend);
            __syncthreads();
            if (to_global) {
                for (int e = threadIdx.x; e < kb * nbat; e += blockDim.x) {
                    const int c = e / nbat, jj = e - c * nbat;
                    __hip_atomic_store(d.csum + (size_t)(cbase + j_lo + jb + jj) * n + i0 + cb + c,
                                       L.csum[(size_t)c * cmax + jj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
__device__ __forceinline__ void accept_proposals(const Dev &d, const SweepLds &L, int x, int i0, int k,
                                                 const Stream &st) {
    const int n = d.nchain, P = d.P;
    const bool traced = subject_traced(d, x);
    for (int c = threadIdx.x; c < k; c += blockDim.x) {
        const int i = i0 + c;
        const double cur = L.lp[i] + L.ll[i];
        const bool a = mh_accept(L.tlp[c] + L.tll[c], cur, L.U[c]);
        if (__builtin_expect(traced, 0))
            trace_put(d, d.subj_job[x], d.subj_unit[x], st, 0, i, i, i, L.tlp[c], L.tll[c], cur, a);
        L.acc[c] = a;
        if (a) {
            L.lp[i] = L.tlp[c];
            L.ll[i] = L.tll[c];
            __hip_atomic_store(d.lp + (size_t)x * n + i, L.tlp[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(d.ll + (size_t)x * n + i, L.tll[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    /* device-scope stores: a later group's item of this subject (or, with
     * several iterations per launch, a later iteration's) may read the rows
     * in the same launch */
    for (int e = threadIdx.x; e < k * P; e += blockDim.x) {
        const int c = e / P;
        if (!L.acc[c]) continue;
        const size_t r = (size_t)(i0 + c) * P + (e - c * P);
        L.th[r] = L.prop[e];
        __hip_atomic_store(d.theta + (size_t)x * n * P + r, L.prop[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
}

/* Migration ring of subject x over all its trials (src/de.cpp:615-665,
 * :157-199), in one workgroup; state updated in LDS and global memory. */
template <int MODE>
__device__ __forceinline__ void subject_migration(const Dev &d, const SweepLds &L, const JobDev &job, int x,
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
            if (__builtin_expect(subject_traced(d, x), 0))
                trace_put(d, d.subj_job[x], d.subj_unit[x], st, 1, k, src, dst, L.tlp[0], L.tll[0], cur, a);
            L.flag[1] = a;
            if (a) {
                L.lp[dst] = L.tlp[0];
                L.ll[dst] = L.tll[0];
                __hip_atomic_store(d.lp + (size_t)x * n + dst, L.tlp[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(d.ll + (size_t)x * n + dst, L.tll[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads();
        if (L.flag[1]) {
            for (int p = threadIdx.x; p < P; p += blockDim.x) {
                L.th[(size_t)dst * P + p] = L.prop[p];
                __hip_atomic_store(d.theta + ((size_t)x * n + dst) * P + p, L.prop[p], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads();
    }
}

/* Thinned store of chains [i0, i0 + k) of subject x (th, lp, ll: the whole
 * subject's rows, in LDS). */
__device__ __forceinline__ void store_chains(const Dev &d, const double *th, const double *lp, const double *ll,
                                             int x, uint32_t iter, int i0, int k) {
    if (iter % (uint32_t)d.thin != 0) return;
    const int col = (int)(iter / (uint32_t)d.thin);
    if (col >= d.nmc || k <= 0) return;
    const int n = d.nchain, P = d.P;
    double *ot = d.out_theta + ((size_t)x * d.nmc + col) * n * P;
    for (int e = threadIdx.x; e < k * P; e += blockDim.x) ot[(size_t)i0 * P + e] = th[(size_t)i0 * P + e];
    for (int c = threadIdx.x; c < k; c += blockDim.x) {
        d.out_lp[((size_t)x * d.nmc + col) * n + i0 + c] = lp[i0 + c];
        d.out_ll[((size_t)x * d.nmc + col) * n + i0 + c] = ll[i0 + c];
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
    double *st;              /* [n*P*4] statistics (mean, M2, min, max), or null: in d.stats */
    double *hc;              /* [2P*6] population prior: p0, p1, lower, upper, cached logs */
    double *sb;              /* [P*2] bounds of the subject prior */
    int *hd;                 /* [2P] population prior distribution */
};

__host__ __device__ inline size_t phi_lds_bytes(int n, int P, int G, bool stats_lds) {
    const size_t ng = (size_t)(n + G - 1) / G;
    const size_t dbl = (size_t)n * 2 * P + 2 * (size_t)n + ng * 2 * P + (ng + 1) * 4 * P +
                       (size_t)n * P * 2 + (ng + 1) * P * 2 + ng + (stats_lds ? (size_t)n * P * 4 : 0) +
                       (size_t)P * 12 + (size_t)P * 2;
    return dbl * sizeof(double) + (3 * ng + n + 4 + 2 * (size_t)P) * sizeof(int);
}

__device__ __forceinline__ PhiLds phi_carve(char *base, int n, int P, int G, bool stats_lds) {
    const int ng = (n + G - 1) / G;
    PhiLds L;
    double *p = reinterpret_cast<double *>(base);
    L.st = nullptr;
    if (stats_lds) {
        L.st = p;
        p += (size_t)n * P * 4;
    }
    L.hc = p; p += (size_t)P * 12;
    L.sb = p; p += (size_t)P * 2;
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
    L.flag = q; q += 4;
    L.hd = q;
    return L;
}

/* Stages population state f, its cached prior logarithms (used for chain c
 * only when d.pcache_valid[f*n + c], those flags in L.sub) and the run's
 * prior constants in LDS; nothing here depends on the subject statistics, so
 * these loads overlap theirs. */
__device__ __forceinline__ void phi_stage(const Dev &d, int f, const JobDev &job, const PhiLds &L) {
    const int n = d.nchain, P = d.P, P2 = 2 * P;
    /* every load of the first pass in flight together, then the stores */
    Piece<double, 2> pph{L.ph, d.phi + (size_t)f * n * P2, n * P2};
    Piece<double, 2> ppc{L.pc, d.pcache + (size_t)f * n * P2, n * P2};   /* n * P * 2 == n * P2 */
    Piece<double, 1> plp{L.plp, d.phi_lp + (size_t)f * n, n};
    Piece<double, 1> pll{L.pll, d.phi_ll + (size_t)f * n, n};
    Piece<int32_t, 1> pva{L.sub, d.pcache_valid + (size_t)f * n, n};
    const int q = min((int)threadIdx.x, P2 - 1), p = min((int)threadIdx.x, P - 1);
    const int k = job.hprior_off + q, kp = job.prior_off + p;
    pph.load();
    ppc.load();
    plp.load();
    pll.load();
    pva.load();
    const double h0 = d.pr_p0[k], h1 = d.pr_p1[k], h2 = d.pr_lo[k], h3 = d.pr_up[k];
    const double h4 = d.scache[2 * k], h5 = d.scache[2 * k + 1];
    const int hd = d.pr_dist[k];
    const double b0 = d.pr_lo[kp], b1 = d.pr_up[kp];
    pph.store();
    ppc.store();
    plp.store();
    pll.store();
    pva.store();
    if ((int)threadIdx.x < P2) {
        double *h = L.hc + (size_t)q * 6;
        h[0] = h0;
        h[1] = h1;
        h[2] = h2;
        h[3] = h3;
        h[4] = h4;
        h[5] = h5;
        L.hd[q] = hd;
    }
    if ((int)threadIdx.x < P) {
        L.sb[2 * p] = b0;
        L.sb[2 * p + 1] = b1;
    }
    /* population designs wider than the workgroup (2P > blockDim) */
    for (int r = threadIdx.x + blockDim.x; r < P2; r += blockDim.x) {
        const int kk = job.hprior_off + r;
        double *h = L.hc + (size_t)r * 6;
        h[0] = d.pr_p0[kk];
        h[1] = d.pr_p1[kk];
        h[2] = d.pr_lo[kk];
        h[3] = d.pr_up[kk];
        h[4] = d.scache[2 * kk];
        h[5] = d.scache[2 * kk + 1];
        L.hd[r] = d.pr_dist[kk];
        if (r < P) {
            L.sb[2 * r] = d.pr_lo[job.prior_off + r];
            L.sb[2 * r + 1] = d.pr_up[job.prior_off + r];
        }
    }
}

/* One population parameter's share of sumloghlike from its sufficient
 * statistics (mean, M2, min, max of the S subject values):
 *   sum_s log tnorm(x_s; mu, sg, lo, up)
 *     = -(M2 + S (mean - mu)^2) / (2 sg^2) - S (log sqrt(2 pi) + log sg + log Z)
 * NaN for sg <= 0 or NaN inputs, -inf when a subject value is outside
 * [lo, up]; summing the P shares in parameter order gives sumloghlike
 * (src/de.cpp:245-270). */
__device__ __forceinline__ double hll_term(const PhiLds &L, int p, double mu, double sg, const double *lc,
                           const double *st, int S) {
    const double mean = st[0], M2 = st[1], mn = st[2], mx = st[3];
    if (!(sg > 0.0) || mean != mean || mu != mu) return __builtin_nan("");
    const double lo = L.sb[2 * p], up = L.sb[2 * p + 1];
    if (mn < lo || mx > up) return -__builtin_inf();
    const double dm = mean - mu;
    return -(M2 + (double)S * dm * dm) / (2.0 * sg * sg) -
           (double)S * (0.918938533204672741780329736406 + lc[0] + lc[1]);
}

/* lc = (log sg, log Z) of parameter p's truncated normal with location mu
 * and scale sg on the subject prior's bounds (Z its mass). */
__device__ __forceinline__ void phi_log_consts(const PhiLds &L, int p, double mu, double sg,
                                               double *lc) {
    const double lo = L.sb[2 * p], up = L.sb[2 * p + 1];
    lc[0] = log_acc(sg);
    lc[1] = log_acc(tnorm_mass((lo - mu) / sg, (up - mu) / sg));
}

__device__ __forceinline__ double hyper_prior_term(const Dev &d, const JobDev &job, int q, double x) {
    const int k = job.hprior_off + q;
    return log_prior1(d.pr_dist[k], d.pr_p0[k], d.pr_p1[k], d.pr_lo[k], d.pr_up[k], x);
}

/* The same from the prior's constants and cached logarithms staged in LDS. */
__device__ __forceinline__ double hyper_prior_term_cached(const PhiLds &L, int q, double x) {
    if (x != x) return x;
    const double *h = L.hc + (size_t)q * 6;
    const int dist = L.hd[q];
    if (dist == GGDMC_DIST_UNIF) {
        if (x < h[0] || x > h[1]) return -__builtin_inf();
        return -h[4];
    }
    if (dist == GGDMC_DIST_TNORM) {
        if (!(h[1] > 0.0)) return __builtin_nan("");
        if (x < h[2] || x > h[3]) return -__builtin_inf();
        const double z = (x - h[0]) / h[1];
        return -0.5 * z * z - 0.918938533204672741780329736406 - h[4] - h[5];
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
/* Loads the 1 <= nb <= kStatBlock values src[i * stride] into registers
 * (clamped indices: the loads issue back to back). */
template <bool AG = false>
__device__ __forceinline__ void block_load(const double *src, size_t stride, int nb, double (&v)[kStatBlock]) {
#pragma unroll
    for (int i = 0; i < kStatBlock; ++i) v[i] = ld<AG>(src + (size_t)min(i, nb - 1) * stride);
}

/* Sum, M2 about the mean sum / nb, min, max and a NaN flag of v[0..nb), in
 * order. */
__device__ __forceinline__ void block_moments(const double (&v)[kStatBlock], int nb, double &sum,
                                              double &m2, double &mn, double &mx, bool &nan) {
    sum = 0.0;
    mn = __builtin_inf();
    mx = -__builtin_inf();
    nan = false;
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
    m2 = 0.0;
#pragma unroll
    for (int i = 0; i < kStatBlock; ++i) {
        if (i < nb) {
            const double dx = v[i] - mean;
            m2 += dx * dx;
        }
    }
}

__device__ __forceinline__ void block_stats(const Dev &d, const JobDev &job, int e0, int e1, double *bst) {
    const int P = d.P, n = d.nchain, S = job.nsubject;
    const int nblk = (S + kStatBlock - 1) / kStatBlock;
    const size_t stride = (size_t)n * P;
    for (int e = e0 + (int)threadIdx.x; e < e1; e += blockDim.x) {
        const int q = e / nblk, j = e - q * nblk;
        const int s0 = j * kStatBlock, nb = min(kStatBlock, S - s0);
        double v[kStatBlock], sum, m2, mn, mx;
        bool nan;
        block_load(d.theta + (size_t)(job.first_subject + s0) * stride + q, stride, nb, v);
        block_moments(v, nb, sum, m2, mn, mx, nan);
        double *o = bst + (size_t)e * 4;
        o[0] = nan ? __builtin_nan("") : sum;
        o[1] = m2;
        o[2] = mn;
        o[3] = mx;
    }
}

/* A run of at most kStatBlock subjects: block_stats and combine_stats in one
 * step, from values already loaded, straight into o[4] (the same values:
 * with one block the combination's correction term is nb * 0^2). */
__device__ __forceinline__ void direct_stats(const double (&v)[kStatBlock], int S, double *o) {
    double sum, m2, mn, mx;
    bool nan;
    block_moments(v, S, sum, m2, mn, mx, nan);
    o[0] = nan ? __builtin_nan("") : sum / (double)S;
    o[1] = m2;
    o[2] = mn;
    o[3] = mx;
}

/* Step 2: combine the blocks of pairs [q0, q1) in order into out[q][4] =
 * (mean, M2, min, max). */
__device__ __forceinline__ void combine_stats(const JobDev &job, int q0, int q1, const double *bst, double *out) {
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
__device__ __forceinline__ void phi_migration(const Dev &d, const JobDev &job, const PhiLds &L, const Stream &sd,
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
                v = hyper_prior_term_cached(L, e, L.prop[e]);
            } else if (e < 3 * P) {
                const int p = e - P2;
                double *lc = L.pp + (size_t)p * 2;
                phi_log_consts(L, p, L.prop[p], L.prop[P + p], lc);
                v = hll_term(L, p, L.prop[p], L.prop[P + p], lc, st_src + p * 4, S);
            } else {
                const int which = e < 4 * P ? src : dst;
                const int p = e - (e < 4 * P ? 3 * P : 4 * P);
                const double *ph = L.ph + (size_t)which * P2;
                v = hll_term(L, p, ph[p], ph[P + p], L.pc + ((size_t)which * P + p) * 2,
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
                if (__builtin_expect(phi_traced(d, job), 0))
                    trace_put(d, d.phi_job[job.phi_index], -1, sd, 1, k, src, dst, tlp, tll, cur, a);
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
__device__ __forceinline__ void phi_crossover(const Dev &d, const JobDev &job, const PhiLds &L, const Stream &st,
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
        L.tm[(size_t)c * 4 * P + q] = hyper_prior_term_cached(L, q, L.prop[e]);
    }
    for (int e = threadIdx.x; e < ng * P; e += blockDim.x) {
        const int c = e / P, p = e - c * P;
        const int i = lo + c;
        const double *x = L.prop + (size_t)c * P2;
        double *lc = L.pp + (size_t)e * 2;
        phi_log_consts(L, p, x[p], x[P + p], lc);
        L.tm[(size_t)c * 4 * P + P2 + p] =
            hll_term(L, p, x[p], x[P + p], lc, stats + ((size_t)(hier ? i : 0) * P + p) * 4, S);
    }
    if (hier) {
        for (int e = threadIdx.x; e < ng * P; e += blockDim.x) {
            const int c = e / P, p = e - c * P;
            const int i = lo + c;
            const double *cur = L.ph + (size_t)i * P2;
            L.tm[(size_t)c * 4 * P + 3 * P + p] = hll_term(L, p, cur[p], cur[P + p],
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
            if (__builtin_expect(phi_traced(d, job), 0))
                trace_put(d, d.phi_job[job.phi_index], -1, st, 0, i, i, i, tlp, tll, cur, a);
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

/* Population step of state f for iteration iter in one workgroup, on the
 * state staged by phi_stage and the per-chain statistics `stats` (LDS or
 * global; src/de.cpp:272-383 decisions). */
__device__ __forceinline__ void phi_update(const Dev &d, int f, const JobDev &job, uint32_t iter, const PhiLds &L,
                           const double *stats) {
    const int n = d.nchain, P = d.P, P2 = 2 * P, G = d.G;
    const bool hier = job.kind == GGDMC_KIND_HIER;
    /* log scale / log mass of every chain's subject prior: kept from the last
     * population step (staged), computed afresh after a start or restore */
    for (int e = threadIdx.x; e < n * P; e += blockDim.x) {
        const int c = e / P, p = e - c * P;
        if (L.sub[c] == 0)
            phi_log_consts(L, p, L.ph[(size_t)c * P2 + p], L.ph[(size_t)c * P2 + P + p], L.pc + (size_t)e * 2);
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
    for (int c = threadIdx.x; c < n; c += blockDim.x) d.pcache_valid[(size_t)f * n + c] = 1;
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
 * With one workgroup per run it stages the population state while the
 * subject values load, and a run of at most kStatBlock subjects needs no
 * block pass. Statistics go to LDS when stats_lds (else to d.stats).
 * Hyper-only runs have fixed statistics: their workgroup b == 0 steps
 * directly. */
__device__ __forceinline__ void phi_block(const Dev &d, uint32_t iter, int nwg, int stats_lds, int bid,
                                          char *lds_raw, bool publish) {
    const int n = d.nchain, P = d.P;
    const int f = bid / nwg, b = bid - f * nwg;
    const JobDev job = d.phi_jobs[f];
    const PhiLds L = phi_carve(lds_raw, n, P, d.G, stats_lds != 0);
    double *gst = d.stats + (size_t)f * n * P * 4;
    double *st = L.st ? L.st : gst;
    const int nblk = (job.nsubject + kStatBlock - 1) / kStatBlock;
    const int nitem = n * P * nblk;
    double *bst = d.bstats + (size_t)f * d.bstats_stride;
    if (job.kind != GGDMC_KIND_HIER) {
        if (b != 0) return;
        phi_stage(d, f, job, L);
        if (L.st) stage_copy(L.st, gst, P * 4);
    } else if (nwg == 1 && nblk == 1) {
        /* this thread's subject values load together with the staging */
        const int nq = n * P, S = job.nsubject;
        const double *base = d.theta + (size_t)job.first_subject * nq;
        const int q0 = min((int)threadIdx.x, nq - 1);
        double v[kStatBlock];
        block_load(base + q0, (size_t)nq, S, v);
        phi_stage(d, f, job, L);
        if ((int)threadIdx.x < nq) direct_stats(v, S, st + (size_t)q0 * 4);
        for (int q = threadIdx.x + blockDim.x; q < nq; q += blockDim.x) {
            block_load(base + q, (size_t)nq, S, v);
            direct_stats(v, S, st + (size_t)q * 4);
        }
    } else if (nwg == 1) {
        phi_stage(d, f, job, L);
        block_stats(d, job, 0, nitem, bst);
        __syncthreads();
        combine_stats(job, 0, n * P, bst, st);
    } else {
        if (b * (int)blockDim.x < nitem)
            block_stats(d, job, b * (int)blockDim.x, min(nitem, (b + 1) * (int)blockDim.x), bst);
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
        phi_stage(d, f, job, L);
        combine_stats(job, 0, n * P, bst, st);
    }
    __syncthreads();
    phi_update(d, f, job, iter, L, st);
    if (publish) {
        /* release: every thread's stores of the state, then the flag */
        __threadfence();
        __syncthreads();
        for (int c = threadIdx.x; c < n; c += blockDim.x)
            __hip_atomic_store(d.phi_ready + (size_t)f * n + c, iter, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ __launch_bounds__(kPhiBlock) void phi_kernel(Dev d, uint32_t iter_off, int nwg, int stats_lds) {
    extern __shared__ __attribute__((aligned(16))) char lds_raw[];
    phi_block(d, *d.iter_base + iter_off, nwg, stats_lds, (int)blockIdx.x, lds_raw, false);
}

/* Whether iteration iter of population state job needs the whole-run
 * population step (phi_block) rather than one workgroup per chain: hyper
 * runs, blocked population sweeps, and iterations that migrate
 * (src/de.cpp:272-383 decisions, the draws phi_update makes). */
__device__ __forceinline__ bool phi_full(const Dev &d, const JobDev &job, uint32_t iter) {
    if (job.kind != GGDMC_KIND_HIER || d.blocked_pop != 0) return true;
    const uint64_t seed = ((uint64_t)job.k1 << 32) | job.k0;
    return decision_u(make_stream(seed, iter + d.philox_offset, kUnitPhi, 0u)) < d.pop_mig;
}

__host__ __device__ inline size_t phi_chain_lds_bytes(int P, int nblk) {
    const size_t dbl = (size_t)P * (6 + 2 + 4 + 2 + 2 + 4 + 12 + 2) + 4 + (nblk > 1 ? (size_t)P * nblk * 4 : 0);
    return dbl * sizeof(double) + ((size_t)2 * P + 4) * sizeof(int);
}

/* Chain i's crossover of population state f in its own workgroup, for an
 * iteration that is a plain crossover of a hierarchical run (phi_full false):
 * the statistics of chain i's subject values, the proposal, its prior and
 * likelihood terms, the refreshed current likelihood and the accept: the
 * operations of block_stats / combine_stats and phi_crossover (the same
 * values), for one chain. Chain i's partners come from the other chain
 * groups, whose states do not change in this launch. Publishes
 * d.phi_ready[f*n + i] = iter after the state is stored. */
__device__ __forceinline__ void wait_chains(const Dev &d, int f, int c_lo, int c_n, uint32_t iter,
                                            bool acquire);

template <bool AG>
__device__ __forceinline__ void phi_chain(const Dev &d, int f, int i, uint32_t iter, char *lds_raw,
                                          bool chained) {
    const int n = d.nchain, P = d.P, P2 = 2 * P, G = d.G;
    const JobDev job = d.phi_jobs[f];
    const int S = job.nsubject, nblk = (S + kStatBlock - 1) / kStatBlock;
    const int t = threadIdx.x, bs = blockDim.x;
    int g = 0;
    while (group_lo(g + 1, n, G) <= i) ++g;
    const uint64_t seed = ((uint64_t)job.k1 << 32) | job.k0;
    const Stream st = make_stream(seed, iter + d.philox_offset, kUnitPhi, 0u);
    int m0, m1;
    double U;
    partners(st, i, group_lo(g, n, G), group_lo(g + 1, n, G), n, m0, m1, U);

    PhiLds L{};   /* the term helpers read hc, hd, sb */
    double *p = reinterpret_cast<double *>(lds_raw);
    double *xi = p; p += P2;      /* rows i, m0, m1 */
    double *x0 = p; p += P2;
    double *x1 = p; p += P2;
    double *prop = p; p += P2;
    double *tm = p; p += 4 * P;
    double *pc = p; p += 2 * P;
    double *pp = p; p += 2 * P;
    double *stt = p; p += 4 * P;
    L.hc = p; p += 12 * P;
    L.sb = p; p += 2 * P;
    double *misc = p; p += 4;     /* lp, ll after the step */
    double *bst = p; p += nblk > 1 ? (size_t)P * nblk * 4 : 0;
    int *q = reinterpret_cast<int *>(p);
    L.hd = q; q += P2;
    int *flag = q;                /* [0] pcache valid, [1] accepted */

    const size_t off = (size_t)f * n;
    const size_t nq = (size_t)n * P;
    const double *tbase = d.theta + (size_t)job.first_subject * nq + (size_t)i * P;
    /* loads in one batch: the subject values of this chain (one block),
     * rows i, m0, m1, chain i's cached logarithms, the prior constants */
    double v[kStatBlock];
    if (nblk == 1) block_load<AG>(tbase + min(t, P - 1), nq, S, v);
    {
        Piece<double, 1, AG> pxi{xi, d.phi + (off + i) * P2, P2};
        Piece<double, 1, AG> ppc{pc, d.pcache + (off + i) * P2, P2};
        const int qq = min(t, P2 - 1), pq = min(t, P - 1);
        const int k = job.hprior_off + qq, kp = job.prior_off + pq;
        pxi.load();
        ppc.load();
        const double h0 = d.pr_p0[k], h1 = d.pr_p1[k], h2 = d.pr_lo[k], h3 = d.pr_up[k];
        const double h4 = d.scache[2 * k], h5 = d.scache[2 * k + 1];
        const int hd = d.pr_dist[k];
        const double b0 = d.pr_lo[kp], b1 = d.pr_up[kp];
        const double lp0 = ld<AG>(d.phi_lp + off + i);
        const int valid = ld<AG>(d.pcache_valid + off + i);
        pxi.store();
        ppc.store();
        if (t < P2) {
            double *h = L.hc + (size_t)qq * 6;
            h[0] = h0;
            h[1] = h1;
            h[2] = h2;
            h[3] = h3;
            h[4] = h4;
            h[5] = h5;
            L.hd[qq] = hd;
        }
        if (t < P) {
            L.sb[2 * pq] = b0;
            L.sb[2 * pq + 1] = b1;
        }
        if (t == 0) {
            misc[0] = lp0;
            flag[0] = valid;
        }
        for (int r = t + bs; r < P2; r += bs) {
            const int kk = job.hprior_off + r;
            double *h = L.hc + (size_t)r * 6;
            h[0] = d.pr_p0[kk];
            h[1] = d.pr_p1[kk];
            h[2] = d.pr_lo[kk];
            h[3] = d.pr_up[kk];
            h[4] = d.scache[2 * kk];
            h[5] = d.scache[2 * kk + 1];
            L.hd[r] = d.pr_dist[kk];
            if (r < P) {
                L.sb[2 * r] = d.pr_lo[job.prior_off + r];
                L.sb[2 * r + 1] = d.pr_up[job.prior_off + r];
            }
        }
    }
    /* statistics of chain i's values of each parameter */
    if (nblk == 1) {
        if (t < P) direct_stats(v, S, stt + (size_t)t * 4);
        for (int r = t + bs; r < P; r += bs) {
            block_load<AG>(tbase + r, nq, S, v);
            direct_stats(v, S, stt + (size_t)r * 4);
        }
    } else {
        for (int e = t; e < P * nblk; e += bs) {
            const int pr = e / nblk, j = e - pr * nblk;
            const int nb = min(kStatBlock, S - j * kStatBlock);
            double sum, m2, mn, mx;
            bool nan;
            block_load<AG>(tbase + (size_t)j * kStatBlock * nq + pr, nq, nb, v);
            block_moments(v, nb, sum, m2, mn, mx, nan);
            double *o = bst + (size_t)e * 4;
            o[0] = nan ? __builtin_nan("") : sum;
            o[1] = m2;
            o[2] = mn;
            o[3] = mx;
        }
    }
    __syncthreads();
    if (nblk > 1) combine_stats(job, 0, P, bst, stt);
    /* chain i's cached logarithms after a start or restore */
    if (flag[0] == 0)
        for (int r = t; r < P; r += bs) phi_log_consts(L, r, xi[r], xi[P + r], pc + (size_t)r * 2);
    /* the partners' rows. Every chain group in one launch: the earlier groups'
     * chains after their step of this iteration (device-scope loads; the
     * statistics above did not need them), and the later groups' chains wait
     * for this one before they change theirs. */
    if (chained && g > 0) wait_chains(d, f, 0, group_lo(g, n, G), iter, false);
    {
        const double *r0 = d.phi + (off + m0) * P2, *r1 = d.phi + (off + m1) * P2;
        for (int r = t; r < P2; r += bs) {
            x0[r] = chained ? __hip_atomic_load(r0 + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : r0[r];
            x1[r] = chained ? __hip_atomic_load(r1 + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : r1[r];
        }
    }
    __syncthreads();
    /* the proposal: jitter of parameter e is word e & 3 of Philox block 1 + e / 4 */
    for (int e = t; e < P2; e += bs) {
        const U4 r = st.block((uint32_t)i, 1u + (uint32_t)(e >> 2));
        const int j = e & 3;   /* selects, not an indexed (private-memory) array */
        const uint32_t wj = j == 0 ? r.x[0] : (j == 1 ? r.x[1] : (j == 2 ? r.x[2] : r.x[3]));
        prop[e] = xi[e] + (jitter_word(wj, d.rp) + job.gamma * (x0[e] - x1[e]));
    }
    __syncthreads();
    /* [0, 2P) prior, [2P, 3P) proposal shares, [3P, 4P) refreshed current shares */
    for (int e = t; e < 4 * P; e += bs) {
        double val;
        if (e < P2) {
            val = hyper_prior_term_cached(L, e, prop[e]);
        } else if (e < 3 * P) {
            const int r = e - P2;
            phi_log_consts(L, r, prop[r], prop[P + r], pp + (size_t)r * 2);
            val = hll_term(L, r, prop[r], prop[P + r], pp + (size_t)r * 2, stt + (size_t)r * 4, S);
        } else {
            const int r = e - 3 * P;
            val = hll_term(L, r, xi[r], xi[P + r], pc + (size_t)r * 2, stt + (size_t)r * 4, S);
        }
        tm[e] = val;
    }
    __syncthreads();
    if (t < 8) {
        const double tlp = group8_sum(tm, P2, t);
        const double tll = group8_sum(tm + P2, P, t);
        const double ref = group8_sum(tm + 3 * P, P, t);
        if (t == 0) {
            const double cur = ref + misc[0];
            const bool a = mh_accept(tlp + tll, cur, U);
            if (__builtin_expect(phi_traced(d, job), 0))
                trace_put(d, d.phi_job[f], -1, st, 0, i, i, i, tlp, tll, cur, a);
            flag[1] = a;
            misc[2] = a ? tlp : misc[0];
            misc[3] = a ? tll : ref;
        }
    }
    __syncthreads();
    const bool acc = flag[1] != 0;
    const bool out = iter % (uint32_t)d.thin == 0 && (int)(iter / (uint32_t)d.thin) < d.nmc;
    const size_t ko = out ? (size_t)f * d.nmc + iter / (uint32_t)d.thin : 0;
    for (int e = t; e < P2; e += bs) {
        const double x = acc ? prop[e] : xi[e];
        d.phi[(off + i) * P2 + e] = x;
        d.pcache[(off + i) * P2 + e] = acc ? pp[e] : pc[e];   /* P * 2 == P2 entries per chain */
        if (out) d.phi_out_theta[(ko * n + i) * P2 + e] = x;
    }
    if (t == 0) {
        d.phi_lp[off + i] = misc[2];
        d.phi_ll[off + i] = misc[3];
        d.pcache_valid[off + i] = 1;
        if (out) {
            d.phi_out_lp[ko * n + i] = misc[2];
            d.phi_out_ll[ko * n + i] = misc[3];
        }
    }
    /* release: every thread's stores, then the flag */
    __threadfence();
    __syncthreads();
    if (t == 0) __hip_atomic_store(d.phi_ready + off + i, iter, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

/* A subject workgroup that needs population chains [c_lo, c_lo + c_n) of
 * state f for iteration iter: threads poll d.phi_ready with device-scope
 * atomic loads a bounded number of times (on giving up they raise d.fault,
 * which the host reports as an error). The writer released the state before
 * the flag; the reader then either reads it with device-scope loads (no
 * fence: late_prior) or, with acquire, fences once and reads it plainly (a
 * fence writes back and invalidates the whole L2 of the XCD, costly when
 * hundreds of workgroups do it). */
constexpr uint32_t kSpinLimit = 1u << 18;

/* A flag holds the last iteration whose step it marks: iteration iter's step
 * is done once flag >= iter (iterations only grow within a session; with
 * several iterations in a launch a flag may already hold a later one). */
__device__ __forceinline__ bool reached(uint32_t v, uint32_t iter) { return (int32_t)(v - iter) >= 0; }

__device__ __forceinline__ void wait_chains(const Dev &d, int f, int c_lo, int c_n, uint32_t iter,
                                            bool acquire) {
    for (int c = threadIdx.x; c < c_n; c += blockDim.x) {
        const uint32_t *flag = d.phi_ready + (size_t)f * d.nchain + c_lo + c;
        uint32_t spins = 0;
        while (!reached(__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), iter)) {
            /* after one failure every later wait gives up at once */
            if (++spins == kSpinLimit ||
                __hip_atomic_load(d.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                atomicOr(d.fault, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(4);
        }
    }
    __syncthreads();
    if (acquire) __threadfence();
}

/* Every chain group in one launch: after an item of (subject, group, range
 * r) has stored its chains (device-scope stores, complete before the
 * barrier), ready[r] = iteration. */
__device__ __forceinline__ void publish_item(uint32_t *ready, int r, uint32_t iter) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(ready + r, iter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* Waits (bounded, as wait_chains) until ready[0..nr) all reach iter. */
__device__ __forceinline__ void wait_items(const Dev &d, const uint32_t *ready, int nr, uint32_t iter) {
    for (int r = threadIdx.x; r < nr; r += blockDim.x) {
        uint32_t spins = 0;
        while (!reached(__hip_atomic_load(ready + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), iter)) {
            if (++spins == kSpinLimit ||
                __hip_atomic_load(d.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                atomicOr(d.fault, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(4);
        }
    }
    __syncthreads();
}

/* Bounded wait (as wait_chains) for one flag; false on giving up. */
__device__ __forceinline__ void wait_flag(const Dev &d, const uint32_t *flag, uint32_t iter) {
    uint32_t spins = 0;
    while (!reached(__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), iter)) {
        if (++spins == kSpinLimit || __hip_atomic_load(d.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            atomicOr(d.fault, 1u);
            break;
        }
        __builtin_amdgcn_s_sleep(4);
    }
}

/* Several iterations per launch: every item of subject x (each group's
 * chain ranges) has finished iteration iter. Items published with
 * device-scope stores; the waiter reads the subject's state with
 * device-scope loads. */
__device__ __forceinline__ void wait_subject(const Dev &d, const SweepArgs &a, int x, uint32_t iter) {
    const int n = d.nchain, G = d.G;
    const uint32_t *ready = d.item_ready + (size_t)x * G * a.rmax;
    for (int e = threadIdx.x; e < G * a.rmax; e += blockDim.x) {
        const int g = e / a.rmax, r = e - g * a.rmax;
        const int ng = group_lo(g + 1, n, G) - group_lo(g, n, G);
        if (r < (ng + a.kmax - 1) / a.kmax) wait_flag(d, ready + e, iter);
    }
    __syncthreads();
}

/* Several iterations per launch: population state f's whole run has finished
 * iteration iter (every chain's population step and, for a hierarchical run,
 * every item of every subject). */
__device__ __forceinline__ void wait_run(const Dev &d, const SweepArgs &a, int f, uint32_t iter) {
    const int n = d.nchain, G = d.G;
    const JobDev job = d.phi_jobs[f];
    for (int c = threadIdx.x; c < n; c += blockDim.x) wait_flag(d, d.phi_ready + (size_t)f * n + c, iter);
    if (job.kind == GGDMC_KIND_HIER) {
        const int per = G * a.rmax;
        const uint32_t *ready = d.item_ready + (size_t)job.first_subject * per;
        for (int e = threadIdx.x; e < job.nsubject * per; e += blockDim.x) {
            const int q = e % per, g = q / a.rmax, r = q - g * a.rmax;
            const int ng = group_lo(g + 1, n, G) - group_lo(g, n, G);
            if (r < (ng + a.kmax - 1) / a.kmax) wait_flag(d, ready + e, iter);
        }
    }
    __syncthreads();
}

/* Several iterations per launch: what chain i's population step of the next
 * iteration reads or overwrites is done with iteration iter: every chain's
 * population step (the partners' rows, and chain i's row that they read) and,
 * for a hierarchical run, in every subject the item holding chain i (chain
 * i's subject values, and the priors it read from chain i's row; a migrating
 * subject's ring marks every range). One flag per thread. */
__device__ __forceinline__ void wait_chain_run(const Dev &d, const SweepArgs &a, int f, int i, uint32_t iter) {
    const int n = d.nchain, G = d.G;
    const JobDev job = d.phi_jobs[f];
    int g = 0;
    while (group_lo(g + 1, n, G) <= i) ++g;
    const int per = G * a.rmax;
    const uint32_t *ready = d.item_ready + (size_t)job.first_subject * per + (size_t)g * a.rmax +
                            (i - group_lo(g, n, G)) / a.kmax;
    const int ns = job.kind == GGDMC_KIND_HIER ? job.nsubject : 0;
    for (int e = threadIdx.x; e < n + ns; e += blockDim.x)
        wait_flag(d, e < n ? d.phi_ready + (size_t)f * n + e : ready + (size_t)(e - n) * per, iter);
    __syncthreads();
}

/* Prior terms of proposals L.prop[0..k) for chains i0 + c under the
 * population state just published (read with device-scope loads): the
 * operations of proposal_stage1/2 (location, scale and cached logarithms per
 * (chain, parameter)), summed per chain in parameter order by 8 lanes into
 * L.tlp, as proposal_stage3. */
__device__ __forceinline__ double load_agent(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void late_prior(const Dev &d, const JobDev &job, const SweepLds &L, int i0, int k) {
    const int P = d.P, n = d.nchain;
    for (int e = threadIdx.x; e < k * P; e += blockDim.x) {
        const int c = e / P, p = e - c * P;
        const size_t r = ((size_t)job.phi_index * n + i0 + c) * P + p;
        const double *ph = d.phi + ((size_t)job.phi_index * n + i0 + c) * (2 * P);
        const double q[4] = {load_agent(ph + p), load_agent(ph + P + p), load_agent(d.pcache + 2 * r),
                             load_agent(d.pcache + 2 * r + 1)};
        L.tm[e] = staged_prior_term(L.pdist[p], q, L.bnd[2 * p], L.bnd[2 * p + 1], L.prop[e]);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ((k + 7) & ~7) * 8; e += blockDim.x) {
        const int c = e >> 3, l = e & 7;
        const double v = group8_sum(L.tm + (size_t)(c < k ? c : 0) * P, P, l);
        if (c < k && l == 0) L.tlp[c] = v;
    }
    __syncthreads();
}

/* ------------------------------------------------------------ sweep kernel */
/* OCC: waves per SIMD the kernel is built for (3: up to 168 VGPRs a lane, 4:
 * 128, with values of the cold paths in scratch). */
template <int MODE, bool MULTI, int OCC>
__global__ __launch_bounds__(kBlock, OCC) void subject_sweep_kernel(Dev d, SweepArgs a) {
    extern __shared__ __attribute__((aligned(16))) char lds_raw[];
    /* several iterations per launch (MULTI): segment seg of seg_wgs
     * workgroups runs iteration base + seg; its workgroups wait only on
     * lower workgroup ids (the previous segment, or earlier in their own),
     * which in-order dispatch has started before them */
    int seg = 0, wid = (int)blockIdx.x;
    if constexpr (MULTI) {
        seg = (int)blockIdx.x / a.seg_wgs;
        wid = (int)blockIdx.x - seg * a.seg_wgs;
        if (wid >= a.phi_wgs + a.n_items) return;   /* padding to a multiple of 8 */
    }
    /* (the population workgroups are few: their code is laid out after the
     * subjects' path) */
    if (__builtin_expect(wid < a.phi_wgs, 0)) {
        /* the population step of the launch's chains, concurrent with the
         * subjects' proposals and likelihoods (they depend on it only through
         * the priors of the hierarchical subjects' proposals): one workgroup
         * per (state, chain); whole-run steps by the first chain's workgroup
         * of the first group */
        const int gb = a.chained ? 0 : a.g_begin, ge = a.chained ? d.G : a.g_end;
        const int c_lo = group_lo(gb, d.nchain, d.G);
        const int nc = group_lo(ge, d.nchain, d.G) - c_lo;
        const int f = wid / nc, i = c_lo + wid % nc;
        const uint32_t it = *d.iter_base + a.iter_off + (uint32_t)seg;
        if (!phi_full(d, d.phi_jobs[f], it)) {
            if (MULTI && seg > 0) wait_chain_run(d, a, f, i, it - 1);
            phi_chain<MULTI>(d, f, i, it, lds_raw, a.chained != 0);
        } else if (gb == 0 && i == 0) {
            if (MULTI && seg > 0) {
                /* acquire: the whole-run step reads the state plainly */
                wait_run(d, a, f, it - 1);
                __threadfence();
            }
            phi_block(d, it, 1, a.phi_stats_lds, f, lds_raw, true);
        }
        return;
    }
    const int ib = wid - a.phi_wgs;
    const int4 item = a.items[ib];
    const int x = item.x;
    const int c0 = item.y >> 16, c1 = item.y & 0xFFFF;
    const int t = item.z >> 16, ntile = item.z & 0xFFFF;
    const int n = d.nchain, P = d.P, G = d.G;
    /* the item's chain groups [g_begin, g_end) and chain range r */
    const int g_begin = a.chained ? item.w >> 16 : a.g_begin;
    const int g_end = a.chained ? (a.span ? G : g_begin + 1) : a.g_end;
    const int rng = item.w & 0xFFFF;
    const SweepLds L = carve(lds_raw, n, P, a.kmax, d.n_cond * d.n_acc, a.chain_batch, a.cmax, a.tstage);
    /* the item's trials; staged in LDS when they fit */
    const int64_t s0 = d.subj_off[x], s1 = d.subj_off[x + 1];
    const int64_t t_lo = ntile > 1 ? s0 + (int64_t)t * d.tpt : s0;
    const int64_t t_hi = ntile > 1 ? min(s1, t_lo + (int64_t)d.tpt) : s1;
    const int nst = a.tstage > 0 ? (int)(t_hi - t_lo) : 0;
    const JobDev job = d.subj_jobs[x];
    {
        /* loads first (they depend on the item only), all in flight together,
         * then the draws: exp table, state, trials, design rows, the subject
         * prior's bounds and families (several iterations per launch: the
         * state after the draws, once the subject's previous iteration, run
         * earlier in this launch, is done) */
        const int per = d.n_cond * d.n_acc * 6;
        Piece<double, 1> pex{g_exp2tab, d.exp2tab, 64};
        Piece<int32_t, 1> psr{L.psrc, d.par_src, per};
        Piece<double, 1> pcs{L.pcst, d.par_const, per};
        const int tp = (int)threadIdx.x, pq = min(tp, P - 1);
        const double blo = d.pr_lo[job.prior_off + pq], bup = d.pr_up[job.prior_off + pq];
        const int pdi = job.kind == GGDMC_KIND_HIER ? GGDMC_DIST_TNORM : d.pr_dist[job.prior_off + pq];
        [[maybe_unused]] Piece<double, 4> pth{L.th, d.theta + (size_t)x * n * P, n * P};
        [[maybe_unused]] Piece<double, 1> plp{L.lp, d.lp + (size_t)x * n, n};
        [[maybe_unused]] Piece<double, 1> pll{L.ll, d.ll + (size_t)x * n, n};
        Piece<double, 4> prt{L.trt, d.rt + t_lo, nst};
        Piece<uint16_t, 4> pcd{L.tcd, d.tcode + t_lo, nst};
        pex.load();
        if constexpr (!MULTI) {
            pth.load();
            plp.load();
            pll.load();
        }
        prt.load();
        pcd.load();
        psr.load();
        pcs.load();
        pex.store();
        if constexpr (!MULTI) {
            pth.store();
            plp.store();
            pll.store();
        }
        prt.store();
        pcd.store();
        psr.store();
        pcs.store();
        if (tp < P) {
            L.bnd[2 * tp] = blo;
            L.bnd[2 * tp + 1] = bup;
            L.pdist[tp] = pdi;
        }
        for (int r = tp + (int)blockDim.x; r < P; r += blockDim.x) {
            L.bnd[2 * r] = d.pr_lo[job.prior_off + r];
            L.bnd[2 * r + 1] = d.pr_up[job.prior_off + r];
            L.pdist[r] = job.kind == GGDMC_KIND_HIER ? GGDMC_DIST_TNORM : d.pr_dist[job.prior_off + r];
        }
    }
    const TrialView tv = nst > 0 ? TrialView{L.trt, L.tcd, t_lo} : TrialView{d.rt, d.tcode, 0};
    const uint32_t iter = *d.iter_base + a.iter_off + (uint32_t)seg;
    const bool single = job.kind != GGDMC_KIND_HIER;
    const bool blocked = d.blocked_sub != 0;
    const bool late = a.phi_wgs > 0 && !single;
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
        /* stage 1 of the first group: partners and draws, no state */
        int lo, hi, i0;
        const int k = group_k(g_begin, lo, hi, i0);
        if (k > 0) proposal_stage1(d, job, L, st, lo, hi, i0, k, para, late);
    }
    if constexpr (MULTI) {
        /* the state, after the subject's previous iteration (earlier in this
         * launch: device-scope loads) */
        if (seg > 0) wait_subject(d, a, x, iter - 1u);
        Piece<double, 4, true> pth{L.th, d.theta + (size_t)x * n * P, n * P};
        Piece<double, 1, true> plp{L.lp, d.lp + (size_t)x * n, n};
        Piece<double, 1, true> pll{L.ll, d.ll + (size_t)x * n, n};
        pth.load();
        plp.load();
        pll.load();
        pth.store();
        plp.store();
        pll.store();
    }
    __syncthreads();
    /* the subject's migration ring runs in this launch, in this item */
    const bool mig_act = blocked && single ? a.sweep == 0 : true;
    const bool mig_here = (late ? g_end == G : g_begin == 0) && c0 == 0 && t == 0 && mig_act;
    /* deferred store: the first sweep's launches of iteration i store
     * iteration i - 1 from the rows staged above. Each item stores the chains
     * it updates, whose rows no other item writes, after its likelihood
     * phase (before its accept changes them; the stores then complete while
     * it waits for the population step); in a migration iteration no item
     * but the ring's writes, and the ring's item stores every chain before
     * it starts. */
    const bool store_due = a.store_prev && t == 0 && iter > 1;
    if (__builtin_expect(store_due && migrate && mig_here, 0)) store_chains(d, L.th, L.lp, L.ll, x, iter - 1, 0, n);
    uint32_t *ready = a.chained ? d.item_ready + ((size_t)x * G) * a.rmax : nullptr;
    uint32_t *staged = a.chained ? d.item_staged + ((size_t)x * G) * a.rmax : nullptr;
    /* this item has read the subject's rows: a later group may change them */
    if (a.chained && t == 0 && threadIdx.x == 0)
        __hip_atomic_store(staged + (size_t)g_begin * a.rmax + rng, iter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    /* group gq > 0 of a subject in one launch: the previous group's items of
     * this subject first; then the rows of the earlier groups' chains, which
     * they changed */
    auto wait_prev_group = [&](int gq) {
        const int glo = group_lo(gq - 1, n, G);
        const int nr = (group_lo(gq, n, G) - glo + a.kmax - 1) / a.kmax;
        if (G == 2 && !a.tiles) {
            /* two groups, whole subjects: every group-0 item of the subject
             * has read its rows, and the items holding this item's partners
             * (all in group 0) are done */
            wait_items(d, staged, nr, iter);
            int lo1, hi1, i01;
            const int k1 = group_k(gq, lo1, hi1, i01);
            for (int e = threadIdx.x; e < 2 * k1; e += blockDim.x) {
                const int m = (e & 1) ? L.m1[e >> 1] : L.m0[e >> 1];
                const uint32_t *f = ready + (m - glo) / a.kmax;
                uint32_t spins = 0;
                while (!reached(__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), iter)) {
                    if (++spins == kSpinLimit ||
                        __hip_atomic_load(d.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                        atomicOr(d.fault, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(4);
                }
            }
            __syncthreads();
        } else {
            wait_items(d, ready + (size_t)(gq - 1) * a.rmax, nr, iter);
        }
        const double *src = d.theta + (size_t)x * n * P;
        for (int e = threadIdx.x; e < group_lo(gq, n, G) * P; e += blockDim.x)
            L.th[e] = __hip_atomic_load(src + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
    };
    if (a.chained && !migrate && g_begin > 0) wait_prev_group(g_begin);

    if (__builtin_expect(migrate, 0)) {
        /* with the population step in these launches, a hierarchical
         * subject migrates in the last group's launch, after every chain's
         * population step */
        if (mig_here) {
            const int para = (blocked && !single) ? a.sweep : -1;
            if (late) wait_chains(d, job.phi_index, 0, n, iter, true);
            subject_migration<MODE>(d, L, job, x, sd, para, a.cmax,
                                    ntile == 1 && a.tstage > 0 ? tv : TrialView{d.rt, d.tcode, 0});
            /* the ring changed chains of every group: it marks every range done
             * (no other item of a migrating subject publishes) */
            if (a.chained) {
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
                for (int e = threadIdx.x; e < G * a.rmax; e += blockDim.x) {
                    const int g = e / a.rmax, r = e - g * a.rmax;
                    const int ng = group_lo(g + 1, n, G) - group_lo(g, n, G);
                    if (r < (ng + a.kmax - 1) / a.kmax)
                        __hip_atomic_store(ready + e, iter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
    } else {
        const int j_lo = (int)((t_lo - s0) / kChunk);
        const int nj = (int)((t_hi - t_lo + kChunk - 1) / kChunk);
        for (int g = g_begin; g < g_end; ++g) {
            int lo, hi, i0;
            const int k = group_k(g, lo, hi, i0);
            if (k <= 0) continue;
            if (g != g_begin) {
                proposal_stage1(d, job, L, st, lo, hi, i0, k, para, late);
                __syncthreads();
                /* an item spanning every group: this group's partners */
                if (a.chained) wait_prev_group(g);
            }
            proposal_stage2(d, job, L, i0, k, para, late);
            __syncthreads();
            const bool built = proposal_stage3(d, L, k, a.chain_batch);
            __syncthreads();
            proposals_loglike<MODE>(d, L, tv, k, a.chain_batch, a.cmax, s0, s1, j_lo, nj, ntile > 1,
                                    d.chunk_off[x], i0, built);
            if (store_due) store_chains(d, L.th, L.lp, L.ll, x, iter - 1, i0, k);
            if (threadIdx.x == 0) {
                int na = 0;
                for (int c = 0; c < k; ++c) na += L.active[c];
                atomicAdd(d.evals + x, (unsigned long long)na * (unsigned long long)(t_hi - t_lo));
            }
            if (__builtin_expect(ntile > 1, 0)) {
                /* the last-arriving tile of (subject, chain range) accepts: the
                 * tiles' chunk sums are device-scope stores, complete before
                 * the counter moves, and read with device-scope loads (no
                 * fence: a fence writes back and invalidates the XCD's L2,
                 * costly when every tile does it) */
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
                uint32_t *ctr = d.tile_ctr + ((size_t)x * G + g) * a.rmax + rng;
                if (threadIdx.x == 0) {
                    const uint32_t old = atomicAdd(ctr, 1u);
                    L.flag[3] = old == (uint32_t)(ntile - 1);
                }
                __syncthreads();
                if (!L.flag[3]) continue;
                if (threadIdx.x == 0) __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const int64_t cb = d.chunk_off[x], ncb = d.chunk_off[x + 1] - cb;
                for (int c = threadIdx.x; c < k; c += blockDim.x) {
                    double s = 0.0;
                    for (int64_t j = 0; j < ncb; ++j) s += ld<true>(d.csum + (size_t)(cb + j) * n + i0 + c);
                    L.tll[c] = s;
                }
                __syncthreads();
            }
            if (late) {
                wait_chains(d, job.phi_index, i0, k, iter, false);
                late_prior(d, job, L, i0, k);
            }
            accept_proposals(d, L, x, i0, k, st);
            if (a.chained) publish_item(ready + (size_t)g * a.rmax, rng, iter);
        }
    }
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
                info[2] = src;
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
        const bool acc = mh_accept(misc[1] + tll, cur, misc[0]);
        if (__builtin_expect(subject_traced(d, s), 0)) {
            /* crossover records carry the sweep of their stream, migration
             * steps the decision's */
            const Stream tr = migrate ? sd
                                      : make_stream(seed, iter + d.philox_offset, (uint32_t)d.subj_unit[s],
                                                    (uint32_t)sweep);
            trace_put(d, d.subj_job[s], d.subj_unit[s], tr, migrate ? 1 : 0, j, migrate ? info[2] : j, dst,
                      misc[1], tll, cur, acc);
        }
        if (acc) {
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
        const bool acc = mh_accept(misc[1] + tll, cur, misc[0]);
        if (__builtin_expect(phi_traced(d, job), 0)) {
            const Stream tr = info[0] == 2 ? sd : make_stream(seed, iter + d.philox_offset, unit, (uint32_t)sweep);
            trace_put(d, d.phi_job[f], -1, tr, info[0] == 2 ? 1 : 0, j, src, dst, misc[1], tll, cur, acc);
        }
        if (acc) {
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
    stage_copy(L.psrc, d.par_src, d.n_cond * d.n_acc * 6);
    stage_copy(L.pcst, d.par_const, d.n_cond * d.n_acc * 6);
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
size_t ggdmc_internal_phi_lds(int n, int P, int G, int stats_lds) {
    return phi_lds_bytes(n, P, G, stats_lds != 0);
}
size_t ggdmc_internal_phi_chain_lds(int P, int max_subjects) {
    return phi_chain_lds_bytes(P, (max_subjects + kStatBlock - 1) / kStatBlock);
}

/* occ (waves per SIMD) 4 exists for several iterations per launch only. */
hipError_t ggdmc_internal_launch_sweep(const Dev *d, const SweepArgs *a, int mode, size_t lds, int occ,
                                       hipStream_t stream) {
    const dim3 block(kBlock);
    const bool mixed = mode == GGDMC_MIXED_BATCHED;
    if (a->iters > 1) {
        const dim3 grid(a->iters * a->seg_wgs);
        if (occ == 4) {
            if (mixed)
                subject_sweep_kernel<MODE_MIXED, true, 4><<<grid, block, lds, stream>>>(*d, *a);
            else
                subject_sweep_kernel<MODE_NAIVE, true, 4><<<grid, block, lds, stream>>>(*d, *a);
        } else {
            if (mixed)
                subject_sweep_kernel<MODE_MIXED, true, 3><<<grid, block, lds, stream>>>(*d, *a);
            else
                subject_sweep_kernel<MODE_NAIVE, true, 3><<<grid, block, lds, stream>>>(*d, *a);
        }
        return hipGetLastError();
    }
    const dim3 grid(a->phi_wgs + a->n_items);
    if (mixed)
        subject_sweep_kernel<MODE_MIXED, false, 3><<<grid, block, lds, stream>>>(*d, *a);
    else
        subject_sweep_kernel<MODE_NAIVE, false, 3><<<grid, block, lds, stream>>>(*d, *a);
    return hipGetLastError();
}

hipError_t ggdmc_internal_launch_phi(const Dev *d, uint32_t iter_off, size_t lds, int stats_lds,
                                     hipStream_t stream) {
    /* items of the largest run: pairs x subject blocks */
    const int nwg = (int)((d->bstats_stride / 4 + kPhiBlock - 1) / kPhiBlock);
    phi_kernel<<<dim3(d->n_phi * std::max(nwg, 1)), dim3(kPhiBlock), lds, stream>>>(
        *d, iter_off, std::max(nwg, 1), stats_lds);
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
