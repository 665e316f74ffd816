/*
 * HIP kernels of the DE-MCMC sampler (gfx950, wave64).
 *
 *  subject_sweep_kernel  DE crossover of one chain group (or all groups) of every
 *                        subject: proposal (src/de.cpp:567-613 / :111-155),
 *                        LBA log-likelihood over a trial tile, last-arriving
 *                        tile accepts (src/de.cpp:81-108); migration ring
 *                        (src/de.cpp:615-665 / :157-199) by the subject's first
 *                        tile.
 *  phi_kernel            per-chain sufficient statistics of the subject states,
 *                        then the population-level step (src/de.cpp:385-465,
 *                        :467-565) with sumloghlike (:245-270) from them.
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

struct DevOps {
    static __device__ __forceinline__ float exp2(float x) { return __builtin_amdgcn_exp2f(x); }
    static __device__ __forceinline__ float log2(float x) { return __builtin_amdgcn_logf(x); }
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ double erfc(double x) { return ::erfc(x); }
    static __device__ __forceinline__ double pnorm(double z) { return 0.5 * ::erfc(-z * kSqrt1_2); }
    static __device__ __forceinline__ double dnorm(double z) {
        return kInvSqrt2Pi * ::exp(-0.5 * z * z);
    }
};

enum { MODE_MIXED = 0, MODE_NAIVE = 1 };

constexpr int kAccKDoubles = sizeof(AccK) / sizeof(double);

/* --------------------------------------------------------------- likelihood */
template <int MODE>
__device__ __forceinline__ double trial_term(const AccK *kc, int n_acc, int w, double rt) {
    if (MODE == MODE_MIXED) {
        const double lp = mixed_winner_logpdf<DevOps>(kc[w], rt);
        double acc = 0.0;
        int neg = 0;
        bool zero = false, nan = lp != lp;
        for (int j = 0; j < n_acc; ++j) {
            if (j == w) continue;
            const double S = mixed_loser_survivor<DevOps>(kc[j], rt);
            nan |= S != S;
            const double aS = fabs(S);
            fold_survivor(S, aS > 0.0 ? mixed_log<DevOps>(aS) : 0.0, acc, neg, zero);
        }
        if (nan) return __builtin_nan("");
        return (zero || neg) ? kLogDblEps : lp + acc;
    } else {
        double pdf, cdf;
        naive_pdf_cdf<DevOps>(kc[w], rt - kc[w].t0, pdf, cdf);
        double dens = pdf != pdf ? pdf : (pdf > kFloor ? pdf : kFloor);
        for (int j = 0; j < n_acc; ++j) {
            if (j == w) continue;
            naive_pdf_cdf<DevOps>(kc[j], rt - kc[j].t0, pdf, cdf);
            const double F = cdf != cdf ? cdf : (cdf > kFloor ? cdf : kFloor);
            dens *= (1.0 - F);
        }
        return ::log(dens <= 0.0 ? 2.220446049250313e-16 : dens);
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
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

/* Evaluates trials [t_lo, t_hi) for chains [0, kb) of a batch whose AccK
 * table is acck[kb][n_cond][n_acc]; adds each wave's per-chain sums to
 * part[wave][cb + c]. Units (chain, block of 64*TPL trials) are split into
 * contiguous ranges per wave, so a wave reduces once per chain it touches. */
template <int MODE, int TPL>
__device__ void eval_units(const Dev &d, int64_t t_lo, int64_t t_hi, int kb, int cb,
                           const int *active, const double *acck, double *part, int ng_stride) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const int64_t span = 64 * TPL;
    const int nblk = (int)((t_hi - t_lo + span - 1) / span);
    const int U = kb * nblk;
    const int u0 = (int)(((int64_t)wave * U) / nw), u1 = (int)(((int64_t)(wave + 1) * U) / nw);
    const int per_chain = d.n_cond * d.n_acc;
    int cur = -1;
    double acc = 0.0;
    for (int u = u0; u < u1; ++u) {
        const int c = u / nblk, blk = u - c * nblk;
        if (c != cur) {
            if (cur >= 0) {
                const double v = wave_sum(acc);
                if (lane == 0) part[wave * ng_stride + cb + cur] += v;
            }
            cur = c;
            acc = 0.0;
        }
        if (!active[cb + c]) continue;
        const AccK *kc = reinterpret_cast<const AccK *>(acck) + (size_t)c * per_chain;
        const int64_t base = t_lo + (int64_t)blk * span + lane;
#pragma unroll
        for (int t = 0; t < TPL; ++t) {
            const int64_t i = base + 64 * t;
            if (i < t_hi) {
                const int code = d.tcode[i];
                acc += trial_term<MODE>(kc + (code >> 3) * d.n_acc, d.n_acc, code & 7, d.rt[i]);
            }
        }
    }
    if (cur >= 0) {
        const double v = wave_sum(acc);
        if (lane == 0) part[wave * ng_stride + cb + cur] += v;
    }
}

/* ------------------------------------------------------------------ priors */
/* Phi(b) - Phi(a) for a <= b; as Q(a) - Q(b) when a > 0 (no cancellation). */
__device__ __forceinline__ double tnorm_mass(double a, double b) {
    return a > 0.0 ? DevOps::pnorm(-a) - DevOps::pnorm(-b) : DevOps::pnorm(b) - DevOps::pnorm(a);
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

/* Subject-level log prior of x for chain `chain`: the run's fixed prior, or
 * for hierarchical runs tnorm(loc = phi head, scale = phi tail) (src/de.cpp:599-600). */
__device__ double subject_logprior(const Dev &d, const JobDev &job, int chain, const double *x) {
    const int P = d.P;
    double s = 0.0;
    if (job.kind == GGDMC_KIND_HIER) {
        const double *ph = d.phi + ((size_t)job.phi_index * d.nchain + chain) * (2 * P);
        for (int p = 0; p < P; ++p)
            s += log_prior1(GGDMC_DIST_TNORM, ph[p], ph[P + p], d.pr_lo[job.prior_off + p],
                            d.pr_up[job.prior_off + p], x[p]);
    } else {
        for (int p = 0; p < P; ++p) {
            const int k = job.prior_off + p;
            s += log_prior1(d.pr_dist[k], d.pr_p0[k], d.pr_p1[k], d.pr_lo[k], d.pr_up[k], x[p]);
        }
    }
    return s;
}

__device__ __forceinline__ bool mh_accept(double tpost, double cur, double U) {
    const double ratio = ::exp(tpost - cur);
    if (ratio != ratio) return false;
    return U < ratio;
}

/* -------------------------------------------------------- LDS carve-up */
struct SweepLds {
    double *th, *lp, *ll, *prop, *tlp, *U, *part, *tile, *acck, *red;
    int *active, *sub, *flag;
};

__host__ __device__ inline size_t sweep_lds_bytes(int n, int P, int ng, int n_cond, int n_acc,
                                                  int chain_batch) {
    size_t dbl = (size_t)n * P + 2 * n + (size_t)ng * P + 3 * ng + (size_t)kWaves * ng +
                 (size_t)chain_batch * n_cond * n_acc * kAccKDoubles + kWaves;
    size_t ints = (size_t)ng + n + 4;
    return dbl * sizeof(double) + ints * sizeof(int);
}

__device__ SweepLds carve(char *base, int n, int P, int ng, int n_cond, int n_acc, int chain_batch) {
    SweepLds L;
    double *p = reinterpret_cast<double *>(base);
    L.th = p; p += (size_t)n * P;
    L.lp = p; p += n;
    L.ll = p; p += n;
    L.prop = p; p += (size_t)ng * P;
    L.tlp = p; p += ng;
    L.U = p; p += ng;
    L.tile = p; p += ng;
    L.part = p; p += (size_t)kWaves * ng;
    L.acck = p; p += (size_t)chain_batch * n_cond * n_acc * kAccKDoubles;
    L.red = p; p += kWaves;
    int *q = reinterpret_cast<int *>(p);
    L.active = q; q += ng;
    L.sub = q; q += n;
    L.flag = q;
    return L;
}

/* Fills the AccK tables of chains [cb, cb+kb) of the proposal batch. */
__device__ void build_acck(const Dev &d, const double *prop, int P, int cb, int kb, double *acck) {
    const int per_chain = d.n_cond * d.n_acc;
    AccK *K = reinterpret_cast<AccK *>(acck);
    for (int e = threadIdx.x; e < kb * per_chain; e += blockDim.x) {
        const int c = e / per_chain, r = e - c * per_chain;
        make_acck<DevOps>(prop + (size_t)(cb + c) * P, d.par_src + r * 6, d.par_const + r * 6, K[e]);
    }
}

/* Likelihood of proposals prop[0..ng) over trials [t_lo, t_hi); result per
 * chain in L.tile (summed over waves in fixed order). */
template <int MODE, int TPL>
__device__ void tile_loglike(const Dev &d, const SweepLds &L, int ng, int chain_batch,
                             int64_t t_lo, int64_t t_hi) {
    const int P = d.P;
    for (int e = threadIdx.x; e < kWaves * ng; e += blockDim.x) L.part[e] = 0.0;
    for (int cb = 0; cb < ng; cb += chain_batch) {
        const int kb = min(chain_batch, ng - cb);
        __syncthreads();
        build_acck(d, L.prop, P, cb, kb, L.acck);
        __syncthreads();
        eval_units<MODE, TPL>(d, t_lo, t_hi, kb, cb, L.active, L.acck, L.part, ng);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < ng; c += blockDim.x) {
        double v = 0.0;
        for (int w = 0; w < kWaves; ++w) v += L.part[w * ng + c];
        L.tile[c] = v;
    }
    __syncthreads();
}

__device__ void store_subject(const Dev &d, const SweepLds &L, int s, uint32_t iter) {
    if (iter % (uint32_t)d.thin != 0) return;
    const int k = (int)(iter / (uint32_t)d.thin);
    if (k >= d.nmc) return;
    const int n = d.nchain, P = d.P;
    double *ot = d.out_theta + ((size_t)s * d.nmc + k) * n * P;
    for (int e = threadIdx.x; e < n * P; e += blockDim.x) ot[e] = L.th[e];
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        d.out_lp[((size_t)s * d.nmc + k) * n + c] = L.lp[c];
        d.out_ll[((size_t)s * d.nmc + k) * n + c] = L.ll[c];
    }
}

/* Migration ring of subject s over all its trials (src/de.cpp:615-665,
 * :157-199). Runs in one workgroup; state in L is updated in place and in
 * global memory. */
template <int MODE, int TPL>
__device__ void subject_migration(const Dev &d, const SweepLds &L, const JobDev &job, int s,
                                  const Stream &st, int para, int chain_batch) {
    const int n = d.nchain, P = d.P;
    if (threadIdx.x == 0) L.flag[0] = subchains(st, n, L.sub);
    __syncthreads();
    const int ns = L.flag[0];
    const int64_t t_lo = d.subj_off[s], t_hi = d.subj_off[s + 1];
    for (int k = 0; k < ns; ++k) {
        const int src = L.sub[k], dst = L.sub[(k + 1) % ns];
        const uint32_t slot = kSlotMigrate | (uint32_t)k;
        if (threadIdx.x == 0) {
            double *x = L.prop;
            for (int p = 0; p < P; ++p) x[p] = L.th[src * P + p];
            if (para >= 0) {
                x[para] += jitter(st, slot, 0, d.rp);
            } else {
                for (int p = 0; p < P; ++p) x[p] += jitter(st, slot, p, d.rp);
            }
            const U4 b = st.block(slot, 0);
            L.U[0] = u53(b.x[0], b.x[1]);
            const double tlp = subject_logprior(d, job, src, x);
            L.tlp[0] = tlp;
            L.active[0] = tlp > -__builtin_inf();
        }
        __syncthreads();
        tile_loglike<MODE, TPL>(d, L, 1, chain_batch, t_lo, t_hi);
        if (threadIdx.x == 0) {
            const double tll = L.tile[0];
            const double cur = L.lp[dst] + L.ll[dst];
            if (L.active[0]) atomicAdd(d.evals, (unsigned long long)(t_hi - t_lo));
            if (mh_accept(L.tlp[0] + tll, cur, L.U[0])) {
                for (int p = 0; p < P; ++p) {
                    L.th[dst * P + p] = L.prop[p];
                    d.theta[((size_t)s * n + dst) * P + p] = L.prop[p];
                }
                L.lp[dst] = L.tlp[0];
                L.ll[dst] = tll;
                d.lp[(size_t)s * n + dst] = L.tlp[0];
                d.ll[(size_t)s * n + dst] = tll;
            }
        }
        __syncthreads();
    }
}

/* ------------------------------------------------------------ sweep kernel */
__device__ void fused_phi_tail(const Dev &d, int s, const JobDev &job, uint32_t iter,
                               char *lds_raw);

template <int MODE, int TPL>
__global__ __launch_bounds__(kBlock) void subject_sweep_kernel(Dev d, SweepArgs a) {
    extern __shared__ __attribute__((aligned(16))) char lds_raw[];
    const int tile = blockIdx.x;
    const int s = d.tile_subj[tile];
    const JobDev job = d.jobs[d.subj_job[s]];
    const int n = d.nchain, P = d.P, G = d.G;
    const int ng_max = (n + G - 1) / G + 1;
    const SweepLds L = carve(lds_raw, n, P, ng_max, d.n_cond, d.n_acc, a.chain_batch);
    const uint32_t iter = *d.iter_base + a.iter_off;
    const bool single = job.kind != GGDMC_KIND_HIER;
    const bool blocked = d.blocked_sub != 0;
    const int ntile = d.subj_ntile[s];
    const bool first_tile = tile == d.subj_tile0[s];

    /* level decision (src/de.cpp:208-236 single level, :333-372 subjects) */
    const int dsweep = (blocked && single) ? 0 : a.sweep;
    const Stream sd = make_stream(((uint64_t)job.k1 << 32) | job.k0, iter + d.philox_offset, (uint32_t)d.subj_unit[s],
                                  (uint32_t)dsweep);
    const bool migrate = decision_u(sd) < d.sub_mig;

    for (int e = threadIdx.x; e < n * P; e += blockDim.x) L.th[e] = d.theta[(size_t)s * n * P + e];
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        L.lp[c] = d.lp[(size_t)s * n + c];
        L.ll[c] = d.ll[(size_t)s * n + c];
    }
    __syncthreads();

    if (migrate) {
        if (!first_tile) return;
        const bool act = blocked && single ? a.sweep == 0 : true;
        if (act && a.g_begin == 0) {
            const int para = (blocked && !single) ? a.sweep : -1;
            subject_migration<MODE, TPL>(d, L, job, s, sd, para, a.chain_batch);
        }
        if (a.final_launch) {
            store_subject(d, L, s, iter);
            if (a.fuse_phi && job.kind == GGDMC_KIND_HIER) fused_phi_tail(d, s, job, iter, lds_raw);
        }
        return;
    }

    const int para = blocked ? a.sweep : -1;
    const Stream st = make_stream(((uint64_t)job.k1 << 32) | job.k0, iter + d.philox_offset, (uint32_t)d.subj_unit[s],
                                  (uint32_t)a.sweep);
    const int64_t t_lo = d.tile_lo[tile], t_hi = d.tile_hi[tile];
    const int64_t ntr = d.subj_off[s + 1] - d.subj_off[s];

    for (int g = a.g_begin; g < a.g_end; ++g) {
        const int lo = group_lo(g, n, G), hi = group_lo(g + 1, n, G);
        const int ng = hi - lo;
        /* proposals (src/de.cpp:583-597) */
        for (int c = threadIdx.x; c < ng; c += blockDim.x) {
            const int i = lo + c;
            int m0, m1;
            double U;
            partners(st, i, lo, hi, n, m0, m1, U);
            L.U[c] = U;
            double *x = L.prop + (size_t)c * P;
            const double *xi = L.th + (size_t)i * P;
            const double *x0 = L.th + (size_t)m0 * P, *x1 = L.th + (size_t)m1 * P;
            for (int p = 0; p < P; ++p) x[p] = xi[p];
            if (para >= 0) {
                x[para] += jitter(st, (uint32_t)i, 0, d.rp) + job.gamma * (x0[para] - x1[para]);
            } else {
                for (int p = 0; p < P; ++p)
                    x[p] += jitter(st, (uint32_t)i, p, d.rp) + job.gamma * (x0[p] - x1[p]);
            }
            const double tlp = subject_logprior(d, job, i, x);
            L.tlp[c] = tlp;
            L.active[c] = tlp > -__builtin_inf();
        }
        __syncthreads();
        tile_loglike<MODE, TPL>(d, L, ng, a.chain_batch, t_lo, t_hi);

        bool last = true;
        if (ntile > 1) {
            for (int c = threadIdx.x; c < ng; c += blockDim.x)
                d.partial[(size_t)tile * n + lo + c] = L.tile[c];
            __threadfence();
            __syncthreads();
            if (threadIdx.x == 0) {
                const uint32_t old = atomicAdd(d.counter + s, 1u);
                L.flag[1] = old == (uint32_t)(ntile - 1);
            }
            __syncthreads();
            last = L.flag[1] != 0;
            if (!last) return;
            __threadfence();
        }
        /* accept (src/de.cpp:81-108) by the last tile of the subject */
        for (int c = threadIdx.x; c < ng; c += blockDim.x) {
            const int i = lo + c;
            double tll = L.tile[c];
            if (ntile > 1) {
                tll = 0.0;
                const int t0 = d.subj_tile0[s];
                for (int t = 0; t < ntile; ++t) tll += d.partial[(size_t)(t0 + t) * n + i];
            }
            const double cur = L.lp[i] + L.ll[i];
            if (mh_accept(L.tlp[c] + tll, cur, L.U[c])) {
                for (int p = 0; p < P; ++p) {
                    const double v = L.prop[(size_t)c * P + p];
                    L.th[(size_t)i * P + p] = v;
                    d.theta[((size_t)s * n + i) * P + p] = v;
                }
                L.lp[i] = L.tlp[c];
                L.ll[i] = tll;
                d.lp[(size_t)s * n + i] = L.tlp[c];
                d.ll[(size_t)s * n + i] = tll;
            }
        }
        if (threadIdx.x == 0) {
            int na = 0;
            for (int c = 0; c < ng; ++c) na += L.active[c];
            atomicAdd(d.evals, (unsigned long long)na * (unsigned long long)ntr);
            if (ntile > 1) d.counter[s] = 0u;
        }
        __syncthreads();
    }
    if (a.final_launch) {
        store_subject(d, L, s, iter);
        if (a.fuse_phi && job.kind == GGDMC_KIND_HIER) fused_phi_tail(d, s, job, iter, lds_raw);
    }
}

/* ------------------------------------------------------------- phi level */
struct PhiLds {
    double *ph, *plp, *pll, *st, *prop, *U, *tlp, *red, *tm;
    int *sub, *flag;
};

__host__ __device__ inline size_t phi_lds_bytes(int n, int P) {
    const size_t dbl = (size_t)n * 2 * P + 2 * n + (size_t)n * P * 4 + (size_t)n * 2 * P + 2 * n +
                       kWaves + (size_t)2 * n * P;
    return dbl * sizeof(double) + (size_t)(n + 4) * sizeof(int);
}

__device__ PhiLds phi_carve(char *base, int n, int P) {
    PhiLds L;
    double *p = reinterpret_cast<double *>(base);
    L.ph = p; p += (size_t)n * 2 * P;
    L.plp = p; p += n;
    L.pll = p; p += n;
    L.st = p; p += (size_t)n * P * 4;
    L.prop = p; p += (size_t)n * 2 * P;
    L.U = p; p += n;
    L.tlp = p; p += n;
    L.red = p; p += kWaves;
    L.tm = p; p += (size_t)2 * n * P;
    int *q = reinterpret_cast<int *>(p);
    L.sub = q; q += n;
    L.flag = q;
    return L;
}

/* One population parameter's share of sumloghlike from its sufficient
 * statistics (mean, M2, min, max of the S subject values):
 *   sum_s log tnorm(x_s; mu, sg, lo, up)
 *     = -(M2 + S (mean - mu)^2) / (2 sg^2) - S (log sqrt(2 pi) + log sg + log Z)
 * NaN for sg <= 0 or NaN inputs, -inf when a subject value is outside
 * [lo, up]; summing the P shares reproduces the per-subject sum's NaN / -inf. */
__device__ double hll_term(const Dev &d, const JobDev &job, int p, double mu, double sg,
                           const double *st, int S) {
    const double mean = st[0], M2 = st[1], mn = st[2], mx = st[3];
    if (!(sg > 0.0) || mean != mean || mu != mu) return __builtin_nan("");
    const double lo = d.pr_lo[job.prior_off + p], up = d.pr_up[job.prior_off + p];
    if (mn < lo || mx > up) return -__builtin_inf();
    const double Z = tnorm_mass((lo - mu) / sg, (up - mu) / sg);
    const double dm = mean - mu;
    return -(M2 + (double)S * dm * dm) / (2.0 * sg * sg) -
           (double)S * (0.918938533204672741780329736406 + ::log(sg) + ::log(Z));
}

/* sumloghlike (src/de.cpp:245-270) from per-parameter sufficient statistics
 * of the S subject values x_s: sum_s log tnorm(x_s; mu, sg, lo, up)
 *   = -(M2 + S (mean - mu)^2) / (2 sg^2) - S (log sqrt(2 pi) + log sg + log Z). */
__device__ double hyper_ll(const Dev &d, const JobDev &job, const double *phi, const double *st,
                           int S) {
    double out = 0.0;
    for (int p = 0; p < d.P; ++p) out += hll_term(d, job, p, phi[p], phi[d.P + p], st + p * 4, S);
    return out;
}

__device__ double hyper_logprior(const Dev &d, const JobDev &job, const double *x) {
    double s = 0.0;
    for (int p = 0; p < 2 * d.P; ++p) {
        const int k = job.hprior_off + p;
        s += log_prior1(d.pr_dist[k], d.pr_p0[k], d.pr_p1[k], d.pr_lo[k], d.pr_up[k], x[p]);
    }
    return s;
}

/* Sufficient statistics of chain c over the run's subjects, two passes;
 * each group of 16 lanes reduces one parameter with a fixed shuffle tree. */
__device__ void chain_stats_range(const Dev &d, const JobDev &job, int c0, int c1, double *out) {
    const int P = d.P, n = d.nchain, S = job.nsubject;
    const int l16 = threadIdx.x & 15, ngrp = blockDim.x >> 4;
    for (int q = threadIdx.x >> 4; q < (c1 - c0) * P; q += ngrp) {
        const int c = c0 + q / P, p = q - (q / P) * P;
        double sum = 0.0, mn = __builtin_inf(), mx = -__builtin_inf();
        int nan = 0;
        for (int s = l16; s < S; s += 16) {
            const double x = d.theta[((size_t)(job.first_subject + s) * n + c) * P + p];
            sum += x;
            mn = fmin(mn, x);
            mx = fmax(mx, x);
            nan |= x != x;
        }
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) {
            sum += __shfl_xor(sum, off, 16);
            mn = fmin(mn, __shfl_xor(mn, off, 16));
            mx = fmax(mx, __shfl_xor(mx, off, 16));
            nan |= __shfl_xor(nan, off, 16);
        }
        const double mean = sum / (double)S;
        double m2 = 0.0;
        for (int s = l16; s < S; s += 16) {
            const double x = d.theta[((size_t)(job.first_subject + s) * n + c) * P + p];
            m2 += (x - mean) * (x - mean);
        }
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) m2 += __shfl_xor(m2, off, 16);
        if (l16 == 0) {
            double *o = out + ((size_t)(c - c0) * P + p) * 4;
            o[0] = nan ? __builtin_nan("") : mean;
            o[1] = m2;
            o[2] = mn;
            o[3] = mx;
        }
    }
}

__device__ void chain_stats(const Dev &d, const JobDev &job, int c, double *out) {
    chain_stats_range(d, job, c, c + 1, out);
}

/* Population step of one run in one workgroup: phi state in LDS, chain i
 * handled by thread i. */
__device__ void phi_update(const Dev &d, int f, const JobDev &job, uint32_t iter, bool store,
                           const PhiLds &L) {
    const int n = d.nchain, P = d.P, P2 = 2 * P, G = d.G;
    const bool hier = job.kind == GGDMC_KIND_HIER;
    const int S = job.nsubject;
    for (int e = threadIdx.x; e < n * P2; e += blockDim.x) L.ph[e] = d.phi[(size_t)f * n * P2 + e];
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        L.plp[c] = d.phi_lp[(size_t)f * n + c];
        L.pll[c] = d.phi_ll[(size_t)f * n + c];
    }
    const int nst = hier ? n : 1;
    for (int e = threadIdx.x; e < nst * P * 4; e += blockDim.x)
        L.st[e] = d.stats[(size_t)f * n * P * 4 + e];
    __syncthreads();
    const uint32_t unit = hier ? kUnitPhi : 0u;
    const double mig = hier ? d.pop_mig : d.sub_mig;
    const bool blocked = hier ? d.blocked_pop != 0 : d.blocked_sub != 0;
    const uint64_t seed = ((uint64_t)job.k1 << 32) | job.k0;
    const int nsweep = blocked ? P2 : 1;
    bool single_mig_done = false;
    for (int sw = 0; sw < nsweep; ++sw) {
        /* run_chains (single level) decides once; run_hchains per parameter */
        const int dsweep = (blocked && !hier) ? 0 : sw;
        const Stream sd = make_stream(seed, iter + d.philox_offset, unit, (uint32_t)dsweep);
        const bool migrate = decision_u(sd) < mig;
        if (migrate) {
            if (!hier && blocked) {
                if (single_mig_done || sw > 0) break;
                single_mig_done = true;
            }
            const int para = (blocked && hier) ? sw : -1;
            if (threadIdx.x == 0) {
                const int ns = subchains(sd, n, L.sub);
                for (int k = 0; k < ns; ++k) {
                    const int src = L.sub[k], dst = L.sub[(k + 1) % ns];
                    if (hier) {
                        L.pll[src] = hyper_ll(d, job, L.ph + (size_t)src * P2, L.st + (size_t)src * P * 4, S);
                        L.pll[dst] = hyper_ll(d, job, L.ph + (size_t)dst * P2, L.st + (size_t)dst * P * 4, S);
                    }
                    const uint32_t slot = kSlotMigrate | (uint32_t)k;
                    double *x = L.prop;
                    for (int p = 0; p < P2; ++p) x[p] = L.ph[(size_t)src * P2 + p];
                    if (para >= 0) {
                        x[para] += jitter(sd, slot, 0, d.rp);
                    } else {
                        for (int p = 0; p < P2; ++p) x[p] += jitter(sd, slot, p, d.rp);
                    }
                    const double tlp = hyper_logprior(d, job, x);
                    const double tll = hyper_ll(d, job, x, L.st + (size_t)(hier ? src : 0) * P * 4, S);
                    const double cur = L.plp[dst] + L.pll[dst];
                    const U4 b = sd.block(slot, 0);
                    if (mh_accept(tlp + tll, cur, u53(b.x[0], b.x[1]))) {
                        for (int p = 0; p < P2; ++p) L.ph[(size_t)dst * P2 + p] = x[p];
                        L.plp[dst] = tlp;
                        L.pll[dst] = tll;
                    }
                }
            }
            __syncthreads();
            if (!hier && blocked) break;
            continue;
        }
        const int para = blocked ? sw : -1;
        const Stream st = make_stream(seed, iter + d.philox_offset, unit, (uint32_t)sw);
        for (int g = 0; g < G; ++g) {
            const int lo = group_lo(g, n, G), hi = group_lo(g + 1, n, G), ng = hi - lo;
            /* proposals from partners outside the group (src/de.cpp:402-423) */
            for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
                int m0, m1;
                double U;
                partners(st, i, lo, hi, n, m0, m1, U);
                L.U[i] = U;
                double *x = L.prop + (size_t)i * P2;
                const double *x0 = L.ph + (size_t)m0 * P2, *x1 = L.ph + (size_t)m1 * P2;
                for (int p = 0; p < P2; ++p) x[p] = L.ph[(size_t)i * P2 + p];
                if (para >= 0) {
                    x[para] += jitter(st, (uint32_t)i, 0, d.rp) + job.gamma * (x0[para] - x1[para]);
                } else {
                    for (int p = 0; p < P2; ++p)
                        x[p] += jitter(st, (uint32_t)i, p, d.rp) + job.gamma * (x0[p] - x1[p]);
                }
                L.tlp[i] = hyper_logprior(d, job, x);
            }
            __syncthreads();
            /* per (chain, parameter): the refreshed current share (src/de.cpp:397)
             * and the proposal's share of sumloghlike */
            const int nt = ng * P;
            for (int e = threadIdx.x; e < (hier ? 2 : 1) * nt; e += blockDim.x) {
                const int which = e / nt, r = e - which * nt;
                const int i = lo + r / P, p = r - (r / P) * P;
                const double *sti = L.st + (size_t)(hier ? i : 0) * P * 4 + p * 4;
                const double *v = which == 0 ? L.prop + (size_t)i * P2 : L.ph + (size_t)i * P2;
                L.tm[(size_t)which * n * P + (size_t)i * P + p] = hll_term(d, job, p, v[p], v[P + p], sti, S);
            }
            __syncthreads();
            for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
                double tll = 0.0, ref = 0.0;
                for (int p = 0; p < P; ++p) {
                    tll += L.tm[(size_t)i * P + p];
                    if (hier) ref += L.tm[(size_t)n * P + (size_t)i * P + p];
                }
                if (hier) L.pll[i] = ref;
                const double cur = L.pll[i] + L.plp[i];
                if (mh_accept(L.tlp[i] + tll, cur, L.U[i])) {
                    for (int p = 0; p < P2; ++p) L.ph[(size_t)i * P2 + p] = L.prop[(size_t)i * P2 + p];
                    L.plp[i] = L.tlp[i];
                    L.pll[i] = tll;
                }
            }
            __syncthreads();
        }
    }
    for (int e = threadIdx.x; e < n * P2; e += blockDim.x) d.phi[(size_t)f * n * P2 + e] = L.ph[e];
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        d.phi_lp[(size_t)f * n + c] = L.plp[c];
        d.phi_ll[(size_t)f * n + c] = L.pll[c];
    }
    if (store && iter % (uint32_t)d.thin == 0) {
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

/* grid = n_phi * nchain: block (f, c) computes chain c's statistics of a
 * hierarchical run; the last block of each run performs its population step.
 * Hyper-only runs have fixed statistics: their block c == 0 steps directly. */
__global__ __launch_bounds__(kBlock) void phi_kernel(Dev d, uint32_t iter_off) {
    extern __shared__ __attribute__((aligned(16))) char lds_raw[];
    const int n = d.nchain, P = d.P;
    const int f = blockIdx.x / n, c = blockIdx.x - f * n;
    const JobDev job = d.jobs[d.phi_job[f]];
    const uint32_t iter = *d.iter_base + iter_off;
    const PhiLds L = phi_carve(lds_raw, n, P);
    if (job.kind != GGDMC_KIND_HIER) {
        if (c != 0) return;
        phi_update(d, f, job, iter, true, L);
        return;
    }
    chain_stats(d, job, c, d.stats + ((size_t)f * n + c) * P * 4);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t old = atomicAdd(d.counter + d.S + f, 1u);
        L.flag[0] = old == (uint32_t)(n - 1);
    }
    __syncthreads();
    if (!L.flag[0]) return;
    __threadfence();
    if (threadIdx.x == 0) d.counter[d.S + f] = 0u;
    phi_update(d, f, job, iter, true, L);
}

/* Called by the workgroup that finished subject s's iteration: the last
 * subject of the run to finish computes every chain's statistics and the
 * population step of the next iteration (the run's subjects are all done, so
 * no workgroup still reads the population state). */
__device__ void fused_phi_tail(const Dev &d, int s, const JobDev &job, uint32_t iter,
                               char *lds_raw) {
    __threadfence();
    __syncthreads();
    int *flag = reinterpret_cast<int *>(lds_raw);
    const int jidx = d.subj_job[s];
    uint32_t *ctr = d.counter + d.S + d.n_phi + jidx;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t old = atomicAdd(ctr, 1u);
        flag[0] = old == (uint32_t)(job.nsubject - 1);
    }
    __syncthreads();
    const bool last = flag[0] != 0;
    __syncthreads();
    if (!last) return;
    __threadfence();
    if (threadIdx.x == 0) *ctr = 0u;
    if (iter + 1 > d.last_iter) return;
    const int f = job.phi_index, n = d.nchain, P = d.P;
    chain_stats_range(d, job, 0, n, d.stats + (size_t)f * n * P * 4);
    __threadfence();
    __syncthreads();
    const PhiLds L = phi_carve(lds_raw, n, P);
    phi_update(d, f, job, iter + 1, true, L);
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
        atomicAdd(d.evals, (unsigned long long)(t_hi - t_lo));
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

/* Thinned store of every subject and population state (plain path). */
__global__ void store_kernel(Dev d, uint32_t iter) {
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
 * (the data of run_hyper do not change, so they are reduced once). */
__global__ void hyper_data_stats_kernel(Dev d, const double *rows, int S, int f) {
    __shared__ double red[kWaves];
    __shared__ double mnv[kBlock], mxv[kBlock];
    const int P = d.P;
    for (int p = 0; p < P; ++p) {
        double sum = 0.0, nanc = 0.0;
        for (int s = threadIdx.x; s < S; s += blockDim.x) {
            const double x = rows[(size_t)s * P + p];
            sum += x;
            nanc += x != x ? 1.0 : 0.0;
        }
        const double mean = block_sum(sum, red) / (double)S;
        const double nanv = block_sum(nanc, red);
        double m2 = 0.0, mn = __builtin_inf(), mx = -__builtin_inf();
        for (int s = threadIdx.x; s < S; s += blockDim.x) {
            const double x = rows[(size_t)s * P + p];
            m2 += (x - mean) * (x - mean);
            mn = fmin(mn, x);
            mx = fmax(mx, x);
        }
        m2 = block_sum(m2, red);
        mnv[threadIdx.x] = mn;
        mxv[threadIdx.x] = mx;
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = __builtin_inf(), b = -__builtin_inf();
            for (int t = 0; t < (int)blockDim.x; ++t) {
                a = fmin(a, mnv[t]);
                b = fmax(b, mxv[t]);
            }
            double *o = d.stats + (size_t)f * d.nchain * P * 4 + p * 4;
            o[0] = nanv > 0.0 ? __builtin_nan("") : mean;
            o[1] = m2;
            o[2] = a;
            o[3] = b;
        }
        __syncthreads();
    }
}

/* ----------------------------------------------------------- sumloglike */
template <int MODE, int TPL>
__global__ __launch_bounds__(kBlock) void loglike_kernel(Dev d, const double *thetas, int s,
                                                        double *out) {
    extern __shared__ __attribute__((aligned(16))) char lds_raw[];
    const int P = d.P;
    const SweepLds L = carve(lds_raw, 1, P, 1, d.n_cond, d.n_acc, 1);
    const int q = blockIdx.x;
    for (int p = threadIdx.x; p < P; p += blockDim.x) L.prop[p] = thetas[(size_t)q * P + p];
    if (threadIdx.x == 0) L.active[0] = 1;
    __syncthreads();
    tile_loglike<MODE, TPL>(d, L, 1, 1, d.subj_off[s], d.subj_off[s + 1]);
    if (threadIdx.x == 0) out[q] = L.tile[0];
}

} // namespace ggdmc

/* ------------------------------------------------------ launch wrappers */
using namespace ggdmc;

extern "C" {

size_t ggdmc_internal_sweep_lds(int n, int P, int ng, int n_cond, int n_acc, int chain_batch) {
    return sweep_lds_bytes(n, P, ng, n_cond, n_acc, chain_batch);
}
size_t ggdmc_internal_phi_lds(int n, int P) { return phi_lds_bytes(n, P); }

hipError_t ggdmc_internal_launch_sweep(const Dev *d, const SweepArgs *a, int mode, int tpl,
                                       size_t lds, hipStream_t stream) {
    dim3 grid(d->n_tiles), block(kBlock);
    if (mode == GGDMC_MIXED_BATCHED) {
        if (tpl == 2)
            subject_sweep_kernel<MODE_MIXED, 2><<<grid, block, lds, stream>>>(*d, *a);
        else
            subject_sweep_kernel<MODE_MIXED, 1><<<grid, block, lds, stream>>>(*d, *a);
    } else {
        subject_sweep_kernel<MODE_NAIVE, 1><<<grid, block, lds, stream>>>(*d, *a);
    }
    return hipGetLastError();
}

hipError_t ggdmc_internal_launch_phi(const Dev *d, uint32_t iter_off, size_t lds, hipStream_t stream) {
    phi_kernel<<<dim3(d->n_phi * d->nchain), dim3(kBlock), lds, stream>>>(*d, iter_off);
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

hipError_t ggdmc_internal_launch_store(const Dev *d, uint32_t iter, hipStream_t stream) {
    store_kernel<<<dim3(256), dim3(kBlock), 0, stream>>>(*d, iter);
    return hipGetLastError();
}

hipError_t ggdmc_internal_launch_hyper_stats(const Dev *d, const double *rows, int S, int f,
                                             hipStream_t stream) {
    hyper_data_stats_kernel<<<dim3(1), dim3(kBlock), 0, stream>>>(*d, rows, S, f);
    return hipGetLastError();
}

hipError_t ggdmc_internal_launch_loglike(const Dev *d, const double *thetas, int nq, int s,
                                         int mode, size_t lds, hipStream_t stream, double *out) {
    if (mode == GGDMC_MIXED_BATCHED)
        loglike_kernel<MODE_MIXED, 2><<<dim3(nq), dim3(kBlock), lds, stream>>>(*d, thetas, s, out);
    else
        loglike_kernel<MODE_NAIVE, 1><<<dim3(nq), dim3(kBlock), lds, stream>>>(*d, thetas, s, out);
    return hipGetLastError();
}

} // extern "C"
