/*
 * Host engine and C ABI of libggdmc_hip (see include/ggdmc_hip.h).
 *
 * A session uploads a batch of runs once (trials, designs, priors, start
 * states), keeps every state in HBM, and advances all runs together:
 *   batched path : per iteration one subject launch per (sweep, chain group)
 *                  over work items (subject, chain range, trial tile), one
 *                  for every group, one for every group and 32 iterations,
 *                  or one per sweep with one item per subject over every group; the
 *                  population step runs inside the subject launches (one
 *                  workgroup per chain of the launch's group) or, with one
 *                  item per subject or blocked subject sweeps, in its own
 *                  launch first; blocks of iterations are captured in a
 *                  hipGraph and replayed.
 *   plain path   : the reference's sequential order, one launch per chain
 *                  step, no graph (the FP64 comparator).
 * The reference's own driver loop is de_class::run_chains / run_hchains
 * (src/de.cpp:201-242, :272-383); the Rcpp glue it replaces is
 * src/de2R.cpp:8-171.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ggdmc_hip.h"
#include "engine.h"
#include "fp64_math.h"
#include "lba_math.h"

using namespace ggdmc;

extern "C" {
size_t ggdmc_internal_sweep_lds(int n, int P, int kmax, int per_chain, int chain_batch, int cmax,
                                int tstage);
size_t ggdmc_internal_phi_lds(int n, int P, int G, int stats_lds);
size_t ggdmc_internal_phi_chain_lds(int P, int max_subjects);
hipError_t ggdmc_internal_launch_sweep(const Dev *d, const SweepArgs *a, int mode, size_t lds, int occ,
                                       hipStream_t stream);
hipError_t ggdmc_internal_launch_phi(const Dev *d, uint32_t iter_off, size_t lds, int stats_lds,
                                     hipStream_t stream);
hipError_t ggdmc_internal_launch_plain_subject(const Dev *d, uint32_t iter, int sweep, int j,
                                               hipStream_t stream);
hipError_t ggdmc_internal_launch_plain_phi(const Dev *d, uint32_t iter, int sweep, int j,
                                           const double *rows, const int64_t *rows_off,
                                           hipStream_t stream);
hipError_t ggdmc_internal_launch_store(const Dev *d, uint32_t iter, int subjects_only,
                                      hipStream_t stream);
hipError_t ggdmc_internal_launch_hyper_stats(const Dev *d, const double *rows, int S, int f,
                                             hipStream_t stream);
hipError_t ggdmc_internal_launch_loglike(const Dev *d, const double *thetas, int nq, int s,
                                         int mode, int cmax, size_t lds, hipStream_t stream,
                                         double *out);
}

namespace {

thread_local std::string g_error;
thread_local std::vector<ggdmc_trace_record> g_trace;   /* of the thread's last one-shot run */
thread_local int64_t g_trace_dropped = 0;

constexpr uint32_t kTraceCap = 1u << 20;   /* decision records kept per advance */

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw Error(e == hipErrorOutOfMemory ? GGDMC_ENOMEM : GGDMC_EHIP,
                    std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHECK(x) hip_check((x), #x)

void require(bool ok, const std::string &msg) {
    if (!ok) throw Error(GGDMC_EINVAL, msg);
}

__global__ void set_base_kernel(uint32_t *p, uint32_t v) { *p = v; }

/* Device memory owned by a session; freed with it. */
struct Pool {
    std::vector<void *> ptrs;
    template <class T>
    T *alloc(size_t n) {
        void *p = nullptr;
        if (n == 0) n = 1;
        HIPCHECK(hipMalloc(&p, n * sizeof(T)));
        ptrs.push_back(p);
        return static_cast<T *>(p);
    }
    /* The copy has completed when this returns: callers pass host vectors
     * that may go out of scope right after. */
    template <class T>
    T *upload(const std::vector<T> &v, hipStream_t st) {
        T *p = alloc<T>(v.size());
        if (!v.empty()) {
            HIPCHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
            HIPCHECK(hipStreamSynchronize(st));
        }
        return p;
    }
    ~Pool() {
        for (void *p : ptrs) (void)hipFree(p);
    }
};

int device_cus() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return cus;
}

/* Items spanning every group wait, in their later groups, on items of their
 * subject with higher workgroup ids as well as lower ones. The items of a
 * subject lie within kXcd * ranges consecutive workgroups (XCD-aware order);
 * while that window fits one workgroup per CU every item it waits on is
 * resident or done, whatever else occupies the chip. */
constexpr int kXcd = 8;
bool span_fits(int ranges) { return kXcd * ranges <= device_cus(); }

int device_lds_limit() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 65536;
    return (int)std::min<size_t>(prop.sharedMemPerBlock, 65536);
}

} // namespace

/* ================================================================ session */
struct ggdmc_session {
    ggdmc_config cfg{};
    std::vector<ggdmc_job> jobs;
    Pool pool;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Dev d{};
    int mode = GGDMC_MIXED_BATCHED;
    /* work decomposition of the subject launches */
    bool fused = false;          /* one item per subject over every chain group */
    int kchain = 0;              /* chains per item */
    int tpt = 0;                 /* trials per item; 0 = whole subject */
    int kmax = 1, cmax = 1, chain_batch = 1, tstage = 0;
    std::vector<const int4 *> lists;   /* item list per launch (per group, or one) */
    std::vector<int> list_n;
    size_t sweep_lds = 0, phi_lds = 0;
    int phi_stats_lds = 0;        /* population kernel keeps the statistics in LDS */
    /* population step inside the subject launches (leading workgroups of
     * kBlock threads, one per (state, chain) of the launch's group) instead
     * of its own kernel: */
    bool overlap = false;
    bool chained = false;         /* every chain group in one subject launch */
    int ipl = 1;                  /* iterations per subject launch (chained only) */
    bool span = false;            /* chained: an item holds its chain range of every group */
    int occ = 3;                  /* waves per SIMD of the subject kernel (4: several
                                     iterations per launch only) */
    bool tiled = false;           /* some subject's trials split over several items */
    int rmax = 1;                 /* chain ranges per group (item_ready stride) */
    int ovl_stats_lds = 0;
    size_t ovl_lds = 0;           /* LDS of those launches: max(sweep, population) */
    int nsweep_sub = 1;
    int64_t total_chunks = 0;
    uint32_t next_iter = 1;      /* next local iteration to run (1-based) */
    uint32_t last_iter = 0;      /* nmc*thin - 1 */
    uint32_t *iter_base = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_iters = 0;
    /* a graph of the remainder of the last advance (fewer iterations than
     * graph_iters), captured before that advance's timed region */
    hipGraphExec_t rest_graph = nullptr;
    int rest_iters = 0;
    float last_ms = 0.f;
    double *hyper_rows = nullptr;
    int64_t *hyper_off = nullptr;
    std::vector<int64_t> subj_ntr;   /* host copy of trial counts */
    std::vector<int> subj_of_job;    /* first subject of each job */
    std::vector<int> phi_of_job;
    int64_t max_ntr = 0;
    std::vector<double> hyper_rows_h;
    std::vector<ggdmc_trace_record> trace;   /* decisions of the last advance, ordered */
    int64_t trace_dropped = 0;

    ~ggdmc_session() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (rest_graph) (void)hipGraphExecDestroy(rest_graph);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }

    void build(const ggdmc_config *c, const ggdmc_model *model, const ggdmc_job *js, int nj);
    void set_decomp(bool fused_new, int k_new, int tpt_new, bool chained_new = false, int ipl_new = 1,
                    bool span_new = false, int occ_new = 3);
    void default_decomp();
    void enqueue_iteration(uint32_t off, std::vector<hipEvent_t> *ev = nullptr, int count = 1);
    void flush_store();
    void run_batched(uint32_t n_iter);
    void run_plain(uint32_t n_iter);
    void autotune();
    void capture(int K);
    hipGraphExec_t make_graph(int K);
    void prepare_rest(uint32_t n_iter);
    /* a subject workgroup gave up waiting for the population step */
    void check_fault() {
        uint32_t f = 0;
        HIPCHECK(hipMemcpy(&f, d.fault, sizeof(f), hipMemcpyDeviceToHost));
        if (f != 0) throw Error(GGDMC_EHIP, "population step not published in time: results invalid");
    }
    /* the decisions recorded by the last advance, in the reference's print
     * order: run, iteration, level (population first), sweep, step */
    void collect_trace() {
        trace.clear();
        trace_dropped = 0;
        if (!d.trace) return;
        uint32_t n = 0;
        HIPCHECK(hipMemcpy(&n, d.trace_n, sizeof(n), hipMemcpyDeviceToHost));
        const uint32_t kept = std::min(n, d.trace_cap);
        trace.resize(kept);
        if (kept)
            HIPCHECK(hipMemcpy(trace.data(), d.trace, sizeof(ggdmc_trace_record) * kept, hipMemcpyDeviceToHost));
        trace_dropped = (int64_t)n - (int64_t)kept;
        std::stable_sort(trace.begin(), trace.end(), [](const ggdmc_trace_record &a, const ggdmc_trace_record &b) {
            if (a.job != b.job) return a.job < b.job;
            if (a.iteration != b.iteration) return a.iteration < b.iteration;
            if (a.subject != b.subject) return a.subject < b.subject;
            if (a.sweep != b.sweep) return a.sweep < b.sweep;
            return a.step < b.step;
        });
    }
    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_iters = 0;
        if (rest_graph) (void)hipGraphExecDestroy(rest_graph);
        rest_graph = nullptr;
        rest_iters = 0;
    }
};

/* Conditions of a design: cells whose per-accumulator parameter rows are
 * identical (in an LBA design they differ only in the response). Trials then
 * carry (condition, responding accumulator) and the per-proposal constants
 * are built once per (condition, accumulator). */
struct Condensed {
    int n_cond = 0;
    std::vector<int> cond_of_cell;
    std::vector<int32_t> src;
    std::vector<double> cst;
};

static Condensed condense(const ggdmc_model *m) {
    Condensed r;
    const int row = m->n_acc * 6;
    for (int c = 0; c < m->n_cell; ++c) {
        const int32_t *s = m->par_src + (size_t)c * row;
        const double *k = m->par_const + (size_t)c * row;
        int found = -1;
        for (int q = 0; q < r.n_cond && found < 0; ++q)
            if (std::equal(s, s + row, r.src.begin() + (size_t)q * row) &&
                std::equal(k, k + row, r.cst.begin() + (size_t)q * row))
                found = q;
        if (found < 0) {
            found = r.n_cond++;
            r.src.insert(r.src.end(), s, s + row);
            r.cst.insert(r.cst.end(), k, k + row);
        }
        r.cond_of_cell.push_back(found);
    }
    return r;
}

static void check_prior(const ggdmc_prior *p, int npar, const char *what) {
    require(p != nullptr, std::string(what) + " is null");
    require(p->npar == npar, std::string(what) + ": npar mismatch");
    require(p->dist && p->p0 && p->p1 && p->lower && p->upper, std::string(what) + ": null array");
    for (int k = 0; k < npar; ++k)
        require(p->dist[k] == GGDMC_DIST_UNIF || p->dist[k] == GGDMC_DIST_TNORM,
                std::string(what) + ": only dist_id 1 (tnorm) and 6 (unif) are implemented");
}

void ggdmc_session::build(const ggdmc_config *c, const ggdmc_model *model, const ggdmc_job *js,
                          int nj) {
    require(c != nullptr && js != nullptr && nj > 0, "config / jobs missing");
    cfg = *c;
    jobs.assign(js, js + nj);
    const int n = cfg.nchain;
    require(n > 2, "Require three or more chains.");
    require(n <= kMaxChains, "at most 256 chains per run");
    require(cfg.nmc >= 1 && cfg.thin >= 1, "nmc and thin must be >= 1");
    mode = cfg.precision;
    require(mode == GGDMC_FP64_PLAIN || mode == GGDMC_MIXED_BATCHED || mode == GGDMC_FP64_BATCHED,
            "unknown precision");
    int G = cfg.sweep_groups;
    if (G == 0) G = mode == GGDMC_FP64_PLAIN ? n : 2;
    require(G >= 2 && G <= n, "sweep_groups must be in [2, nchain]");
    require(mode != GGDMC_FP64_PLAIN || G == n, "the plain path runs the reference order (sweep_groups = nchain)");
    require(n - (n + G - 1) / G >= 2, "each chain group must leave two partners");

    bool any_lba = false;
    int P = -1;
    for (const auto &j : jobs) {
        require(j.kind >= GGDMC_KIND_SUBJECT && j.kind <= GGDMC_KIND_HIER, "unknown job kind");
        if (j.kind != GGDMC_KIND_HYPER) any_lba = true;
    }
    if (any_lba) {
        require(model != nullptr && model->winner && model->par_src && model->par_const, "model missing");
        require(model->n_acc >= 1 && model->n_cell >= 1 && model->npar >= 1, "empty model");
        require(model->n_acc <= 8, "at most 8 accumulators");
        require(model->n_cell <= 4096, "at most 4096 cells");
        require(model->npar <= 64, "at most 64 parameters");
        P = model->npar;
        for (int e = 0; e < model->n_cell * model->n_acc; ++e) {
            for (int f = 0; f < 6; ++f) {
                const int src = model->par_src[e * 6 + f];
                require(src >= -1 && src < P, "par_src out of range");
                if (f == 4) require(src == -1 && model->par_const[e * 6 + 4] == 0.0,
                                    "st0 must be the constant 0");
            }
        }
        for (int cidx = 0; cidx < model->n_cell; ++cidx)
            require(model->winner[cidx] >= 0 && model->winner[cidx] < model->n_acc, "winner out of range");
    }
    for (const auto &j : jobs) {
        if (j.kind == GGDMC_KIND_HYPER) {
            require(j.p_prior != nullptr, "hyper run needs p_prior");
            if (P < 0) P = j.p_prior->npar;
            require(j.p_prior->npar == P, "all runs must share the parameter count");
        }
    }
    require(P >= 1 && 2 * P <= 128, "parameter count out of range");

    HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHECK(hipEventCreate(&ev0));
    HIPCHECK(hipEventCreate(&ev1));

    /* ---- host-side flattening */
    std::vector<double> rt;
    std::vector<uint16_t> tcode;
    Condensed cond;
    if (any_lba) {
        cond = condense(model);
        require(cond.n_cond * model->n_acc <= 256, "design has more than 256 condition x accumulator pairs");
    }
    std::vector<int64_t> subj_off{0};
    std::vector<int32_t> subj_job, subj_unit;
    std::vector<JobDev> jd;
    std::vector<int32_t> phi_job;
    std::vector<int32_t> pr_dist;
    std::vector<double> pr_p0, pr_p1, pr_lo, pr_up;
    std::vector<double> th0, lp0, ll0, ph0, plp0, pll0;
    std::vector<int64_t> hyper_off_h;
    auto add_prior = [&](const ggdmc_prior *p) {
        const int off = (int)pr_dist.size();
        for (int k = 0; k < p->npar; ++k) {
            pr_dist.push_back(p->dist[k]);
            pr_p0.push_back(p->p0[k]);
            pr_p1.push_back(p->p1[k]);
            pr_lo.push_back(p->lower[k]);
            pr_up.push_back(p->upper[k]);
        }
        return off;
    };
    auto add_state = [&](const ggdmc_state *st, int np, std::vector<double> &th,
                         std::vector<double> &lp, std::vector<double> &ll) {
        require(st && st->theta && st->lp && st->ll, "start state missing");
        th.insert(th.end(), st->theta, st->theta + (size_t)np * n);
        lp.insert(lp.end(), st->lp, st->lp + n);
        ll.insert(ll.end(), st->ll, st->ll + n);
    };
    auto add_subject = [&](const ggdmc_subject *sb, int job, int unit) {
        require(sb && sb->cell_off && (sb->rt || sb->ntrial == 0), "subject data missing");
        require(sb->cell_off[0] == 0 && sb->cell_off[model->n_cell] == sb->ntrial,
                "cell_off must start at 0 and end at ntrial");
        for (int cidx = 0; cidx < model->n_cell; ++cidx) {
            require(sb->cell_off[cidx + 1] >= sb->cell_off[cidx], "cell_off must be non-decreasing");
            for (int64_t i = sb->cell_off[cidx]; i < sb->cell_off[cidx + 1]; ++i) {
                rt.push_back(sb->rt[i]);
                tcode.push_back((uint16_t)((cond.cond_of_cell[cidx] << 3) | model->winner[cidx]));
            }
        }
        subj_off.push_back((int64_t)rt.size());
        subj_job.push_back(job);
        subj_unit.push_back(unit);
        subj_ntr.push_back(sb->ntrial);
        max_ntr = std::max<int64_t>(max_ntr, sb->ntrial);
    };

    for (int ji = 0; ji < nj; ++ji) {
        const ggdmc_job &j = jobs[ji];
        JobDev J{};
        J.kind = j.kind;
        J.k0 = (uint32_t)j.seed;
        J.k1 = (uint32_t)(j.seed >> 32);
        const int npde = j.nparameter_de > 0 ? j.nparameter_de : (j.kind == GGDMC_KIND_SUBJECT ? P : 2 * P);
        J.gamma = cfg.gamma_precursor / std::sqrt(2.0 * npde);
        J.first_subject = (int)subj_job.size();
        J.phi_index = -1;
        subj_of_job.push_back(J.first_subject);
        if (j.kind == GGDMC_KIND_SUBJECT) {
            check_prior(j.p_prior, P, "p_prior");
            require(j.subjects && j.subject_starts && j.subject_outs, "subject run needs data, start, out");
            J.prior_off = add_prior(j.p_prior);
            J.hprior_off = 0;
            J.nsubject = 1;
            add_subject(&j.subjects[0], ji, 0);
            add_state(&j.subject_starts[0], P, th0, lp0, ll0);
        } else {
            check_prior(j.p_prior, P, "p_prior");
            check_prior(j.h_prior, 2 * P, "h_prior");
            for (int k = 0; k < P; ++k)
                require(j.p_prior->dist[k] == GGDMC_DIST_TNORM, "population likelihood needs tnorm p_prior");
            require(j.nsubject >= 1, "nsubject must be >= 1");
            require(j.phi_start && j.phi_out, "population start / out missing");
            J.prior_off = add_prior(j.p_prior);
            J.hprior_off = add_prior(j.h_prior);
            J.nsubject = j.nsubject;
            J.phi_index = (int)phi_job.size();
            phi_job.push_back(ji);
            add_state(j.phi_start, 2 * P, ph0, plp0, pll0);
            if (j.kind == GGDMC_KIND_HIER) {
                require(j.subjects && j.subject_starts && j.subject_outs, "hierarchical run needs subjects");
                for (int s = 0; s < j.nsubject; ++s) {
                    add_subject(&j.subjects[s], ji, s);
                    add_state(&j.subject_starts[s], P, th0, lp0, ll0);
                }
                hyper_off_h.push_back(0);
            } else {
                require(j.hyper_data != nullptr, "hyper run needs hyper_data");
                hyper_off_h.push_back((int64_t)hyper_rows_h.size());
                hyper_rows_h.insert(hyper_rows_h.end(), j.hyper_data, j.hyper_data + (size_t)j.nsubject * P);
            }
        }
        phi_of_job.push_back(J.phi_index);
        jd.push_back(J);
    }

    const int S = (int)subj_job.size();
    const int F = (int)phi_job.size();
    d.P = P;
    d.n_acc = any_lba ? model->n_acc : 1;
    d.n_cond = any_lba ? cond.n_cond : 1;
    d.nchain = n;
    d.G = G;
    d.nmc = cfg.nmc;
    d.thin = cfg.thin;
    d.blocked_sub = cfg.is_pblocked != 0;
    d.blocked_pop = cfg.is_hblocked != 0;
    d.rp = cfg.rp;
    d.sub_mig = cfg.sub_migration_prob;
    d.pop_mig = cfg.pop_migration_prob;
    d.S = S;
    d.n_phi = F;
    d.n_jobs = nj;
    d.philox_offset = (uint32_t)cfg.iter_offset;

    if (any_lba) {
        d.par_src = pool.upload(cond.src, stream);
        d.par_const = pool.upload(cond.cst, stream);
    }
    {
        std::vector<double> tab(64);
        fill_exp2_table(tab.data());
        d.exp2tab = pool.upload(tab, stream);
    }
    d.rt = pool.upload(rt, stream);
    d.tcode = pool.upload(tcode, stream);
    d.subj_off = pool.upload(subj_off, stream);
    d.subj_job = pool.upload(subj_job, stream);
    d.subj_unit = pool.upload(subj_unit, stream);
    d.jobs = pool.upload(jd, stream);
    d.phi_job = pool.upload(phi_job, stream);
    {
        std::vector<JobDev> sj, pj;
        for (int32_t j : subj_job) sj.push_back(jd[j]);
        for (int32_t j : phi_job) pj.push_back(jd[j]);
        d.subj_jobs = pool.upload(sj, stream);
        d.phi_jobs = pool.upload(pj, stream);
    }
    d.pr_dist = pool.upload(pr_dist, stream);
    d.pr_p0 = pool.upload(pr_p0, stream);
    d.pr_p1 = pool.upload(pr_p1, stream);
    d.pr_lo = pool.upload(pr_lo, stream);
    d.pr_up = pool.upload(pr_up, stream);
    {
        /* cached logarithms of the fixed priors: unif log(upper - lower);
         * tnorm log(scale), log(mass on [lower, upper]) with Phi as R's pnorm;
         * the kernels' own functions, so cached and direct terms agree */
        auto pnorm = [](double z) {
            return z == HUGE_VAL ? 1.0 : (z == -HUGE_VAL ? 0.0 : pnorm_acc(z));
        };
        std::vector<double> sc(2 * std::max<size_t>(pr_dist.size(), 1), 0.0);
        for (size_t k = 0; k < pr_dist.size(); ++k) {
            if (pr_dist[k] == GGDMC_DIST_UNIF) {
                sc[2 * k] = log_acc(pr_p1[k] - pr_p0[k]);
            } else {
                const double a = (pr_lo[k] - pr_p0[k]) / pr_p1[k], b = (pr_up[k] - pr_p0[k]) / pr_p1[k];
                const double Z = a > 0.0 ? pnorm(-a) - pnorm(-b) : pnorm(b) - pnorm(a);
                sc[2 * k] = log_acc(pr_p1[k]);
                sc[2 * k + 1] = log_acc(Z);
            }
        }
        d.scache = pool.upload(sc, stream);
    }
    d.theta = pool.upload(th0, stream);
    d.lp = pool.upload(lp0, stream);
    d.ll = pool.upload(ll0, stream);
    d.phi = pool.upload(ph0, stream);
    d.phi_lp = pool.upload(plp0, stream);
    d.phi_ll = pool.upload(pll0, stream);
    d.stats = pool.alloc<double>((size_t)std::max(F, 1) * n * P * 4);
    d.pcache = pool.alloc<double>((size_t)std::max(F, 1) * n * P * 2);
    d.pcache_valid = pool.alloc<int32_t>((size_t)std::max(F, 1) * n);
    HIPCHECK(hipMemsetAsync(d.pcache_valid, 0, sizeof(int32_t) * std::max(F, 1) * n, stream));
    {
        int max_sub = 1;
        for (const auto &j : jobs)
            if (j.kind == GGDMC_KIND_HIER) max_sub = std::max(max_sub, (int)j.nsubject);
        d.bstats_stride = (int64_t)n * P * ((max_sub + kStatBlock - 1) / kStatBlock) * 4;
        d.bstats = pool.alloc<double>((size_t)std::max(F, 1) * d.bstats_stride);
    }
    d.stats_ctr = pool.alloc<uint32_t>((size_t)std::max(F, 1));
    HIPCHECK(hipMemsetAsync(d.stats_ctr, 0, sizeof(uint32_t) * std::max(F, 1), stream));
    d.phi_ready = pool.alloc<uint32_t>((size_t)std::max(F, 1) * n);
    HIPCHECK(hipMemsetAsync(d.phi_ready, 0, sizeof(uint32_t) * std::max(F, 1) * n, stream));
    d.fault = pool.alloc<uint32_t>(1);
    HIPCHECK(hipMemsetAsync(d.fault, 0, sizeof(uint32_t), stream));
    d.trace = nullptr;
    d.trace_n = nullptr;
    d.trace_cap = 0;
    d.trace_pop = cfg.pop_debug != 0;
    d.trace_sub = cfg.sub_debug != 0;
    if (d.trace_pop || d.trace_sub) {
        d.trace = pool.alloc<ggdmc_trace_record>(kTraceCap);
        d.trace_n = pool.alloc<uint32_t>(1);
        d.trace_cap = kTraceCap;
        HIPCHECK(hipMemsetAsync(d.trace_n, 0, sizeof(uint32_t), stream));
    }
    d.evals = pool.alloc<unsigned long long>((size_t)std::max(S, 1));
    HIPCHECK(hipMemsetAsync(d.evals, 0, sizeof(unsigned long long) * std::max(S, 1), stream));
    {
        std::vector<int64_t> chunk_off{0};
        for (int x = 0; x < S; ++x) chunk_off.push_back(chunk_off.back() + (subj_ntr[x] + kChunk - 1) / kChunk);
        total_chunks = chunk_off.back();
        d.chunk_off = pool.upload(chunk_off, stream);
        d.csum = pool.alloc<double>((size_t)std::max<int64_t>(total_chunks, 1) * n);
    }
    iter_base = pool.alloc<uint32_t>(1);
    d.iter_base = iter_base;
    hyper_rows = pool.upload(hyper_rows_h, stream);
    hyper_off = pool.upload(hyper_off_h, stream);

    /* outputs, column 0 = start state */
    const size_t nmc = (size_t)cfg.nmc;
    d.out_theta = pool.alloc<double>(std::max<size_t>((size_t)S * nmc * n * P, 1));
    d.out_lp = pool.alloc<double>(std::max<size_t>((size_t)S * nmc * n, 1));
    d.out_ll = pool.alloc<double>(std::max<size_t>((size_t)S * nmc * n, 1));
    d.phi_out_theta = pool.alloc<double>(std::max<size_t>((size_t)F * nmc * n * 2 * P, 1));
    d.phi_out_lp = pool.alloc<double>(std::max<size_t>((size_t)F * nmc * n, 1));
    d.phi_out_ll = pool.alloc<double>(std::max<size_t>((size_t)F * nmc * n, 1));
    for (int s = 0; s < S; ++s) {
        HIPCHECK(hipMemcpyAsync(d.out_theta + (size_t)s * nmc * n * P, d.theta + (size_t)s * n * P,
                                sizeof(double) * n * P, hipMemcpyDeviceToDevice, stream));
        HIPCHECK(hipMemcpyAsync(d.out_lp + (size_t)s * nmc * n, d.lp + (size_t)s * n, sizeof(double) * n,
                                hipMemcpyDeviceToDevice, stream));
        HIPCHECK(hipMemcpyAsync(d.out_ll + (size_t)s * nmc * n, d.ll + (size_t)s * n, sizeof(double) * n,
                                hipMemcpyDeviceToDevice, stream));
    }
    for (int f = 0; f < F; ++f) {
        HIPCHECK(hipMemcpyAsync(d.phi_out_theta + (size_t)f * nmc * n * 2 * P, d.phi + (size_t)f * n * 2 * P,
                                sizeof(double) * n * 2 * P, hipMemcpyDeviceToDevice, stream));
        HIPCHECK(hipMemcpyAsync(d.phi_out_lp + (size_t)f * nmc * n, d.phi_lp + (size_t)f * n, sizeof(double) * n,
                                hipMemcpyDeviceToDevice, stream));
        HIPCHECK(hipMemcpyAsync(d.phi_out_ll + (size_t)f * nmc * n, d.phi_ll + (size_t)f * n, sizeof(double) * n,
                                hipMemcpyDeviceToDevice, stream));
    }
    for (int f = 0; f < F; ++f) {
        const ggdmc_job &j = jobs[phi_job[f]];
        if (j.kind == GGDMC_KIND_HYPER)
            HIPCHECK(ggdmc_internal_launch_hyper_stats(&d, hyper_rows + hyper_off_h[f], j.nsubject, f, stream));
    }

    nsweep_sub = cfg.is_pblocked ? P : 1;
    last_iter = (uint32_t)(cfg.nmc * cfg.thin - 1);
    d.last_iter = last_iter;
    next_iter = 1;
    phi_stats_lds = ggdmc_internal_phi_lds(n, P, G, 1) <= (size_t)device_lds_limit() ? 1 : 0;
    phi_lds = ggdmc_internal_phi_lds(n, P, G, phi_stats_lds);
    require(F == 0 || phi_lds <= (size_t)device_lds_limit(), "population state too large for LDS");
    default_decomp();
    HIPCHECK(hipStreamSynchronize(stream));
}

/* The decomposition the configuration asks for, or (auto) enough items to
 * fill the chip: chains per item so that subjects x chain ranges >= 1024. */
void ggdmc_session::default_decomp() {
    const int ngmax = (d.nchain + d.G - 1) / d.G;
    require(cfg.launch_mode >= 0 && cfg.launch_mode <= 4, "launch_mode must be 0, 1, 2, 3 or 4");
    const bool one = cfg.launch_mode >= 2;
    const int ipl_cfg = cfg.launch_mode >= 3 ? kItersPerLaunch : 1;
    const bool span_cfg = cfg.launch_mode == 4;
    require(cfg.waves_per_simd == 0 || cfg.waves_per_simd == 3 || cfg.waves_per_simd == 4,
            "waves_per_simd must be 0, 3 or 4");
    const int occ_cfg = cfg.waves_per_simd == 4 && ipl_cfg > 1 ? 4 : 3;
    if (cfg.chains_per_item < 0) {
        require(cfg.trials_per_tile == 0, "one item per subject needs trials_per_tile = 0");
        require(!one, "one item per subject already covers every group in one launch");
        set_decomp(true, 0, 0);
    } else if (cfg.chains_per_item > 0 || cfg.trials_per_tile > 0) {
        set_decomp(false, cfg.chains_per_item, cfg.trials_per_tile, one, ipl_cfg, span_cfg, occ_cfg);
    } else {
        int k = ngmax;
        while (k > 1 && (int64_t)d.S * ((ngmax + k - 1) / k) < 1024) k = (k + 1) / 2;
        set_decomp(false, k, 0, one, ipl_cfg, span_cfg, occ_cfg);
    }
}

/* Builds the item lists of the subject launches: per chain group, every
 * subject's chain ranges of k_new chains x trial tiles of tpt_new trials
 * (rounded up to a multiple of kChunk; 0 = whole subject); or, fused, one
 * item per subject that sweeps every group in one launch. Sizes the LDS. */
void ggdmc_session::set_decomp(bool fused_new, int k_new, int tpt_new, bool chained_new, int ipl_new,
                               bool span_new, int occ_new) {
    if (tpt_new > 0) tpt_new = ((tpt_new + kChunk - 1) / kChunk) * kChunk;
    require(!fused_new || tpt_new == 0, "one item per subject needs whole subjects");
    require(!(fused_new && chained_new), "one item per subject already covers every group");
    require(!chained_new || nsweep_sub == 1, "every group in one launch needs unblocked subject sweeps");
    require(ipl_new == 1 || chained_new, "several iterations per launch need every group in one launch");
    require(!span_new || (chained_new && tpt_new == 0),
            "items spanning every group need every group in one launch and whole subjects");
    const int n = d.nchain, G = d.G;
    const int ngmax = (n + G - 1) / G;
    fused = fused_new;
    chained = chained_new;
    ipl = ipl_new;
    span = span_new;
    occ = ipl_new > 1 ? occ_new : 3;
    tiled = false;
    tpt = tpt_new;
    d.tpt = tpt;
    kchain = fused ? ngmax : std::min(k_new > 0 ? k_new : ngmax, ngmax);
    lists.clear();
    list_n.clear();
    int64_t max_chunks_item = 1;
    const int nl = (fused || span) ? 1 : G;
    std::vector<int4> all_items;   /* every group's items, in group order (chained) */
    rmax = 1;
    /* XCD-aware order: workgroup w runs on XCD w % 8, so the items of subject
     * x go to workgroups w = x (mod 8) where possible, and share that XCD's
     * L2 for the subject's trials and state. The population workgroups come
     * first in a launch when the population step runs inside it. */
    const bool phi_in = d.n_phi > 0 && d.S > 0 && !fused && nsweep_sub == 1;
    int64_t wg_base = phi_in && chained ? (int64_t)d.n_phi * n : 0;
    auto xcd_order = [&](std::vector<int4> &v, int64_t base) {
        std::vector<std::vector<int4>> b(kXcd);
        for (const int4 &it : v) b[it.x % kXcd].push_back(it);
        std::vector<size_t> pos(kXcd, 0);
        std::vector<int4> out;
        out.reserve(v.size());
        for (size_t p = 0; p < v.size(); ++p) {
            const int want = (int)((base + (int64_t)p) % kXcd);
            int pick = -1;
            for (int dx = 0; dx < kXcd && pick < 0; ++dx) {
                const int c = (want + dx) % kXcd;
                if (pos[c] < b[c].size()) pick = c;
            }
            out.push_back(b[pick][pos[pick]++]);
        }
        v.swap(out);
    };
    for (int g = 0; g < nl; ++g) {
        const int ng = (fused || span) ? ngmax : (((g + 1) * n) / G - (g * n) / G);
        const int nch = fused ? 1 : (ng + kchain - 1) / kchain;
        std::vector<int4> items;
        for (int x = 0; x < d.S; ++x) {
            const int64_t ntr = subj_ntr[x];
            const int64_t nt = (tpt == 0 || ntr <= tpt) ? 1 : (ntr + tpt - 1) / tpt;
            require(nt <= 65535, "trials_per_tile too small for a subject's trial count");
            tiled = tiled || nt > 1;
            require((int64_t)nch * nt < (int64_t)1 << 30, "too many work items per subject");
            for (int ch = 0; ch < nch; ++ch) {
                const int c0 = ch * kchain;
                const int c1 = (fused || ch == nch - 1) ? kAllChains : c0 + kchain;
                for (int64_t t = 0; t < nt; ++t)
                    items.push_back(make_int4(x, (c0 << 16) | c1, ((int)t << 16) | (int)nt, (g << 16) | ch));
            }
            max_chunks_item = std::max<int64_t>(max_chunks_item,
                                                nt == 1 ? (ntr + kChunk - 1) / kChunk : tpt / kChunk);
        }
        rmax = std::max(rmax, nch);
        require(!span || span_fits(nch), "too many chain ranges per subject for items spanning every group");
        if (chained) {
            xcd_order(items, wg_base);
            wg_base += (int64_t)items.size();
            all_items.insert(all_items.end(), items.begin(), items.end());
            continue;
        }
        xcd_order(items, phi_in ? (int64_t)d.n_phi * ng : 0);
        require(items.size() < ((size_t)1 << 31), "too many work items");
        list_n.push_back((int)items.size());
        lists.push_back(pool.upload(items, stream));
    }
    if (chained) {
        require(all_items.size() < ((size_t)1 << 31), "too many work items");
        list_n.push_back((int)all_items.size());
        lists.push_back(pool.upload(all_items, stream));
        const size_t nf = (size_t)std::max(d.S, 1) * G * rmax;
        d.item_ready = pool.alloc<uint32_t>(nf);
        HIPCHECK(hipMemsetAsync(d.item_ready, 0, sizeof(uint32_t) * nf, stream));
        d.item_staged = pool.alloc<uint32_t>(nf);
        HIPCHECK(hipMemsetAsync(d.item_staged, 0, sizeof(uint32_t) * nf, stream));
    }
    {
        const size_t nctr = (size_t)std::max(d.S, 1) * G * rmax;
        d.tile_ctr = pool.alloc<uint32_t>(nctr);
        HIPCHECK(hipMemsetAsync(d.tile_ctr, 0, sizeof(uint32_t) * nctr, stream));
    }
    /* LDS: accumulator constants of chain_batch chains and cmax chunk sums
     * per chain at once, within 48 KB when possible */
    kmax = kchain;
    cmax = (int)std::min<int64_t>(max_chunks_item, std::max(1, 4096 / kmax));
    chain_batch = kmax;
    const int per_chain = d.n_cond * d.n_acc;
    /* the item's trials in LDS when they fit in 24 KB */
    int64_t max_item_trials = 0;
    for (int x = 0; x < d.S; ++x)
        max_item_trials = std::max<int64_t>(max_item_trials,
                                            tpt == 0 ? subj_ntr[x] : std::min<int64_t>(tpt, subj_ntr[x]));
    tstage = max_item_trials * 10 <= 24 * 1024 ? (int)max_item_trials : 0;
    auto bytes = [&]() {
        return ggdmc_internal_sweep_lds(n, d.P, kmax, per_chain, chain_batch, cmax, tstage);
    };
    const size_t soft = 48 * 1024, limit = (size_t)device_lds_limit();
    while (chain_batch > 1 && bytes() > soft) --chain_batch;
    while (cmax > 1 && bytes() > soft) cmax = (cmax + 1) / 2;
    if (bytes() > soft) tstage = 0;
    sweep_lds = bytes();
    require(sweep_lds <= limit, "design too large for LDS (chains x parameters)");
    /* the population step of each chain group runs in that group's subject
     * launch (not with one item per subject over every group, nor with
     * blocked subject sweeps: those keep the population kernel) */
    overlap = false;
    if (d.n_phi > 0 && d.S > 0 && !fused && nsweep_sub == 1) {
        const int max_sub = (int)(d.bstats_stride / 4 / ((int64_t)n * d.P)) * kStatBlock;
        /* the whole-run population step (its LDS holds the whole population
         * state) runs in these launches only for hyper-only runs, blocked
         * population sweeps and population migration; without them every
         * workgroup needs only the per-chain step's or an item's LDS */
        bool whole_run = d.blocked_pop != 0 || d.pop_mig > 0.0;
        for (const ggdmc_job &j : jobs) whole_run = whole_run || j.kind == GGDMC_KIND_HYPER;
        ovl_stats_lds = ggdmc_internal_phi_lds(n, d.P, G, 1) <= limit ? 1 : 0;
        ovl_lds = std::max({sweep_lds, whole_run ? ggdmc_internal_phi_lds(n, d.P, G, ovl_stats_lds) : (size_t)0,
                            ggdmc_internal_phi_chain_lds(d.P, max_sub)});
        overlap = ovl_lds <= limit;
    }
    require(ipl == 1 || d.n_phi == 0 || overlap,
            "several iterations per launch need the population step inside the subject launch");
    drop_graph();
}

/* One iteration's launches: the population step of the iteration (from the
 * subject states of the previous one), then one subject launch per (sweep,
 * chain group), or per sweep with one item per subject over every group. The
 * first subject launch stores the previous iteration's subject states when
 * due. ev, if given, receives an event pair around every subject launch. */
void ggdmc_session::enqueue_iteration(uint32_t off, std::vector<hipEvent_t> *ev, int count) {
    if (d.n_phi > 0 && !overlap) HIPCHECK(ggdmc_internal_launch_phi(&d, off, phi_lds, phi_stats_lds, stream));
    if (d.S == 0) return;
    const int lk = mode == GGDMC_FP64_BATCHED ? GGDMC_FP64_BATCHED : GGDMC_MIXED_BATCHED;
    const int ngl = (fused || chained) ? 1 : d.G;
    for (int sw = 0; sw < nsweep_sub; ++sw) {
        for (int g = 0; g < ngl; ++g) {
            /* every group's launch of the first sweep stores its own chains */
            const bool store = sw == 0;
            const bool with_phi = overlap && sw == 0;
            /* chains of the launch: group g, or every group */
            const int nc = chained ? d.nchain : ((g + 1) * d.nchain) / d.G - (g * d.nchain) / d.G;
            const int pw = with_phi ? d.n_phi * nc : 0;
            /* several iterations (count) per launch: segments of a multiple of
             * 8 workgroups, so that every segment maps items to XCDs alike */
            const int multi = chained && ipl > 1 ? count : 1;
            const int seg_wgs = multi > 1 ? (pw + list_n[g] + 7) / 8 * 8 : pw + list_n[g];
            SweepArgs a{off, sw, (fused || chained) ? 0 : g, (fused || chained) ? d.G : g + 1, (int)store,
                        chain_batch, kmax, cmax, tstage, list_n[g], lists[g],
                        pw, ovl_stats_lds, (int)chained, rmax, (int)tiled, multi, seg_wgs, (int)span};
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (ev) {
                HIPCHECK(hipEventCreate(&e0));
                HIPCHECK(hipEventCreate(&e1));
                ev->push_back(e0);
                ev->push_back(e1);
                HIPCHECK(hipEventRecord(e0, stream));
            }
            HIPCHECK(ggdmc_internal_launch_sweep(&d, &a, lk, with_phi ? ovl_lds : sweep_lds, occ, stream));
            if (ev) HIPCHECK(hipEventRecord(e1, stream));
        }
    }
}

/* Stores the subject states of the last iteration run (deferred in the
 * kernels to the next iteration's first-sweep launches) when that column is due. */
void ggdmc_session::flush_store() {
    const uint32_t it = next_iter - 1;
    if (it >= 1 && it % (uint32_t)cfg.thin == 0 && d.S > 0)
        HIPCHECK(ggdmc_internal_launch_store(&d, it, 1, stream));
}

/* The launches of K iterations (iteration base from the device word) as an
 * instantiated graph. */
hipGraphExec_t ggdmc_session::make_graph(int K) {
    hipGraph_t g = nullptr;
    HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
        const int step = chained ? ipl : 1;
        for (int k = 0; k < K; k += step) enqueue_iteration((uint32_t)k, nullptr, std::min(step, K - k));
    } catch (...) {
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    HIPCHECK(hipStreamEndCapture(stream, &g));
    hipGraphExec_t x = nullptr;
    const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHECK(e);
    return x;
}

void ggdmc_session::capture(int K) {
    drop_graph();
    if (K < 2) return;
    graph = make_graph(K);
    graph_iters = K;
}

/* An advance of n_iter iterations replays the K-iteration graph while it
 * can; the rest (fewer than K iterations, as in a short run or a restart's
 * pieces) gets a graph of its own, captured here, before the advance's timed
 * region, and kept for the next advance of the same remainder. */
void ggdmc_session::prepare_rest(uint32_t n_iter) {
    if (!graph) return;
    const int r = (int)(n_iter % (uint32_t)graph_iters);
    if (r < 2 || r == rest_iters) return;
    if (rest_graph) (void)hipGraphExecDestroy(rest_graph);
    rest_graph = nullptr;
    rest_iters = 0;
    rest_graph = make_graph(r);
    rest_iters = r;
}

void ggdmc_session::run_batched(uint32_t n_iter) {
    if (n_iter == 0) return;
    while (n_iter > 0) {
        if (graph && n_iter >= (uint32_t)graph_iters) {
            set_base_kernel<<<1, 1, 0, stream>>>(iter_base, next_iter);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipGraphLaunch(graph, stream));
            next_iter += graph_iters;
            n_iter -= graph_iters;
        } else if (rest_graph && n_iter == (uint32_t)rest_iters) {
            set_base_kernel<<<1, 1, 0, stream>>>(iter_base, next_iter);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipGraphLaunch(rest_graph, stream));
            next_iter += rest_iters;
            n_iter = 0;
        } else {
            set_base_kernel<<<1, 1, 0, stream>>>(iter_base, next_iter);
            HIPCHECK(hipGetLastError());
            const int cnt = (int)std::min<uint32_t>(chained ? (uint32_t)ipl : 1u, n_iter);
            enqueue_iteration(0, nullptr, cnt);
            next_iter += (uint32_t)cnt;
            n_iter -= (uint32_t)cnt;
        }
    }
    flush_store();
}

void ggdmc_session::run_plain(uint32_t n_iter) {
    int nsw_phi = 0;
    if (d.n_phi > 0) {
        const bool any_blocked = cfg.is_hblocked || cfg.is_pblocked;
        nsw_phi = any_blocked ? 2 * d.P : 1;
    }
    for (uint32_t k = 0; k < n_iter; ++k, ++next_iter) {
        for (int sw = 0; sw < nsw_phi; ++sw)
            for (int j = 0; j < d.nchain; ++j)
                HIPCHECK(ggdmc_internal_launch_plain_phi(&d, next_iter, sw, j, hyper_rows, hyper_off, stream));
        if (d.S > 0)
            for (int sw = 0; sw < nsweep_sub; ++sw)
                for (int j = 0; j < d.nchain; ++j)
                    HIPCHECK(ggdmc_internal_launch_plain_subject(&d, next_iter, sw, j, stream));
        HIPCHECK(ggdmc_internal_launch_store(&d, next_iter, 0, stream));
    }
}

/* Picks the work decomposition by timing graph replays of a few real
 * iterations of each candidate from a snapshot of the state, which is then
 * restored (results do not depend on the decomposition, only the time does). */
void ggdmc_session::autotune() {
    if (d.S == 0) return;
    struct Cand {
        bool fused;
        int k, tpt;
        bool chained;
        int ipl;
        bool span;
        int occ;
    };
    const int ngmax = (d.nchain + d.G - 1) / d.G;
    std::vector<Cand> cands{{true, 0, 0, false, 1, false, 3}};
    std::vector<int> ks;
    for (int k : {1, 2, 4, 8, ngmax})
        if (k <= ngmax && std::find(ks.begin(), ks.end(), k) == ks.end()) ks.push_back(k);
    const bool can_chain = nsweep_sub == 1 && cfg.launch_mode != 1;
    for (int one = 0; one < (can_chain ? 2 : 1); ++one)
        for (int k : ks)
            for (int t : {0, 256, 1024})
                if (t == 0 || t < max_ntr) {
                    const int64_t nt = t == 0 ? 1 : (max_ntr + t - 1) / t;
                    if ((int64_t)d.S * ((ngmax + k - 1) / k) * nt <= ((int64_t)1 << 22)) {
                        cands.push_back({false, k, t, one == 1, 1, false, 3});
                        /* several iterations per launch: whole subjects, the
                         * population step inside the launch; items of one
                         * group, or spanning every group */
                        if (one == 1 && t == 0 && cfg.launch_mode != 2) {
                            /* the subject kernel at 3 and at 4 waves per SIMD */
                            for (int oc : {3, 4}) {
                                if (cfg.waves_per_simd != 0 && oc != cfg.waves_per_simd) continue;
                                cands.push_back({false, k, t, true, kItersPerLaunch, false, oc});
                                if (span_fits((ngmax + k - 1) / k))
                                    cands.push_back({false, k, t, true, kItersPerLaunch, true, oc});
                            }
                        }
                    }
                }
    const size_t nth = (size_t)d.S * d.nchain * d.P, nch = (size_t)d.S * d.nchain;
    const size_t nph = (size_t)d.n_phi * d.nchain * 2 * d.P, npc = (size_t)d.n_phi * d.nchain;
    const size_t nev = (size_t)std::max(d.S, 1);
    double *snap = pool.alloc<double>(nth + 2 * nch + nph + 2 * npc);
    unsigned long long *ev_snap = pool.alloc<unsigned long long>(nev);
    auto copy = [&](void *dst, const void *src, size_t bytes) {
        if (bytes) HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream));
    };
    auto save = [&]() {
        copy(snap, d.theta, nth * 8);
        copy(snap + nth, d.lp, nch * 8);
        copy(snap + nth + nch, d.ll, nch * 8);
        copy(snap + nth + 2 * nch, d.phi, nph * 8);
        copy(snap + nth + 2 * nch + nph, d.phi_lp, npc * 8);
        copy(snap + nth + 2 * nch + nph + npc, d.phi_ll, npc * 8);
        copy(ev_snap, d.evals, nev * 8);
        HIPCHECK(hipStreamSynchronize(stream));
    };
    auto restore = [&]() {
        copy(d.theta, snap, nth * 8);
        copy(d.lp, snap + nth, nch * 8);
        copy(d.ll, snap + nth + nch, nch * 8);
        copy(d.phi, snap + nth + 2 * nch, nph * 8);
        copy(d.phi_lp, snap + nth + 2 * nch + nph, npc * 8);
        copy(d.phi_ll, snap + nth + 2 * nch + nph + npc, npc * 8);
        copy(d.evals, ev_snap, nev * 8);
        if (d.n_phi > 0) {
            const size_t nf = (size_t)d.n_phi * d.nchain;
            HIPCHECK(hipMemsetAsync(d.pcache_valid, 0, sizeof(int32_t) * nf, stream));
            /* the iterations run again: their population steps are not done */
            HIPCHECK(hipMemsetAsync(d.phi_ready, 0, sizeof(uint32_t) * nf, stream));
        }
        if (chained) {
            const size_t nf = (size_t)std::max(d.S, 1) * d.G * rmax;
            HIPCHECK(hipMemsetAsync(d.item_ready, 0, sizeof(uint32_t) * nf, stream));
            HIPCHECK(hipMemsetAsync(d.item_staged, 0, sizeof(uint32_t) * nf, stream));
        }
        HIPCHECK(hipStreamSynchronize(stream));
    };
    save();
    const uint32_t it0 = next_iter;
    float best = 1e30f;
    Cand pick = cands[0];
    /* iterations per timed replay: one launch of launch modes 3 and 4 when
     * the run has room for three (warm-up and two timed) */
    const int K = last_iter >= 3 * (uint32_t)kItersPerLaunch ? kItersPerLaunch : 16;
    for (const Cand &c : cands) {
        try {
            set_decomp(c.fused, c.k, c.tpt, c.chained, c.ipl, c.span, c.occ);
        } catch (const Error &) {
            if (c.ipl == 1) throw;
            continue;   /* no population step inside the launch for this design */
        }
        next_iter = it0;
        capture(K);
        run_batched(K); /* warm-up */
        /* the faster of two timed replays */
        float ms = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
            HIPCHECK(hipEventRecord(ev0, stream));
            run_batched(K);
            HIPCHECK(hipEventRecord(ev1, stream));
            HIPCHECK(hipEventSynchronize(ev1));
            float m = 0.f;
            HIPCHECK(hipEventElapsedTime(&m, ev0, ev1));
            ms = std::min(ms, m);
        }
        drop_graph();
        restore();
        if (ms < best) {
            best = ms;
            pick = c;
        }
    }
    next_iter = it0;
    set_decomp(pick.fused, pick.k, pick.tpt, pick.chained, pick.ipl, pick.span, pick.occ);
}

static int fail(const std::exception &e) {
    if (const Error *er = dynamic_cast<const Error *>(&e)) {
        g_error = er->what();
        return er->code;
    }
    g_error = e.what();
    return GGDMC_EINVAL;
}

static int invalid(const char *msg) {
    g_error = msg;
    return GGDMC_EINVAL;
}

/* Staging array of ggdmc_loglike_noalloc: a fixed device array (every GPU the
 * library runs on holds its own copy until the process ends), so that the
 * call needs no device allocation. In host code `scratch` names a host-side
 * stand-in, not a device address: the device address of each GPU's copy comes
 * from hipGetSymbolAddress, taken once per device with that device current,
 * and only that address is used. Calls on one GPU share the array; a
 * per-device lock serialises them (overlapping calls would race on it). */
constexpr size_t kStagingBytes = (size_t)16 << 20;
static __device__ __align__(256) char scratch[kStagingBytes];

struct Staging {
    std::mutex mu;       /* held for a whole call */
    char *ptr = nullptr; /* this device's copy of scratch */
};

static Staging &staging_for_current_device() {
    static std::mutex table_mu;
    static std::vector<std::unique_ptr<Staging>> table;
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(table_mu);
    if ((int)table.size() <= dev) table.resize(dev + 1);
    if (!table[dev]) {
        auto st = std::make_unique<Staging>();
        void *p = nullptr;
        HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, "staging array has no device address");
        st->ptr = static_cast<char *>(p);
        table[dev] = std::move(st);
    }
    return *table[dev];
}

static int loglike_impl(const ggdmc_model *model, const ggdmc_subject *data, const double *theta,
                        int32_t n, int32_t precision, double *ll_out, bool noalloc);

/* ================================================================= C ABI */
extern "C" {

const char *ggdmc_last_error(void) { return g_error.c_str(); }

int ggdmc_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    int ok = 0;
    for (int i = 0; i < n; ++i) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, i) == hipSuccess && std::strncmp(p.gcnArchName, "gfx950", 6) == 0) ++ok;
    }
    return ok;
}

int ggdmc_session_create(const ggdmc_config *cfg, const ggdmc_model *model, const ggdmc_job *jobs,
                         int32_t njobs, ggdmc_session **out) {
    if (!out) {
        g_error = "out is null";
        return GGDMC_EINVAL;
    }
    *out = nullptr;
    if (ggdmc_device_count() < 1) {
        g_error = "no gfx950 device visible";
        return GGDMC_ENODEV;
    }
    auto s = std::make_unique<ggdmc_session>();
    try {
        s->build(cfg, model, jobs, njobs);
        if (s->mode != GGDMC_FP64_PLAIN) {
            /* tune the tile shape when the run is long enough to repay it */
            if (cfg->trials_per_tile == 0 && cfg->chains_per_item == 0 && cfg->launch_mode < 2 &&
                s->last_iter >= 64)
                s->autotune();
            const uint32_t total = s->last_iter;
            const int K = (int)std::min<uint32_t>(64u, total);
            s->capture(K);
        }
        HIPCHECK(hipStreamSynchronize(s->stream));
    } catch (const std::exception &e) {
        return fail(e);
    }
    *out = s.release();
    return GGDMC_OK;
}

int ggdmc_session_advance(ggdmc_session *s, int32_t n_iter) {
    if (!s || n_iter < 0) {
        g_error = "bad session / n_iter";
        return GGDMC_EINVAL;
    }
    try {
        const uint32_t left = s->next_iter <= s->last_iter ? s->last_iter - s->next_iter + 1 : 0;
        const uint32_t n = std::min<uint32_t>((uint32_t)n_iter, left);
        if (s->d.trace) HIPCHECK(hipMemsetAsync(s->d.trace_n, 0, sizeof(uint32_t), s->stream));
        if (s->mode != GGDMC_FP64_PLAIN) s->prepare_rest(n);
        HIPCHECK(hipEventRecord(s->ev0, s->stream));
        if (s->mode == GGDMC_FP64_PLAIN) s->run_plain(n);
        else s->run_batched(n);
        HIPCHECK(hipEventRecord(s->ev1, s->stream));
        HIPCHECK(hipEventSynchronize(s->ev1));
        HIPCHECK(hipEventElapsedTime(&s->last_ms, s->ev0, s->ev1));
        s->check_fault();
        s->collect_trace();
    } catch (const std::exception &e) {
        return fail(e);
    }
    return GGDMC_OK;
}

int ggdmc_session_elapsed_ms(ggdmc_session *s, float *ms) {
    if (!s || !ms) return invalid("session / ms is null");
    *ms = s->last_ms;
    return GGDMC_OK;
}

int ggdmc_session_waves_per_simd(ggdmc_session *s, int32_t *waves_per_simd) {
    if (!s || !waves_per_simd) return invalid("session / waves_per_simd is null");
    *waves_per_simd = s->occ;
    return GGDMC_OK;
}

int ggdmc_session_decomposition(ggdmc_session *s, int32_t *chains_per_item,
                                int32_t *trials_per_tile, int32_t *items_per_launch,
                                int32_t *launch_mode) {
    if (!s) return invalid("session is null");
    if (launch_mode) *launch_mode = s->span ? 4 : (s->ipl > 1 ? 3 : (s->chained ? 2 : 1));
    if (chains_per_item) *chains_per_item = s->fused ? -1 : s->kchain;
    if (trials_per_tile) *trials_per_tile = s->tpt;
    if (items_per_launch)
        *items_per_launch = s->list_n.empty() ? 0 : *std::max_element(s->list_n.begin(), s->list_n.end());
    return GGDMC_OK;
}

int ggdmc_session_trial_evals(ggdmc_session *s, double *evals) {
    if (!s || !evals) return invalid("session / evals is null");
    try {
        const size_t ns = (size_t)std::max(s->d.S, 1);
        std::vector<unsigned long long> v(ns);
        HIPCHECK(hipMemcpyAsync(v.data(), s->d.evals, 8 * ns, hipMemcpyDeviceToHost, s->stream));
        HIPCHECK(hipStreamSynchronize(s->stream));
        double tot = 0.0;
        for (unsigned long long e : v) tot += (double)e;
        *evals = tot;
    } catch (const std::exception &e) {
        return fail(e);
    }
    return GGDMC_OK;
}

int ggdmc_session_profile(ggdmc_session *s, int32_t n_iter, double *subject_kernel_ms,
                          int32_t *n_launches) {
    if (!s || !subject_kernel_ms || n_iter < 1) return invalid("bad session / output / n_iter");
    try {
        if (s->mode == GGDMC_FP64_PLAIN || s->d.S == 0) {
            g_error = "profile needs a batched session with subjects";
            return GGDMC_EINVAL;
        }
        const uint32_t left = s->next_iter <= s->last_iter ? s->last_iter - s->next_iter + 1 : 0;
        const uint32_t n = std::min<uint32_t>((uint32_t)n_iter, left);
        std::vector<hipEvent_t> ev;
        /* launches as the run makes them (several iterations each in launch mode 3) */
        const uint32_t step = s->chained ? (uint32_t)s->ipl : 1u;
        for (uint32_t k = 0; k < n; k += step) {
            const int cnt = (int)std::min(step, n - k);
            set_base_kernel<<<1, 1, 0, s->stream>>>(s->iter_base, s->next_iter);
            HIPCHECK(hipGetLastError());
            s->enqueue_iteration(0, &ev, cnt);
            s->next_iter += (uint32_t)cnt;
        }
        s->flush_store();
        const size_t q = ev.size();
        HIPCHECK(hipStreamSynchronize(s->stream));
        s->check_fault();
        double tot = 0.0;
        for (size_t i = 0; i < q; i += 2) {
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
            tot += ms;
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        *subject_kernel_ms = q ? tot / (double)(q / 2) : 0.0;
        if (n_launches) *n_launches = (int32_t)(q / 2);
    } catch (const std::exception &e) {
        return fail(e);
    }
    return GGDMC_OK;
}

int ggdmc_session_download(ggdmc_session *s) {
    if (!s) return invalid("session is null");
    try {
        const size_t nmc = (size_t)s->cfg.nmc, n = (size_t)s->d.nchain, P = (size_t)s->d.P;
        HIPCHECK(hipStreamSynchronize(s->stream));
        for (size_t ji = 0; ji < s->jobs.size(); ++ji) {
            const ggdmc_job &j = s->jobs[ji];
            const int f = s->phi_of_job[ji];
            if (f >= 0) {
                ggdmc_samples *o = j.phi_out;
                HIPCHECK(hipMemcpy(o->theta, s->d.phi_out_theta + (size_t)f * nmc * n * 2 * P,
                                   sizeof(double) * nmc * n * 2 * P, hipMemcpyDeviceToHost));
                HIPCHECK(hipMemcpy(o->lp, s->d.phi_out_lp + (size_t)f * nmc * n, sizeof(double) * nmc * n,
                                   hipMemcpyDeviceToHost));
                HIPCHECK(hipMemcpy(o->ll, s->d.phi_out_ll + (size_t)f * nmc * n, sizeof(double) * nmc * n,
                                   hipMemcpyDeviceToHost));
            }
            if (j.kind == GGDMC_KIND_HYPER) continue;
            const int ns = j.kind == GGDMC_KIND_SUBJECT ? 1 : j.nsubject;
            for (int k = 0; k < ns; ++k) {
                const size_t sg = (size_t)s->subj_of_job[ji] + k;
                ggdmc_samples *o = &j.subject_outs[k];
                HIPCHECK(hipMemcpy(o->theta, s->d.out_theta + sg * nmc * n * P, sizeof(double) * nmc * n * P,
                                   hipMemcpyDeviceToHost));
                HIPCHECK(hipMemcpy(o->lp, s->d.out_lp + sg * nmc * n, sizeof(double) * nmc * n,
                                   hipMemcpyDeviceToHost));
                HIPCHECK(hipMemcpy(o->ll, s->d.out_ll + sg * nmc * n, sizeof(double) * nmc * n,
                                   hipMemcpyDeviceToHost));
            }
        }
    } catch (const std::exception &e) {
        return fail(e);
    }
    return GGDMC_OK;
}

void ggdmc_session_destroy(ggdmc_session *s) { delete s; }

int ggdmc_session_trace(ggdmc_session *s, const ggdmc_trace_record **records, int64_t *n,
                        int64_t *dropped) {
    if (!s || !records || !n) return invalid("session / records / n is null");
    *records = s->trace.empty() ? nullptr : s->trace.data();
    *n = (int64_t)s->trace.size();
    if (dropped) *dropped = s->trace_dropped;
    return GGDMC_OK;
}

int ggdmc_last_trace(const ggdmc_trace_record **records, int64_t *n, int64_t *dropped) {
    if (!records || !n) return invalid("records / n is null");
    *records = g_trace.empty() ? nullptr : g_trace.data();
    *n = (int64_t)g_trace.size();
    if (dropped) *dropped = g_trace_dropped;
    return GGDMC_OK;
}

int ggdmc_run_batch(const ggdmc_config *cfg, const ggdmc_model *model, const ggdmc_job *jobs,
                    int32_t njobs) {
    ggdmc_session *s = nullptr;
    g_trace.clear();
    g_trace_dropped = 0;
    int rc = ggdmc_session_create(cfg, model, jobs, njobs, &s);
    if (rc != GGDMC_OK) return rc;
    rc = ggdmc_session_advance(s, (int32_t)s->last_iter);
    if (rc == GGDMC_OK) rc = ggdmc_session_download(s);
    g_trace.swap(s->trace);
    g_trace_dropped = s->trace_dropped;
    ggdmc_session_destroy(s);
    return rc;
}

int ggdmc_run_subject(const ggdmc_config *cfg, const ggdmc_model *model, const ggdmc_prior *p_prior,
                      const ggdmc_subject *data, const ggdmc_state *start, uint64_t seed,
                      ggdmc_samples *out) {
    ggdmc_job j{};
    j.kind = GGDMC_KIND_SUBJECT;
    j.seed = seed;
    j.p_prior = p_prior;
    j.nsubject = 1;
    j.subjects = data;
    j.subject_starts = start;
    j.subject_outs = out;
    return ggdmc_run_batch(cfg, model, &j, 1);
}

int ggdmc_run_hyper(const ggdmc_config *cfg, const ggdmc_prior *p_prior, const ggdmc_prior *h_prior,
                    int32_t nsubject, const double *hyper_data, const ggdmc_state *start, uint64_t seed,
                    ggdmc_samples *out) {
    ggdmc_job j{};
    j.kind = GGDMC_KIND_HYPER;
    j.seed = seed;
    j.p_prior = p_prior;
    j.h_prior = h_prior;
    j.nsubject = nsubject;
    j.hyper_data = hyper_data;
    j.phi_start = start;
    j.phi_out = out;
    return ggdmc_run_batch(cfg, nullptr, &j, 1);
}

int ggdmc_run(const ggdmc_config *cfg, const ggdmc_model *model, const ggdmc_prior *p_prior,
              const ggdmc_prior *h_prior, int32_t nsubject, const ggdmc_subject *subjects,
              const ggdmc_state *phi_start, const ggdmc_state *subject_starts, uint64_t seed,
              ggdmc_samples *phi_out, ggdmc_samples *subject_outs) {
    ggdmc_job j{};
    j.kind = GGDMC_KIND_HIER;
    j.seed = seed;
    j.p_prior = p_prior;
    j.h_prior = h_prior;
    j.nsubject = nsubject;
    j.subjects = subjects;
    j.phi_start = phi_start;
    j.subject_starts = subject_starts;
    j.phi_out = phi_out;
    j.subject_outs = subject_outs;
    return ggdmc_run_batch(cfg, model, &j, 1);
}

int ggdmc_loglike(const ggdmc_model *model, const ggdmc_subject *data, const double *theta, int32_t n,
                  int32_t precision, double *ll_out) {
    return loglike_impl(model, data, theta, n, precision, ll_out, false);
}

int ggdmc_loglike_noalloc(const ggdmc_model *model, const ggdmc_subject *data, const double *theta,
                          int32_t n, int32_t precision, double *ll_out) {
    return loglike_impl(model, data, theta, n, precision, ll_out, true);
}

} // extern "C"

/* ggdmc_loglike and its allocation-free variant: the design, trials, thetas
 * and output are packed into one device buffer, allocated per call or (noalloc)
 * the device's staging array, the only difference between the two. */
static int loglike_impl(const ggdmc_model *model, const ggdmc_subject *data, const double *theta,
                        int32_t n, int32_t precision, double *ll_out, bool noalloc) {
    if (ggdmc_device_count() < 1) {
        g_error = "no gfx950 device visible";
        return GGDMC_ENODEV;
    }
    void *dbuf = nullptr;
    std::unique_lock<std::mutex> staging_lock;
    try {
        require(n >= 1 && theta && ll_out, "theta / ll_out missing");
        require(precision == GGDMC_FP64_PLAIN || precision == GGDMC_MIXED_BATCHED,
                "loglike precision must be FP64_PLAIN or MIXED_BATCHED");
        require(model && model->winner && model->par_src && model->par_const, "model missing");
        require(model->n_cell >= 1 && model->n_cell <= 4096 && model->n_acc >= 1 &&
                    model->n_acc <= 8 && model->npar >= 1 && model->npar <= 64,
                "design out of range");
        require(data && data->cell_off && (data->rt || data->ntrial == 0), "subject data missing");
        require(data->cell_off[0] == 0 && data->cell_off[model->n_cell] == data->ntrial,
                "cell_off must start at 0 and end at ntrial");
        const int nc = model->n_cell, na = model->n_acc, P = model->npar;
        for (int e = 0; e < nc * na; ++e) {
            require(model->par_src[e * 6 + 4] == -1 && model->par_const[e * 6 + 4] == 0.0,
                    "st0 must be the constant 0");
            for (int f = 0; f < 6; ++f)
                require(model->par_src[e * 6 + f] >= -1 && model->par_src[e * 6 + f] < P,
                        "par_src out of range");
        }
        for (int c = 0; c < nc; ++c) {
            require(model->winner[c] >= 0 && model->winner[c] < na, "winner out of range");
            require(data->cell_off[c + 1] >= data->cell_off[c], "cell_off must be non-decreasing");
        }
        /* one packed device buffer: design, trials, thetas, output */
        const Condensed cond = condense(model);
        require(cond.n_cond * na <= 256, "design has more than 256 condition x accumulator pairs");
        const int64_t T = data->ntrial;
        std::vector<uint16_t> tcode((size_t)T);
        for (int c = 0; c < nc; ++c)
            for (int64_t i = data->cell_off[c]; i < data->cell_off[c + 1]; ++i)
                tcode[i] = (uint16_t)((cond.cond_of_cell[c] << 3) | model->winner[c]);
        const int nq = cond.n_cond;
        const int64_t soff[2] = {0, T};
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        const size_t o_src = 0, o_const = o_src + al(sizeof(int32_t) * nq * na * 6);
        const size_t o_rt = o_const + al(sizeof(double) * nq * na * 6);
        const size_t o_cell = o_rt + al(sizeof(double) * (size_t)T);
        const size_t o_off = o_cell + al(sizeof(uint16_t) * (size_t)T);
        const size_t o_tab = o_off + al(sizeof(soff));
        const size_t o_th = o_tab + al(sizeof(double) * 64);
        const size_t o_out = o_th + al(sizeof(double) * (size_t)n * P);
        const size_t total = o_out + al(sizeof(double) * (size_t)n);
        std::vector<char> h(total);
        std::memcpy(&h[o_src], cond.src.data(), sizeof(int32_t) * nq * na * 6);
        std::memcpy(&h[o_const], cond.cst.data(), sizeof(double) * nq * na * 6);
        if (T) std::memcpy(&h[o_rt], data->rt, sizeof(double) * (size_t)T);
        if (T) std::memcpy(&h[o_cell], tcode.data(), sizeof(uint16_t) * (size_t)T);
        std::memcpy(&h[o_off], soff, sizeof(soff));
        fill_exp2_table(reinterpret_cast<double *>(&h[o_tab]));
        std::memcpy(&h[o_th], theta, sizeof(double) * (size_t)n * P);
        char *b = nullptr;
        if (noalloc) {
            require(total <= kStagingBytes,
                    "ggdmc_loglike_noalloc: " + std::to_string(total) + " bytes of design, trials, "
                    "thetas and output exceed the " + std::to_string(kStagingBytes) +
                    "-byte staging array (use ggdmc_loglike)");
            Staging &st = staging_for_current_device();
            staging_lock = std::unique_lock<std::mutex>(st.mu);
            b = st.ptr;
        } else {
            HIPCHECK(hipMalloc(&dbuf, total));
            b = static_cast<char *>(dbuf);
        }
        HIPCHECK(hipMemcpy(b, h.data(), o_out, hipMemcpyHostToDevice));
        Dev d{};
        d.P = P;
        d.n_acc = na;
        d.n_cond = nq;
        d.nchain = 1;
        d.par_src = reinterpret_cast<const int32_t *>(b + o_src);
        d.par_const = reinterpret_cast<const double *>(b + o_const);
        d.rt = reinterpret_cast<const double *>(b + o_rt);
        d.tcode = reinterpret_cast<const uint16_t *>(b + o_cell);
        d.subj_off = reinterpret_cast<const int64_t *>(b + o_off);
        d.exp2tab = reinterpret_cast<const double *>(b + o_tab);
        d.S = 1;
        const int cmax = (int)std::max<int64_t>(1, std::min<int64_t>((T + kChunk - 1) / kChunk, 1024));
        const size_t lds = ggdmc_internal_sweep_lds(1, P, 1, nq * na, 1, cmax, 0);
        HIPCHECK(ggdmc_internal_launch_loglike(&d, reinterpret_cast<const double *>(b + o_th), n, 0,
                                               precision, cmax, lds, nullptr,
                                               reinterpret_cast<double *>(b + o_out)));
        HIPCHECK(hipMemcpy(ll_out, b + o_out, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
        if (dbuf) HIPCHECK(hipFree(dbuf));
        dbuf = nullptr;
    } catch (const std::exception &e) {
        if (dbuf) (void)hipFree(dbuf);
        return fail(e);
    }
    return GGDMC_OK;
}
