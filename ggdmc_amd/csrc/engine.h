/*
 * Internal device-side problem description shared by the kernels
 * (sampler_kernels.hip) and the host engine (engine.hip).
 *
 * HBM layout (all arrays device-resident for the whole run):
 *   rt      double [T]          trials of every subject, subject-major, cell-sorted
 *   tcode   uint16 [T]          condition << 3 | responding accumulator of each trial
 *                               (a condition = cells that differ only in the response:
 *                               every accumulator's parameters are the same)
 *   theta   double [S][n][P]    current subject states (chain-major, P contiguous)
 *   lp, ll  double [S][n]
 *   phi     double [F][n][2P]   population states of hierarchical / hyper runs
 *   stats   double [F][n][P][4] per-chain sufficient statistics of the subject states
 *   bstats  double [F][n*P][blocks][4] the same per block of 32 subjects
 *   csum    double [chunks][n]  chunk sums of trial-split work items
 *   out_*   [S][nmc][n][P] etc. stored samples, R's column-major order per subject
 *
 * Log-likelihood sums are canonical: chunk j of a subject is its trials
 * [j * kChunk, (j + 1) * kChunk); a 64-lane wave sums a chunk with lane l
 * holding term(j*kChunk + l) + term(j*kChunk + 64 + l) and a fixed xor
 * butterfly; a chain's log-likelihood is the sequential sum of its chunk sums
 * in chunk order. How chains and trials are split over workgroups therefore
 * never changes a result bit.
 */
#ifndef GGDMC_ENGINE_H
#define GGDMC_ENGINE_H

#include <cstdint>

#include "../../include/ggdmc_hip.h"

namespace ggdmc {

constexpr int kBlock = 256;        /* threads per workgroup of the sweep kernels */
constexpr int kPhiBlock = 512;     /* threads per workgroup of the population kernel */
constexpr int kWaves = kBlock / 64;
constexpr int kMaxChains = 256;    /* chains per run */
constexpr int kChunk = 128;        /* trials per canonical chunk sum */
constexpr int kStatBlock = 32;     /* subjects per block of the population statistics */
constexpr int kItersPerLaunch = 32; /* iterations per subject launch in launch mode 3 */

struct JobDev {
    int32_t kind;           /* GGDMC_KIND_* */
    uint32_t k0, k1;        /* Philox key */
    double gamma;
    int32_t first_subject, nsubject;
    int32_t phi_index;      /* population state index, -1 for SUBJECT runs */
    int32_t prior_off;      /* offset of the P-entry subject prior (or tnorm bounds) */
    int32_t hprior_off;     /* offset of the 2P-entry population prior */
};

struct Dev {
    int32_t P, n_acc, n_cond, nchain, G, nmc, thin;
    int32_t blocked_sub, blocked_pop;
    double rp, sub_mig, pop_mig;
    int32_t S, n_phi, n_jobs;
    int32_t tpt;                /* trials per tile of trial-split items (multiple of kChunk) */
    uint32_t philox_offset;     /* Philox iteration = local iteration + this */
    uint32_t last_iter;         /* last local iteration of the run (nmc*thin - 1) */

    const double *exp2tab;      /* [64] 2^(j/64) */
    const int32_t *par_src;     /* [n_cond][n_acc][6] */
    const double *par_const;

    const double *rt;
    const uint16_t *tcode;
    const int64_t *subj_off;    /* [S+1] trial offsets */
    const int64_t *chunk_off;   /* [S+1] chunk offsets (rows of csum) */
    const int32_t *subj_job;    /* [S] */
    const int32_t *subj_unit;   /* [S] Philox unit (index within its run) */

    const JobDev *jobs;
    const int32_t *phi_job;     /* [F] job of each population state */
    const JobDev *subj_jobs;    /* [S] the job of each subject (one load from the subject) */
    const JobDev *phi_jobs;     /* [F] the job of each population state */

    const int32_t *pr_dist;
    const double *pr_p0, *pr_p1, *pr_lo, *pr_up;

    double *theta, *lp, *ll;
    double *phi, *phi_lp, *phi_ll;
    double *stats;              /* [F][n][P][4]: mean, M2, min, max over subjects */
    double *bstats;             /* [F][bstats_stride]: per (pair, subject block) statistics */
    int64_t bstats_stride;      /* doubles per run: n * P * blocks of the largest run * 4 */
    double *csum;
    double *pcache;             /* [F][n][P][2]: log scale, log mass of each chain's subject prior */
    int32_t *pcache_valid;      /* [F][n]: pcache of a chain matches its phi (cleared on start and restore) */
    const double *scache;       /* [prior entries][2]: the same for fixed priors (unif: log width) */

    uint32_t *stats_ctr;        /* [F] statistics workgroups finished */
    uint32_t *tile_ctr;         /* [S][G][rmax] tiles of a (subject, group, chain range) finished */
    uint32_t *phi_ready;        /* [F][n] last iteration whose population step of a chain is stored */
    uint32_t *fault;            /* [1] nonzero: a wait for phi_ready gave up (results invalid) */
    uint32_t *item_ready;       /* [S][G][rmax] last iteration whose (subject, group, chain range)
                                   step is stored (all groups in one launch) */
    uint32_t *item_staged;      /* [S][G][rmax] last iteration whose (subject, group, chain range)
                                   item has read the subject's rows */
    unsigned long long *evals;  /* [S] trial-likelihood evaluations per subject */
    const uint32_t *iter_base;  /* device iteration base (graph replays) */

    double *out_theta, *out_lp, *out_ll;
    double *phi_out_theta, *phi_out_lp, *phi_out_ll;

    /* decision trace (debug flags): records appended at trace_n, at most
     * trace_cap kept; null when no level is traced */
    ggdmc_trace_record *trace;
    uint32_t *trace_n;
    uint32_t trace_cap;
    int32_t trace_pop, trace_sub;
};

/* A work item of a subject launch: subject x, chains [c0, c1) of chain group
 * g (c1 = kAllChains: to the group's end), trial tile t of ntile, and r = the
 * index of the chain range in the group (c0 / chains per item). Packed as
 * int4: {x, c0 << 16 | c1, t << 16 | ntile, g << 16 | r}. The tiles of one
 * (x, g, c0) are consecutive, in tile order. */
constexpr int kAllChains = 0xFFFF;

/* Per-launch arguments of the subject sweep kernel. */
struct SweepArgs {
    uint32_t iter_off;    /* iteration = *iter_base + iter_off */
    int32_t sweep;        /* blocked parameter index or 0 */
    int32_t g_begin, g_end;
    int32_t store_prev;   /* first sweep's launches: store iteration - 1 if due (each item its own chains) */
    int32_t chain_batch;  /* chains whose accumulator constants fit in LDS at once */
    int32_t kmax;         /* most chains one item proposes */
    int32_t cmax;         /* chunk sums per chain held in LDS at once */
    int32_t tstage;       /* trials per item staged in LDS (0: read from HBM) */
    int32_t n_items;
    const int4 *items;    /* grid = phi_wgs + one workgroup per item */
    int32_t phi_wgs;      /* leading workgroups that run the population step of the launch's
                             chains, one per (state, chain) (0: none) */
    int32_t phi_stats_lds; /* a whole-run step keeps the statistics in LDS */
    int32_t chained;      /* every chain group in this launch: items carry their group, an
                             item waits for the previous group's items of its subject */
    int32_t rmax;         /* chain ranges per group (item_ready stride) */
    int32_t tiles;        /* some subject's trials are split over several items */
    int32_t iters;        /* consecutive iterations in this launch (chained only; else 1): the
                             grid is iters segments of seg_wgs workgroups, segment s running
                             iteration *iter_base + iter_off + s */
    int32_t seg_wgs;      /* phi_wgs + n_items rounded up to a multiple of 8 (iters > 1) */
    int32_t span;         /* chained: an item holds its chain range of every group, in group
                             order (items carry group 0); else one group per item */
};

} // namespace ggdmc

#endif
