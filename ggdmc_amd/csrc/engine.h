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
 *   partial double [tiles][n]   per-tile partial log-likelihoods of the current group
 *   out_*   [S][nmc][n][P] etc. stored samples, R's column-major order per subject
 */
#ifndef GGDMC_ENGINE_H
#define GGDMC_ENGINE_H

#include <cstdint>

namespace ggdmc {

constexpr int kBlock = 256;        /* threads per workgroup of the sweep kernels */
constexpr int kWaves = kBlock / 64;
constexpr int kMaxChains = 256;    /* chains per run handled by one workgroup */

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
    int32_t S, n_tiles, n_phi, n_jobs;
    uint32_t philox_offset;     /* Philox iteration = local iteration + this */
    uint32_t last_iter;         /* last local iteration of the run (nmc*thin - 1) */

    const int32_t *par_src;     /* [n_cond][n_acc][6] */
    const double *par_const;

    const double *rt;
    const uint16_t *tcode;
    const int64_t *subj_off;    /* [S+1] trial offsets */
    const int32_t *subj_job;    /* [S] */
    const int32_t *subj_unit;   /* [S] Philox unit (index within its run) */
    const int32_t *subj_tile0;  /* [S] first tile */
    const int32_t *subj_ntile;  /* [S] */
    const int32_t *tile_subj;   /* [tiles] */
    const int64_t *tile_lo, *tile_hi;

    const JobDev *jobs;
    const int32_t *phi_job;     /* [F] job of each population state */

    const int32_t *pr_dist;
    const double *pr_p0, *pr_p1, *pr_lo, *pr_up;

    double *theta, *lp, *ll;
    double *phi, *phi_lp, *phi_ll;
    double *stats;              /* [F][n][P][4]: mean, M2, min, max over subjects */

    double *partial;
    uint32_t *counter;          /* [S] tile arrivals, [F] stats arrivals, [jobs] subject arrivals */
    unsigned long long *evals;  /* trial-likelihood evaluations */
    const uint32_t *iter_base;  /* device iteration base (graph replays) */

    double *out_theta, *out_lp, *out_ll;
    double *phi_out_theta, *phi_out_lp, *phi_out_ll;
};

/* Per-launch arguments of the subject sweep kernel. */
struct SweepArgs {
    uint32_t iter_off;   /* iteration = *iter_base + iter_off */
    int32_t sweep;       /* blocked parameter index or 0 */
    int32_t g_begin, g_end;
    int32_t final_launch; /* last subject launch of the iteration: store if due */
    int32_t chain_batch;  /* chains whose constants fit in LDS at once */
    int32_t fuse_phi;     /* final launch: the last subject of a hierarchical run
                             performs the next iteration's population step */
};

} // namespace ggdmc

#endif
