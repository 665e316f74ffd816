/*
 * CPU FP64 oracle for the DE-MCMC hot path of ggdmc.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the timed CPU baseline. The product (ggdmc_amd/libggdmc_hip.so) never
 * links or calls it.
 *
 * What it restates (reference file:line):
 *   - LBA trial log-likelihood: external ggdmcHeaders/likelihood.h
 *     `Likelihood::sumloglike`, called at src/de.cpp:145,187,604,653; the
 *     per-trial floor of R/phi.R:3-13 (.sumlog). Pinned by the reference's
 *     fixtures tests/testthat/Group1/data/lba_data{2..6}.rda.
 *   - Prior log densities (uniform, lower/upper-truncated normal): external
 *     ggdmcPrior `sumlogprior`, called at src/de.cpp:144,186,267,426,518,603,652.
 *   - Hyper log-likelihood: src/de.cpp:245-270 (sumloghlike).
 *   - DE sampler: src/de.cpp:111-155 (crossover), :157-199 (migration),
 *     :201-242 (run_chains), :272-383 (run_hchains), :385-465 / :467-565 (phi
 *     crossover / migration), :567-613 / :615-665 (subject|phi crossover /
 *     migration), with Philox4x32-10 streams in place of R's Mersenne-Twister
 *     and a chain-group sweep whose group count nchain reproduces the
 *     reference's sequential in-place order.
 */
#ifndef GGDMC_ORACLE_H
#define GGDMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flattened LBA design: for cell c and accumulator j, family f (A, B, mean_v,
 * sd_v, st0, t0) comes from theta[par_src[(c*n_acc+j)*6+f]] or, when that is
 * -1, from par_const[(c*n_acc+j)*6+f]. winner[c] is the responding
 * accumulator of cell c. */
typedef struct {
    int32_t n_acc, n_cell, npar;
    const int32_t *winner;
    const int32_t *par_src;
    const double *par_const;
} oracle_model;

/* One prior per parameter: dist 6 = uniform(p0, p1), dist 1 = normal(p0, p1)
 * truncated to [lower, upper]. All densities are on the log scale. */
typedef struct {
    int32_t npar;
    const int32_t *dist;
    const double *p0, *p1, *lower, *upper;
} oracle_prior;

typedef struct {
    int32_t nmc, thin, nchain;
    int32_t nparameter_de;      /* de_input@nparameter: gamma = precursor/sqrt(2*this) */
    double pop_migration_prob, sub_migration_prob;
    double gamma_precursor, rp;
    int32_t is_hblocked, is_pblocked;
    int32_t sweep_groups;       /* chain groups per sweep, 2..nchain */
    uint64_t seed;
    uint64_t iter_offset;       /* added to the Philox iteration counter */
} oracle_config;

/* Model kinds for the single-level sampler. */
enum { ORACLE_MODEL_LBA = 0, ORACLE_MODEL_NORMAL = 1 };

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

double oracle_lba_loglike(const oracle_model *m, const double *rt,
                          const int64_t *cell_off, const double *theta);
/* Per-trial log density (after the floors), length cell_off[n_cell]. */
void oracle_lba_trial_terms(const oracle_model *m, const double *rt,
                            const int64_t *cell_off, const double *theta,
                            double *out);
double oracle_sumlogprior(const oracle_prior *p, const double *x);
/* sum_s sum_p log tnorm(theta_s[p]; loc = phi[p], scale = phi[P+p],
 * lower/upper of p_prior), theta_s = subj + s*stride. */
double oracle_sumloghlike(const oracle_prior *p_prior, const double *phi,
                          int32_t nsubject, const double *subj, int64_t stride);

/* Single-level run (run_subject, src/de2R.cpp:8-23 + run_chains).
 * model_kind LBA: likelihood over (rt, cell_off) with model m.
 * model_kind NORMAL: cell c's observations ~ N(theta[c], 1) (closed-form tests).
 * init_*: start state [P x nchain], [nchain]. Outputs are R column-major
 * [P x nchain x nmc] and [nchain x nmc]; column 0 holds the start state.
 * Returns 0 or a negative error code. */
int oracle_run_subject(const oracle_config *cfg, int32_t model_kind,
                       const oracle_model *m, const oracle_prior *prior,
                       const double *rt, const int64_t *cell_off,
                       const double *init_theta, const double *init_lp,
                       const double *init_ll, double *theta_out,
                       double *lp_out, double *ll_out);

/* Hyper-only run (run_hyper, src/de2R.cpp:30-47): the data are subject
 * parameter vectors data[s*P + p]; the likelihood of phi is sumloghlike. */
int oracle_run_hyper(const oracle_config *cfg, const oracle_prior *p_prior,
                     const oracle_prior *h_prior, int32_t nsubject,
                     const double *data, const double *init_theta,
                     const double *init_lp, const double *init_ll,
                     double *theta_out, double *lp_out, double *ll_out);

/* Hierarchical run (run, src/de2R.cpp:123-171 + run_hchains).
 * subject_off[S+1] indexes rt; cell_off is [S x (n_cell+1)] relative to the
 * subject's first trial. phi_*: [2P x nchain] / [nchain]; subj_init_*:
 * [S][P x nchain] / [S][nchain]. Outputs: phi [2P x nchain x nmc],
 * subjects [S][P x nchain x nmc] and [S][nchain x nmc]. */
int oracle_run_hier(const oracle_config *cfg, const oracle_model *m,
                    const oracle_prior *p_prior, const oracle_prior *h_prior,
                    int32_t nsubject, const double *rt,
                    const int64_t *subject_off, const int64_t *cell_off,
                    const double *phi_init_theta, const double *phi_init_lp,
                    const double *phi_init_ll, const double *subj_init_theta,
                    const double *subj_init_lp, const double *subj_init_ll,
                    double *phi_theta_out, double *phi_lp_out,
                    double *phi_ll_out, double *subj_theta_out,
                    double *subj_lp_out, double *subj_ll_out);

#ifdef __cplusplus
}
#endif
#endif
