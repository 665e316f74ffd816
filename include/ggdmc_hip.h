/*
 * libggdmc_hip: MI355X (gfx950) DE-MCMC engine for LBA models.
 *
 * C ABI of the drop-in boundary that replaces the Rcpp exports of ggdmc's
 * src/de2R.cpp (run_subject :8-23, run_hyper :30-47, run :123-171), which
 * R reaches through .Call stubs in R/RcppExports.R:6-7,12-13,81-82 under
 * StartSampling* / RestartSampling* (R/sampling.R:247,308,446,507,615,681).
 *
 * Conventions
 *  - Plain pointers and sizes; every array is caller-owned host memory in R's
 *    column-major layout (theta [P x nchain x nmc] is theta[p + P*(c + nchain*k)]).
 *  - Start states are one column: theta [P x nchain], lp [nchain], ll [nchain]
 *    (the Rcpp shim passes column 1 of initialise_theta/initialise_phi output,
 *    or the last stored column of a previous run to restart).
 *  - Outputs: column 0 is the start state; column k holds the state after
 *    iteration k*thin (nmc*thin - 1 iterations are run).
 *  - Return value 0 on success, a negative GGDMC_E* code otherwise; the
 *    message of the last failure on the calling thread is ggdmc_last_error().
 *  - The library initialises HIP on first use and is not fork-safe afterwards:
 *    run independent runs through ggdmc_run_batch instead of forked processes.
 */
#ifndef GGDMC_HIP_H
#define GGDMC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GGDMC_OK 0
#define GGDMC_EINVAL -1   /* invalid argument / inconsistent sizes */
#define GGDMC_EHIP -2     /* HIP runtime failure */
#define GGDMC_ENOMEM -3   /* device allocation failed */
#define GGDMC_ENODEV -4   /* no usable gfx950 device */

/* Arithmetic of the likelihood. */
#define GGDMC_FP64_PLAIN 0   /* reference arithmetic and sweep order; one launch per chain step */
#define GGDMC_MIXED_BATCHED 1 /* FP32 transcendentals + FP64 sums; batched, graph-captured */
#define GGDMC_FP64_BATCHED 2 /* reference arithmetic inside the batched kernels */

/* Run kinds (which reference export a job mirrors). */
#define GGDMC_KIND_SUBJECT 0 /* run_subject: one subject, fixed prior */
#define GGDMC_KIND_HYPER 1   /* run_hyper: population parameters of fixed subject vectors */
#define GGDMC_KIND_HIER 2    /* run: population level + subjects */

/* Prior distributions (ggdmcPrior dist_id). */
#define GGDMC_DIST_TNORM 1
#define GGDMC_DIST_UNIF 6

/* S4 de_input + theta_input fields (R/model-class.R:36-57, :1288-1313)
 * plus the engine's own knobs. */
typedef struct {
    int32_t nmc, thin, nchain;
    double pop_migration_prob, sub_migration_prob;
    double gamma_precursor; /* 2.38 */
    double rp;              /* 0.001 */
    int32_t is_hblocked, is_pblocked;
    int32_t precision;      /* GGDMC_FP64_PLAIN / GGDMC_MIXED_BATCHED / GGDMC_FP64_BATCHED */
    int32_t sweep_groups;   /* chain groups per sweep: 0 = default (nchain for
                               FP64_PLAIN, 2 otherwise); nchain = the reference's
                               sequential in-place order */
    int32_t trials_per_tile; /* trials per work item: 0 = whole subject (autotuned when
                                chains_per_item is 0 too); rounded up to a multiple of 128 */
    int32_t chains_per_item; /* chains of a chain group per work item: 0 = all (autotuned
                                when trials_per_tile is 0 too); -1 = one item per subject
                                covering every group in one launch */
    int32_t launch_mode;    /* subject launches per sweep: 0 = autotuned with the two above
                               (else 1), 1 = one per chain group, 2 = one for every chain
                               group (an item waits for the previous group's items of its
                               subject), 3 = as 2 with 32 iterations per launch (an item
                               also waits for its subject's previous iteration), 4 = as 3
                               with an item holding its chain range of every group (its
                               later groups wait for the earlier groups' items); never
                               changes results */
    uint64_t iter_offset;   /* Philox iteration offset (restart continues a stream) */
    int32_t pop_debug;      /* de_input@pop_debug: record the population level's decisions */
    int32_t sub_debug;      /* de_input@sub_debug: record the decisions of single-subject and
                               hyper-only runs and of the first three subjects of hierarchical
                               runs (src/de.cpp:85-106, :327) */
    int32_t waves_per_simd; /* launch modes 3 and 4: the subject kernel built for 3 waves per
                               SIMD (up to 168 registers a lane) or 4 (128, some values kept
                               in scratch); 0 = autotuned with the decomposition, else 3;
                               never changes results */
} ggdmc_config;

/* One Metropolis decision of a traced level (ggdmc_config pop_debug /
 * sub_debug), what src/de.cpp prints with debug on (:85-106 update_theta,
 * :441-463 population crossover, :533-558 population migration). */
typedef struct {
    int32_t job;         /* run index in the batch */
    int32_t subject;     /* subject index within the run; -1: the population level */
    uint32_t iteration;  /* 1 .. nmc*thin - 1 within this call or session */
    int32_t sweep;       /* blocked parameter index, 0 for joint updates */
    int32_t kind;        /* 0 crossover, 1 migration */
    int32_t step;        /* crossover: the proposing chain; migration: the ring step */
    int32_t chain;       /* proposing chain (crossover) / source chain (migration) */
    int32_t target;      /* chain whose state the proposal would replace */
    double tmp_lp, tmp_ll; /* the proposal's log prior and log likelihood */
    double cur_post;     /* the target's cached log posterior it is compared with */
    int32_t accepted;    /* 1 accepted, 0 rejected, -1 rejected (NaN ratio) */
    int32_t reserved;
} ggdmc_trace_record;

/* Flattened LBA design: family f (0 A, 1 B, 2 mean_v, 3 sd_v, 4 st0, 5 t0) of
 * accumulator j in cell c is theta[par_src[(c*n_acc+j)*6+f]], or
 * par_const[...] when par_src is -1 (derived from model@model_boolean,
 * model@constants; winner[c] = dmi@node_1_index[c, 1] - 0-based). */
typedef struct {
    int32_t n_acc, n_cell, npar;
    const int32_t *winner;
    const int32_t *par_src;
    const double *par_const;
} ggdmc_model;

typedef struct { /* one entry per parameter, p_prior / h_prior of an S4 prior */
    int32_t npar;
    const int32_t *dist;
    const double *p0, *p1, *lower, *upper;
} ggdmc_prior;

typedef struct { /* dmi@data: cell-sorted RTs, cell_off[n_cell + 1] */
    int64_t ntrial;
    const double *rt;
    const int64_t *cell_off;
} ggdmc_subject;

typedef struct {
    const double *theta, *lp, *ll;
} ggdmc_state;

typedef struct {
    double *theta, *lp, *ll;
} ggdmc_samples;

typedef struct {
    int32_t kind;            /* GGDMC_KIND_* */
    uint64_t seed;           /* Philox key of this run */
    int32_t nparameter_de;   /* de_input@nparameter (gamma = precursor/sqrt(2*this)); 0 = P (subject) / 2P */
    const ggdmc_prior *p_prior; /* SUBJECT: the prior; HIER/HYPER: tnorm bounds of the population */
    const ggdmc_prior *h_prior; /* HIER/HYPER: prior of the 2P population parameters */
    int32_t nsubject;
    const ggdmc_subject *subjects;   /* SUBJECT (1) / HIER (nsubject) */
    const double *hyper_data;        /* HYPER: [nsubject x P] row-major subject vectors */
    const ggdmc_state *phi_start;    /* HYPER / HIER */
    const ggdmc_state *subject_starts; /* SUBJECT / HIER: [nsubject] */
    ggdmc_samples *phi_out;          /* HYPER / HIER */
    ggdmc_samples *subject_outs;     /* SUBJECT / HIER: [nsubject] */
} ggdmc_job;

/* Replaces run_subject (src/de2R.cpp:8-23 -> de_class::run_chains, src/de.cpp:201-242). */
int ggdmc_run_subject(const ggdmc_config *cfg, const ggdmc_model *model,
                      const ggdmc_prior *p_prior, const ggdmc_subject *data,
                      const ggdmc_state *start, uint64_t seed, ggdmc_samples *out);

/* Replaces run_hyper (src/de2R.cpp:30-47): the data are subject parameter
 * vectors hyper_data [nsubject x P] (row-major), the likelihood of the 2P
 * population parameters is de_class::sumloghlike (src/de.cpp:245-270). */
int ggdmc_run_hyper(const ggdmc_config *cfg, const ggdmc_prior *p_prior,
                    const ggdmc_prior *h_prior, int32_t nsubject, const double *hyper_data,
                    const ggdmc_state *start, uint64_t seed, ggdmc_samples *out);

/* Replaces run (src/de2R.cpp:123-171 -> de_class::run_hchains, src/de.cpp:272-383). */
int ggdmc_run(const ggdmc_config *cfg, const ggdmc_model *model, const ggdmc_prior *p_prior,
              const ggdmc_prior *h_prior, int32_t nsubject, const ggdmc_subject *subjects,
              const ggdmc_state *phi_start, const ggdmc_state *subject_starts, uint64_t seed,
              ggdmc_samples *phi_out, ggdmc_samples *subject_outs);

/* Many independent runs (parallel_lapply's ncore runs, R/sampling.R:13-121;
 * recovery fits) in one launch sequence. All jobs share cfg and model. */
int ggdmc_run_batch(const ggdmc_config *cfg, const ggdmc_model *model, const ggdmc_job *jobs,
                    int32_t njobs);

/* Likelihood::sumloglike (external ggdmcHeaders, called at src/de.cpp:145) /
 * ggdmcLikelihood::compute_subject_likelihood + .sumlog (R/phi.R:3-13,176) for
 * n parameter vectors theta [P x n] of one subject. precision: GGDMC_FP64_PLAIN
 * (reference arithmetic) or GGDMC_MIXED_BATCHED. */
int ggdmc_loglike(const ggdmc_model *model, const ggdmc_subject *data, const double *theta,
                  int32_t n, int32_t precision, double *ll_out);

/* No counterpart in the reference. ggdmc_loglike for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): same arguments and results (bit for bit),
 * but the design, trials, thetas and output are staged in a fixed 16 MiB
 * device array per GPU instead of a per-call allocation; it calls no
 * hipMalloc / hipFree. Calls whose staging exceeds 16 MiB (about 1.5 M trials)
 * fail with GGDMC_EINVAL. Calls on one GPU share the array and are serialised
 * by the library. */
int ggdmc_loglike_noalloc(const ggdmc_model *model, const ggdmc_subject *data, const double *theta,
                          int32_t n, int32_t precision, double *ll_out);

/* Sessions: a batch whose inputs and state stay resident in HBM, advanced in
 * steps (restart without re-upload; timing of the device-resident loop). */
typedef struct ggdmc_session ggdmc_session;
int ggdmc_session_create(const ggdmc_config *cfg, const ggdmc_model *model,
                         const ggdmc_job *jobs, int32_t njobs, ggdmc_session **out);
/* Runs the next n_iter iterations; stored columns follow the thinning rule. */
int ggdmc_session_advance(ggdmc_session *s, int32_t n_iter);
/* Device time in ms of the last advance (HIP events on the session stream). */
int ggdmc_session_elapsed_ms(ggdmc_session *s, float *ms);
/* Copies every job's stored samples into its *_out buffers. */
int ggdmc_session_download(ggdmc_session *s);
/* Average device time (ms) of the subject-level kernel over n_iter
 * iterations run launch by launch with events around each launch, and the
 * number of those launches. */
int ggdmc_session_profile(ggdmc_session *s, int32_t n_iter, double *subject_kernel_ms,
                          int32_t *n_launches);
/* Trial-likelihood evaluations (trial x proposal) done by the subject-level
 * kernels since creation, counting every proposal whose prior is finite. */
int ggdmc_session_trial_evals(ggdmc_session *s, double *evals);
void ggdmc_session_destroy(ggdmc_session *s);

/* Work decomposition of a session's subject launches (after create, so after
 * autotuning): chains per item (-1: one item per subject over every group),
 * trials per item (0: whole subject), the items of the largest launch and the
 * launch mode (1 to 4, as ggdmc_config.launch_mode). The decomposition never
 * changes results: log-likelihoods are summed in a fixed chunk order whatever
 * the split. */
int ggdmc_session_decomposition(ggdmc_session *s, int32_t *chains_per_item,
                                int32_t *trials_per_tile, int32_t *items_per_launch,
                                int32_t *launch_mode);

/* Waves per SIMD the session's subject kernel was built for (3 or 4, as
 * ggdmc_config.waves_per_simd). No counterpart in the reference. */
int ggdmc_session_waves_per_simd(ggdmc_session *s, int32_t *waves_per_simd);

/* Decisions recorded during the last advance of a session (pop_debug /
 * sub_debug), in the reference's print order: by run, iteration, level
 * (population first, then subjects in order), sweep and step. *records stays
 * valid until the next advance or destroy. At most 2^20 records are kept per
 * advance; *dropped counts the rest. */
int ggdmc_session_trace(ggdmc_session *s, const ggdmc_trace_record **records, int64_t *n,
                        int64_t *dropped);
/* The same for the last ggdmc_run_subject / ggdmc_run_hyper / ggdmc_run /
 * ggdmc_run_batch call on the calling thread (valid until its next call). */
int ggdmc_last_trace(const ggdmc_trace_record **records, int64_t *n, int64_t *dropped);

const char *ggdmc_last_error(void);
/* Number of visible gfx950 devices (0 without a GPU; never fails). */
int ggdmc_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
