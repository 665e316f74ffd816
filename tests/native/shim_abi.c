/*
 * The ABI side of the Rcpp shim that replaces src/de2R.cpp, in plain C, for tests: what
 * the shim's run_subject / run_hyper / run do with R's buffers before and
 * while calling libggdmc_hip, without R.
 *
 *  - samples objects arrive as R arrays: theta [P x nchain x nmc_in] and
 *    lp / ll [nchain x nmc_in], column-major; the start column is the last
 *    column whose parameters are all finite (a fresh object from
 *    set_up_new_samples has only column 1 filled, R/phi.R:72-87; a restart
 *    passes a finished posterior);
 *  - hyper-only data arrive as an R matrix [S x P], column-major, and go to
 *    the ABI row-major;
 *  - the run goes through ggdmc_run_batch with one job (the shim's path) or
 *    through ggdmc_run_subject / ggdmc_run_hyper / ggdmc_run (direct = 1).
 *
 * Test infrastructure (tests/test_gpu_native.py loads it); not product code.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/ggdmc_hip.h"

/* 0-based index of the last column of theta [P x n x nmc] whose P*n
 * parameters are all finite (0 if none is). */
int shim_start_column(const double *theta, int P, int n, int nmc) {
    const size_t col = (size_t)P * n;
    for (int k = nmc - 1; k > 0; --k) {
        size_t i = 0;
        while (i < col && isfinite(theta[k * col + i])) ++i;
        if (i == col) return k;
    }
    return 0;
}

static ggdmc_state start_of(const double *th, const double *lp, const double *ll, int P, int n,
                            int nmc) {
    const int k = shim_start_column(th, P, n, nmc);
    ggdmc_state s = {th + (size_t)k * P * n, lp + (size_t)k * n, ll + (size_t)k * n};
    return s;
}

/* kind: GGDMC_KIND_SUBJECT (subjects[0], sub_* [1]), GGDMC_KIND_HYPER
 * (hyper_matrix [S x P] column-major, nsubject = S), GGDMC_KIND_HIER. */
int shim_run(const ggdmc_config *cfg, const ggdmc_model *model, const ggdmc_prior *p_prior,
             const ggdmc_prior *h_prior, int32_t kind, int32_t nsubject,
             const ggdmc_subject *subjects, const double *hyper_matrix, int32_t P,
             const double *phi_th, const double *phi_lp, const double *phi_ll,
             const double *const *sub_th, const double *const *sub_lp,
             const double *const *sub_ll, int32_t nmc_in, uint64_t seed, int32_t nparameter_de,
             int32_t direct, ggdmc_samples *phi_out, ggdmc_samples *sub_outs) {
    const int n = cfg->nchain;
    const int ns = kind == GGDMC_KIND_HYPER ? 0 : (kind == GGDMC_KIND_SUBJECT ? 1 : nsubject);
    ggdmc_state *starts = (ggdmc_state *)calloc((size_t)(ns > 0 ? ns : 1), sizeof(ggdmc_state));
    double *rows = NULL;
    ggdmc_state phi_start;
    int rc;
    if (!starts) return GGDMC_ENOMEM;
    for (int s = 0; s < ns; ++s) starts[s] = start_of(sub_th[s], sub_lp[s], sub_ll[s], P, n, nmc_in);
    if (kind != GGDMC_KIND_SUBJECT) phi_start = start_of(phi_th, phi_lp, phi_ll, 2 * P, n, nmc_in);
    if (kind == GGDMC_KIND_HYPER) {
        rows = (double *)malloc(sizeof(double) * (size_t)nsubject * P);
        if (!rows) {
            free(starts);
            return GGDMC_ENOMEM;
        }
        for (int s = 0; s < nsubject; ++s)
            for (int q = 0; q < P; ++q) rows[(size_t)s * P + q] = hyper_matrix[s + (size_t)nsubject * q];
    }
    if (direct) {
        if (kind == GGDMC_KIND_SUBJECT)
            rc = ggdmc_run_subject(cfg, model, p_prior, subjects, starts, seed, sub_outs);
        else if (kind == GGDMC_KIND_HYPER)
            rc = ggdmc_run_hyper(cfg, p_prior, h_prior, nsubject, rows, &phi_start, seed, phi_out);
        else
            rc = ggdmc_run(cfg, model, p_prior, h_prior, nsubject, subjects, &phi_start, starts, seed,
                           phi_out, sub_outs);
    } else {
        ggdmc_job job;
        memset(&job, 0, sizeof(job));
        job.kind = kind;
        job.seed = seed;
        job.nparameter_de = nparameter_de;
        job.p_prior = p_prior;
        job.h_prior = h_prior;
        job.nsubject = kind == GGDMC_KIND_SUBJECT ? 1 : nsubject;
        job.subjects = subjects;
        job.hyper_data = rows;
        job.phi_start = kind == GGDMC_KIND_SUBJECT ? NULL : &phi_start;
        job.subject_starts = kind == GGDMC_KIND_HYPER ? NULL : starts;
        job.phi_out = phi_out;
        job.subject_outs = sub_outs;
        rc = ggdmc_run_batch(cfg, kind == GGDMC_KIND_HYPER ? NULL : model, &job, 1);
    }
    free(rows);
    free(starts);
    return rc;
}
