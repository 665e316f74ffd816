// Host instantiation of ggdmc_amd/csrc/lba_math.h (test infrastructure).
// The transcendental policy uses libm in place of the FP32 hardware
// instructions, so the formulation can be checked against the CPU oracle
// without a GPU.
#include "../../ggdmc_amd/csrc/fp64_math.h"
#include "../../ggdmc_amd/csrc/lba_math.h"

#include <cmath>
#include <vector>

namespace {
struct Exp2Table {
    double t[64];
    Exp2Table() { ggdmc::fill_exp2_table(t); }
};
const Exp2Table g_exp2tab;

struct HostOps {
    static double exp2_frac(int j) { return g_exp2tab.t[j]; }
    static float exp2(float x) { return std::exp2(x); }
    static float log2(float x) { return std::log2(x); }
    static float rcp(float x) { return 1.0f / x; }
    /* the kernels' own FP64 functions, as on the device */
    static double qtail(double a) { return ggdmc::qtail_acc(a); }
    static double pnorm(double z) { return ggdmc::pnorm_acc(z); }
    static double dnorm(double z) { return ggdmc::kInvSqrt2Pi * ggdmc::exp_acc(-0.5 * z * z); }
    static double log(double x) { return ggdmc::log_acc(x); }
};
} // namespace

extern "C" {

// x[i] / d[i] through the batched path's division from the reciprocal.
void host_div_rn(const double *x, const double *d, int64_t n, double *out) {
    for (int64_t i = 0; i < n; ++i) {
        const double y = 1.0 / d[i];
        out[i] = ggdmc::div_rn(x[i], d[i], y);
    }
}

// One accumulator at one response time: the mixed-precision winner log density
// and loser survivor, and the naive FP64 (reference formula) pdf and CDF.
void host_acc_eval(const double *pars, double rt, double *out4) { /* out4: 6 values */
    using namespace ggdmc;
    const double theta[6] = {pars[0], pars[1], pars[2], pars[3], 0.0, pars[4]};
    const int32_t src[6] = {0, 1, 2, 3, 4, 5};
    const double cst[6] = {0, 0, 0, 0, 0, 0};
    AccK k;
    make_acck<HostOps>(theta, src, cst, k);
    bool ill_w = false, ill_l = false;
    out4[0] = mixed_winner_logpdf<HostOps>(k, rt, ill_w);
    out4[1] = mixed_loser_survivor<HostOps>(k, rt, ill_l);
    naive_pdf_cdf<HostOps>(k, rt - k.t0, out4[2], out4[3]);
    out4[4] = ill_w;
    out4[5] = ill_l;
}

// The batched path's normal tail: e[i] = exp(-z^2/2), q[i] = Q(|z[i]|).
void host_mixed_tail(const double *z, int64_t n, double *e, double *q) {
    for (int64_t i = 0; i < n; ++i) ggdmc::mixed_tail<HostOps>(z[i], e[i], q[i]);
}

// Per-trial log terms (mixed precision) for one subject; returns the sum.
double host_mixed_terms(const int32_t *winner, const int32_t *par_src, const double *par_const,
                        int32_t n_cell, int32_t n_acc, const double *rt, const int64_t *cell_off,
                        const double *theta, double *out) {
    using namespace ggdmc;
    std::vector<AccK> ks(n_acc);
    double sum = 0.0;
    for (int c = 0; c < n_cell; ++c) {
        for (int j = 0; j < n_acc; ++j)
            make_acck<HostOps>(theta, par_src + (c * n_acc + j) * 6,
                               par_const + (c * n_acc + j) * 6, ks[j]);
        const int w = winner[c];
        for (int64_t i = cell_off[c]; i < cell_off[c + 1]; ++i) {
            const double term = mixed_trial_term<HostOps>(ks.data(), n_acc, w, rt[i]);
            if (out) out[i] = term;
            sum += term;
        }
    }
    return sum;
}

// Per-trial flags: 1 where the batched path evaluates the trial with the
// reference formula (an ill-conditioned accumulator); returns their count.
int64_t host_mixed_flags(const int32_t *winner, const int32_t *par_src, const double *par_const,
                         int32_t n_cell, int32_t n_acc, const double *rt, const int64_t *cell_off,
                         const double *theta, int32_t *flags) {
    using namespace ggdmc;
    std::vector<AccK> ks(n_acc);
    int64_t n = 0;
    for (int c = 0; c < n_cell; ++c) {
        for (int j = 0; j < n_acc; ++j)
            make_acck<HostOps>(theta, par_src + (c * n_acc + j) * 6,
                               par_const + (c * n_acc + j) * 6, ks[j]);
        const int w = winner[c];
        for (int64_t i = cell_off[c]; i < cell_off[c + 1]; ++i) {
            bool ill = false;
            (void)mixed_winner_logpdf<HostOps>(ks[w], rt[i], ill);
            for (int j = 0; j < n_acc; ++j)
                if (j != w) (void)mixed_loser_survivor<HostOps>(ks[j], rt[i], ill);
            if (flags) flags[i] = ill;
            n += ill;
        }
    }
    return n;
}
}
