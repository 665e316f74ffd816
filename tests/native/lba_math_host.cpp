// Host instantiation of ggdmc_amd/csrc/lba_math.h (test infrastructure).
// The transcendental policy uses libm in place of the FP32 hardware
// instructions, so the formulation can be checked against the CPU oracle
// without a GPU.
#include "../../ggdmc_amd/csrc/lba_math.h"

#include <cmath>
#include <vector>

namespace {
struct HostOps {
    static float exp2(float x) { return std::exp2(x); }
    static float log2(float x) { return std::log2(x); }
    static float rcp(float x) { return 1.0f / x; }
    static double erfc(double x) { return std::erfc(x); }
    static double pnorm(double z) { return 0.5 * std::erfc(-z * ggdmc::kSqrt1_2); }
    static double dnorm(double z) { return ggdmc::kInvSqrt2Pi * std::exp(-0.5 * z * z); }
};
} // namespace

extern "C" {

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
            double lp = mixed_winner_logpdf<HostOps>(ks[w], rt[i]);
            double acc = 0.0;
            int neg = 0;
            bool zero = false, nan = lp != lp;
            for (int j = 0; j < n_acc; ++j) {
                if (j == w) continue;
                const double S = mixed_loser_survivor<HostOps>(ks[j], rt[i]);
                if (S != S) nan = true;
                const double aS = std::fabs(S);
                fold_survivor(S, aS > 0.0 ? mixed_log<HostOps>(aS) : 0.0, acc, neg, zero);
            }
            double term;
            if (nan) term = NAN;
            else if (zero || neg) term = kLogDblEps;
            else term = lp + acc;
            if (out) out[i] = term;
            sum += term;
        }
    }
    return sum;
}
}
