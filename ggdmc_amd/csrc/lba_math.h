/*
 * LBA trial log-likelihood arithmetic for the HIP sampler.
 *
 * Reference semantics (external ggdmcHeaders/likelihood.h `sumloglike`, called
 * at src/de.cpp:145,187,604,653; per-trial floor R/phi.R:3-13): for a trial of
 * cell c with responding accumulator w and each accumulator j with
 * t = rt - t0_j > 0,
 *     zs = t*s, chi = b - t*v, z1 = chi/zs, z2 = (chi - A)/zs, den = max(Phi(v/s), 1e-10)
 *     pdf_j = (v*(Phi(z1)-Phi(z2)) + s*(phi(z2)-phi(z1))) / (A*den)
 *     F_j   = (1 + (zs*(phi(z2)-phi(z1)) + (chi-A)*Phi(z2) - chi*Phi(z1)) / A) / den
 *     density = max(pdf_w, 1e-10) * prod_{j != w} (1 - max(F_j, 1e-10))
 *     term = log(density > 0 ? density : DBL_EPSILON)
 * (t <= 0 gives pdf = F = 0.)
 *
 * Two evaluations are provided:
 *
 *  - `naive_*`: the formula above, term by term, in FP64 (the plain FP64 path).
 *
 *  - `mixed_*`: the batched path. CDNA4 runs FP64 FMA at the rate of
 *    non-packed FP32, so all polynomial / rational arithmetic stays FP64 and
 *    only the transcendental instructions run in FP32 hardware (v_exp_f32,
 *    v_log_f32, v_rcp_f32) with split arguments:
 *      * exp(-z^2/2) = 2^y, y = -z^2/2*log2(e): 2^n * 2^(j/64) (table) * 2^f
 *        (short polynomial), all FP64;
 *      * 1/x from v_rcp_f32 plus one FP64 Newton step;
 *      * Q(|z|) = exp(-z^2/2) * R(|z|), R the Mills-ratio factor
 *        Q(z)*exp(z^2/2), approximated as t*sum_k c_k t^k, k = 0..10, with
 *        t = 1/(1+0.2625|z|) (coefficients fitted for this library to
 *        0.5*erfcx(z/sqrt(2)) on [0, 14] by least squares reweighted toward
 *        the minimax relative error; max relative error 2.9e-10);
 *      * the small tail of Phi is always the one computed, so
 *        Phi(z1) - Phi(z2) is taken as Q(z2) - Q(z1) when z2 >= 0; where
 *        the reference's own FP64 value RN(1 - Q(z1)) - RN(1 - Q(z2)) could
 *        differ from that by more than the tolerance (the winner's far right
 *        tail), that rounded difference is reproduced from the computed
 *        tails whenever their error cannot move either rounding;
 *      * the survivor of a loser is formed from whichever of the untruncated
 *        CDF G or its complement 1 - G is small (both written without
 *        cancellation, see below) and then passed through the reference's own
 *        FP64 operations G = 1 + (G - 1), F = G/den, 1 - max(F, 1e-10). Where
 *        the reference's survivor is quantised by FP64 rounding (survivors of
 *        1e-16 .. 1e-10 of sharp accumulators, whose F sits within a few ulps
 *        of 1), this reproduces the rounding itself, not just its scale; the
 *        DBL_EPSILON replacement and its sign rule follow from the same FP64
 *        values. Phi follows R's pnorm: for z > 0 it is 1 - Q(z) rounded in
 *        FP64, Q the accurate upper tail.
 *    Untruncated CDF and complement, with Q = 1 - Phi, Pc = Phi(-v/s):
 *        G     = (zs*(phi2 - phi1) + chi*Q1 - (chi - A)*Q2) / A
 *        1 - G = (Phi2 - Pc) + (chi*(Phi1 - Phi2) + zs*(phi1 - phi2)) / A + Pc
 *    (the first two terms from integrating the drift density over the start
 *    point: P(0 < drift < (b - a)/t) averaged over a ~ U(0, A)).
 *    The trial's density is the reference's own FP64 product of the floored
 *    winner pdf and the survivors, in accumulator order, logged once:
 *    log(x) = e*ln2 + log2(m)*ln2 with the FP32 log2 applied to the mantissa
 *    only, so its error is absolute ~1e-8 whatever the magnitude of log(x).
 */
#ifndef GGDMC_LBA_MATH_H
#define GGDMC_LBA_MATH_H

#include <cmath>
#include <cstdint>

#ifdef __HIPCC__
#define GGDMC_FN __host__ __device__ __forceinline__
#else
#define GGDMC_FN inline
#endif

namespace ggdmc {

constexpr double kFloor = 1e-10;
constexpr double kLogFloor = -23.025850929940457;          /* log(1e-10)           */
constexpr double kLogDblEps = -36.04365338911715;          /* log(2.220446e-16)   */
constexpr double kInvSqrt2Pi = 0.3989422804014327;
constexpr double kLog2e = 1.4426950408889634;
constexpr double kLn2 = 0.6931471805599453;
constexpr double kSqrt1_2 = 0.7071067811865476;
constexpr double kMillsK = 0.2625;
constexpr int kMillsDeg = 10;
/* R(z) ~= t * sum_{k=0}^{10} c_k t^k, t = 1/(1 + 0.2625 z), z >= 0 */
constexpr double kMills[kMillsDeg + 1] = {
    0.10471908204997292, 0.10480896722129296, 0.09649084604609494,  0.090014727337448,
    0.032247173067295534, 0.13170626451118267, -0.1694328833813132, 0.2609796199857909,
    -0.23411156866613111, 0.09871999173116763, -0.016142220044834274};

/* Per (chain, cell, accumulator) constants, computed once per proposal. */
struct AccK {
    double A, b, v, s, t0;
    double invA, den, inv_den, Pc;
};

/* ---------------------------------------------------------------- host/device
 * Policy interface (Ops): float exp2(float), float log2(float), float rcp(float).
 */

template <class Ops>
GGDMC_FN double rcp64(double x) {
    float xf = (float)x;
    xf = xf < 1e-30f ? 1e-30f : xf;
    double r = (double)Ops::rcp(xf);
    return r * (2.0 - x * r);
}

/* rcp64 for x >= 1 (or NaN): no clamp of the FP32 seed needed. */
template <class Ops>
GGDMC_FN double rcp64_ge1(double x) {
    const double r = (double)Ops::rcp((float)x);
    return r * (2.0 - x * r);
}

/* 2^y for y <= 0 in FP64 arithmetic: y = n + j/64 + f with integers n, j
 * (0 <= j < 64) and |f| <= 1/128; 2^(j/64) from a table of correctly rounded
 * values (Ops::exp2_frac), 2^f = exp(f ln 2) from its Taylor series to degree
 * 5 (truncation below 4e-17 for |f ln 2| <= 0.0055). y <= -1075 (and -inf)
 * is taken as -1075, whose 2^-1075 rounds to 0; a NaN y also gives 0 (fmax),
 * so callers carry NaN through another factor (mixed_tail: the Mills ratio
 * of the same argument). */
template <class Ops>
GGDMC_FN double exp2_neg(double y) {
    const double yc = std::fmax(y, -1075.0);
    const double k = std::rint(yc * 64.0);
    const double x = (yc - k * 0.015625) * kLn2;
    const int ki = (int)k;
    /* Horner: one constant per multiply-add, which GFX9's single constant-bus
     * read per instruction takes straight from a scalar register */
    const double p =
        1.0 + x * (1.0 + x * (0.5 + x * (0.16666666666666666 + x * (0.041666666666666664 +
                                                                  x * 0.008333333333333333))));
    return std::ldexp(Ops::exp2_frac(ki & 63) * p, ki >> 6);
}

/* 2^(j/64), j = 0..63, correctly rounded to FP64 (the table of exp2_neg). */
inline void fill_exp2_table(double *t) {
    for (int j = 0; j < 64; ++j) t[j] = (double)std::exp2((long double)j / 64.0L);
}

/* e = exp(-z^2/2), qa = Q(|z|) = upper normal tail beyond |z| (NaN for a NaN
 * z, through t; e is then 0). */
template <class Ops>
GGDMC_FN void mixed_tail(double z, double &e, double &qa) {
    e = exp2_neg<Ops>((z * z) * (-0.5 * kLog2e));
    const double a = std::fabs(z);
    const double t = rcp64_ge1<Ops>(1.0 + kMillsK * a);
    /* sum_k c_k t^k, k = 0..10, by Horner's rule: each multiply-add reads one
     * coefficient (a scalar register), so no coefficient is first moved into
     * vector registers; the two trials and two tails a lane evaluates give the
     * dependency chains to interleave */
    double h = kMills[kMillsDeg];
#pragma unroll
    for (int j = kMillsDeg - 1; j >= 0; --j) h = kMills[j] + t * h;
    qa = e * t * h;
}

/* log(x) for x > 0, FP32 log2 on the mantissa in [sqrt(1/2), sqrt(2)). */
template <class Ops>
GGDMC_FN double mixed_log(double x) {
    int ex;
    double m = std::frexp(x, &ex); /* x = m * 2^ex, m in [0.5, 1) */
    const bool lo = m < kSqrt1_2;
    m = lo ? 2.0 * m : m;
    ex = lo ? ex - 1 : ex;
    const float mf = (float)m;
    /* log(m) = log(mf) + r with r = (m - mf)/mf, |r| < 2^-24: an FP32 reciprocal suffices */
    const double r = (m - (double)mf) * (double)Ops::rcp(mf);
    return ((double)Ops::log2(mf) + (double)ex) * kLn2 + r;
}

/* Shared core: Phi differences and the z's of one accumulator at time t > 0;
 * q1, q2, e1, e2 are the computed tails and exp(-z^2/2) (error bounds). */
struct Core {
    double zs, chi, z1, z2, dP, dphi, P2, Q1, Q2, q1, q2, e1, e2;
};

template <class Ops>
GGDMC_FN Core mixed_core(const AccK &k, double t) {
    Core c;
    c.zs = t * k.s;
    const double r = rcp64<Ops>(c.zs);
    c.chi = k.b - t * k.v;
    const double z1 = c.chi * r;
    const double z2 = (c.chi - k.A) * r;
    c.z1 = z1;
    c.z2 = z2;
    double e1, q1, e2, q2;
    mixed_tail<Ops>(z1, e1, q1);
    mixed_tail<Ops>(z2, e2, q2);
    c.q1 = q1;
    c.q2 = q2;
    c.e1 = e1;
    c.e2 = e2;
    c.Q1 = z1 >= 0.0 ? q1 : 1.0 - q1;
    const double P1 = z1 >= 0.0 ? 1.0 - q1 : q1;
    c.Q2 = z2 >= 0.0 ? q2 : 1.0 - q2;
    c.P2 = z2 >= 0.0 ? 1.0 - q2 : q2;
    c.dP = z2 >= 0.0 ? c.Q2 - c.Q1 : P1 - c.P2; /* Phi(z1) - Phi(z2) */
    c.dphi = (e2 - e1) * kInvSqrt2Pi;            /* phi(z2) - phi(z1) */
    return c;
}

/* ------------------------------------------------------------- naive FP64 */
template <class Ops>
GGDMC_FN void naive_pdf_cdf(const AccK &k, double t, double &pdf, double &cdf) {
    if (!(t > 0.0)) {
        pdf = t == t ? 0.0 : t;
        cdf = pdf;
        return;
    }
    const double zs = t * k.s;
    const double chi = k.b - t * k.v;
    const double chimA = chi - k.A;
    const double z1 = chi / zs, z2 = chimA / zs;
    const double P1 = Ops::pnorm(z1), P2 = Ops::pnorm(z2);
    const double d1 = Ops::dnorm(z1), d2 = Ops::dnorm(z2);
    pdf = (k.v * (P1 - P2) + k.s * (d2 - d1)) / (k.A * k.den);
    cdf = (1.0 + (zs * (d2 - d1) + chimA * P2 - chi * P1) / k.A) / k.den;
}

/* Error model of the mixed evaluation: every computed tail and exp(-z^2/2)
 * carries a relative error of up to kTailErr (FP32 exp2); a result whose
 * bound exceeds kCondTol relative, or whose reference FP64 value carries more
 * rounding noise than that where the rounding is not reproduced (Phi rounded
 * near 1, ulp 1.1e-16, scaled by the formula's coefficients), is flagged
 * ill-conditioned and the trial is evaluated with the reference's FP64
 * formula instead. */
constexpr double kTailErr = 1e-9;   /* Mills-ratio fit (2.9e-10) and the FP32 seeds */
constexpr double kExpErr = 1e-15;   /* exp(-z^2/2) in FP64 arithmetic */
constexpr double kCondTol = 1e-5;
constexpr double kUlp1 = 1.1102230246251565e-16;

/* x / d correctly rounded, given y = 1/d correctly rounded: q = x*y, then
 * r = x - q*d exactly (one FMA) and RN(q + r*y) = RN(x/d) (Markstein's
 * correction step) for finite x; x = +-inf takes x*y, which is x/d there.
 * Three FP64 operations instead of a division sequence. */
GGDMC_FN double div_rn(double x, double d, double y) {
    const double q = x * y;
    const double c = std::fma(std::fma(-q, d, x), y, q);
    return std::isinf(x) ? q : c;
}

/* Winner: max(pdf, 1e-10), the reference's floored density; NaN propagates. */
template <class Ops>
GGDMC_FN double mixed_winner_pdf(const AccK &k, double rt, bool &ill) {
    /* branch-free: t <= 0 (or NaN) is evaluated at t = 1 and replaced */
    const double t = rt - k.t0;
    const bool pos = t > 0.0;
    const Core c = mixed_core<Ops>(k, pos ? t : 1.0);
    /* |v| (kTailErr (q1 + q2) + [z2 > 0] 2 ulp) + s (e1 + e2) kExpErr / sqrt(2 pi) */
    const double tv = std::fma(kTailErr, c.q1 + c.q2, c.z2 > 0.0 ? 2.0 * kUlp1 : 0.0);
    const double ex = k.s * (kExpErr * kInvSqrt2Pi) * (c.e1 + c.e2);
    double num = k.v * c.dP + k.s * c.dphi;
    double err = std::fma(std::fabs(k.v), tv, ex);
    if (__builtin_expect(c.z2 > 0.0 && err > kCondTol * std::fabs(num), 0)) {
        /* the far right tail, where the reference's Phi(z1) - Phi(z2) is the
         * difference of two values rounded near 1: take that difference,
         * RN(1 - Q1) - RN(1 - Q2), itself. 1 - q = p + r exactly (Fast2Sum,
         * p = RN(1 - q)); when r lies further than q's error from the rounding
         * midpoint p -+ ulp/2 (ulp = 2^-53 on [1/2, 1]), the reference's 1 - Q
         * rounds to p as well and the quantised difference is reproduced
         * exactly; otherwise it may be one ulp off */
        const double p1 = 1.0 - c.q1, p2 = 1.0 - c.q2;
        const double r1 = -c.q1 - (p1 - 1.0), r2 = -c.q2 - (p2 - 1.0);
        const double eq1 = kTailErr * c.q1, eq2 = kTailErr * c.q2;
        const double dP = p1 - p2;
        /* |v| (error of dP + 4 ulp of it for the products) + the exp term */
        const double tr = (0.5 * kUlp1 - std::fabs(r1) > eq1 ? 0.0 : eq1 + kUlp1) +
                          (0.5 * kUlp1 - std::fabs(r2) > eq2 ? 0.0 : eq2 + kUlp1) + 4.0 * kUlp1 * dP;
        num = k.v * dP + k.s * c.dphi;
        err = std::fma(std::fabs(k.v), tr, ex);
    }
    const double pdf = num * k.invA * k.inv_den;
    /* both floored unless the pdf or its uncertainty reaches the floor */
    ill |= pos && err > kCondTol * std::fabs(num) && (std::fabs(num) + err) * k.invA * k.inv_den > 0.5 * kFloor;
    const double pf = pdf < kFloor ? kFloor : pdf;
    return pos ? pf : (t == t ? kFloor : t);
}

/* Winner: log(max(pdf, 1e-10)); NaN propagates. */
template <class Ops>
GGDMC_FN double mixed_winner_logpdf(const AccK &k, double rt, bool &ill) {
    const double pf = mixed_winner_pdf<Ops>(k, rt, ill);
    return pf == kFloor ? kLogFloor : mixed_log<Ops>(pf);
}

/* Loser: the reference's FP64 survivor 1 - max(F, 1e-10) (may be <= 0). */
template <class Ops>
GGDMC_FN double mixed_loser_survivor(const AccK &k, double rt, bool &ill) {
    /* branch-free: t <= 0 (or NaN) is evaluated at t = 1 and replaced */
    const double t = rt - k.t0;
    const bool pos = t > 0.0;
    const Core c = mixed_core<Ops>(k, pos ? t : 1.0);
    const double achi = std::fabs(c.chi), achiA = std::fabs(c.chi - k.A);
    const double G = (c.zs * c.dphi + c.chi * c.Q1 - (c.chi - k.A) * c.Q2) * k.invA;
    /* F >= 1/2: 1 - G without cancellation, then the reference's own FP64
     * steps F = (1 + (G - 1)) / den: the rounding of 1 + (G - 1) is what
     * quantises the reference's survivor, and it is reproduced exactly
     * whenever 1 - G is accurate to well under half an ulp of 1 */
    const double omG = (c.P2 - k.Pc) + (c.chi * c.dP - c.zs * c.dphi) * k.invA + k.Pc;
    const double F = (G * k.inv_den < 0.5) ? G * k.inv_den : div_rn(1.0 - omG, k.den, k.inv_den);
    const double S0 = 1.0 - (F > kFloor ? F : (F == F ? kFloor : F));
    /* error of F from the tails, plus the reference's own noise from Phi
     * values rounded near 1 (z > 0), against the survivor */
    const double u1 = std::fma(kTailErr, c.q1, c.z1 > 0.0 ? kUlp1 : 0.0);
    const double u2 = std::fma(kTailErr, c.q2, c.z2 > 0.0 ? kUlp1 : 0.0);
    const double errF =
        std::fma(achi, u1, std::fma(achiA, u2, c.zs * (kExpErr * kInvSqrt2Pi) * (c.e1 + c.e2))) * k.invA * k.inv_den;
    /* below a hundredth of an ulp of 1 the reference's own rounding of F is
     * reproduced exactly, whatever the survivor */
    ill |= pos && errF > kCondTol * std::fabs(S0) && errF > 1e-18;
    return pos ? S0 : (t == t ? 1.0 - kFloor : t);
}

/* Reference FP64 trial term (src formula above), for the plain path and for
 * the batched path's ill-conditioned trials. */
template <class Ops>
GGDMC_FN double naive_trial_term(const AccK *kc, int n_acc, int w, double rt) {
    double pdf, cdf;
    naive_pdf_cdf<Ops>(kc[w], rt - kc[w].t0, pdf, cdf);
    double dens = pdf != pdf ? pdf : (pdf > kFloor ? pdf : kFloor);
    /* losers in accumulator order; indexing the l-th loser (rather than
     * skipping j == w) keeps every lane of a wave on the same iteration */
    for (int l = 0; l + 1 < n_acc; ++l) {
        const int j = l + (l >= w);
        naive_pdf_cdf<Ops>(kc[j], rt - kc[j].t0, pdf, cdf);
        const double F = cdf != cdf ? cdf : (cdf > kFloor ? cdf : kFloor);
        dens *= (1.0 - F);
    }
    return Ops::log(dens <= 0.0 ? 2.220446049250313e-16 : dens);
}

/* Batched-path trial term, mixed evaluation: the reference's density
 * max(pdf_w, 1e-10) * prod_{j != w} (1 - max(F_j, 1e-10)), multiplied in FP64
 * in accumulator order as the reference does, and one log of it (that of
 * DBL_EPSILON for a density <= 0). A survivor is 0, or at least 2^-53 in
 * magnitude, so the product of n_acc - 1 <= 18 of them and the floored pdf
 * stays normal. ill is set when any accumulator's evaluation is
 * ill-conditioned (the caller then takes naive_trial_term instead). NaN
 * propagates without the flag. */
template <class Ops>
GGDMC_FN double mixed_trial_term_flagged(const AccK *kc, int n_acc, int w, double rt, bool &ill) {
    double dens = mixed_winner_pdf<Ops>(kc[w], rt, ill);
    for (int l = 0; l + 1 < n_acc; ++l) {
        const int j = l + (l >= w);
        dens *= mixed_loser_survivor<Ops>(kc[j], rt, ill);
    }
    const bool nan = dens != dens;
    ill = ill && !nan;
    return nan ? dens : (dens > 0.0 ? mixed_log<Ops>(dens) : kLogDblEps);
}

/* Two independent trials' terms (A and B) in one body: both winner terms,
 * then each loser step of both, so that the two dependency chains interleave
 * in a wave that has few neighbours to hide latency with. Per trial the same
 * operations as mixed_trial_term_flagged, hence the same values. */
template <class Ops>
GGDMC_FN void mixed_trial_term_flagged2(const AccK *kA, int wA, double rtA, const AccK *kB, int wB,
                                        double rtB, int n_acc, double &outA, double &outB, bool &illA,
                                        bool &illB) {
    double dA = mixed_winner_pdf<Ops>(kA[wA], rtA, illA);
    double dB = mixed_winner_pdf<Ops>(kB[wB], rtB, illB);
    for (int l = 0; l + 1 < n_acc; ++l) {
        const int jA = l + (l >= wA), jB = l + (l >= wB);
        dA *= mixed_loser_survivor<Ops>(kA[jA], rtA, illA);
        dB *= mixed_loser_survivor<Ops>(kB[jB], rtB, illB);
    }
    const bool nanA = dA != dA, nanB = dB != dB;
    illA = illA && !nanA;
    illB = illB && !nanB;
    outA = nanA ? dA : (dA > 0.0 ? mixed_log<Ops>(dA) : kLogDblEps);
    outB = nanB ? dB : (dB > 0.0 ? mixed_log<Ops>(dB) : kLogDblEps);
}

/* Batched-path trial term: mixed evaluation, or the reference formula when
 * any accumulator's evaluation is ill-conditioned. */
template <class Ops>
GGDMC_FN double mixed_trial_term(const AccK *kc, int n_acc, int w, double rt) {
    bool ill = false;
    const double v = mixed_trial_term_flagged<Ops>(kc, n_acc, w, rt, ill);
    return ill ? naive_trial_term<Ops>(kc, n_acc, w, rt) : v;
}


/* ----------------------------------------------------------- per (cell, acc)
 * Resolves the six families from theta / constants and fills AccK.
 * par_src / par_const are [n_acc * 6] for the cell. Returns false if st0 != 0
 * (the path implements st0 = 0 only).
 */
template <class Ops>
GGDMC_FN bool make_acck(const double *theta, const int32_t *src6, const double *const6,
                        AccK &k) {
    double f[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) f[i] = src6[i] >= 0 ? theta[src6[i]] : const6[i];
    k.A = f[0];
    k.b = f[0] + f[1];
    k.v = f[2];
    k.s = f[3];
    k.t0 = f[5];
    k.invA = 1.0 / k.A;
    const double x = k.v / k.s;
    /* one tail gives both: q = Phi(-|x|), the other is 1 - q */
    const double q = Ops::qtail(std::fabs(x));
    const double Pv = x >= 0.0 ? 1.0 - q : (x == x ? q : x);   /* Phi(v/s)  */
    k.Pc = x >= 0.0 ? q : (x == x ? 1.0 - q : x);              /* Phi(-v/s) */
    k.den = Pv > kFloor ? Pv : (Pv == Pv ? kFloor : Pv);
    k.inv_den = 1.0 / k.den;
    return f[4] == 0.0;
}

} // namespace ggdmc

#endif
