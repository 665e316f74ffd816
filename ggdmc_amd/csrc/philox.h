/*
 * Counter-based random streams of the DE sampler.
 *
 * The reference draws from R's Mersenne-Twister in a fixed sequential order
 * (Rf_runif at src/de.cpp:65,88,131,139,176,182,210,296,310,348,362 and
 * arma::shuffle at :58,73). Here every draw is a pure function of
 * (seed, iteration, unit, sweep, slot, block), so any thread can produce the
 * draws of any chain without a shared generator state:
 *
 *   Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), key = seed (lo, hi),
 *   counter = (iteration, unit, sweep << 16 | slot, block)
 *     unit  : subject index within its run (0 for single-level runs),
 *             0xFFFFFFFF for the population (phi) level
 *     sweep : parameter index of a blocked sweep, 0 for joint updates
 *     slot  : chain index i for the crossover of chain i;
 *             0x8000 | k for migration ring step k;
 *             0x7FFF for the migration subset; 0x7FFE for the level decision
 *   words of a slot are read as one sequence (block = word / 4):
 *     crossover : w0, w1 -> the two partners; w2, w3 -> accept U (53 bits);
 *                 w4 + k -> jitter of the k-th perturbed parameter
 *     migration : w0, w1 -> accept U; w4 + k -> jitter
 *     subset    : w0, w1 -> U for the size; w2 + k -> Fisher-Yates draw k
 *     decision  : w0, w1 -> U
 */
#ifndef GGDMC_PHILOX_H
#define GGDMC_PHILOX_H

#include <cstdint>

#ifdef __HIPCC__
#define GGDMC_RFN __host__ __device__ __forceinline__
#else
#define GGDMC_RFN inline
#endif

namespace ggdmc {

constexpr uint32_t kUnitPhi = 0xFFFFFFFFu;
constexpr uint32_t kSlotDecision = 0x7FFEu;
constexpr uint32_t kSlotSubset = 0x7FFFu;
constexpr uint32_t kSlotMigrate = 0x8000u;

struct U4 {
    uint32_t x[4];
};

GGDMC_RFN U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                           uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
    }
    U4 o;
    o.x[0] = c0;
    o.x[1] = c1;
    o.x[2] = c2;
    o.x[3] = c3;
    return o;
}

struct Stream {
    uint32_t k0, k1, iter, unit, sweep;
    GGDMC_RFN U4 block(uint32_t slot, uint32_t b) const {
        return philox4x32_10(iter, unit, (sweep << 16) | slot, b, k0, k1);
    }
    GGDMC_RFN uint32_t word(uint32_t slot, uint32_t w) const {
        const U4 b = block(slot, w >> 2);
        const uint32_t k = w & 3u; /* select, not an indexed (private-memory) load */
        return k == 0u ? b.x[0] : (k == 1u ? b.x[1] : (k == 2u ? b.x[2] : b.x[3]));
    }
};

GGDMC_RFN Stream make_stream(uint64_t seed, uint32_t iter, uint32_t unit, uint32_t sweep) {
    Stream s;
    s.k0 = (uint32_t)seed;
    s.k1 = (uint32_t)(seed >> 32);
    s.iter = iter;
    s.unit = unit;
    s.sweep = sweep;
    return s;
}

/* U[0,1) with 53 random bits. */
GGDMC_RFN double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}
/* U(0,1) from 32 bits, never 0 or 1. */
GGDMC_RFN double u32_open(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386962890625e-10; }
/* floor(w * n / 2^32): uniform on 0..n-1. */
GGDMC_RFN uint32_t below(uint32_t w, uint32_t n) { return (uint32_t)(((uint64_t)w * n) >> 32); }

/* Jitter of a perturbed parameter from its word: U(-rp, rp). */
GGDMC_RFN double jitter_word(uint32_t w, double rp) { return rp * (2.0 * u32_open(w) - 1.0); }

/* Jitter k of a slot: word 4 + k, i.e. word k & 3 of block 1 + k / 4. */
GGDMC_RFN double jitter(const Stream &st, uint32_t slot, uint32_t k, double rp) {
    return jitter_word(st.word(slot, 4 + k), rp);
}

GGDMC_RFN double decision_u(const Stream &st) {
    const U4 b = st.block(kSlotDecision, 0);
    return u53(b.x[0], b.x[1]);
}

/* Chain groups of a sweep: group g = chains [g*n/G, (g+1)*n/G). */
GGDMC_RFN int group_lo(int g, int n, int G) { return (g * n) / G; }

/* Two distinct partners of chain i drawn from the chains outside [lo, hi),
 * plus the accept U (src/de.cpp:54-60 draws a random ordered pair from the
 * chains other than i; with one chain per group this is the same law). */
GGDMC_RFN void partners(const Stream &st, int i, int lo, int hi, int n, int &m0, int &m1,
                        double &U) {
    const U4 b = st.block((uint32_t)i, 0);
    const uint32_t pool = (uint32_t)(n - (hi - lo));
    uint32_t r0 = below(b.x[0], pool);
    uint32_t r1 = below(b.x[1], pool - 1);
    if (r1 >= r0) ++r1;
    m0 = (int)r0 < lo ? (int)r0 : (int)r0 + (hi - lo);
    m1 = (int)r1 < lo ? (int)r1 : (int)r1 + (hi - lo);
    U = u53(b.x[2], b.x[3]);
}

/* Migration subset (src/de.cpp:62-78): size clamp(ceil(n*U), 2, n), a uniform
 * random subset of that size, sorted ascending. out needs n entries. */
GGDMC_RFN int subchains(const Stream &st, int n, int *out) {
    const double U = u53(st.word(kSlotSubset, 0), st.word(kSlotSubset, 1));
    double c = U * (double)n;
    int ns = (int)c;
    if ((double)ns < c) ++ns;
    ns = ns < 2 ? 2 : (ns > n ? n : ns);
    for (int k = 0; k < n; ++k) out[k] = k;
    for (int k = 0; k < ns; ++k) {
        const int j = k + (int)below(st.word(kSlotSubset, 2 + k), (uint32_t)(n - k));
        const int t = out[k];
        out[k] = out[j];
        out[j] = t;
    }
    for (int a = 1; a < ns; ++a) { /* insertion sort of the first ns */
        const int v = out[a];
        int b = a - 1;
        while (b >= 0 && out[b] > v) {
            out[b + 1] = out[b];
            --b;
        }
        out[b + 1] = v;
    }
    return ns;
}

} // namespace ggdmc

#endif
